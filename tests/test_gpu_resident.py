"""GPU parity of the resident (tiled, fused) path against the FP64 oracle and against the
drop-in path.  Tolerances (FP32 GPU vs FP64 oracle):

* cell indices: identical for every particle that does not end within FP32 rounding of
  a cell edge during the compared steps;
* energy histories: 1e-4 relative (the FP32 path's stated tolerance against FP64) over 20-50 steps;
* Gauss law: the change of div E - rho stays at FP32 round-off of E (1e-5 of max|rho|).
"""
import numpy as np
import pytest

import osiris_amd as oa
from oracle import oracle as O
from tests.pic_inputs import periodic_fields, random_particles, thermal_plasma

pytestmark = pytest.mark.gpu


def _wrap(ix, nx):
    ix = ix.copy()
    for d in range(ix.shape[1]):
        ix[:, d] = (ix[:, d] - 1) % nx[d] + 1
    return ix


@pytest.mark.parametrize("dims,order,pusher", [(2, 1, 0), (2, 2, 0), (2, 3, 0), (2, 2, 1),
                                               (3, 1, 0), (3, 2, 1), (3, 3, 1)])
def test_one_step_matches_dropin(dims, order, pusher):
    """First step of the resident path = drop-in kernel on the same inputs (bit-exact
    particles, same arithmetic)."""
    rng = np.random.default_rng(30 + dims * 7 + order + pusher)
    nx = (32, 24) if dims == 2 else (16, 12, 8)
    dx = (0.1,) * dims
    dt = 0.05 if dims == 2 else 0.04
    g = O.make_grid(dims, nx, order)
    e, b = periodic_fields(rng, g, 0.05, 0.3, np.float32)
    n = int(np.prod(nx)) * 3
    ix, x, u, q = random_particles(rng, dims, nx, n, uth=0.2)
    print(f"case dims={dims} order={order} pusher={pusher} n={n}", flush=True)
    with oa.Simulation(dims, nx, dx, dt, order, tile=(8, 8, 4)[:dims]) as sim:
        sid = sim.add_species(-1.0, pusher)
        sim.set_particles(sid, ix, x, u, q)
        sim.set_fields(e, b, gcl=g.gcl)
        sim.run(1)
        rix, rx, ru, rq = sim.get_particles(sid)
        fld, par = sim.history()
    lo = g.gcl[0]
    j = g.zeros(3, np.float32)
    dix, dxx, duu, dene = oa.advance_deposit(dims, order, ix.copy(), x.copy(), u.copy(), q, -1.0,
                                             e, b, lo, j, lo, dx, dt, pusher=pusher)
    assert np.array_equal(rix, _wrap(dix, nx))
    assert np.array_equal(rx.view(np.uint32), dxx.view(np.uint32))
    assert np.array_equal(ru.view(np.uint32), duu.view(np.uint32))
    assert np.array_equal(rq, q)
    dvol = float(np.prod(dx))
    assert abs(par[0, 0] - dene * -1.0 * dvol) <= 1e-12 * abs(par[0, 0])
    # field energy of the input fields
    fe = O.field_energy(g, e.astype(np.float64), b.astype(np.float64), dvol)
    assert np.allclose(fld[0], fe, rtol=1e-12, atol=0)


def _compare_with_oracle(dims, nx, dx, dt, order, species, steps, tile=None, fields=None,
                         e_tol=1e-4, check_ix_steps=None):
    """Run the same seeded input on the GPU and in the FP64 oracle; compare."""
    with oa.Simulation(dims, nx, dx, dt, order, tile=tile) as sim:
        ref = O.Sim(dims, nx, dx, dt, order)
        for (rqm, pusher, ix, x, u, q) in species:
            sid = sim.add_species(rqm, pusher)
            sim.set_particles(sid, ix, x, u, q)
            ref.add_species(rqm, pusher, ix, x, u, q)
        if fields is not None:
            e, b = fields
            sim.set_fields(e, b, gcl=ref.grid.gcl)
            ref.e[...] = e
            ref.b[...] = b
        g0 = sim.gauss()
        ix_steps = check_ix_steps or 0
        if ix_steps:
            sim.run(ix_steps)
            ref.step(ix_steps)
            for s in range(len(species)):
                gix = sim.get_particles(s)[0]
                rix = ref.species[s].ix
                rx = ref.species[s].x
                bad = np.any(gix != rix, axis=1)
                # any mismatch must be a particle sitting within FP32 rounding of an edge
                edge = np.any(np.abs(np.abs(rx) - 0.5) < 2e-5, axis=1)
                assert np.count_nonzero(bad & ~edge) == 0, f"species {s}: {np.count_nonzero(bad)} cells differ"
        sim.run(steps - ix_steps)
        ref.step(steps - ix_steps)
        fld, par = sim.history()
        rfld = np.array(ref.fld_hist)
        rpar = np.array(ref.par_hist)
        tot_g = fld.sum(1) + par.sum(1)
        tot_r = rfld.sum(1) + rpar.sum(1)
        # particle energies and total energy: relative to themselves
        assert np.max(np.abs(par - rpar) / np.abs(rpar)) <= e_tol
        assert np.max(np.abs(tot_g - tot_r) / np.abs(tot_r)) <= e_tol
        # field energy components: relative to the total field energy of the row
        fscale = np.maximum(rfld.sum(1, keepdims=True), 1e-300)
        assert np.max(np.abs(fld - rfld) / fscale) <= e_tol
        # particle moments: mean and rms momentum per component, relative to the rms
        for s in range(len(species)):
            gu = sim.get_particles(s)[2].astype(np.float64)
            ru = ref.species[s].u.astype(np.float64)
            rms = np.sqrt((ru ** 2).mean(0))
            assert np.all(np.abs(gu.mean(0) - ru.mean(0)) <= e_tol * rms)
            assert np.all(np.abs(np.sqrt((gu ** 2).mean(0)) - rms) <= e_tol * rms)
        g1 = sim.gauss()
        return sim.relayouts, g0, g1


@pytest.mark.parametrize("order", (1, 2, 3))
def test_thermal_2d_vs_oracle(order):
    rng = np.random.default_rng(40 + order)
    nx, dx, dt = (32, 32), (0.1, 0.1), 0.07
    ix, x, u, q = thermal_plasma(rng, 2, nx, 16, 0.05)
    _compare_with_oracle(2, nx, dx, dt, order, [(-1.0, 0, ix, x, u, q)], 30, tile=(16, 16),
                         check_ix_steps=3)


def test_two_stream_2d_vs_oracle():
    """Counter-streaming beams (the 1024x1024 two-stream workload's physics at 64x64):
    energy histories and moments."""
    rng = np.random.default_rng(44)
    nx, dx, dt = (64, 64), (0.05, 0.05), 0.035
    sp = []
    for ufl in (0.2, -0.2):
        ix, x, u, q = thermal_plasma(rng, 2, nx, 8, 0.01, ufl=(ufl, 0, 0), density=0.5)
        sp.append((-1.0, 0, ix, x, u, q))
    _compare_with_oracle(2, nx, dx, dt, 2, sp, 40, tile=(16, 16), check_ix_steps=3)


def test_c1_thermal_128_full_run_vs_fp64_oracle():
    """The reference's CPU-runnable workload at full size: 2D periodic thermal
    plasma, 128x128 cells, dx = 0.1, dt = 0.07, 16 ppc at uniform random positions,
    u_th = 0.05, linear splines, Boris, 500 steps.  FP32 on the GPU against the FP64
    oracle: energy histories within 1e-4 (the FP32 restatement of the oracle stays within
    1.2e-7 of FP64 here), cell indices identical over the first steps but for particles
    within FP32 rounding of a cell edge, Gauss law conserved."""
    rng = np.random.default_rng(47)
    nx, dx, dt = (128, 128), (0.1, 0.1), 0.07
    ix, x, u, q = thermal_plasma(rng, 2, nx, 16, 0.05)
    relayouts, g0, g1 = _compare_with_oracle(2, nx, dx, dt, 1, [(-1.0, 0, ix, x, u, q)], 500,
                                             check_ix_steps=3)
    assert abs(g1 - g0) < 1e-3 * max(1.0, abs(g0))


def test_magnetized_3d_vay_vs_oracle():
    """The 256^3 magnetized thermal plasma's physics at 16^3: cubic, Vay, uniform
    B0 = 0.5 z."""
    rng = np.random.default_rng(45)
    nx, dx, dt = (16, 16, 16), (0.1, 0.1, 0.1), 0.05
    ix, x, u, q = thermal_plasma(rng, 3, nx, 2, 0.1)
    g = O.make_grid(3, nx, 3)
    e = g.zeros(3, np.float32)
    b = g.zeros(3, np.float32)
    b[..., 2] = 0.5
    _compare_with_oracle(3, nx, dx, dt, 3, [(-1.0, 1, ix, x, u, q)], 20, tile=(8, 8, 4),
                         fields=(e, b), check_ix_steps=3)


def test_gauss_law_and_relayout():
    """A beam crossing into empty tiles overflows tile capacities (forcing re-layouts);
    results must still match the oracle, and div E - rho must not change beyond round-off
    (charge conservation of the deposit)."""
    rng = np.random.default_rng(46)
    nx, dx, dt = (32, 32), (0.1, 0.1), 0.07
    # dense beam in the lower-left tile moving diagonally
    ix, x, u, q = random_particles(rng, 2, nx, 6000, uth=0.05, ufl=(2.0, 1.5, 0.0), q=-0.02,
                                   lo=[1, 1], hi=[8, 8])
    nrel, g0, g1 = _compare_with_oracle(2, nx, dx, dt, 2, [(-1.0, 0, ix, x, u, q)], 25,
                                        tile=(8, 8), e_tol=1e-4)
    assert nrel >= 1
    with oa.Simulation(2, nx, dx, dt, 2, tile=(8, 8)) as sim:
        sid = sim.add_species(-1.0)
        sim.set_particles(sid, ix, x, u, q)
        m0, r0 = sim.gauss(True)
        sim.run(25)
        m1, r1 = sim.gauss(True)
        rho_scale = np.abs(r0).max()
        assert np.abs(r1 - r0).max() <= 1e-5 * rho_scale


def test_determinism():
    """Two identical runs give bit-identical particles, fields and histories (integer
    current accumulation, deterministic tile maintenance, fixed-order reductions)."""
    rng = np.random.default_rng(47)
    nx, dx, dt = (64, 48), (0.1, 0.1), 0.07
    ix, x, u, q = thermal_plasma(rng, 2, nx, 8, 0.1)
    out = []
    for _ in range(2):
        with oa.Simulation(2, nx, dx, dt, 2, tile=(16, 16)) as sim:
            sid = sim.add_species(-1.0)
            sim.set_particles(sid, ix, x, u, q)
            sim.run(20)
            out.append((sim.get_particles(sid), sim.get_fields(), sim.history()))
    (p0, f0, h0), (p1, f1, h1) = out
    for a, b in zip(p0 + f0 + h0, p1 + f1 + h1):
        assert np.array_equal(a, b)


def test_generate_statistics():
    with oa.Simulation(2, (64, 64), (0.1, 0.1), 0.07, 2) as sim:
        sid = sim.add_species(-1.0)
        sim.generate(sid, (4, 4), density=1.0, ufl=(0.2, 0, 0), uth=(0.05, 0.05, 0.05), seed=3)
        ix, x, u, q = sim.get_particles(sid)
    assert len(q) == 64 * 64 * 16
    assert np.all(q == np.float32(-1.0 / 16))
    assert np.all((x >= -0.5) & (x < 0.5))
    counts = np.bincount((ix[:, 0] - 1) + 64 * (ix[:, 1] - 1), minlength=4096)
    assert np.all(counts == 16)
    assert abs(u[:, 0].mean() - 0.2) < 1e-3 and abs(u[:, 0].std() - 0.05) < 1e-3
    assert abs(u[:, 2].mean()) < 1e-3 and abs(u[:, 1].std() - 0.05) < 1e-3


def test_plain_path_matches_fused_first_step_and_oracle():
    """The one-kernel-per-phase path: same particle arithmetic (bit-identical first step),
    current via global FP32 atomics (order not reproducible), energies vs the oracle."""
    rng = np.random.default_rng(48)
    nx, dx, dt = (48, 32), (0.1, 0.1), 0.07
    ix, x, u, q = thermal_plasma(rng, 2, nx, 8, 0.1)
    res = []
    for plain in (False, True):
        with oa.Simulation(2, nx, dx, dt, 2, plain=plain) as sim:
            sid = sim.add_species(-1.0)
            sim.set_particles(sid, ix, x, u, q)
            sim.run(1)
            res.append(sim.get_particles(sid))
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)
    with oa.Simulation(2, nx, dx, dt, 2, plain=True) as sim:
        sid = sim.add_species(-1.0)
        sim.set_particles(sid, ix, x, u, q)
        sim.run(20)
        fld, par = sim.history()
    ref = O.Sim(2, nx, dx, dt, 2)
    ref.add_species(-1.0, 0, ix, x, u, q)
    ref.step(20)
    rpar = np.array(ref.par_hist)
    assert np.max(np.abs(par - rpar) / np.abs(rpar)) <= 1e-4


def test_determinism_large():
    """Bit-reproducibility of the fused kernel (per-wave FP32 LDS tiles + fixed-order sum
    + integer global current) on a larger two-stream run with many tile crossings."""
    rng = np.random.default_rng(49)
    nx, dx, dt = (256, 256), (0.05, 0.05), 0.035
    out = []
    for _ in range(2):
        with oa.Simulation(2, nx, dx, dt, 2, tile=(16, 16)) as sim:
            for k, ufl in enumerate((0.2, -0.2)):
                sid = sim.add_species(-1.0)
                sim.generate(sid, (4, 4), density=0.5, ufl=(ufl, 0, 0), uth=(0.05,) * 3,
                             seed=77 + k)
            sim.run(40)
            out.append((sim.get_fields(), sim.history(), sim.get_particles(0)))
    (f0, h0, p0), (f1, h1, p1) = out
    for a, b in zip(f0 + h0 + p0, f1 + h1 + p1):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("dims,order", [(2, 1), (2, 2), (2, 3), (3, 2), (3, 3)])
def test_charge_density_vs_oracle(dims, order):
    """The FP64 rho of the Gauss-law diagnostic equals DEPOSIT_RHO of the oracle."""
    rng = np.random.default_rng(50 + 3 * dims + order)
    nx = (32, 32) if dims == 2 else (16, 16, 8)
    n = 4000
    ix, x, u, q = random_particles(rng, dims, nx, n, uth=0.1, q=-0.02)
    g = O.make_grid(dims, nx, order)
    rho = O.deposit_rho(g, order, ix, x.astype(np.float64), q.astype(np.float64))
    O.update_periodic(g, rho, 1)
    ro = g.interior(rho)[..., 0]
    with oa.Simulation(dims, nx, (0.1,) * dims, 0.04, order) as sim:
        sid = sim.add_species(-1.0)
        sim.set_particles(sid, ix, x, u, q)
        rg = sim.charge_density()
    assert np.abs(rg - ro).max() <= 1e-12 * np.abs(ro).max()


def test_tile_shape_and_kernel_independence():
    """Particles and fields do not depend on the tile shape, the workgroup size (4 waves
    for 8x8 and 16x16 tiles, 8 waves for 32x16) or the kernel variant (compile-time tile vs
    generic): the current is summed in integers, so every path gives the same bits.  The
    FP64 energy sums are per tile and per wave, so they agree to round-off only."""
    nx, dx, dt = (128, 96), (0.05, 0.05), 0.035
    out = []
    for tile, kernel in (((32, 16), 0), ((16, 16), 0), ((8, 8), 0), ((32, 16), 1)):
        print(f"case tile={tile} kernel={kernel}", flush=True)
        with oa.Simulation(2, nx, dx, dt, 2, tile=tile, kernel=kernel) as sim:
            for k, ufl in enumerate((0.2, -0.2)):
                sid = sim.add_species(-1.0)
                sim.generate(sid, (4, 4), density=0.5, ufl=(ufl, 0, 0), uth=(0.05,) * 3,
                             seed=91 + k)
            sim.run(30)
            out.append((sim.get_fields(), sim.get_particles(0), sim.get_particles(1),
                        sim.history()))
    f0, a0, b0, h0 = out[0]
    for f, a, b, h in out[1:]:
        for x, y in zip(f0 + a0 + b0, f + a + b):
            assert np.array_equal(x, y)
        assert np.array_equal(h0[0], h[0])  # field energies: FP64 sums over the same grid
        assert np.allclose(h0[1], h[1], rtol=1e-12, atol=0)


def test_tile_shape_independence_3d():
    """3D (cubic, Vay, magnetized): particles and fields do not depend on the tile shape or
    the kernel variant (compile-time 8x8x4 tile with padded LDS pitches and blocked cell
    order, or the generic runtime-shape kernel); the current is summed in integers
    whatever order the queued virtual particles are deposited in."""
    nx, dx, dt = (24, 16, 16), (0.1, 0.1, 0.1), 0.05
    out = []
    for tile, kernel in (((8, 8, 4), 0), ((8, 8, 4), 1), ((4, 4, 4), 0), ((8, 4, 8), 0)):
        print(f"case tile={tile} kernel={kernel}", flush=True)
        with oa.Simulation(3, nx, dx, dt, 3, tile=tile, kernel=kernel) as sim:
            sid = sim.add_species(-1.0, oa.OSR_VAY)
            sim.generate(sid, (2, 2, 2), density=1.0, uth=(0.1,) * 3, seed=5)
            sim.fill_fields(b0=(0.0, 0.0, 0.5))
            sim.run(12)
            out.append((sim.get_fields(), sim.get_particles(0), sim.history()))
    f0, p0, h0 = out[0]
    for f, p, h in out[1:]:
        for x, y in zip(f0 + p0, f + p):
            assert np.array_equal(x, y)
        assert np.array_equal(h0[0], h[0])
        assert np.allclose(h0[1], h[1], rtol=1e-12, atol=0)


@pytest.mark.parametrize("dims,period", [(3, 1), (3, 5), (2, 3)])
def test_reorder_changes_no_result(dims, period):
    """Restoring the in-tile particle order (tile_reorder) moves particles between slots
    only: fields and particles (returned in upload order) are bit-identical with and
    without it; the FP64 energy sums change only by round-off."""
    if dims == 3:
        nx, dx, dt, tile, order = (32, 24, 16), (0.1,) * 3, 0.05, (8, 8, 4), 3
    else:
        nx, dx, dt, tile, order = (96, 64), (0.05, 0.05), 0.035, (16, 16), 2
    out = []
    for reorder in (-1, period):
        print(f"case dims={dims} reorder={reorder}", flush=True)
        with oa.Simulation(dims, nx, dx, dt, order, tile=tile, reorder=reorder) as sim:
            sid = sim.add_species(-1.0, oa.OSR_VAY if dims == 3 else oa.OSR_BORIS)
            sim.generate(sid, (2,) * dims, density=1.0, uth=(0.1,) * 3, seed=11)
            if dims == 3:
                sim.fill_fields((0, 0, 0), (0, 0, 0.5))
            sim.run(12)
            out.append((sim.get_fields(), sim.get_particles(sid), sim.history()))
    (f0, p0, h0), (f1, p1, h1) = out
    for a, b in zip(f0 + p0, f1 + p1):
        assert np.array_equal(a, b)
    assert np.allclose(h0[0], h1[0], rtol=1e-12, atol=0)
    assert np.allclose(h0[1], h1[1], rtol=1e-12, atol=0)
