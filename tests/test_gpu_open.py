"""GPU parity of open x walls (Lindman absorbing fields, emf/boundary/os-emf-lindman.f90,
and absorbing particles, spec/os-spec-boundary.f03:476-513) in the resident 2D path.

* Fields only: the FP32 GPU step equals the FP32 oracle bit for bit on every interior
  and guard cell, with and without a uniform background field in the guard cells; the
  field energies agree to the FP32 rounding of the oracle's squares.
* Plasma: against the FP64 oracle, the particle counts agree to the few particles that
  leave within FP32 rounding of a wall, and the energy histories to 1e-4 plus the share
  of those particles.
* Reproducibility: tile shape, workgroup size and kernel variant change no bit of the
  fields, the surviving particles or their number.
"""
import numpy as np
import pytest

import osiris_amd as oa
from oracle import oracle as O
from osiris_amd._lib import OsrError
from tests.pic_inputs import random_particles, thermal_plasma, vacuum_pulse

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("direction,b0", [(1, None), (-1, None), (1, (0.0, 0.1, 0.3))])
def test_lindman_fields_bitexact_vs_fp32_oracle(direction, b0):
    nx, ny, dx, dt, steps = 64, 16, 0.2, 0.12, 90
    print(f"case direction={direction} b0={b0}", flush=True)
    e, b = vacuum_pulse(nx, ny, dx, direction=direction)
    if b0 is not None:
        b = b + np.array(b0)
    e, b = e.astype(np.float32), b.astype(np.float32)
    ref = O.Sim(2, (nx, ny), (dx, dx), dt, 2, prec="f32", open_x=True)
    g = ref.grid
    with oa.Simulation(2, (nx, ny), (dx, dx), dt, 2, tile=(16, 16), xbnd="open") as sim:
        if b0 is not None:
            sim.fill_fields(b0=b0)
            ref.fill_uniform(b0=b0)
        sim.set_fields(e, b)
        ref.set_interior("e", e)
        ref.set_interior("b", b)
        sim.run(steps)
        ref.step(steps)
        ref.wall_refresh()  # the wall values the next push would read
        ge, gb = sim.get_fields(gcl=g.gcl, gcu=g.gcu)
        fld, _ = sim.history()
    assert np.array_equal(ge.view(np.uint32), ref.e.view(np.uint32))
    assert np.array_equal(gb.view(np.uint32), ref.b.view(np.uint32))
    # the FP32 oracle squares in FP32 like the reference's single-precision total_vdf;
    # the GPU squares the same FP32 values in FP64
    assert np.allclose(fld, np.array(ref.fld_hist), rtol=2e-7, atol=0)
    assert fld[-1].sum() < 0.5 * fld[0].sum()  # part of the pulse has left


def test_open_plasma_vs_fp64_oracle():
    rng = np.random.default_rng(71)
    nx, dx, dt, steps = (64, 32), (0.1, 0.1), 0.07, 40
    ix, x, u, q = thermal_plasma(rng, 2, nx, 8, 0.2)
    ref = O.Sim(2, nx, dx, dt, 2, open_x=True)
    ref.add_species(-1.0, 0, ix, x, u, q)
    with oa.Simulation(2, nx, dx, dt, 2, tile=(16, 16), xbnd="open") as sim:
        sid = sim.add_species(-1.0)
        sim.set_particles(sid, ix, x, u, q)
        sim.run(steps)
        ref.step(steps)
        n_gpu = sim.count(sid)
        gix, gx, gu, gq = sim.get_particles(sid)
        fld, par = sim.history()
    n0, n_ref = len(q), len(ref.species[0].q)
    print(f"particles {n0} -> gpu {n_gpu}, oracle {n_ref}", flush=True)
    assert n_gpu == len(gq) and n_ref < n0 - 200
    dn = abs(n_gpu - n_ref)
    assert dn <= 3
    assert np.all((gix[:, 0] >= 1) & (gix[:, 0] <= nx[0]))
    tol = 1e-4 + 2.0 * dn / n_ref
    rfld, rpar = np.array(ref.fld_hist), np.array(ref.par_hist)
    assert np.max(np.abs(par - rpar) / np.abs(rpar)) <= tol
    fscale = np.maximum(rfld.sum(1, keepdims=True), 1e-300)
    assert np.max(np.abs(fld - rfld) / fscale) <= tol
    ru = ref.species[0].u.astype(np.float64)
    rms = np.sqrt((ru ** 2).mean(0))
    assert np.all(np.abs(gu.astype(np.float64).mean(0) - ru.mean(0)) <= tol * rms)


@pytest.mark.parametrize("order,pusher", [(1, 0), (3, 0), (3, 1)])
def test_open_orders_vs_oracles(order, pusher):
    """Linear and cubic interpolation reach other guard columns (x = 0 only, or
    x = -1 .. nx+2 with the cubic stencil) and need other guard bands (8 columns here):
    fields bit-exact against the FP32 oracle, then a leaking plasma against FP64."""
    print(f"case order={order} pusher={pusher}", flush=True)
    nx, ny, dx, dt = 64, 16, 0.2, 0.12
    e, b = vacuum_pulse(nx, ny, dx, direction=-1)
    e, b = e.astype(np.float32), b.astype(np.float32)
    ref = O.Sim(2, (nx, ny), (dx, dx), dt, order, prec="f32", open_x=True)
    g = ref.grid
    with oa.Simulation(2, (nx, ny), (dx, dx), dt, order, tile=(8, 8), xbnd="open") as sim:
        sim.set_fields(e, b)
        ref.set_interior("e", e)
        ref.set_interior("b", b)
        sim.run(70)
        ref.step(70)
        ref.wall_refresh()
        ge, gb = sim.get_fields(gcl=g.gcl, gcu=g.gcu)
    assert np.array_equal(ge.view(np.uint32), ref.e.view(np.uint32))
    assert np.array_equal(gb.view(np.uint32), ref.b.view(np.uint32))

    rng = np.random.default_rng(90 + order)
    nxp, dxp, dtp, steps = (48, 24), (0.1, 0.1), 0.06, 30
    ix, x, u, q = thermal_plasma(rng, 2, nxp, 8, 0.2)
    ref = O.Sim(2, nxp, dxp, dtp, order, open_x=True)
    ref.add_species(-1.0, pusher, ix, x, u, q)
    with oa.Simulation(2, nxp, dxp, dtp, order, tile=(8, 8), xbnd="open") as sim:
        sid = sim.add_species(-1.0, pusher)
        sim.set_particles(sid, ix, x, u, q)
        sim.run(steps)
        ref.step(steps)
        n_gpu = sim.count(sid)
        fld, par = sim.history()
    n_ref = len(ref.species[0].q)
    assert n_ref < len(q) and abs(n_gpu - n_ref) <= 2
    tol = 1e-4 + 2.0 * abs(n_gpu - n_ref) / n_ref
    rfld, rpar = np.array(ref.fld_hist), np.array(ref.par_hist)
    assert np.max(np.abs(par - rpar) / np.abs(rpar)) <= tol
    fscale = np.maximum(rfld.sum(1, keepdims=True), 1e-300)  # row 0: no field yet
    assert np.max(np.abs(fld - rfld) / fscale) <= tol


@pytest.mark.parametrize("xbnd", ["open", "vpml"])
def test_open_walls_with_relayouts(xbnd):
    """A dense beam crossing into empty tiles overflows tile capacities, so steps halt and
    the host re-lays the particles out from the tiles and the pending movers; movers
    pushed through a wall in those steps must stay removed, and the wall fields advance
    once per step.  Counts and energies against the FP64 oracle, as without overflows."""
    rng = np.random.default_rng(47)
    nx, dx, dt, steps = (32, 32), (0.1, 0.1), 0.07, 30
    ix, x, u, q = random_particles(rng, 2, nx, 6000, uth=0.05, ufl=(2.0, 1.5, 0.0), q=-0.02,
                                   lo=[17, 1], hi=[24, 8])
    ref = O.Sim(2, nx, dx, dt, 2, open_x=True, xwall="vpml" if xbnd == "vpml" else "lindman")
    ref.add_species(-1.0, 0, ix, x, u, q)
    with oa.Simulation(2, nx, dx, dt, 2, tile=(8, 8), xbnd=xbnd) as sim:
        sid = sim.add_species(-1.0)
        sim.set_particles(sid, ix, x, u, q)
        sim.run(steps)
        ref.step(steps)
        n_gpu, rel = sim.count(sid), sim.relayouts
        gix = sim.get_particles(sid)[0]
        fld, par = sim.history()
    n_ref = len(ref.species[0].q)
    print(f"relayouts {rel}, particles {len(q)} -> gpu {n_gpu}, oracle {n_ref}", flush=True)
    assert rel >= 1 and n_ref < len(q) // 2
    assert abs(n_gpu - n_ref) <= 2 and len(gix) == n_gpu
    assert np.all((gix[:, 0] >= 1) & (gix[:, 0] <= nx[0]))
    tol = 1e-4 + 2.0 * abs(n_gpu - n_ref) / n_ref
    rfld, rpar = np.array(ref.fld_hist), np.array(ref.par_hist)
    assert np.max(np.abs(par - rpar) / np.abs(rpar)) <= tol
    fscale = np.maximum(rfld.sum(1, keepdims=True), 1e-300)
    assert np.max(np.abs(fld - rfld) / fscale) <= tol


def test_open_tile_and_kernel_independence():
    nx, dx, dt = (128, 64), (0.05, 0.05), 0.035
    out = []
    for tile, kernel in (((32, 16), 0), ((16, 16), 0), ((8, 8), 0), ((16, 16), 1)):
        print(f"case tile={tile} kernel={kernel}", flush=True)
        with oa.Simulation(2, nx, dx, dt, 2, tile=tile, kernel=kernel, xbnd="open") as sim:
            for k, ufl in enumerate((0.3, -0.3)):
                sid = sim.add_species(-1.0)
                sim.generate(sid, (4, 4), density=0.5, ufl=(ufl, 0, 0), uth=(0.05,) * 3,
                             seed=17 + k)
            sim.run(30)
            out.append((sim.get_fields(gcl=(2, 2), gcu=(3, 3)), sim.get_particles(0),
                        sim.get_particles(1), (sim.count(0), sim.count(1)), sim.history()))
    f0, a0, b0, c0, h0 = out[0]
    assert c0[0] < 128 * 64 * 16 and c0[1] < 128 * 64 * 16  # both beams lost particles
    for f, a, b, c, h in out[1:]:
        assert c == c0
        for x, y in zip(f0 + a0 + b0, f + a + b):
            assert np.array_equal(x, y)
        # the guard band is a whole number of tiles, so the FP64 field energy sums run
        # over differently sized rings: equal to round-off
        assert np.allclose(h0[0], h[0], rtol=1e-12, atol=0)
        assert np.allclose(h0[1], h[1], rtol=1e-12, atol=0)


def test_open_rejects_unsupported():
    with pytest.raises(OsrError):
        oa.Simulation(2, (40, 32), (0.1, 0.1), 0.05, 2, tile=(16, 16), xbnd="open")
    with pytest.raises(OsrError):
        oa.Simulation(2, (32, 32), (0.1, 0.1), 0.05, 2, plain=True, xbnd="open")
    with oa.Simulation(2, (32, 32), (0.1, 0.1), 0.05, 2, tile=(16, 16), xbnd="open") as sim:
        with pytest.raises(OsrError):
            sim.gauss()
        n, c = oa.smooth_setup(2, ("binomial", "none"), (1, 0))
        with pytest.raises(OsrError):
            sim.set_smooth(n, c)


def _pulse3(nx, ny, nz, dx, k0=4.0):
    xi = np.arange(1, nx + 1) * dx
    x0 = nx * dx / 2
    f = np.exp(-((xi - x0) / 2.0) ** 2) * np.cos(k0 * (xi - x0))
    fh = np.exp(-((xi + dx / 2 - x0) / 2.0) ** 2) * np.cos(k0 * (xi + dx / 2 - x0))
    e = np.zeros((nz, ny, nx, 3), np.float32)
    b = np.zeros_like(e)
    e[..., 1] = f  # a y-polarised pulse along +x, with an E3 part along -x
    b[..., 2] = fh
    e[..., 2] = 0.5 * f
    b[..., 1] = 0.5 * fh
    return e, b


@pytest.mark.parametrize("b0", [None, (0.2, 0.0, 0.4)])
def test_open_3d_fields_bitexact_vs_fp32_oracle(b0):
    """3D Lindman walls (update_lindman_3d_e/b): the buffer's B2 terms come in another
    order than the Yee step's, and two guard reads of the second B half step are stale."""
    nx, ny, nz, dx, dt, steps = 48, 8, 8, 0.2, 0.1, 70
    print(f"case 3d b0={b0}", flush=True)
    e, b = _pulse3(nx, ny, nz, dx)
    if b0 is not None:
        b = (b + np.array(b0, np.float32)).astype(np.float32)
    ref = O.Sim(3, (nx, ny, nz), (dx,) * 3, dt, 2, prec="f32", open_x=True)
    g = ref.grid
    with oa.Simulation(3, (nx, ny, nz), (dx,) * 3, dt, 2, tile=(8, 8, 4), xbnd="open") as sim:
        if b0 is not None:
            sim.fill_fields(b0=b0)
            ref.fill_uniform(b0=b0)
        sim.set_fields(e, b)
        ref.set_interior("e", e)
        ref.set_interior("b", b)
        sim.run(steps)
        ref.step(steps)
        ref.wall_refresh()
        ge, gb = sim.get_fields(gcl=g.gcl, gcu=g.gcu)
        fld, _ = sim.history()
    assert np.array_equal(ge.view(np.uint32), ref.e.view(np.uint32))
    assert np.array_equal(gb.view(np.uint32), ref.b.view(np.uint32))
    assert np.allclose(fld, np.array(ref.fld_hist), rtol=2e-7, atol=0)
    assert fld[-1].sum() < 0.7 * fld[0].sum()


@pytest.mark.parametrize("order,pusher", [(2, 0), (3, 1)])
def test_open_3d_plasma_vs_fp64_oracle(order, pusher):
    print(f"case 3d order={order} pusher={pusher}", flush=True)
    rng = np.random.default_rng(61 + order)
    nx, dx, dt, steps = (24, 8, 8), (0.1,) * 3, 0.05, 25
    ix, x, u, q = thermal_plasma(rng, 3, nx, 8, 0.2)
    ref = O.Sim(3, nx, dx, dt, order, open_x=True)
    ref.add_species(-1.0, pusher, ix, x, u, q)
    with oa.Simulation(3, nx, dx, dt, order, tile=(8, 8, 4), xbnd="open") as sim:
        sid = sim.add_species(-1.0, pusher)
        sim.set_particles(sid, ix, x, u, q)
        sim.run(steps)
        ref.step(steps)
        n_gpu = sim.count(sid)
        fld, par = sim.history()
    n_ref = len(ref.species[0].q)
    assert n_ref < len(q) and abs(n_gpu - n_ref) <= 2
    tol = 1e-4 + 2.0 * abs(n_gpu - n_ref) / n_ref
    rfld, rpar = np.array(ref.fld_hist), np.array(ref.par_hist)
    assert np.max(np.abs(par - rpar) / np.abs(rpar)) <= tol
    fscale = np.maximum(rfld.sum(1, keepdims=True), 1e-300)
    assert np.max(np.abs(fld - rfld) / fscale) <= tol
