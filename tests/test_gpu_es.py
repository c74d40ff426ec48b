"""GPU parity of the initial electrostatic field (osr_sim_init_fields: init_fields,
spec/os-spec-initfields.f90, es_solver_2d, emf/os-emf-es-solver.f03:125-236).

* Given the same charge density, the GPU's direct sum equals the FP32 oracle's bit for
  bit: the density the device deposits (exact fixed-point sum, read back through
  osr_sim_gauss) is fed to the oracle's solver.
* Against the FP64 oracle end to end (its own FP64 deposit): 1e-5 of max |E|, periodic
  and open x walls (Lindman or VPML; the reference's update_beam leaves x walls as they
  are, os-emf-vpml.f03:3693-3750); and 20-30 steps after the set-up, energies to 1e-4.
* The same for the beam branch (es_solver_2d_beam, a species drifting along x), E and B.
"""
import ctypes as C

import numpy as np
import pytest

import osiris_amd as oa
from oracle import oracle as O
from tests.pic_inputs import thermal_plasma

pytestmark = pytest.mark.gpu


def _plasma(rng, nx, ppc=4):
    ix, x, u, q = thermal_plasma(rng, 2, nx, ppc, 0.05)
    # a density bump so the electron charge has structure
    xc = (ix[:, 0] - 0.5 + x[:, 0]) / nx[0]
    q = (q * (1.0 + 0.5 * np.sin(2 * np.pi * xc))).astype(np.float32)
    return ix, x, u, q


def test_es_field_bitexact_given_the_charge():
    rng = np.random.default_rng(3)
    nx, dx = (48, 40), (0.1, 0.15)
    ix, x, u, q = _plasma(rng, nx)
    with oa.Simulation(2, nx, dx, 0.05, 2, tile=(16, 8)) as sim:
        sid = sim.add_species(-1.0)
        sim.set_particles(sid, ix, x, u, q)
        rho = sim.charge_density()  # the device's fixed-point density, FP64 (ny, nx)
        sim.init_fields(sid)
        ge, _ = sim.get_fields()
    g = O.make_grid(2, nx, 2)
    L = O._lib("f32")
    r = g.zeros(1, np.float32)
    g.interior(r)[..., 0] = rho.astype(np.float32)
    ef = g.zeros(3, np.float32)
    L.orc_es_solver_2d(C.byref(g.c()), O._p(np.array([dx[0], dx[1], 1.0])), O._p(r), O._p(ef))
    want = g.interior(ef)
    assert np.array_equal(ge.view(np.uint32), want.view(np.uint32))
    assert np.abs(want[..., :2]).max() > 0


@pytest.mark.parametrize("xbnd", ["periodic", "open", "vpml"])
def test_init_fields_vs_fp64_oracle(xbnd):
    rng = np.random.default_rng(4)
    nx, dx, dt, steps = (64, 32), (0.1, 0.1), 0.05, 30
    ix, x, u, q = _plasma(rng, nx)
    ref = O.Sim(2, nx, dx, dt, 2, open_x=(xbnd != "periodic"),
                xwall="vpml" if xbnd == "vpml" else "lindman")
    ref.add_species(-1.0, 0, ix, x, u, q)
    ref.init_fields(0)
    with oa.Simulation(2, nx, dx, dt, 2, tile=(16, 16), xbnd=xbnd) as sim:
        sid = sim.add_species(-1.0)
        sim.set_particles(sid, ix, x, u, q)
        sim.init_fields(sid)
        ge, _ = sim.get_fields()
        want = ref.grid.interior(ref.e)
        assert np.abs(ge - want).max() <= 1e-5 * np.abs(want).max()
        sim.run(steps)
        ref.step(steps)
        fld, par = sim.history()
    rfld, rpar = np.array(ref.fld_hist), np.array(ref.par_hist)
    assert rfld[0].sum() > 0  # the ES field is in the energy history from the first row
    assert np.max(np.abs(par - rpar) / np.abs(rpar)) <= 1e-4
    fscale = rfld.sum(1, keepdims=True)
    assert np.max(np.abs(fld - rfld) / fscale) <= 1e-4


def test_beam_field_bitexact_given_the_charge():
    """es_solver_2d_beam: from the device's charge density the FP32 oracle's beam field
    added to E and B equals the GPU's bit for bit."""
    rng = np.random.default_rng(5)
    nx, dx, u1 = (48, 40), (0.1, 0.15), 0.8
    ix, x, u, q = _plasma(rng, nx)
    with oa.Simulation(2, nx, dx, 0.05, 2, tile=(16, 8)) as sim:
        sid = sim.add_species(-1.0)
        sim.set_particles(sid, ix, x, u, q)
        rho = sim.charge_density()
        sim.init_fields(sid, u1)
        ge, gb = sim.get_fields()
    g = O.make_grid(2, nx, 2)
    L = O._lib("f32")
    r = g.zeros(1, np.float32)
    g.interior(r)[..., 0] = rho.astype(np.float32)
    ef = g.zeros(3, np.float32)
    bf = g.zeros(3, np.float32)
    L.orc_es_solver_2d_beam(C.byref(g.c()), O._p(np.array([dx[0], dx[1], 1.0])), O._p(r),
                            float(np.float32(u1)), 0, O._p(ef), O._p(bf))
    we = np.zeros_like(ge) + g.interior(ef)  # added to zero fields
    wb = np.zeros_like(gb) + g.interior(bf)
    assert np.array_equal(ge.view(np.uint32), we.view(np.uint32))
    assert np.array_equal(gb.view(np.uint32), wb.view(np.uint32))
    assert np.abs(wb[..., 2]).max() > 0


@pytest.mark.parametrize("xbnd", ["periodic", "open", "vpml"])
def test_beam_init_fields_vs_fp64_oracle(xbnd):
    rng = np.random.default_rng(6)
    nx, dx, dt, steps, u1 = (64, 32), (0.1, 0.1), 0.05, 20, 0.6
    ix, x, u, q = _plasma(rng, nx)
    u = u.copy()
    u[:, 0] += u1
    ref = O.Sim(2, nx, dx, dt, 2, open_x=(xbnd != "periodic"),
                xwall="vpml" if xbnd == "vpml" else "lindman")
    ref.add_species(-1.0, 0, ix, x, u, q)
    ref.init_fields(0, u1)
    with oa.Simulation(2, nx, dx, dt, 2, tile=(16, 16), xbnd=xbnd) as sim:
        sid = sim.add_species(-1.0)
        sim.set_particles(sid, ix, x, u, q)
        sim.init_fields(sid, u1)
        ge, gb = sim.get_fields()
        we, wb = ref.grid.interior(ref.e), ref.grid.interior(ref.b)
        assert np.abs(ge - we).max() <= 1e-5 * np.abs(we).max()
        assert np.abs(gb - wb).max() <= 1e-5 * np.abs(wb).max()
        sim.run(steps)
        ref.step(steps)
        fld, par = sim.history()
    rfld, rpar = np.array(ref.fld_hist), np.array(ref.par_hist)
    assert np.max(np.abs(par - rpar) / np.abs(rpar)) <= 1e-4
    assert np.max(np.abs(fld - rfld) / rfld.sum(1, keepdims=True)) <= 1e-4


def test_init_fields_rejects_3d_at_rest():
    """The reference's es_solver has no 3D form for a species at rest (ufl(1) = 0)."""
    with oa.Simulation(3, (8, 8, 8), (0.1,) * 3, 0.05, 2, tile=(8, 8, 4)) as sim:
        sid = sim.add_species(-1.0)
        with pytest.raises(oa._lib.OsrError):
            sim.init_fields(sid)


def _plasma3(rng, nx, ppc=8):
    ix, x, u, q = thermal_plasma(rng, 3, nx, ppc, 0.05)
    yc = (ix[:, 1] - 0.5 + x[:, 1]) / nx[1]
    zc = (ix[:, 2] - 0.5 + x[:, 2]) / nx[2]
    q = (q * (1.0 + 0.4 * np.sin(2 * np.pi * yc) * np.cos(2 * np.pi * zc))).astype(np.float32)
    return ix, x, u, q


def test_beam3d_field_bitexact_given_the_charge():
    """es_solver_3d_beam: from the device's charge density the FP32 oracle's beam field
    added to E and B equals the GPU's bit for bit."""
    rng = np.random.default_rng(15)
    nx, dx, u1 = (16, 24, 16), (0.1, 0.12, 0.15), 0.8
    ix, x, u, q = _plasma3(rng, nx)
    with oa.Simulation(3, nx, dx, 0.04, 2, tile=(8, 8, 4)) as sim:
        sid = sim.add_species(-1.0)
        sim.set_particles(sid, ix, x, u, q)
        rho = sim.charge_density()
        sim.init_fields(sid, u1)
        ge, gb = sim.get_fields()
    g = O.make_grid(3, nx, 2)
    L = O._lib("f32")
    r = g.zeros(1, np.float32)
    g.interior(r)[..., 0] = rho.astype(np.float32)
    ef = g.zeros(3, np.float32)
    bf = g.zeros(3, np.float32)
    L.orc_es_solver_3d_beam(C.byref(g.c()), O._p(np.array(dx, np.float64)), O._p(r),
                            float(np.float32(u1)), 0, O._p(ef), O._p(bf))
    we = np.zeros_like(ge) + g.interior(ef)
    wb = np.zeros_like(gb) + g.interior(bf)
    assert np.array_equal(ge.view(np.uint32), we.view(np.uint32))
    assert np.array_equal(gb.view(np.uint32), wb.view(np.uint32))
    assert np.abs(wb[..., 1]).max() > 0 and np.abs(wb[..., 2]).max() > 0


@pytest.mark.parametrize("xbnd", ["periodic", "open"])
def test_beam3d_init_fields_vs_fp64_oracle(xbnd):
    rng = np.random.default_rng(16)
    nx, dx, dt, steps, u1 = (16, 16, 16), (0.1, 0.1, 0.1), 0.04, 12, 0.6
    ix, x, u, q = _plasma3(rng, nx)
    u = u.copy()
    u[:, 0] += u1
    ref = O.Sim(3, nx, dx, dt, 2, open_x=(xbnd != "periodic"))
    ref.add_species(-1.0, 0, ix, x, u, q)
    ref.init_fields(0, u1)
    with oa.Simulation(3, nx, dx, dt, 2, tile=(8, 8, 4), xbnd=xbnd) as sim:
        sid = sim.add_species(-1.0)
        sim.set_particles(sid, ix, x, u, q)
        sim.init_fields(sid, u1)
        ge, gb = sim.get_fields()
        we, wb = ref.grid.interior(ref.e), ref.grid.interior(ref.b)
        assert np.abs(ge - we).max() <= 1e-5 * np.abs(we).max()
        assert np.abs(gb - wb).max() <= 1e-5 * np.abs(wb).max()
        sim.run(steps)
        ref.step(steps)
        fld, par = sim.history()
    rfld, rpar = np.array(ref.fld_hist), np.array(ref.par_hist)
    assert np.max(np.abs(par - rpar) / np.abs(rpar)) <= 1e-4
    assert np.max(np.abs(fld - rfld) / rfld.sum(1, keepdims=True)) <= 1e-4
