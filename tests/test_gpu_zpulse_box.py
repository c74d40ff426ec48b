"""GPU parity of the in-box laser pulse (osr_sim_add_zpulse: launch_zpulse,
zpulse/os-zpulse-std.f03:2350-2469).

* The library builds the pulse on the host in the reference's FP64 operation order: after
  the launch, the device's E and B equal the FP32 oracle's bit for bit on every interior
  cell (2D and 3D, orders 1-3, plane and Hermite-Gaussian profiles, the three envelopes,
  both propagation directions, linear and circular polarisation, with and without the
  divergence correction), and with Lindman x walls also on the x guard cells the guard
  band holds.
* The launched pulse then crossing a plasma: energy histories within 1e-4 of the FP64
  oracle.
"""
import numpy as np
import pytest

import osiris_amd as oa
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    (2, 1, dict(a0=1.0, omega0=3.0, pol=0.0, lon_type="polynomial", lon_start=9.0,
                lon_rise=2.0, lon_flat=1.5, lon_fall=2.0)),
    (2, 2, dict(a0=1.3, omega0=2.5, pol=30.0, phase=15.0, lon_type="gaussian",
                lon_duration=2.0, lon_x0=6.0, lon_range=8.0, per_type="gaussian",
                per_w0=2.0, per_focus=5.0, per_tem_mode=1)),
    (2, 3, dict(a0=0.8, omega0=4.0, pol_type=1, propagation="backward", lon_type="sin2",
                lon_start=3.0, lon_fwhm=4.0, per_type="hermite", per_fwhm=3.0,
                per_focus=6.0, per_center=4.0)),
    (2, 2, dict(a0=1.0, omega0=3.0, pol_type=-1, lon_type="gaussian", lon_fwhm=3.0,
                lon_x0=7.0, lon_range=9.0, per_type="gaussian", per_w0=2.5, per_focus=7.0,
                no_div_corr=True)),
    (3, 2, dict(a0=1.1, omega0=2.5, pol=60.0, lon_type="gaussian", lon_duration=1.5,
                lon_x0=4.0, lon_range=6.0, per_type="gaussian", per_w0=1.5, per_focus=3.0,
                per_tem_mode=(1, 0))),
    (3, 1, dict(a0=0.6, omega0=3.0, pol_type=1, lon_type="polynomial", lon_start=6.0,
                lon_fwhm=3.0)),
]


@pytest.mark.parametrize("k", range(len(CASES)))
def test_box_pulse_equals_fp32_oracle(k):
    dims, order, kw = CASES[k]
    nx = (64, 40) if dims == 2 else (40, 16, 12)
    dx = (0.2, 0.3) if dims == 2 else (0.2, 0.3, 0.35)
    xmin = (0.5, -1.0, 0.25)[:dims]
    print(f"case box pulse {k} dims={dims} order={order}", flush=True)
    g = O.make_grid(dims, nx, order)
    e = g.zeros(3, np.float32)
    b = g.zeros(3, np.float32)
    p = O.zpulse_box_params(nx, dx, xmin, **kw)
    O.zpulse_box_add(g, p, order, dx, e, b, "f32")
    tile = (16, 8) if dims == 2 else (8, 8, 4)
    with oa.Simulation(dims, nx, dx, 0.05, order, tile=tile) as sim:
        sim.add_zpulse(xmin=xmin, **kw)
        ge, gb = sim.get_fields()
    we, wb = g.interior(e), g.interior(b)
    assert np.array_equal(ge.view(np.uint32), we.view(np.uint32))
    assert np.array_equal(gb.view(np.uint32), wb.view(np.uint32))
    assert np.abs(we).max() > 0.1


def test_box_pulse_guard_band_with_lindman_walls():
    """With open x walls the pulse's x guard cells land in the guard band: the fields read
    back with the reference's guards equal the oracle's on interior rows, x guards
    included (the walls' restore then sets E, B at x = 0 and E at x = nx + 1 as the
    reference's next boundary update does, so those are compared after it on both sides)."""
    dims, order, nx, dx = 2, 2, (64, 24), (0.2, 0.3)
    kw = dict(a0=1.0, omega0=3.0, pol=0.0, lon_type="gaussian", lon_duration=3.0,
              lon_x0=11.0, lon_range=10.0, per_type="gaussian", per_w0=2.0, per_focus=11.0)
    ref = O.Sim(dims, nx, dx, 0.05, order, prec="f32", open_x=True)
    ref.add_zpulse(**kw)
    ref.wall_refresh()
    with oa.Simulation(dims, nx, dx, 0.05, order, tile=(16, 8), xbnd="lindman") as sim:
        sim.add_zpulse(**kw)
        gcl, gcu = ref.grid.gcl, ref.grid.gcu
        ge, gb = sim.get_fields(gcl=gcl[:dims], gcu=gcu[:dims])
    ys = slice(gcl[1], gcl[1] + nx[1])
    for got, want in ((ge, ref.e), (gb, ref.b)):
        assert np.array_equal(got[ys].view(np.uint32), want[ys].view(np.uint32))
    assert np.abs(ref.e[ys]).max() > 0.1


def test_box_pulse_through_plasma_vs_fp64_oracle():
    from tests.pic_inputs import thermal_plasma

    rng = np.random.default_rng(12)
    nx, dx, dt, steps = (128, 32), (0.2, 0.3), 0.12, 40
    ix, x, u, q = thermal_plasma(rng, 2, nx, 4, 0.01, density=0.05)
    kw = dict(a0=1.5, omega0=3.0, pol=0.0, lon_type="gaussian", lon_duration=2.0,
              lon_x0=8.0, lon_range=8.0, per_type="gaussian", per_w0=2.5, per_focus=10.0)
    ref = O.Sim(2, nx, dx, dt, 2)
    ref.add_species(-1.0, 0, ix, x, u, q)
    ref.add_zpulse(**kw)
    with oa.Simulation(2, nx, dx, dt, 2, tile=(16, 16)) as sim:
        sid = sim.add_species(-1.0)
        sim.set_particles(sid, ix, x, u, q)
        sim.add_zpulse(**kw)
        sim.run(steps)
        fld, par = sim.history()
    ref.step(steps)
    rfld, rpar = np.array(ref.fld_hist), np.array(ref.par_hist)
    assert np.max(np.abs(par - rpar) / np.abs(rpar)) <= 1e-4
    assert np.max(np.abs(fld - rfld) / rfld.sum(1, keepdims=True)) <= 1e-4
