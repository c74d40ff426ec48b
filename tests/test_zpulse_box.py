"""The in-box laser pulse (zpulse, zpulse/os-zpulse-std.f03: launch_zpulse_2d / 3d
:1072-1387, div_corr_zpulse_2d / 3d :1681-2205) in the CPU oracle, against what the
physics fixes:

* after the divergence correction the pulse's E and B are divergence-free on the Yee
  lattice to round-off wherever the envelope is on (the correction integrates the
  longitudinal components from the upper x edge); without it the divergence is of the
  order of k E;
* a flat-top plane pulse has amplitude a0 omega0, its carrier wavenumber is omega0 and
  B3 is E2 half a cell ahead (a forward plane wave); a circular pulse has |E_perp| =
  a0 omega0 throughout its flat top;
* a gaussian beam at its focus has the TEM-mode profile H_m(sqrt(2) rho / w0) exp(-rho^2
  / w0^2) across the beam (2D and 3D);
* launched into a vacuum box, the pulse's energy moves at c in the direction of
  propagation (Yee steps of the oracle).
"""
import numpy as np
import pytest

from oracle import oracle as O


def _launch(nx, dx, order, prec="f64", **kw):
    dims = len(nx)
    g = O.make_grid(dims, nx, order)
    real = np.float64 if prec == "f64" else np.float32
    e = g.zeros(3, real)
    b = g.zeros(3, real)
    p = O.zpulse_box_params(nx, dx, **kw)
    O.zpulse_box_add(g, p, order, dx, e, b, prec)
    return g, e, b, p


def _div(g, f, dx, bfield):
    """Discrete divergence on the full guarded arrays (Fortran index i at array i-1+gcl):
    E at the nodes (i, j, k): sum_d (E_d(i) - E_d(i - e_d)) / dx_d; B at the cell centres:
    sum_d (B_d(i + e_d) - B_d(i)) / dx_d."""
    dims = len(g.nx)
    f = np.moveaxis(f, -1, 0)  # (3, [z,] y, x)
    out = 0.0
    for d in range(dims):
        ax = dims - d  # array axis of dimension d in f[c]
        a = f[d]
        if bfield:
            diff = np.roll(a, -1, axis=ax - 1) - a
        else:
            diff = a - np.roll(a, 1, axis=ax - 1)
        out = out + diff / dx[d]
    return out


@pytest.mark.parametrize("dims,order", [(2, 1), (2, 2), (3, 2), (3, 3)])
def test_corrected_pulse_is_divergence_free(dims, order):
    nx = (96, 48) if dims == 2 else (100, 24, 20)
    dx = (0.2, 0.5) if dims == 2 else (0.2, 0.5, 0.6)
    kw = dict(a0=1.5, omega0=2.0, pol=30.0, lon_type="gaussian", lon_duration=2.5,
              lon_x0=9.6, lon_range=10.0, per_type="gaussian", per_w0=3.0, per_focus=8.0)
    g, e, b, _ = _launch(nx, dx, order, **kw)
    _, e0, b0, _ = _launch(nx, dx, order, no_div_corr=True, **kw)
    gl = g.gcl
    # nodes well inside the pulse (envelope on) and inside the transverse interior
    x = (np.arange(e.shape[dims - 1]) - gl[0]) * dx[0]  # array index -> x of its cell
    sel_x = np.abs(x - 9.6) < 3.5
    for f, f0, isb in ((e, e0, False), (b, b0, True)):
        dv = _div(g, f, dx, isb)
        dv0 = _div(g, f0, dx, isb)
        scale = np.abs(f).max() / min(dx)
        if dims == 2:
            core = dv[gl[1] + 2:gl[1] + nx[1] - 2][:, sel_x]
            core0 = dv0[gl[1] + 2:gl[1] + nx[1] - 2][:, sel_x]
        else:
            core = dv[gl[2] + 2:gl[2] + nx[2] - 2, gl[1] + 2:gl[1] + nx[1] - 2][..., sel_x]
            core0 = dv0[gl[2] + 2:gl[2] + nx[2] - 2, gl[1] + 2:gl[1] + nx[1] - 2][..., sel_x]
        assert np.abs(core).max() <= 1e-12 * scale, np.abs(core).max() / scale
        assert np.abs(core0).max() >= 1e-2 * scale  # the uncorrected pulse is not


def test_plane_pulse_amplitude_wavenumber_and_b_field():
    nx, dx, order = (400, 8), (0.1, 0.5), 2
    g, e, b, _ = _launch(nx, dx, order, a0=2.0, omega0=3.0, pol=0.0, lon_type="polynomial",
                         lon_start=35.0, lon_rise=5.0, lon_flat=20.0, lon_fall=5.0)
    E = g.interior(e)[0]  # one row (plane wave)
    B = g.interior(b)[0]
    x = (np.arange(nx[0]) + 0.5) * dx[0]  # order 2: cell centres
    flat = (x > 12.0) & (x < 28.0)  # lon_start - z in [rise, rise + flat)
    assert abs(np.abs(E[flat, 1]).max() - 2.0 * 3.0) < 2e-3 * 6.0
    assert not E[:, 0].any() and np.abs(E[:, 2]).max() < 1e-12
    spec = np.abs(np.fft.rfft(E[flat, 1] * np.hanning(flat.sum())))
    k = 2 * np.pi * np.fft.rfftfreq(flat.sum(), dx[0])
    assert abs(k[np.argmax(spec)] - 3.0) < 2 * np.pi / (flat.sum() * dx[0])
    mid = 0.5 * (E[flat, 1] + np.roll(E[:, 1], -1)[flat])
    assert np.abs(B[flat, 2] - mid).max() < 0.02 * 6.0  # B3 = E2 half a cell ahead


@pytest.mark.parametrize("pol_type", [1, -1])
def test_circular_pulse_has_constant_transverse_amplitude(pol_type):
    nx, dx, order = (400, 8), (0.1, 0.5), 1
    g, e, b, _ = _launch(nx, dx, order, a0=1.0, omega0=2.0, pol_type=pol_type,
                         lon_type="sin2", lon_start=35.0, lon_rise=5.0, lon_flat=20.0,
                         lon_fall=5.0)
    E = g.interior(e)[0]
    x = np.arange(nx[0]) * dx[0]
    flat = (x > 12.0) & (x < 28.0)
    amp = np.sqrt(E[flat, 1] ** 2 + E[flat, 2] ** 2)
    assert np.abs(amp - 2.0).max() < 1e-12
    # the two halves rotate in opposite senses: E3 leads or lags E2 by a quarter period
    lag = np.sign(np.mean(E[flat, 1] * np.roll(E[:, 2], -4)[flat]))
    assert lag != 0  # a definite sense ...
    g2, e2, _, _ = _launch(nx, dx, order, a0=1.0, omega0=2.0, pol_type=-pol_type,
                           lon_type="sin2", lon_start=35.0, lon_rise=5.0, lon_flat=20.0,
                           lon_fall=5.0)
    lag2 = np.sign(np.mean(g2.interior(e2)[0][flat, 1] * np.roll(g2.interior(e2)[0][:, 2], -4)[flat]))
    assert lag2 == -lag  # ... reversed by the opposite pol_type


@pytest.mark.parametrize("dims,mode", [(2, 0), (2, 1), (3, (0, 0)), (3, (1, 2))])
def test_focused_beam_is_the_tem_mode(dims, mode):
    nx = (80, 64) if dims == 2 else (40, 36, 32)
    dx = (0.25, 0.3) if dims == 2 else (0.25, 0.3, 0.35)
    w0, focus = 2.5, 5.0
    g, e, b, p = _launch(nx, dx, 1, a0=1.0, omega0=2.0, pol=0.0, lon_type="gaussian",
                         lon_duration=3.0, lon_x0=5.0, lon_range=9.0, per_type="gaussian",
                         per_w0=w0, per_focus=focus, per_tem_mode=mode, no_div_corr=True)
    i = int(round(focus / dx[0]))  # order 1: E2 of column i sits at x = i dx
    E = g.interior(e)
    from numpy.polynomial.hermite import hermval

    def h(m, u):
        return hermval(u, [0] * m + [1])

    if dims == 2:
        y = (np.arange(nx[1]) + 0.5) * dx[1] - p.center[0]  # E2 at y + 1/2
        col = E[:, i, 1]
        want = h(mode, np.sqrt(2) * y / w0) * np.exp(-(y / w0) ** 2)
    else:
        y = (np.arange(nx[1]) + 0.5) * dx[1] - p.center[0]
        z = np.arange(nx[2]) * dx[2] - p.center[1]
        col = E[:, :, i, 1]  # (z, y)
        want = (h(mode[0], np.sqrt(2) * y[None, :] / w0) * h(mode[1], np.sqrt(2) * z[:, None] / w0) *
                np.exp(-(y[None, :] ** 2 + z[:, None] ** 2) / w0 ** 2))
    # at the focus column the phase is common: the profile is the column up to one factor
    c = np.sum(col * want) / np.sum(want * want)
    assert abs(c) > 1e-3
    assert np.abs(col - c * want).max() <= 1e-10 * np.abs(col).max()


@pytest.mark.parametrize("propagation", ["forward", "backward"])
def test_pulse_moves_at_c_in_vacuum(propagation):
    nx, dx, dt, steps = (320, 4), (0.1, 0.5), 0.09, 100
    sim = O.Sim(2, nx, dx, dt, 2)
    sim.add_zpulse(a0=1.0, omega0=4.0, pol=0.0, propagation=propagation,
                   lon_type="gaussian", lon_duration=2.0, lon_x0=16.0, lon_range=12.0)

    def centroid():
        E = sim.grid.interior(sim.e)[0]
        w = E[:, 1] ** 2 + E[:, 2] ** 2
        x = (np.arange(nx[0]) + 0.5) * dx[0]
        return np.sum(w * x) / np.sum(w)

    c0 = centroid()
    sim.step(steps)
    shift = centroid() - c0
    want = steps * dt * (1 if propagation == "forward" else -1)
    assert abs(shift - want) < 0.03 * abs(want), (shift, want)
