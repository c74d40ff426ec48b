"""Laser injection from an x wall (zpulse_wall, zpulse/os-zpulse-wall.f03) in the CPU
oracle, against what the physics and the closed forms fix:

* a flat-top plane wave launched with amplitude a0 carries E = a0 omega0 (normalised
  units) into the box, polarised along z for pol = 90 degrees and along y for pol = 0,
  travelling away from the wall it is launched from at the speed of light;
* a beam focused on the wall adds a column whose transverse profile is the TEM mode
  H_m(sqrt(2) rho / w0) exp(-rho^2 / w0^2), and nothing outside the launch window.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

NX, NY, DX, DT = 400, 4, 0.1, 0.06
FLAT = dict(tenv_type="polynomial", tenv_rise=3.0, tenv_flat=20.0, tenv_fall=3.0)


def _plane_run(pol, propagation="forward", steps=350, a0=0.1, omega0=4.0, refine=1):
    s = O.Sim(2, (NX * refine, NY), (DX / refine, DX / refine), DT / refine, 2, open_x=True)
    s.add_zpulse_wall(a0=a0, omega0=omega0, pol=pol, propagation=propagation, **FLAT)
    s.step(steps * refine)
    s.wall_refresh()
    return s, s.grid.interior(s.e)[0]


@pytest.mark.parametrize("pol,comp", [(90.0, 2), (0.0, 1)])
def test_plane_wave_amplitude_and_polarisation(pol, comp):
    """The amplitude converges to a0 omega0 at second order in dx, dt (the lattice
    dispersion of the wall source)."""
    a0, omega0 = 0.1, 4.0
    dev = []
    for refine in (1, 2):
        s, e = _plane_run(pol, a0=a0, omega0=omega0, refine=refine)
        # the front has travelled c t = 21 from x = 0: sample the flat part, x in 5..15
        amp = np.abs(e[50 * refine:150 * refine, comp]).max()
        dev.append(amp / (a0 * omega0) - 1.0)
    assert abs(dev[1]) < 5e-3 and abs(dev[0]) > 3.5 * abs(dev[1]), dev
    s, e = _plane_run(pol, a0=a0, omega0=omega0)
    amp = np.abs(e[50:150, comp]).max()
    other = 3 - comp  # the other transverse component
    assert np.abs(e[:, other]).max() < 1e-6 * amp
    assert np.abs(e[:, 0]).max() < 1e-6 * amp  # no longitudinal field in vacuum
    # nothing ahead of the front (x > 21 + a few cells of dispersion tail)
    assert np.abs(e[int(23.0 / DX):, comp]).max() < 1e-3 * amp


def test_backward_pulse_enters_from_the_upper_wall():
    s, e = _plane_run(90.0, propagation="backward")
    amp = np.abs(e[:, 2]).max()
    assert np.abs(e[NX - 150:NX - 50, 2]).max() > 0.95 * amp
    assert np.abs(e[:NX - int(23.0 / DX), 2]).max() < 1e-3 * amp


def _column(order=2, **kw):
    g = O.make_grid(2, (8, 64), order)
    p = O.zpulse_wall_params((8, 64), (0.2, 0.25), **kw)
    b = g.zeros(3, np.float64)
    L = O._lib("f64")
    dxa = np.array([0.2, 0.25, 1.0])
    return g, p, b, L, dxa


@pytest.mark.parametrize("modes", [(0, 0), (1, 2)])
def test_focused_beam_profile_3d_is_the_tem_mode(modes):
    """3D (launch_em_wall_3d): B2 at (y, z + 1/2) across the wall plane of a beam focused
    there is H_m1(sqrt(2) oy / w0) H_m2(sqrt(2) oz / w0) exp(-(oy^2 + oz^2) / w0^2)."""
    n, d = (8, 24, 20), (0.2, 0.25, 0.3)
    g = O.make_grid(3, n, 2)
    w0 = 1.5
    p = O.zpulse_wall_params(n, d, per_type="hermite", per_w0=w0, per_focus=0.1,
                             per_tem_mode=modes, omega0=5.0, pol=90.0, **FLAT)
    b = g.zeros(3, np.float64)
    L = O._lib("f64")
    L.orc_zpulse_wall_2d(C.byref(g.c()), C.byref(p), 2, O._p(np.array(d)), 0.05, 10.0, O._p(b))
    col = b[g.gcl[2]:g.gcl[2] + n[2], g.gcl[1]:g.gcl[1] + n[1], g.gcl[0], 1]  # (z, y)
    oy = d[1] * np.arange(n[1]) + d[1] / 2 - n[1] * d[1] / 2
    oz = d[2] * np.arange(n[2]) + d[2] - n[2] * d[2] / 2  # z + 1/2 of the quadratic shift
    H = [lambda u: np.ones_like(u), lambda u: 2 * u, lambda u: 4 * u ** 2 - 2]
    want = (H[modes[0]](np.sqrt(2) * oy / w0)[None, :] * H[modes[1]](np.sqrt(2) * oz / w0)[:, None] *
            np.exp(-(oy[None, :] ** 2 + oz[:, None] ** 2) / w0 ** 2))
    k = np.unravel_index(np.argmax(np.abs(want)), want.shape)
    assert np.allclose(col / col[k], want / want[k], rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_focused_beam_profile_is_the_tem_mode(mode):
    w0 = 2.0
    # quadratic interpolation launches at z = xmin + dx/2: focus there
    g, p, b, L, dxa = _column(per_type="hermite", per_w0=w0, per_focus=0.1, per_tem_mode=mode,
                              omega0=5.0, pol=90.0, **FLAT)
    t = 10.0  # inside the flat part
    L.orc_zpulse_wall_2d(C.byref(g.c()), C.byref(p), 2, O._p(dxa), 0.05, t, O._p(b))
    col = b[g.gcl[1]:g.gcl[1] + 64, g.gcl[0], 1]  # B2 at x = 1, y rows 1..64
    r = 0.25 * np.arange(64) + 0.125 - 8.0  # rho at the B2 points (rmin + dy/2)
    u = np.sqrt(2.0) * r / w0
    h = [np.ones_like(u), 2 * u, 4 * u ** 2 - 2][mode]
    want = h * np.exp(-r ** 2 / w0 ** 2)
    k = np.argmax(np.abs(want))
    assert np.allclose(col / col[k], want / want[k], rtol=1e-12, atol=1e-14)


def test_nothing_outside_the_launch_window():
    g, p, b, L, dxa = _column(launch_time=2.0, **FLAT)
    for t in (1.99, 2.0 + 26.0 + 1e-9):
        L.orc_zpulse_wall_2d(C.byref(g.c()), C.byref(p), 2, O._p(dxa), 0.05, t, O._p(b))
    assert not b.any()
    L.orc_zpulse_wall_2d(C.byref(g.c()), C.byref(p), 2, O._p(dxa), 0.05, 12.0, O._p(b))
    assert b.any()


def test_zpulse_needs_open_walls():
    s = O.Sim(2, (16, 8), (0.1, 0.1), 0.05, 2)
    with pytest.raises(ValueError):
        s.add_zpulse_wall(a0=1.0)


def test_library_defaults_are_the_reference_defaults():
    """osr_zpulse_wall_defaults (host code) = the nl_zpulse_wall defaults
    (os-zpulse-wall.f03:196-239) the oracle resolves from."""
    import osiris_amd._lib as L
    z = L.ZpulseWall()
    L.lib().osr_zpulse_wall_defaults(C.byref(z))
    d = O.ZPULSE_DEFAULTS
    for k in ("a0", "omega0", "phase", "pol", "pol_type", "tenv_fwhm", "tenv_rise", "tenv_flat",
              "tenv_fall", "tenv_duration", "per_w0", "per_fwhm", "launch_time"):
        assert getattr(z, k) == d[k], k
    assert list(z.per_tem_mode) == [d["per_tem_mode"]] * 2
    unset = -np.finfo(np.float64).max  # the reference's -huge(1.0)
    for k in ("tenv_range", "per_focus"):
        assert d[k] is None and getattr(z, k) == unset, k
    assert d["per_center"] is None and list(z.per_center) == [unset] * 2
    assert (z.propagation, z.tenv_type, z.per_type) == (L.OSR_PROP[d["propagation"]],
                                                        L.OSR_TENV[d["tenv_type"]],
                                                        L.OSR_PER[d["per_type"]])


def test_fortran_binding_layout():
    """The bind(c) osr_zpulse_wall and osr_zpulse of m_osiris_amd have the C sizes and,
    with their default initialisers, the values osr_zpulse_wall_defaults and
    osr_zpulse_defaults write (host code, no GPU)."""
    import os
    import subprocess

    import osiris_amd._lib as L
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "osiris_amd", "lib", "zpulse_layout_example")
    if not os.path.exists(exe):
        pytest.skip("Fortran example not built (no Fortran compiler)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == f"size {C.sizeof(L.ZpulseWall)}"
    assert f"size2 {C.sizeof(L.ZpulseBox)}" in lines
    for ft, ct in (("fortran ", "c "), ("fortran2 ", "c2 ")):
        f = [ln.split(None, 1)[1] for ln in lines if ln.startswith(ft)]
        c = [ln.split(None, 1)[1] for ln in lines if ln.startswith(ct)]
        assert len(f) == 2 and f == c
