"""Open x walls on the CPU: the oracle's Lindman absorbing field boundary
(emf/boundary/os-emf-lindman.f90) and absorbing particles (spec/os-spec-boundary.f03:
476-513) against what the mathematics fixes.

* The energy a light pulse leaves behind after crossing a wall is the spectral average
  of |R(k)|^2, R being the closed-form reflection coefficient of the discrete Mur wall on
  the Yee lattice (tests/pic_inputs.mur_reflection); a mistake planted in the oracle's
  wall coefficient moves the result far from it.
* Particles of negligible charge stream freely; the number left after every step is
  fixed by the step at which each one's cell index leaves 1..nx.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.pic_inputs import mur_reflection, vacuum_pulse

NX, NY, DX, DT = 240, 4, 0.2, 0.12


def _pulse_run(direction, prec="f64", plant=None, k0=4.0):
    s = O.Sim(2, (NX, NY), (DX, DX), DT, 2, prec=prec, open_x=True)
    if plant == "sign":
        s.walls.c.disp_corr = -s.walls.c.disp_corr
    e, b = vacuum_pulse(NX, NY, DX, k0=k0, direction=direction)
    s.set_interior("e", e)
    s.set_interior("b", b)
    steps = int(1.1 * NX * DX / DT)  # the pulse has left; the reflection has not come back
    s.step(steps)
    fe = np.array(s.fld_hist).sum(1)
    return fe[0], fe[-1], s


def _predicted_fraction(k0):
    e, _ = vacuum_pulse(NX, NY, DX, k0=k0)
    F = np.abs(np.fft.fft(e[0, :, 1])) ** 2
    k = np.fft.fftfreq(NX) * 2 * np.pi
    return float((np.abs(mur_reflection(k, DT / DX)) ** 2 * F).sum() / F.sum())


@pytest.mark.parametrize("k0", [4.0, 2.5])
@pytest.mark.parametrize("direction", [1, -1])
def test_lindman_reflection_is_the_mur_closed_form(direction, k0):
    e0, e1, _ = _pulse_run(direction, k0=k0)
    got = e1 / e0
    want = _predicted_fraction(k0)
    assert want < 2e-3
    assert abs(got / want - 1.0) < 0.02, (got, want)


def test_lindman_3d_reflection_is_the_mur_closed_form():
    """A y-z uniform pulse on the 3D lattice obeys the same 1D Mur reflection."""
    nx, dx, dt, k0 = 160, 0.2, 0.1, 4.0
    s = O.Sim(3, (nx, 4, 4), (dx, dx, dx), dt, 2, open_x=True)
    xi = np.arange(1, nx + 1) * dx
    x0 = nx * dx / 2

    def f(x):
        return np.exp(-((x - x0) / 2.0) ** 2) * np.cos(k0 * (x - x0))

    s.grid.interior(s.e)[..., 1] = f(xi)
    s.grid.interior(s.b)[..., 2] = f(xi + dx / 2)
    s.step(int(1.1 * nx * dx / dt))
    fe = np.array(s.fld_hist).sum(1)
    F = np.abs(np.fft.fft(f(xi))) ** 2
    k = np.fft.fftfreq(nx) * 2 * np.pi
    want = float((np.abs(mur_reflection(k, dt / dx)) ** 2 * F).sum() / F.sum())
    assert abs(fe[-1] / fe[0] / want - 1.0) < 0.02, (fe[-1] / fe[0], want)


def test_lindman_planted_sign_mistake_is_caught():
    e0, e1, _ = _pulse_run(1, plant="sign")
    assert e1 / e0 > 50 * _predicted_fraction(4.0)


def test_lindman_fp32_restatement_tracks_fp64():
    a = _pulse_run(1, prec="f64")
    b = _pulse_run(1, prec="f32")
    assert abs(b[1] / b[0] - a[1] / a[0]) < 0.02 * a[1] / a[0]


def test_wall_guard_values():
    """The values update_emfbound_lindman puts on the walls: E1 = 0 at x = 0 and
    x = nx+1 (never written into the wall buffers), E2/E3 at x = 0 and B at x = 0 from
    the lower wall buffers, whose E entries hold the last Mur extrapolation."""
    _, _, s = _pulse_run(1)
    s.wall_refresh()
    g = s.grid
    e, b = s.e, s.b
    i0, j0 = g.gcl[0] - 1, g.gcl[1]  # array column of x = 0, row of y = 1
    rows = slice(j0, j0 + NY)
    assert not e[:, i0, 0].any() and not e[:, i0 + NX + 1, 0].any()
    assert np.array_equal(s.walls.ebl[0, rows, 0, 1:], e[rows, i0, 1:])
    assert np.array_equal(s.walls.bbl[0, rows, 0, :], b[rows, i0, :])
    assert np.abs(e[rows, i0, 1:]).max() > 0


def _free_streaming(seed, steps=40, n=400):
    rng = np.random.default_rng(seed)
    nx, ny, dx, dt = 24, 6, 0.2, 0.1
    ix = np.stack([rng.integers(1, nx + 1, n), rng.integers(1, ny + 1, n)], 1).astype(np.int32)
    x = rng.uniform(-0.5, 0.5, (n, 2))
    u = np.zeros((n, 3))
    u[:, 0] = rng.uniform(-1.5, 1.5, n)
    u[:, 1] = rng.uniform(-0.5, 0.5, n)
    q = np.full(n, -1e-12)
    s = O.Sim(2, (nx, ny), (dx, dx), dt, 2, open_x=True)
    s.add_species(-1.0, 0, ix, x, u, q)
    # X = ix + x in cell units; the particle is removed after the first push that takes
    # its cell index out of 1..nx, i.e. X out of [0.5, nx + 0.5)
    v = u[:, 0] / np.sqrt(1.0 + (u ** 2).sum(1)) * dt / dx
    X = ix[:, 0] + x[:, 0]
    k = np.arange(1, steps + 1)[:, None]
    pos = X[None, :] + k * v[None, :]
    gone = (pos < 0.5) | (pos >= nx + 0.5)
    first = np.where(gone.any(0), gone.argmax(0) + 1, steps + 1)
    margin = np.min(np.minimum(np.abs(pos - 0.5), np.abs(pos - nx - 0.5)))
    want = [(first > s_).sum() for s_ in range(1, steps + 1)]
    got = []
    for _ in range(steps):
        s.step(1)
        got.append(len(s.species[0].q))
    return want, got, margin


@pytest.mark.parametrize("seed", [1, 2])
def test_absorbing_particles_leave_at_the_predicted_step(seed):
    want, got, margin = _free_streaming(seed)
    assert margin > 1e-6  # no particle sits on a wall within round-off
    assert got == want
    assert got[-1] < got[0] < 400


def test_open_sim_rejects_smoothing():
    with pytest.raises(ValueError):
        O.Sim(2, (8, 8), (0.2, 0.2), 0.1, 2, open_x=True, smooth_order=(1, 0),
              smooth_coef=np.zeros((3, 31)))
