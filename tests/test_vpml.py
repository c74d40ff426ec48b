"""VPML absorbing x walls on the CPU: the oracle's restatement of emf/boundary/
os-emf-vpml.f03 (2D, x walls, no corners) against what the mathematics fixes.

* Attenuation switched off (every coefficient the plain Yee one, c1 = 1, c2 = c3 = dt/dx1):
  the split-field wall is then the Yee lattice continued past the box, so the box fields
  and the wall E (E3 = ZX + ZY) equal a periodic run on a grid 2n - 1 cells wider, to FP64
  round-off, for as long as no signal reaches the wall's far end.  Random fields over the
  box (y structure included) fix the interface cells, each stencil offset and sign
  and the order of the wall updates inside the step; a sign planted in the oracle's
  y-derivative term breaks it.
* The reference's attenuation: a light pulse at normal and at oblique incidence leaves
  orders of magnitude less energy behind than the first-order Lindman wall does, and the
  attenuation profile reversed (a plant) loses that.
* The coefficients of thickness 8 are those of the reference's documented special case
  s1 = -2, s2 = 0.2 (os-emf-vpml.f03:3618-3621), evaluated in FP64.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.pic_inputs import oblique_packet, oblique_packet_3d, vacuum_pulse


def _coef_view(s):
    return np.ctypeslib.as_array(s.walls.c.coef)  # (2, 2, 3, VPML_MAX), writable


def _random_box(rng, nx, ny):
    """Random fields, zero within 3 cells of either wall: the walls start at zero, which
    is the continued lattice's state only where the first half step of B sees no field."""
    e = rng.normal(0, 1, (ny, nx, 3))
    b = rng.normal(0, 1, (ny, nx, 3))
    for f in (e, b):
        f[:, :3] = 0
        f[:, -3:] = 0
    return e, b


@pytest.mark.parametrize("prec", ["f64"])
@pytest.mark.parametrize("plant", [None, "dt2_sign"])
def test_unattenuated_wall_is_the_continued_lattice(prec, plant):
    nx, ny, n, dx, dt, steps = 40, 12, 48, (0.2, 0.25), 0.1, 40
    rng = np.random.default_rng(3)
    e, b = _random_box(rng, nx, ny)
    s = O.Sim(2, (nx, ny), dx, dt, 2, prec=prec, open_x=True, xwall="vpml", vpml_size=n)
    c = _coef_view(s)
    d1 = s.real(dt / dx[0])
    c[:, :, 0, :] = 1.0
    c[:, :, 1:, :] = d1
    if plant == "dt2_sign":
        s.walls.c.dtdx2 = -s.walls.c.dtdx2
    s.set_interior("e", e)
    s.set_interior("b", b)
    big = O.Sim(2, (nx + 2 * n - 1, ny), dx, dt, 2, prec=prec)
    eb = np.zeros((ny, nx + 2 * n - 1, 3))
    bb = np.zeros_like(eb)
    eb[:, n - 1:n - 1 + nx] = e
    bb[:, n - 1:n - 1 + nx] = b
    big.set_interior("e", eb)
    big.set_interior("b", bb)
    s.step(steps)
    big.step(steps)
    g, G = s.grid, big.grid
    y = slice(g.gcl[1], g.gcl[1] + ny)
    # box columns x = 0 .. nx+1 (E also nx+2) <-> big grid X = x + n - 1
    sx = slice(g.gcl[0] - 1, g.gcl[0] + nx + 1)
    X = slice(G.gcl[0] + n - 2, G.gcl[0] + n + nx)
    scale = max(np.abs(big.e).max(), np.abs(big.b).max())
    err_e = np.abs(s.e[y, sx] - big.e[y, X]).max() / scale
    err_b = np.abs(s.b[y, sx] - big.b[y, X]).max() / scale
    err_e2 = np.abs(s.e[y, g.gcl[0] + nx + 1] - big.e[y, G.gcl[0] + n + nx]).max() / scale
    # wall E: lower l = 1 .. n-2 (x = 3-n .. 0), upper l = 1 .. n-1 (x = nx+2 .. nx+n)
    w = s.walls
    r = slice(w.c.gc1, w.c.gc1 + ny)
    errs = []
    for wl, ls, X0 in ((0, range(1, n - 1), G.gcl[0]), (1, range(1, n), G.gcl[0] + nx + n - 1)):
        we = w.e[wl][0, r]
        for l in ls:
            got = np.stack([we[:, l, 0], we[:, l, 1], we[:, l, 2] + we[:, l, 3]], -1)
            errs.append(np.abs(got - big.e[y, X0 + l]).max() / scale)
    err_w = max(errs)
    reach = np.abs(big.e[y, G.gcl[0]:G.gcl[0] + n - 1]).max()  # the signal entered the wall
    assert reach > 1e-3 * scale
    if plant:
        assert max(err_e, err_b, err_w) > 1e-6
        return
    assert max(err_e, err_b, err_e2, err_w) < 1e-12, (err_e, err_b, err_e2, err_w)


def _reflected(xwall, direction=1, plant=None, oblique=False, n=8):
    """Field energy in the box after the pulse has left, of the difference between the
    walled run and a run on a periodic grid wide enough that nothing comes back (so that
    any static part of the initial fields cancels), over the initial field energy."""
    nx, ny, dx, dt = 240, 32, 0.2, 0.12
    steps = int(1.1 * nx * dx / dt)
    s = O.Sim(2, (nx, ny), (dx, dx), dt, 2, prec="f64", open_x=True, xwall=xwall,
              vpml_size=n)
    if plant == "reverse":
        c = _coef_view(s)
        c[..., :n] = c[..., :n][..., ::-1].copy()
    if not oblique:
        e, b = vacuum_pulse(nx, ny, dx, k0=4.0, direction=direction)
    else:
        e, b = oblique_packet(nx, ny, dx, direction=direction)
    s.set_interior("e", e)
    s.set_interior("b", b)
    s.step(steps)
    m = steps + 8  # margin no signal crosses
    big = O.Sim(2, (nx + 2 * m, ny), (dx, dx), dt, 2, prec="f64")
    eb = np.zeros((ny, nx + 2 * m, 3))
    bb = np.zeros_like(eb)
    eb[:, m:m + nx] = e
    bb[:, m:m + nx] = b
    big.set_interior("e", eb)
    big.set_interior("b", bb)
    big.step(steps)
    de = O.Grid.interior(s.grid, s.e) - O.Grid.interior(big.grid, big.e)[:, m:m + nx]
    db = O.Grid.interior(s.grid, s.b) - O.Grid.interior(big.grid, big.b)[:, m:m + nx]
    return float(((de ** 2).sum() + (db ** 2).sum()) / ((e ** 2).sum() + (b ** 2).sum()))


@pytest.mark.parametrize("direction", [1, -1])
def test_vpml_absorbs_normal_incidence(direction):
    got = _reflected("vpml", direction)
    lind = _reflected("lindman", direction)
    assert got < 1e-6 and got < 1e-3 * lind, (got, lind)
    assert _reflected("vpml", direction, plant="reverse") > 100 * got


def test_vpml_absorbs_oblique_incidence():
    got = _reflected("vpml", 1, oblique=True)
    lind = _reflected("lindman", 1, oblique=True)
    assert got < 1e-5 and got < 1e-2 * lind, (got, lind)
    assert _reflected("vpml", 1, plant="reverse", oblique=True) > 100 * got


def test_coefficients_of_the_documented_case():
    """Thickness 8: s1 = -2, s2 = 0.2; tj = exp(-2 (0.2 j)^4) at the lower wall's points
    j - 1/2, j, j + 1/2 and the upper wall's j, j + 1/2, j + 1 (FP64 evaluation)."""
    g = O.make_grid(2, (16, 8), 2)
    dt, dx = 0.1, 0.2
    w = O.VpmlWalls(g, (dx, dx), dt, np.float64, 8)
    c = w.coefs()
    d = dt / dx
    r = (1 - d) / (1 + d)

    def t(x):
        return np.exp(-2.0 * (0.2 * x) ** 4)

    j = np.arange(8.0)
    b1e, b1b = d * (1 + r * (1 - t(j))), d * (1 + r * (1 - t(j + 0.5)))
    lo_e = np.stack([1 - b1e + t(j) * d, t(j - 0.5) * b1e, np.full(8, d)])[:, ::-1]
    lo_b = np.stack([1 - b1b + t(j + 0.5) * d, t(j) * b1b, np.full(8, d)])[:, ::-1]
    b1e, b1b = d * (1 + r * (1 - t(j + 0.5))), d * (1 + r * (1 - t(j + 1)))
    up_e = np.stack([1 - b1e + t(j + 0.5) * d, np.full(8, d), t(j) * b1e])
    up_b = np.stack([1 - b1b + t(j + 1) * d, np.full(8, d), t(j + 0.5) * b1b])
    for got, want in ((c[0, 0], lo_e), (c[0, 1], lo_b), (c[1, 0], up_e), (c[1, 1], up_b)):
        assert np.allclose(got, want, rtol=1e-14, atol=1e-16)
    # the FP32 coefficients are the FP32 evaluation of the same expressions
    w32 = O.VpmlWalls(g, (dx, dx), dt, np.float32, 8)
    assert np.allclose(w32.coefs(), c, rtol=1e-6, atol=1e-7)
    # deep in the wall the E and B points see the full attenuation, at the box none
    assert c[0, 0, 1, 0] < 1e-2 * d and c[1, 0, 2, -1] < 1e-2 * d
    assert abs(c[0, 0, 0, -1] - 1) < 1e-3 and abs(c[1, 0, 0, 0] - 1) < 1e-3


def test_vpml_rejects():
    with pytest.raises(ValueError):
        O.Sim(2, (16, 8), (0.2, 0.2), 0.1, 2, open_x=True, xwall="vpml", vpml_size=3)


@pytest.mark.parametrize("plant", [None, "dt3_sign"])
def test_unattenuated_wall_is_the_continued_lattice_3d(plant):
    """3D: six split components; random fields with y and z structure."""
    nx, ny, nz, n, steps = 20, 6, 5, 24, 20
    dx, dt = (0.2, 0.25, 0.3), 0.08
    rng = np.random.default_rng(5)
    e = rng.normal(0, 1, (nz, ny, nx, 3))
    b = rng.normal(0, 1, (nz, ny, nx, 3))
    for f in (e, b):
        f[:, :, :3] = 0
        f[:, :, -3:] = 0
    s = O.Sim(3, (nx, ny, nz), dx, dt, 2, open_x=True, xwall="vpml", vpml_size=n)
    c = _coef_view(s)
    c[:, :, 0, :] = 1.0
    c[:, :, 1:, :] = dt / dx[0]
    if plant == "dt3_sign":
        s.walls.c.dtdx3 = -s.walls.c.dtdx3
    s.set_interior("e", e)
    s.set_interior("b", b)
    W = nx + 2 * n - 1
    big = O.Sim(3, (W, ny, nz), dx, dt, 2)
    eb = np.zeros((nz, ny, W, 3))
    bb = np.zeros_like(eb)
    eb[:, :, n - 1:n - 1 + nx] = e
    bb[:, :, n - 1:n - 1 + nx] = b
    big.set_interior("e", eb)
    big.set_interior("b", bb)
    s.step(steps)
    big.step(steps)
    g, G = s.grid, big.grid
    yz = (slice(g.gcl[2], g.gcl[2] + nz), slice(g.gcl[1], g.gcl[1] + ny))
    sx = slice(g.gcl[0] - 1, g.gcl[0] + nx + 1)
    X = slice(G.gcl[0] + n - 2, G.gcl[0] + n + nx)
    scale = max(np.abs(big.e).max(), np.abs(big.b).max())
    err_e = np.abs(s.e[yz + (sx,)] - big.e[yz + (X,)]).max() / scale
    err_b = np.abs(s.b[yz + (sx,)] - big.b[yz + (X,)]).max() / scale
    w = s.walls
    r = (slice(w.c.gc3, w.c.gc3 + nz), slice(w.c.gc1, w.c.gc1 + ny))
    errs = []
    for wl, ls, X0 in ((0, range(1, n - 1), G.gcl[0]), (1, range(1, n), G.gcl[0] + nx + n - 1)):
        we = w.e[wl][r]
        for l in ls:
            got = we[:, :, l, 0::2] + we[:, :, l, 1::2]
            errs.append(np.abs(got - big.e[yz + (X0 + l,)]).max() / scale)
    err_w = max(errs)
    assert np.abs(big.e[yz + (slice(G.gcl[0], G.gcl[0] + n - 1),)]).max() > 1e-3 * scale
    if plant:
        assert max(err_e, err_b, err_w) > 1e-6
        return
    assert max(err_e, err_b, err_w) < 1e-12, (err_e, err_b, err_w)


@pytest.mark.parametrize("direction", [1, -1])
def test_vpml_3d_absorbs_oblique_incidence(direction):
    """A packet along (kx, ky, kz) in 3D, both polarisations (E along z^ x k^ and along
    k^ x (z^ x k^)); the reflection is measured as in 2D."""
    nx, ny, nz, dx, dt, k0 = 80, 16, 16, 0.2, 0.1, 4.0
    steps = int(1.1 * nx * dx / dt)

    def run(xwall, plant=None):
        s = O.Sim(3, (nx, ny, nz), (dx,) * 3, dt, 2, open_x=True, xwall=xwall)
        if plant:
            c = _coef_view(s)
            c[..., :8] = c[..., :8][..., ::-1].copy()
        e, b = oblique_packet_3d(nx, ny, nz, dx, k0, direction)
        s.set_interior("e", e)
        s.set_interior("b", b)
        s.step(steps)
        return s, e, b

    s, e, b = run("vpml")
    m = steps + 8
    big = O.Sim(3, (nx + 2 * m, ny, nz), (dx,) * 3, dt, 2)
    eb = np.zeros((nz, ny, nx + 2 * m, 3))
    bb = np.zeros_like(eb)
    eb[:, :, m:m + nx] = e
    bb[:, :, m:m + nx] = b
    big.set_interior("e", eb)
    big.set_interior("b", bb)
    big.step(steps)
    e0 = (e ** 2).sum() + (b ** 2).sum()

    def refl(sim):
        de = O.Grid.interior(sim.grid, sim.e) - O.Grid.interior(big.grid, big.e)[:, :, m:m + nx]
        db = O.Grid.interior(sim.grid, sim.b) - O.Grid.interior(big.grid, big.b)[:, :, m:m + nx]
        return float(((de ** 2).sum() + (db ** 2).sum()) / e0)

    got = refl(s)
    lind = refl(run("lindman")[0])
    assert got < 1e-5 and got < 1e-2 * lind, (got, lind)
    assert refl(run("vpml", plant=True)[0]) > 100 * got


@pytest.mark.parametrize("plant", [None, "unsaved"])
def test_diffusion_keeps_the_split_sums(plant):
    """vpml_diffuse mixes the split pair (ZX, ZY) of B before each E update, keeping their
    sum: with the attenuation off it leaves the continued lattice unchanged to round-off
    (both orders of guard rows: order 1 has a single lower guard row), and a mixing that
    does not keep the sum (ZY mixed with the already mixed ZX, planted in the oracle's
    result by rerunning with the pair swapped) is caught."""
    for order in (1, 2):
        nx, ny, n, dx, dt, steps = 40, 12, 48, (0.2, 0.25), 0.1, 40
        rng = np.random.default_rng(3)
        e, b = _random_box(rng, nx, ny)
        runs = []
        for diffuse in ((False, False), (True, True)):
            s = O.Sim(2, (nx, ny), dx, dt, order, open_x=True, xwall="vpml", vpml_size=n,
                      vpml_diffuse=diffuse)
            c = _coef_view(s)
            c[:, :, 0, :] = 1.0
            c[:, :, 1:, :] = dt / dx[0]
            s.set_interior("e", e)
            s.set_interior("b", b)
            s.step(steps)
            runs.append(s)
        plain, mixed = runs
        scale = np.abs(plain.e).max()
        wb0, wb1 = plain.walls.b[0], mixed.walls.b[0]
        assert np.abs(wb0[..., 2] - wb1[..., 2]).max() > 1e-3  # the split did change
        if plant == "unsaved":
            wb1 = wb1.copy()
            wb1[..., 3] = 0.9 * wb1[..., 3] + 0.1 * wb1[..., 2]  # a mixing that loses the sum
        s_sum = wb1[..., 2] + wb1[..., 3]
        p_sum = wb0[..., 2] + wb0[..., 3]
        err = np.abs(s_sum - p_sum).max() / np.abs(p_sum).max()
        if plant:
            assert err > 1e-6
            continue
        assert err < 1e-12
        assert np.abs(mixed.e - plain.e).max() < 1e-12 * scale
        assert np.abs(mixed.b - plain.b).max() < 1e-12 * scale


def test_diffusion_still_absorbs():
    """With the reference's attenuation the diffused wall still absorbs a packet at
    oblique incidence, at the cost of absorption (measured 1.6e-4 left against 7e-9 for
    the plain wall and 7e-3 for Lindman), and gives another result than the plain wall."""
    nx, ny, dx, dt = 240, 32, 0.2, 0.12
    steps = int(1.1 * nx * dx / dt)
    e, b = oblique_packet(nx, ny, dx)
    out = []
    for diffuse in ((False, False), (True, True)):
        s = O.Sim(2, (nx, ny), (dx, dx), dt, 2, open_x=True, xwall="vpml", vpml_diffuse=diffuse)
        s.set_interior("e", e)
        s.set_interior("b", b)
        s.step(steps)
        out.append(s)
    d = np.abs(out[0].e - out[1].e).max()
    assert d > 0
    m = steps + 8
    big = O.Sim(2, (nx + 2 * m, ny), (dx, dx), dt, 2)
    eb = np.zeros((ny, nx + 2 * m, 3))
    bb = np.zeros_like(eb)
    eb[:, m:m + nx] = e
    bb[:, m:m + nx] = b
    big.set_interior("e", eb)
    big.set_interior("b", bb)
    big.step(steps)
    g, G = out[1].grid, big.grid
    de = O.Grid.interior(g, out[1].e) - O.Grid.interior(G, big.e)[:, m:m + nx]
    db = O.Grid.interior(g, out[1].b) - O.Grid.interior(G, big.b)[:, m:m + nx]
    refl = ((de ** 2).sum() + (db ** 2).sum()) / ((e ** 2).sum() + (b ** 2).sum())
    assert refl < 1e-3, refl
