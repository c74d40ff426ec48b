"""The initial electrostatic field (init_fields -> es_solver_2d, emf/os-emf-es-solver.f03:
125-236) in the CPU oracle, against what the mathematics fixes.  The reference sums the
free-space field of line charges, charge * r / r^2 without the 1/(2 pi) of the 2D Green
function, so its field satisfies div E = 2 pi rho:

* a smooth charge blob: the discrete divergence at the cell corners equals 2 pi rho to
  the lattice error (1%), and the discrete curl vanishes;
* a point charge: E1 and E2 are exactly antisymmetric about it on the Yee lattice, and
  far away the field is the monopole Q r / r^2;
* a deliberately shifted lattice (the E1 offset applied to E2) fails the Gauss check.

The 3D beam branch (es_solver_3d_beam, :795-1003) sums the fields of infinite wires
along x in each x slice, lambda / (2 pi r) with lambda = rho dy dz: its boosted field
obeys the discrete Gauss law of the (y, z) plane to the lattice error, B = beta x E, a
single wire gives the exact mirror symmetries and the 1/r far field, and the sum with an
unshifted E2 misses the Gauss law by several times that error.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

N, DX = 64, (0.1, 0.1)
DXA = np.array([DX[0], DX[1], 1.0])


def _solve(rho_int, prec="f64"):
    g = O.make_grid(2, (N, N), 2)
    L = O._lib(prec)
    r = g.zeros(1, L._real)
    g.interior(r)[..., 0] = rho_int
    ef = g.zeros(3, L._real)
    L.orc_es_solver_2d(C.byref(g.c()), O._p(DXA), O._p(r), O._p(ef))
    return g.interior(ef)


def _blob(sigma=4.0):
    x = np.arange(1, N + 1)
    X, Y = np.meshgrid(x, x, indexing="xy")
    c = N / 2 + 0.5
    return np.exp(-((X - c) ** 2 + (Y - c) ** 2) / (2 * sigma ** 2))


def _div(E):
    return ((E[1:, 1:, 0] - E[1:, :-1, 0]) / DX[0] + (E[1:, 1:, 1] - E[:-1, 1:, 1]) / DX[1])


def test_gauss_law_and_zero_curl():
    rho = _blob()
    E = _solve(rho)
    core = rho[1:, 1:] > 0.5
    ratio = _div(E)[core] / (2 * np.pi * rho[1:, 1:][core])
    assert np.all(np.abs(ratio - 1) < 0.01), (ratio.min(), ratio.max())
    curl = ((E[:-1, 1:, 1] - E[:-1, :-1, 1]) / DX[0] - (E[1:, :-1, 0] - E[:-1, :-1, 0]) / DX[1])
    assert np.abs(curl).max() < 1e-3 * np.abs(E).max() / DX[0]
    assert not E[..., 2].any()


def test_shifted_lattice_fails_gauss():
    """The same sum with E2 sampled at E1's offsets: the check above rejects it."""
    rho = _blob()
    x = np.arange(1, N + 1)
    ei, ri = np.meshgrid(x, x, indexing="ij")
    d = (x[:, None] - x[None, :]).astype(np.float64)  # target - source, per axis
    E1 = np.zeros((N, N))
    E2 = np.zeros((N, N))
    for (sj, si) in zip(*np.nonzero(rho > 1e-12)):
        q = rho[sj, si]
        a1 = (d[:, si] + 0.5) * DX[0]  # x offsets of E1 (x + 1/2)
        b1 = d[:, sj] * DX[1]
        r2 = a1[None, :] ** 2 + b1[:, None] ** 2
        E1 += q * a1[None, :] / r2
        E2 += q * b1[:, None] / r2  # wrong: E2 belongs at (x, y + 1/2)
    E = np.stack([E1, E2], -1) * DX[0] * DX[1]
    core = rho[1:, 1:] > 0.5
    ratio = _div(E)[core] / (2 * np.pi * rho[1:, 1:][core])
    assert np.abs(ratio - 1).max() > 0.05


def test_point_charge_symmetry_and_far_field():
    rho = np.zeros((N, N))
    s = 20  # source at x = y = 21 (1-based)
    rho[s, s] = 3.0
    E = _solve(rho)
    # E1 lives at x + 1/2: E1(s + k) = -E1(s - 1 - k) along x, even along y about y = s
    k = np.arange(0, 15)
    assert np.array_equal(E[s, s + k, 0], -E[s, s - 1 - k, 0])
    assert np.array_equal(E[s + k, s, 1], -E[s - 1 - k, s, 1])
    assert np.array_equal(E[s + 5, :, 0], E[s - 5, :, 0])
    # far field: E = Q dV r / r^2 exactly for one source (the sum has one term)
    Q = 3.0 * DX[0] * DX[1]
    j, i = 60, 50  # target E1 point (x + 1/2, y)
    rx, ry = (i - s + 0.5) * DX[0], (j - s) * DX[1]
    assert abs(E[j, i, 0] / (Q * rx / (rx * rx + ry * ry)) - 1) < 1e-12


def test_fp32_accumulation_tracks_fp64():
    rho = _blob(3.0)
    a = _solve(rho, "f64")
    b = _solve(rho.astype(np.float32), "f32")
    assert np.abs(b - a).max() < 2e-5 * np.abs(a).max()


@pytest.mark.parametrize("u1", [0.5, -2.0])
def test_beam_field_is_gauss_and_beta_e(u1):
    """es_solver_2d_beam: the field of charged sheets drifting along x.  After the
    solve, (E2(j) - E2(j-1)) / dy = rho(j) up to rounding between interior rows (discrete
    Gauss law along y), and B3 = beta * (E2(i) + E2(i+1)) / 2 with beta = u1 / gamma."""
    nx, ny = 16, 40
    g = O.make_grid(2, (nx, ny), 2)
    L = O._lib("f64")
    rng = np.random.default_rng(9)
    r = g.zeros(1, np.float64)
    rho = np.zeros((ny, nx))
    rho[10:30, :] = rng.uniform(0.5, 1.5, (20, nx))
    g.interior(r)[..., 0] = rho
    ef = g.zeros(3, np.float64)
    bf = g.zeros(3, np.float64)
    L.orc_es_solver_2d_beam(C.byref(g.c()), O._p(DXA), O._p(r), float(np.float32(u1)), 0,
                            O._p(ef), O._p(bf))
    E = g.interior(ef)
    B = g.interior(bf)
    # interior rows (the guard rows hold periodic images, which a finite column of sheets
    # does not have)
    gauss = (E[1:, :, 1] - E[:-1, :, 1]) / DX[1]
    assert np.allclose(gauss, rho[1:], rtol=0, atol=1e-12 * np.abs(rho).max())
    gamma = np.sqrt(float(np.float32(u1) ** 2) + 1.0)  # u1^2 in float32, the rest in FP64
    e2p = np.concatenate([E[:, 1:, 1], E[:, :1, 1]], axis=1)  # periodic x + 1
    want = np.float32(u1) / gamma * 0.5 * (E[..., 1] + e2p)
    assert np.allclose(B[..., 2], want, rtol=1e-12, atol=1e-14)
    assert not E[..., 0].any() and not E[..., 2].any() and not B[..., :2].any()


def _beam3d(rho_int, u1, dx, prec="f64", fn="orc_es_solver_3d_beam"):
    nz, ny, nx = rho_int.shape
    g = O.make_grid(3, (nx, ny, nz), 2)
    L = O._lib(prec)
    r = g.zeros(1, L._real)
    g.interior(r)[..., 0] = rho_int
    ef = g.zeros(3, L._real)
    bf = g.zeros(3, L._real)
    getattr(L, fn)(C.byref(g.c()), O._p(np.array(dx, np.float64)), O._p(r),
                   float(np.float32(u1)), 0, O._p(ef), O._p(bf))
    return g.interior(ef), g.interior(bf)


def _blob3d(nx, ny, nz, dx, width=6.0):
    """A smooth charge column profile in (y, z), the same in every x slice but scaled per
    slice, so each slice is its own 2D problem."""
    y = (np.arange(ny) - ny / 2 + 0.25) * dx[1]
    z = (np.arange(nz) - nz / 2 - 0.25) * dx[2]
    w = width * min(dx[1], dx[2])
    prof = np.exp(-(y[None, :] ** 2 + z[:, None] ** 2) / (2 * w * w))
    scale = 1.0 + 0.5 * np.arange(nx) / nx
    return prof[:, :, None] * scale[None, None, :]


@pytest.mark.parametrize("u1", [0.4, 2.5])
def test_beam3d_gauss_law_in_each_slice_and_b_is_beta_cross_e(u1):
    """es_solver_3d_beam: each x slice holds the field of infinite wires along x.  The
    boosted transverse field satisfies the discrete Gauss law of the (y, z) plane,
    (E2(j) - E2(j-1)) / dy + (E3(k) - E3(k-1)) / dz = rho, to the lattice error of a
    smooth blob (the wire field lambda / (2 pi r) with lambda = rho dy dz, charge_norm =
    dy dz / gamma / 2 / pi, times gamma); its discrete curl vanishes to the same order;
    B = beta x E with beta = u1 / gamma: B2 = -beta (E3(i) + E3(i+1)) / 2 and
    B3 = beta (E2(i) + E2(i+1)) / 2 to rounding; E1 = B1 = 0."""
    nx, ny, nz, dx = 4, 48, 40, (0.3, 0.1, 0.12)
    rho = _blob3d(nx, ny, nz, dx)
    E, B = _beam3d(rho, u1, dx)
    gamma = np.sqrt(float(np.float32(u1) ** 2) + 1.0)
    div = ((E[1:, 1:, :, 1] - E[1:, :-1, :, 1]) / dx[1] +
           (E[1:, 1:, :, 2] - E[:-1, 1:, :, 2]) / dx[2])
    want = rho[1:, 1:]
    err = np.abs(div - want).max() / np.abs(rho).max()
    assert err < 5e-3, err
    curl = ((E[:-1, 1:, :, 2] - E[:-1, :-1, :, 2]) / dx[1] -
            (E[1:, :-1, :, 1] - E[:-1, :-1, :, 1]) / dx[2])
    assert np.abs(curl).max() < 5e-3 * np.abs(E[..., 1:]).max() / min(dx[1:])
    beta = np.float32(u1) / gamma
    xp = lambda a: np.concatenate([a[..., 1:], a[..., :1]], axis=-1)  # noqa: E731 periodic i+1
    assert np.allclose(B[..., 2], beta * 0.5 * (E[..., 1] + xp(E[..., 1])), rtol=1e-12, atol=1e-14)
    assert np.allclose(B[..., 1], -beta * 0.5 * (E[..., 2] + xp(E[..., 2])), rtol=1e-12,
                       atol=1e-14)
    assert not E[..., 0].any() and not B[..., 0].any()


def test_beam3d_point_wire_symmetry_and_slice_locality():
    """One charged cell: E2 is antisymmetric about the wire along y and symmetric along z
    (E3 the other way round) on the staggered points, exactly; far from the wire the
    field is gamma * lambda / (2 pi r) with lambda = q dy dz / gamma, the 1/r of a line
    charge; and only the source's x slice carries E (B averages it with the slice below)."""
    nx, ny, nz, dx, u1 = 5, 41, 41, (0.2, 0.1, 0.1), 0.7
    rho = np.zeros((nz, ny, nx))
    rho[20, 20, 2] = 3.0
    E, B = _beam3d(rho, u1, dx)
    e2, e3 = E[..., 2, 1], E[..., 2, 2]
    # E2 at y + 1/2: mirror j -> 39 - j about the wire (y = 20), k -> 40 - k symmetric
    assert np.array_equal(e2[:, :40], -e2[:, :40][:, ::-1])
    assert np.array_equal(e2, e2[::-1, :])
    assert np.array_equal(e3[:40, :], -e3[:40, :][::-1, :])
    assert np.array_equal(e3, e3[:, ::-1])
    # far field along +y at z = 20: E2(j) sits at (j + 1/2 - 20) dy from the wire
    j = np.arange(30, 40)
    r = (j + 0.5 - 20) * dx[1]
    far = 3.0 * dx[1] * dx[2] / (2 * np.pi * r)
    assert np.allclose(e2[20, j], far, rtol=1e-12)
    others = [k for k in range(nx) if k != 2]
    assert not E[..., others, :].any()
    assert B[..., 1, 2].any() and B[..., 2, 2].any()
    assert not B[..., [0, 3, 4], :].any()


def test_beam3d_planted_unshifted_e2_fails_gauss():
    """The same sum with E2 on the cell-centred y points (no half shift) misses the Gauss
    law by several times the lattice error of the right sum (2.9% against 0.3% here): the
    check sees the staggering."""
    nx, ny, nz, dx, u1 = 2, 48, 40, (0.3, 0.1, 0.12), 0.4
    rho = _blob3d(nx, ny, nz, dx)
    gamma = np.sqrt(float(np.float32(u1) ** 2) + 1.0)
    cn = dx[1] * dx[2] / gamma / 2 / np.pi
    e2 = np.zeros((nz, ny))
    e3 = np.zeros((nz, ny))
    jj, kk = np.meshgrid(np.arange(ny), np.arange(nz))
    for rk in range(nz):
        for rj in range(ny):
            q = rho[rk, rj, 0]
            dy = (jj - rj) * dx[1]  # planted: should be (jj - rj + 0.5) dy
            dz = (kk - rk) * dx[2]
            r2 = dy * dy + dz * dz
            with np.errstate(divide="ignore", invalid="ignore"):
                e2 += np.where(r2 > 0, q * cn * dy / r2, 0.0)
            dzh = (kk - rk + 0.5) * dx[2]
            e3 += q * cn * dzh / (dy * dy + dzh * dzh)
    e2 *= gamma
    e3 *= gamma
    div = (e2[1:, 1:] - e2[1:, :-1]) / dx[1] + (e3[1:, 1:] - e3[:-1, 1:]) / dx[2]
    err = np.abs(div - rho[1:, 1:, 0]).max() / np.abs(rho).max()
    E, _ = _beam3d(rho, u1, dx)
    good = np.abs((E[1:, 1:, 0, 1] - E[1:, :-1, 0, 1]) / dx[1] +
                  (E[1:, 1:, 0, 2] - E[:-1, 1:, 0, 2]) / dx[2] - rho[1:, 1:, 0]).max()
    good /= np.abs(rho).max()
    assert good < 5e-3 and err > 0.02 and err > 5 * good, (good, err)


def test_beam3d_fp32_tracks_fp64():
    nx, ny, nz, dx, u1 = 3, 24, 20, (0.3, 0.1, 0.12), 0.9
    rho = _blob3d(nx, ny, nz, dx)
    E64, B64 = _beam3d(rho, u1, dx, "f64")
    E32, B32 = _beam3d(rho.astype(np.float32), u1, dx, "f32")
    assert np.abs(E32 - E64).max() <= 2e-5 * np.abs(E64).max()
    assert np.abs(B32 - B64).max() <= 2e-5 * np.abs(B64).max()
