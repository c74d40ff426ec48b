"""Current smoothing (vdf/os-vdf-smooth.f90) on the CPU: the oracle's coefficients and
filter against closed forms and independent constructions, and the library's coefficient
set-up (host code, no GPU) against the oracle bit for bit."""
import math

import numpy as np
import pytest

import osiris_amd as oa
from osiris_amd._lib import OsrError
from oracle import oracle as O
from oracle import smooth as S


def _taps(coef, d, n):
    return np.asarray(coef[d, S.ORDER_MAX - n:S.ORDER_MAX + n + 1], np.float64)


@pytest.mark.parametrize("n", range(1, 16))
def test_binomial_is_pascal_row(n):
    """n passes of (1,2,1)/4 = binomial coefficients C(2n, n+k) / 4^n."""
    order, coef = S.coefficients(2, (S.BINOMIAL, S.NONE), (n, 0), (1.0, 1.0))
    assert order == [n, 0, 0]
    want = np.array([math.comb(2 * n, n + k) for k in range(-n, n + 1)], np.float64) / 4.0 ** n
    assert np.allclose(_taps(coef, 0, n), want, rtol=2e-7, atol=0)
    assert not coef[1].any()


@pytest.mark.parametrize("n", range(2, 9))
def test_compensated_has_flat_response(n):
    """The compensator cancels the k^2 term of the transfer function: T = 1 + O(k^4)."""
    order, coef = S.coefficients(2, (S.COMPENSATED, S.COMPENSATED), (n, 1), (1.0, 1.0))
    assert order[:2] == [n, 2]  # order < 2 is raised to 2
    s = _taps(coef, 0, n)
    k = np.arange(-n, n + 1)
    assert abs(s.sum() - 1.0) < 1e-6 and np.array_equal(s, s[::-1])
    assert abs((k ** 2 * s).sum()) < 1e-6 * (k ** 2 * np.abs(s)).sum()
    # and the uncompensated binomial of the same passes does have a k^2 term
    sb = _taps(S.coefficients(2, (S.BINOMIAL, 0), (n - 1, 0), (1.0, 1.0))[1], 0, n - 1)
    kb = np.arange(-(n - 1), n)
    assert (kb ** 2 * sb).sum() > 0.1


def test_5pass_tri6_tri8_tables():
    _, c5 = S.coefficients(3, (S.PASS5, S.TRI6, S.TRI8), (0, 0, 0), (1.0,) * 3)
    s5 = _taps(c5, 0, 5)
    k = np.arange(-5, 6)
    assert abs(s5.sum() - 1) < 1e-6 and abs((k ** 2 * s5).sum()) < 1e-6
    assert np.array_equal(_taps(c5, 1, 3), np.array([1, -6, 15, 44, 15, -6, 1]) / 64.0)
    assert np.array_equal(_taps(c5, 2, 4), np.array([-1, 8, -28, 56, 186, 56, -28, 8, -1]) / 256.0)


@pytest.mark.parametrize("n", [1, 2, 4])
def test_custom_is_kernel_convolution(n):
    """combine_kernel3 multiplies the generating polynomials of the 3-point kernels; the
    first entry of kern3 weighs the upper neighbour (offset +1)."""
    rng = np.random.default_rng(n)
    swfj = np.ones((2, 15, 3))
    swfj[0, :n] = rng.uniform(0.2, 2.0, (n, 3))
    _, coef = S.coefficients(2, (S.CUSTOM, S.NONE), (n, 0), (1.0, 1.0), swfj=swfj)
    want = np.array([1.0])
    for k in range(n):
        a, b, c = swfj[0, k]
        want = np.convolve(want, [c, b, a])  # offsets -1, 0, +1
    want = want / want.sum()
    assert np.allclose(_taps(coef, 0, n), want, rtol=3e-7, atol=0)


def test_besseli0_polynomial_fit():
    x = np.linspace(0.0, 25.0, 501)
    got = np.array([S.besseli0(v) for v in x])
    assert np.max(np.abs(got / np.i0(x) - 1.0)) < 5e-7


@pytest.mark.parametrize("A,fc,n", [(40.0, 0.3, 5), (60.0, 0.5, 8), (10.0, 0.2, 3)])
def test_digital_is_kaiser_windowed_sinc(A, fc, n):
    """digital_filter = fc sinc(fc i) times a Kaiser window of shape alpha(A)."""
    _, coef = S.coefficients(2, (S.DIGITAL, S.NONE), (n, 0), (1.0, 1.0),
                             digital_A=(A, 0.0), digital_fc=(fc, 0.0))
    alpha = 0.0 if A <= 21 else (0.1102 * (A - 8.7) if A >= 50 else
                                 0.5842 * (A - 21) ** 0.4 + 0.07886 * (A - 21))
    i = np.arange(-n, n + 1)
    want = fc * np.sinc(fc * i) * np.i0(alpha * np.sqrt(1 - (i / n) ** 2)) / np.i0(alpha)
    assert np.allclose(_taps(coef, 0, n), want / want.sum(), rtol=2e-6, atol=1e-8)


@pytest.mark.parametrize("gamma", [1.0, 3.0])
def test_laserkdx_passes_the_laser_mode(gamma):
    """The laserkdx filter has unit response at k dx = omega0 dx / (gamma (1 + beta))."""
    dx = (0.2, 0.5)
    omega0 = 2.5
    order, coef = S.coefficients(2, (S.LASERKDX, S.NONE), (0, 0), dx, gamma=gamma,
                                 laserkdx_omega0=omega0, laserkdx_dir=1)
    assert order == [2, 0, 0]
    kap = omega0 * dx[0] / (gamma + math.sqrt(gamma ** 2 - 1))
    s = _taps(coef, 0, 2)
    t = sum(s[k + 2] * math.cos(kap * k) for k in range(-2, 3))
    assert abs(t - 1.0) < 1e-6
    assert abs(s.sum() - 1.0) < 1e-6


def test_filter_acts_as_convolution_on_fourier_modes():
    """orc_smooth on a periodic field: each Fourier mode is multiplied by
    H(kappa) = sum_k s(k) exp(i kappa k); an asymmetric kernel pins the orientation."""
    nx, ny = 24, 16
    n = 2
    swfj = np.ones((2, 15, 3))
    swfj[0, :2] = [[0.3, 1.0, 0.7], [1.1, 0.9, 0.2]]
    order, coef = S.coefficients(2, (S.CUSTOM, S.BINOMIAL), (n, 1), (1.0, 1.0), swfj=swfj,
                                 kfld=np.float64)
    g = O.Grid(2, (nx, ny), (n, n), (n + 1, n + 1))
    mx, my = 3, 2
    kx, ky = 2 * np.pi * mx / nx, 2 * np.pi * my / ny
    xi = np.arange(1 - g.gcl[0], nx + g.gcu[0] + 1)
    yi = np.arange(1 - g.gcl[1], ny + g.gcu[1] + 1)
    ph = kx * xi[None, :] + ky * yi[:, None]
    f = np.zeros(tuple(reversed(g.ld)) + (2,))
    f[..., 0], f[..., 1] = np.cos(ph), np.sin(ph)
    O.smooth(g, f, order, coef)
    hx = sum(coef[0, 15 + k] * np.exp(1j * kx * k) for k in range(-n, n + 1))
    hy = sum(coef[1, 15 + k] * np.exp(1j * ky * k) for k in range(-1, 2))
    z = hx * hy * np.exp(1j * ph)
    inner = g.interior(f)
    zi = g.interior(z[..., None])[..., 0]
    assert np.allclose(inner[..., 0], zi.real, atol=1e-13)
    assert np.allclose(inner[..., 1], zi.imag, atol=1e-13)


def test_filter_keeps_line_ends():
    """Cells closer than n to either end of a line keep their values (int_lb..int_ub)."""
    g = O.Grid(2, (10, 6), (1, 1), (2, 2))
    rng = np.random.default_rng(3)
    f = rng.standard_normal(tuple(reversed(g.ld)) + (3,))
    h = f.copy()
    order, coef = S.coefficients(2, (S.BINOMIAL, S.NONE), (2, 0), (1.0, 1.0), kfld=np.float64)
    O.smooth(g, h, order, coef)
    assert np.array_equal(h[:, :2], f[:, :2]) and np.array_equal(h[:, -2:], f[:, -2:])
    assert not np.array_equal(h[:, 2:-2], f[:, 2:-2])


CASES = [
    (2, ("binomial", "binomial"), (1, 3), {}),
    (2, ("compensated", "none"), (4, 0), {}),
    (3, ("compensated", "binomial", "tri6"), (2, 2, 0), {}),
    (2, ("5pass", "tri8"), (0, 0), {}),
    (3, ("digital", "digital", "none"), (6, 3, 0),
     {"digital_A": (45.0, 12.0, 0.0), "digital_fc": (0.4, 0.25, 0.0)}),
    (2, ("custom", "custom"), (3, 1), {"swfj": None}),
    (2, ("laserkdx", "none"), (0, 0), {"laserkdx_omega0": 3.0, "laserkdx_dir": 1}),
    (3, ("binomial", "none", "binomial"), (15, 0, 7), {}),
]


@pytest.mark.parametrize("dims,types,order,kw", CASES)
def test_library_setup_equals_oracle(dims, types, order, kw):
    """osr_smooth_setup (host code of the library) reproduces the oracle's FP32 table."""
    kw = dict(kw)
    if "swfj" in kw:
        kw["swfj"] = np.random.default_rng(5).uniform(0.1, 2.0, (3, 15, 3))
    dx = (0.2, 0.3, 0.4)[:dims]
    n, c = oa.smooth_setup(dims, types, order, dx=dx, gamma=1.0, **kw)
    codes = [oa._lib.OSR_SMOOTH[t] for t in types]
    kw2 = {k: (v[:dims] if isinstance(v, (tuple, np.ndarray)) else v) for k, v in kw.items()}
    n2, c2 = S.coefficients(dims, codes, order, dx, **kw2)
    assert n == n2
    assert np.array_equal(c.astype(np.float64), c2)


def test_setup_rejects_what_the_reference_aborts_on():
    with pytest.raises(OsrError):
        oa.smooth_setup(2, ("binomial", "none"), (0, 0))
    with pytest.raises(OsrError):
        oa.smooth_setup(2, ("binomial", "none"), (16, 0))
    with pytest.raises(OsrError):
        oa.smooth_setup(2, ("laserkdx", "none"), laserkdx_omega0=-1.0)
    with pytest.raises(OsrError):
        oa.smooth_setup(2, (7, "none"), (1, 0))
    with pytest.raises(ValueError):
        S.coefficients(2, (S.BINOMIAL, S.NONE), (0, 0), (1.0, 1.0))


def test_fortran_binding_setup():
    """The bind(c) osr_smooth_params of m_osiris_amd matches the C layout: the Fortran
    example's coefficients equal the closed forms (host code only, no GPU)."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "osiris_amd", "lib", "smooth_setup_example")
    if not os.path.exists(exe):
        pytest.skip("Fortran example not built (no Fortran compiler)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.split()
    vals = [float(v) for v in lines if "E" in v]
    assert lines[:4] == ["dir", "1", "order", "2"]
    assert np.allclose(vals[:5], np.array([-1, 4, 10, 4, -1]) / 16.0, rtol=0, atol=0)
    assert np.allclose(vals[5:], np.array([1, 6, 15, 20, 15, 6, 1]) / 64.0, rtol=0, atol=0)
