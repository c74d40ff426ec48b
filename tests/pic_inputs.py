"""Seeded synthetic inputs shared by the tests (reference array layout)."""
import numpy as np

from oracle import oracle as O


def random_particles(rng, dims, nx, n, uth=0.3, ufl=(0.0, 0.0, 0.0), q=-0.01, lo=None, hi=None):
    """Uniform random cells/positions (cells lo..hi, 1-based) and Maxwellian momenta."""
    lo = lo or [1] * dims
    hi = hi or list(nx)
    ix = np.stack([rng.integers(lo[d], hi[d] + 1, n) for d in range(dims)], 1).astype(np.int32)
    x = rng.uniform(-0.5, 0.5, (n, dims)).astype(np.float32)
    u = (np.array(ufl) + rng.normal(0, uth, (n, 3))).astype(np.float32)
    qa = np.full(n, q, np.float32)
    return ix, x, u, qa


def random_fields(rng, grid: "O.Grid", e_amp=0.1, b_amp=0.5, dtype=np.float32):
    e = grid.zeros(3, dtype)
    b = grid.zeros(3, dtype)
    e[...] = rng.normal(0, e_amp, e.shape)
    b[...] = rng.normal(0, b_amp, b.shape)
    return e, b


def periodic_fields(rng, grid: "O.Grid", e_amp=0.1, b_amp=0.5, dtype=np.float64):
    """Random interior fields with guard cells holding the periodic images."""
    e, b = random_fields(rng, grid, e_amp, b_amp, dtype)
    O.update_periodic(grid, e, 0, "f64" if dtype == np.float64 else "f32")
    O.update_periodic(grid, b, 0, "f64" if dtype == np.float64 else "f32")
    return e, b


def thermal_plasma(rng, dims, nx, ppc, uth, ufl=(0.0, 0.0, 0.0), density=1.0, rqm=-1.0):
    """ppc particles per cell at uniform random in-cell positions, Maxwellian momenta,
    charge sign(rqm)*density/ppc (spec/psource/os-psource-std.f03:846)."""
    cells = np.stack(np.meshgrid(*[np.arange(1, nx[d] + 1) for d in range(dims)],
                                 indexing="ij"), -1).reshape(-1, dims)
    ix = np.repeat(cells, ppc, axis=0).astype(np.int32)
    n = len(ix)
    x = rng.uniform(-0.5, 0.5, (n, dims)).astype(np.float32)
    u = (np.array(ufl) + uth * rng.normal(0, 1, (n, 3))).astype(np.float32)
    q = np.full(n, np.sign(rqm) * density / ppc, np.float32)
    return ix, x, u, q


def vacuum_pulse(nx, ny, dx, k0=4.0, width=2.0, direction=1, dtype=np.float64):
    """Interior E, B (shape (ny, nx, 3)) of a y-uniform light pulse travelling along
    +x (direction 1) or -x (-1): E2 = f(x) on the Yee E2 points x_i, B3 = +-f on the B3
    points x_{i+1/2}, f = exp(-((x - x0)/width)^2) cos(k0 (x - x0)) centred in the box."""
    xi = np.arange(1, nx + 1) * dx
    x0 = nx * dx / 2

    def f(x):
        return np.exp(-((x - x0) / width) ** 2) * np.cos(k0 * (x - x0))

    e = np.zeros((ny, nx, 3), dtype)
    b = np.zeros((ny, nx, 3), dtype)
    e[..., 1] = f(xi)[None, :]
    b[..., 2] = (1.0 if direction > 0 else -1.0) * f(xi + dx / 2)[None, :]
    return e, b


def oblique_packet(nx, ny, dx, k0=4.0, m=2, direction=1):
    """Interior E, B (shape (ny, nx, 3)) of a light packet along (kx, ky), ky = 2 pi m /
    (ny dx) periodic on the box, |k| = k0, Gaussian envelope along x centred in the box,
    both polarisations: E3 with B = k^ x z^ f, and in-plane E = z^ x k^ f with B3 = f, each
    component sampled at its Yee point."""
    ky = 2 * np.pi * m / (ny * dx)
    kx = np.sqrt(k0 ** 2 - ky ** 2) * direction
    kk = np.hypot(kx, ky)
    x0 = nx * dx / 2
    X, Y = np.meshgrid(np.arange(1, nx + 1) * dx, np.arange(1, ny + 1) * dx)
    h = dx / 2

    def f(x, y):
        return np.exp(-((x - x0) / 2.0) ** 2) * np.cos(kx * (x - x0) + ky * y)

    e = np.zeros((ny, nx, 3))
    b = np.zeros((ny, nx, 3))
    e[..., 2] = f(X, Y)
    b[..., 0] = (ky / kk) * f(X, Y + h)
    b[..., 1] = -(kx / kk) * f(X + h, Y)
    e[..., 0] = -(ky / kk) * f(X + h, Y)
    e[..., 1] = (kx / kk) * f(X, Y + h)
    b[..., 2] = f(X + h, Y + h)
    return e, b


def oblique_packet_3d(nx, ny, nz, dx, k0=4.0, direction=1):
    """Interior E, B (shape (nz, ny, nx, 3)) of a light packet along (kx, ky, kz) with
    ky = 2 pi / (ny dx), kz = 2 pi / (nz dx), |k| = k0, Gaussian envelope along x centred in
    the box, E along a1 + a2 (a1 = z^ x k^ normalised, a2 = k^ x a1) and B = k^ x E, each
    component sampled at its Yee point."""
    ky, kz = 2 * np.pi / (ny * dx), 2 * np.pi / (nz * dx)
    kx = np.sqrt(k0 ** 2 - ky ** 2 - kz ** 2) * direction
    k = np.array([kx, ky, kz]) / k0
    a1 = np.cross([0, 0, 1.0], k)
    a1 /= np.linalg.norm(a1)
    a2 = np.cross(k, a1)
    x0 = nx * dx / 2
    idx = [np.arange(1, m + 1) * dx for m in (nx, ny, nz)]
    h = dx / 2

    def f(sh):
        Z, Y, X = np.meshgrid(idx[2] + sh[2], idx[1] + sh[1], idx[0] + sh[0], indexing="ij")
        return np.exp(-((X - x0) / 2.0) ** 2) * np.cos(kx * (X - x0) + ky * Y + kz * Z)

    e = np.zeros((nz, ny, nx, 3))
    b = np.zeros_like(e)
    bvec = np.cross(k, a1 + a2)
    for c in range(3):
        se = [0.0, 0.0, 0.0]
        se[c] = h
        sb = [h, h, h]
        sb[c] = 0.0
        e[..., c] = (a1[c] + a2[c]) * f(se)
        b[..., c] = bvec[c] * f(sb)
    return e, b


def mur_reflection(kdx, courant):
    """Reflection coefficient of the first-order Mur wall E_0^{n+1} = E_1^n +
    c (E_0^n - E_1^{n+1}), c = (1 - S)/(1 + S), on the 1D Yee lattice (waves along x on
    the 2D lattice with k_y = 0), for a monochromatic wave of wavenumber k dx: with
    E_i^n = z^n (e^{i k i} + R e^{-i k i}), z = e^{i w dt} and the lattice dispersion
    sin(w dt/2) = S sin(k dx/2), the wall condition gives R = (e^{ik}(1 - cz) - (z - c)) /
    ((z - c) - e^{-ik}(1 - cz))."""
    k = np.asarray(kdx, np.float64)
    S = courant
    z = np.exp(2j * np.arcsin(S * np.sin(k / 2)))
    c = (1 - S) / (1 + S)
    with np.errstate(invalid="ignore", divide="ignore"):
        r = (np.exp(1j * k) * (1 - c * z) - (z - c)) / ((z - c) - np.exp(-1j * k) * (1 - c * z))
    return np.nan_to_num(r)
