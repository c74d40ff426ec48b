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
