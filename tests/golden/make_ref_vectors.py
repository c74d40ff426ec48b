"""Generate tests/golden/ref_simd_vectors.npz: input/output vectors of the reference's
own fused AVX2 push+deposit kernels (FNAME(vadvance_deposit_{2d,3d}_s{1,2,3}) of
simd/os-spec-push-avx2{,d}.c), compiled from the reference sources into oracle/_ref by
oracle/Makefile.

For every (dims, order) a small seeded set of particles and fields is pushed once by the
single (AVX2) and double (AVX2d) precision kernels; the file stores the inputs and the
outputs (ix, x, u, j, energy).  Run from the repository root:

    make -C oracle ref && python tests/golden/make_ref_vectors.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_simd_vectors.npz")


def main():
    data = {}
    for dims in (2, 3):
        for order in (1, 2, 3):
            rng = np.random.default_rng(1000 + 10 * dims + order)
            nx = (6,) * dims if dims == 2 else (4,) * dims
            g = O.make_grid(dims, nx, order)
            lo = g.gcl[0]
            n = 96
            e = g.zeros(3, np.float64)
            b = g.zeros(3, np.float64)
            e[...] = rng.normal(0, 0.1, e.shape)
            b[...] = rng.normal(0, 0.5, b.shape)
            ix = np.stack([rng.integers(1, nx[d] + 1, n) for d in range(dims)], 1).astype(np.int32)
            x = rng.uniform(-0.5, 0.5, (n, dims))
            u = rng.normal(0, 0.3, (n, 3))
            q = rng.uniform(-0.02, -0.005, n)
            dx = (0.1,) * dims
            dt = 0.05 if dims == 2 else 0.04
            key = f"{dims}d_s{order}"
            data[key + "_in_ix"] = ix
            data[key + "_in_x"] = x
            data[key + "_in_u"] = u
            data[key + "_in_q"] = q
            data[key + "_in_e"] = e
            data[key + "_in_b"] = b
            data[key + "_par"] = np.array([dt, dx[0], lo])
            for prec, dtp in (("s", np.float32), ("d", np.float64)):
                rix, rx, ru, rj, rene = O.ref_vadvance_deposit(
                    dims, order, ix, x.astype(dtp), u.astype(dtp), q.astype(dtp), -1.0,
                    e.astype(dtp), b.astype(dtp), lo, g.zeros(3, dtp), lo, dx, dt, prec=prec)
                data[f"{key}_{prec}_ix"] = rix
                data[f"{key}_{prec}_x"] = rx
                data[f"{key}_{prec}_u"] = ru
                data[f"{key}_{prec}_j"] = rj
                data[f"{key}_{prec}_ene"] = np.array(rene)
    np.savez_compressed(OUT, **data)
    print(OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
