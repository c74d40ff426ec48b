// k_smooth.hip -- current smoothing (digital filtering), vdf/os-vdf-smooth.f90.
//
// The filter is separable: one pass per direction with order n > 0, x before y before z
// (filter_2d :635-800, filter_3d :806-1100).  A pass replaces each value by
//   s(-n) o(i-n) + s(-n+1) o(i-n+1) + ... + s(n) o(i+n)
// summed left to right in FP32 with every product rounded (no contraction), o being the
// values before the pass -- the reference's single-precision arithmetic and order.
//
// Two forms:
//  * smooth_pass: the resident simulation, planar interior arrays with periodic index
//    wraps (equal to the reference on the interior, whose guard cells hold the periodic
//    images after the current's boundary update, os-simulation.f03:650-657); the first
//    pass reads the folded current (fixed-point or the plain path's float) and converts it
//    as the E step would, later passes read the previous output;
//  * smooth_vdf_pass: the stateless drop-in, a component-interleaved array with guard
//    cells, filtered exactly like smooth_vdf: cells closer than n to either end of a line
//    keep their values.
// Both are memory-bound gathers along one direction: one thread per cell, consecutive
// threads on consecutive x, so every tap of a wave is a coalesced row segment.
#include "osr_device.hpp"
#include "osr_types.hpp"

namespace osr {

OSR_DEV int wrapn(int i, int n) {
  int w = i % n;
  return w < 0 ? w + n : w;
}

__global__ __launch_bounds__(256) void smooth_pass(const FieldDesc* __restrict__ fds,
                                                   const int2* __restrict__ work, int pass) {
  const int2 wk = work[blockIdx.x];
  const FieldDesc& fd = fds[wk.x];
  const GridDesc& g = fd.g;
  if (g.ctrl->halt >= 2 || pass >= fd.sm_np) return;
  const long long nc = g.ncell;
  const long long idx = (long long)wk.y * 256 + threadIdx.x;
  if (idx >= nc) return;
  const int nx = g.nx, ny = g.ny;
  const int dir = fd.sm_dir[pass], n = fd.sm_n[pass];
  const float* s = fd.sm_coef[pass] + 15;
  int ci, cj, ck;
  cell_xyz(g, idx, ci, cj, ck);
  int c, len;
  long long st;
  if (dir == 0) {
    c = ci;
    len = nx;
    st = 1;
  } else if (dir == 1) {
    c = cj;
    len = ny;
    st = nx;
  } else {
    c = ck;
    len = g.nz;
    st = (long long)nx * ny;
  }
  const long long row = idx - (long long)c * st;  // the line's first cell
  const int src = pass == 0 ? fd.sm_src : 2;
  const unsigned long long* Jq = g.J;
  const float* Jp = (const float*)g.J;
  const float* P = fd.sm_buf[(pass + 1) & 1];  // previous pass's output
  float* dst = fd.sm_buf[pass & 1];
  const double jinv = g.jinv;
#pragma unroll
  for (int comp = 0; comp < 3; comp++) {
    const long long off = comp * nc + row;
    auto val = [&](int k) -> float {
      const long long id = off + (long long)wrapn(c + k, len) * st;
      if (src == 0) return (float)((double)(long long)Jq[id] * jinv);
      if (src == 1) return Jp[id];
      return P[id];
    };
    float acc = val(-n) * s[-n];
    for (int k = -n + 1; k <= n; k++) acc = acc + val(k) * s[k];
    dst[comp * nc + idx] = acc;
  }
}

struct SmoothCoef {
  float s[31];  // s(k) at [15 + k]
};

__global__ __launch_bounds__(256) void smooth_vdf_pass(const float* __restrict__ src,
                                                       float* __restrict__ dst, int fdim,
                                                       int ld0, int ld1, int ld2, int dir,
                                                       int n, SmoothCoef coef) {
  const long long total = (long long)fdim * ld0 * ld1 * ld2;
  const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
  if (id >= total) return;
  const long long cell = id / fdim;
  int c, len;
  long long st;
  if (dir == 0) {
    c = (int)(cell % ld0);
    len = ld0;
    st = fdim;
  } else if (dir == 1) {
    c = (int)((cell / ld0) % ld1);
    len = ld1;
    st = (long long)fdim * ld0;
  } else {
    c = (int)(cell / ((long long)ld0 * ld1));
    len = ld2;
    st = (long long)fdim * ld0 * ld1;
  }
  if (c < n || c > len - 1 - n) {
    dst[id] = src[id];
    return;
  }
  const float* s = coef.s + 15;
  float acc = src[id - n * st] * s[-n];
  for (int k = -n + 1; k <= n; k++) acc = acc + src[id + k * st] * s[k];
  dst[id] = acc;
}

}  // namespace osr
