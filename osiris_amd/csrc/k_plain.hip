// k_plain.hip -- the plain one-kernel-per-phase GPU path (comparison baseline).
//
// What a direct GPU port of the reference's per-phase loop looks like: one kernel per
// phase of advance_deposit (spec/os-spec-push.f03:1438-1518), with intermediate arrays
// in HBM between them, fields read straight from global memory, no tiles, no LDS, and
// the current accumulated with global FP32 atomics (not reproducible run to run):
//   plain_gather : E, B at every particle          -> ep, bp      (GET_EMF_2D)
//   plain_push   : u += Boris/Vay(ep, bp)          -> u, energy   (dudt_boris/dudt_vay)
//   plain_move   : x += v dt, trim, split          -> x, cell, VP buffer (split_2d)
//   plain_deposit: every VP                         -> J (global atomics) (DEP_CURRENT_2D)
// followed by the same Yee kernels.  2D only.
#include "osr_device.hpp"
#include "osr_types.hpp"

namespace osr {

struct PlainBufs {
  float* ep;   // [3][cap]
  float* bp;   // [3][cap]
  float* vp;   // [7][3*cap]: x0, y0, x1, y1, q, vz, node (bit cast int)
  int* nvp;    // [cap]
  float* Jf;   // [3][ncell]
  long long cap;
};

template <int ORDER>
__global__ __launch_bounds__(256) void plain_gather(const SpecDesc* __restrict__ descs, int sid,
                                                    PlainBufs b) {
  using O = Ord<ORDER>;
  constexpr int NP = O::NP;
  const SpecDesc& d = descs[sid];
  const GridDesc& g = d.g;
  if (g.ctrl->halt != 0) return;
  const int n = d.cnt[0];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int ix, iy;
  unpack2(d.cell[i], ix, iy);
  const float x = d.x[i], y = d.y[i];
  const int nx = g.nx, ny = g.ny;
  const long long nc = g.ncell;
  float wx[NP], wy[NP], wxh[NP], wyh[NP];
  spline<ORDER>(x, wx);
  spline<ORDER>(y, wy);
  splineh<ORDER>(x, x < 0.f ? 1.f : 0.f, wxh);
  splineh<ORDER>(y, y < 0.f ? 1.f : 0.f, wyh);
  const int a1 = ix + O::LP, a2 = iy + O::LP;
  const int h1 = a1 - (x < 0.f), h2 = a2 - (y < 0.f);
  auto at = [&](const float* f, int c, int i1, int i2) {
    return f[c * nc + (long long)wrapi(i2, ny) * nx + wrapi(i1, nx)];
  };
  float e1 = 0.f, e2 = 0.f, e3 = 0.f, b1 = 0.f, b2 = 0.f, b3 = 0.f;
#pragma unroll
  for (int k2 = 0; k2 < NP; k2++) {
    float l1 = 0.f, l2 = 0.f, l3 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f;
#pragma unroll
    for (int k1 = 0; k1 < NP; k1++) {
      l1 = fma_(at(g.E, 0, h1 + k1, a2 + k2), wxh[k1], l1);
      l2 = fma_(at(g.E, 1, a1 + k1, h2 + k2), wx[k1], l2);
      l3 = fma_(at(g.E, 2, a1 + k1, a2 + k2), wx[k1], l3);
      m1 = fma_(at(g.B, 0, a1 + k1, h2 + k2), wx[k1], m1);
      m2 = fma_(at(g.B, 1, h1 + k1, a2 + k2), wxh[k1], m2);
      m3 = fma_(at(g.B, 2, h1 + k1, h2 + k2), wxh[k1], m3);
    }
    e1 = fma_(l1, wy[k2], e1);
    e2 = fma_(l2, wyh[k2], e2);
    e3 = fma_(l3, wy[k2], e3);
    b1 = fma_(m1, wyh[k2], b1);
    b2 = fma_(m2, wy[k2], b2);
    b3 = fma_(m3, wyh[k2], b3);
  }
  const long long cap = b.cap;
  b.ep[i] = e1;
  b.ep[cap + i] = e2;
  b.ep[2 * cap + i] = e3;
  b.bp[i] = b1;
  b.bp[cap + i] = b2;
  b.bp[2 * cap + i] = b3;
}

template <int PUSHER>
__global__ __launch_bounds__(256) void plain_push(const SpecDesc* __restrict__ descs, int sid,
                                                  PlainBufs b) {
  __shared__ double s_e[4];
  const SpecDesc& d = descs[sid];
  if (d.g.ctrl->halt != 0) return;
  const int n = d.cnt[0];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double ene = 0.0;
  if (i < n) {
    const long long cap = b.cap;
    float u1 = d.ux[i], u2 = d.uy[i], u3 = d.uz[i];
    const float ke = push<PUSHER>(d.tem, b.ep[i], b.ep[cap + i], b.ep[2 * cap + i], b.bp[i],
                                  b.bp[cap + i], b.bp[2 * cap + i], u1, u2, u3);
    d.ux[i] = u1;
    d.uy[i] = u2;
    d.uz[i] = u3;
    ene = (double)d.q[i] * (double)ke;
  }
  ene = wave_sum(ene);
  if ((threadIdx.x & 63) == 0) s_e[threadIdx.x >> 6] = ene;
  __syncthreads();
  if (threadIdx.x == 0) d.epart[blockIdx.x] = ((s_e[0] + s_e[1]) + s_e[2]) + s_e[3];
}

__global__ __launch_bounds__(256) void plain_move(const SpecDesc* __restrict__ descs, int sid,
                                                  PlainBufs b) {
  const SpecDesc& d = descs[sid];
  const GridDesc& g = d.g;
  if (g.ctrl->halt != 0) return;
  const int n = d.cnt[0];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int ix, iy;
  unpack2(d.cell[i], ix, iy);
  const float x = d.x[i], y = d.y[i];
  const float u1 = d.ux[i], u2 = d.uy[i], u3 = d.uz[i];
  const float rg = rgamma(u1, u2, u3);
  const float x1 = fma_(u1 * rg, g.dt_dx[0], x), y1 = fma_(u2 * rg, g.dt_dx[1], y);
  const float tx = trim(x1), ty = trim(y1);
  VP2 vp[3];
  const int nv = split2d(x, y, x1, y1, (int)tx, (int)ty, d.q[i], u3 * rg, vp);
  const long long c3 = 3 * b.cap;
#pragma unroll
  for (int v = 0; v < 3; v++) {
    if (v < nv) {
      const long long s = 3LL * i + v;
      b.vp[s] = vp[v].x0;
      b.vp[c3 + s] = vp[v].y0;
      b.vp[2 * c3 + s] = vp[v].x1;
      b.vp[3 * c3 + s] = vp[v].y1;
      b.vp[4 * c3 + s] = vp[v].q;
      b.vp[5 * c3 + s] = vp[v].vz;
      // global cell of the VP (may lie one cell outside the box; wrapped at deposit)
      b.vp[6 * c3 + s] = __int_as_float(pack2((ix + vp[v].di + 1) & 0xFFFF, (iy + vp[v].dj + 1) & 0xFFFF));
    }
  }
  b.nvp[i] = nv;
  d.x[i] = x1 - tx;
  d.y[i] = y1 - ty;
  d.cell[i] = pack2(wrapi(ix + (int)tx, g.nx), wrapi(iy + (int)ty, g.ny));
}

template <int ORDER>
__global__ __launch_bounds__(256) void plain_deposit(const SpecDesc* __restrict__ descs, int sid,
                                                     PlainBufs b) {
  using O = Ord<ORDER>;
  constexpr int NP = O::NP;
  const SpecDesc& d = descs[sid];
  const GridDesc& g = d.g;
  if (g.ctrl->halt != 0) return;
  const int n = d.cnt[0];
  const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
  const int i = (int)(s / 3);
  if (i >= n || (int)(s % 3) >= b.nvp[i]) return;
  const long long c3 = 3 * b.cap;
  const float x0 = b.vp[s], y0 = b.vp[c3 + s], x1 = b.vp[2 * c3 + s], y1 = b.vp[3 * c3 + s];
  const float q = b.vp[4 * c3 + s], vz = b.vp[5 * c3 + s];
  int cx, cy;
  unpack2((uint32_t)__float_as_int(b.vp[6 * c3 + s]), cx, cy);
  cx -= 1;
  cy -= 1;
  float S0x[NP], S1x[NP], S0y[NP], S1y[NP], wl1[NP], wl2[NP];
  spline<ORDER>(x0, S0x);
  spline<ORDER>(x1, S1x);
  spline<ORDER>(y0, S0y);
  spline<ORDER>(y1, S1y);
  wl<ORDER>(q * g.jnorm[0], x0, x1, wl1);
  wl<ORDER>(q * g.jnorm[1], y0, y1, wl2);
  const float qvz = (q * vz) * (1.0f / 3.0f);
  const long long nc = g.ncell;
#pragma unroll
  for (int k2 = 0; k2 < NP; k2++) {
    const float wp1 = S0y[k2] + S1y[k2];
    const long long row = (long long)wrapi(cy + O::LP + k2, g.ny) * g.nx;
#pragma unroll
    for (int k1 = 0; k1 < NP; k1++) {
      const long long gi = row + wrapi(cx + O::LP + k1, g.nx);
      const float wp2 = S0x[k1] + S1x[k1];
      float t1 = fma_(S1y[k2], S1x[k1], S0x[k1] * S0y[k2]);
      const float t2 = fma_(S0y[k2], S1x[k1], S0x[k1] * S1y[k2]);
      t1 = fma_(t2, 0.5f, t1);
      if (k1 < NP - 1) atomicAdd(&b.Jf[gi], wl1[k1] * wp1);
      if (k2 < NP - 1) atomicAdd(&b.Jf[nc + gi], wl2[k2] * wp2);
      atomicAdd(&b.Jf[2 * nc + gi], qvz * t1);
    }
  }
}

template __global__ void plain_gather<1>(const SpecDesc*, int, PlainBufs);
template __global__ void plain_gather<2>(const SpecDesc*, int, PlainBufs);
template __global__ void plain_gather<3>(const SpecDesc*, int, PlainBufs);
template __global__ void plain_push<0>(const SpecDesc*, int, PlainBufs);
template __global__ void plain_push<1>(const SpecDesc*, int, PlainBufs);
template __global__ void plain_deposit<1>(const SpecDesc*, int, PlainBufs);
template __global__ void plain_deposit<2>(const SpecDesc*, int, PlainBufs);
template __global__ void plain_deposit<3>(const SpecDesc*, int, PlainBufs);

}  // namespace osr
