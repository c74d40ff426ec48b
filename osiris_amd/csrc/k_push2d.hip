// k_push2d.hip -- fused 2D push + deposit with compile-time tile shape.
//
// One workgroup (4 waves) per (species, tile).  Compared with the generic kernel in
// k_tiles.hip:
//  * TX/TY are template parameters, so every LDS stencil offset is an immediate;
//  * each wave appends the virtual particles (VPs) of its particles to a private LDS
//    queue and deposits them 64 at a time, so the deposit runs on full waves instead of
//    looping max(#VP) times over partly idle lanes (a trajectory splits into 1..3 VPs);
//  * the current of all waves is accumulated into one int64 fixed-point LDS tile; integer
//    adds are exact and order independent, so the deposit is bit-reproducible run to run
//    (FP32 LDS atomics were measured to be far slower here: a 40x longer LDS wait).
// The arithmetic of gather, push, position advance and split is the one of the generic
// kernel (the reference's AVX2 kernel, simd/os-spec-push-avx2.c:471-760).
#include "osr_device.hpp"
#include "osr_types.hpp"

namespace osr {

template <int ORDER, int PUSHER, int TX, int TY>
__global__ __launch_bounds__(256) void push_deposit_2d_t(const SpecDesc* __restrict__ descs,
                                                         const int2* __restrict__ work) {
  using O = Ord<ORDER>;
  constexpr int NP = O::NP;
  constexpr int NT = 256, NW = 4;
  constexpr int GX = TX + O::GLO + O::GHI, GY = TY + O::GLO + O::GHI, GN = GX * GY;
  constexpr int JX = TX + O::JLO + O::JHI, JY = TY + O::JLO + O::JHI, JN = JX * JY;
  constexpr int QC = 128;  // VP queue capacity per wave (< 64 pending + 64 appended)
  __shared__ float sF[6 * GN];
  __shared__ unsigned long long sJ[3 * JN];
  __shared__ float sQ[NW][6][QC];  // x0, y0, x1, y1, q, vz
  __shared__ int sQn[NW][QC];      // LDS node index of the VP's first stencil point
  __shared__ double s_ene[NW];

  const int2 wk = work[blockIdx.x];
  const SpecDesc& d = descs[wk.x];
  const GridDesc& g = d.g;
  if (g.ctrl->halt != 0) return;
  const int tile = wk.y;
  const int cx0 = (tile % g.ntx) * TX, cy0 = (tile / g.ntx) * TY;
  const int nx = g.nx, ny = g.ny;
  const long long nc = g.ncell;

  for (int i = threadIdx.x; i < GN; i += NT) {
    const int gx = wrapi(cx0 + (i % GX) - O::GLO, nx), gy = wrapi(cy0 + (i / GX) - O::GLO, ny);
    const long long gi = gx + (long long)nx * gy;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      sF[c * GN + i] = g.E[c * nc + gi];
      sF[(3 + c) * GN + i] = g.B[c * nc + gi];
    }
  }
  for (int i = threadIdx.x; i < 3 * JN; i += NT) sJ[i] = 0ull;
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int base = d.off[tile];
  const int n = d.cnt[tile];
  const int per_wave = ((n + NT - 1) / NT) * 64;
  const int w0 = wave * per_wave;
  const int w1 = min(w0 + per_wave, n);
  const float tem = d.tem;
  const float dtdx1 = g.dt_dx[0], dtdx2 = g.dt_dx[1];
  const float jn1 = g.jnorm[0], jn2 = g.jnorm[1];
  const double jscale = g.jscale;
  unsigned long long* const J1 = sJ;
  unsigned long long* const J2 = sJ + JN;
  unsigned long long* const J3 = sJ + 2 * JN;
  float(*const Q)[QC] = sQ[wave];
  int* const QN = sQn[wave];

  // deposit of the VPs in queue slots [head, head + cnt) (cnt <= 64, one per lane)
  auto deposit = [&](int head, int cnt) {
    if (lane < cnt) {
      const int s = (head + lane) & (QC - 1);
      const float x0 = Q[0][s], y0 = Q[1][s], x1 = Q[2][s], y1 = Q[3][s];
      const float q = Q[4][s], vz = Q[5][s];
      const int nb = QN[s];
      float S0x[NP], S1x[NP], S0y[NP], S1y[NP], wl1[NP], wl2[NP];
      spline<ORDER>(x0, S0x);
      spline<ORDER>(x1, S1x);
      spline<ORDER>(y0, S0y);
      spline<ORDER>(y1, S1y);
      wl<ORDER>(q * jn1, x0, x1, wl1);
      wl<ORDER>(q * jn2, y0, y1, wl2);
      const float qvz = (q * vz) * (1.0f / 3.0f);
#pragma unroll
      for (int k2 = 0; k2 < NP; k2++) {
        const float wp1 = S0y[k2] + S1y[k2];
#pragma unroll
        for (int k1 = 0; k1 < NP; k1++) {
          const int idx = nb + k2 * JX + k1;
          const float wp2 = S0x[k1] + S1x[k1];
          float t1 = fma_(S1y[k2], S1x[k1], S0x[k1] * S0y[k2]);
          const float t2 = fma_(S0y[k2], S1x[k1], S0x[k1] * S1y[k2]);
          t1 = fma_(t2, 0.5f, t1);
          if (k1 < NP - 1) atomicAdd(&J1[idx], (unsigned long long)to_fix(wl1[k1] * wp1, jscale));
          if (k2 < NP - 1) atomicAdd(&J2[idx], (unsigned long long)to_fix(wl2[k2] * wp2, jscale));
          atomicAdd(&J3[idx], (unsigned long long)to_fix(qvz * t1, jscale));
        }
      }
    }
  };

  double ene = 0.0;
  int mcount = 0;
  int qhead = 0, qlen = 0;
  bool bad = false;
  for (int p0 = w0; p0 < w1; p0 += 64) {
    const int p = p0 + lane;
    bool act = p < w1;
    float x = 0.f, y = 0.f, u1 = 0.f, u2 = 0.f, u3 = 0.f, q = 0.f;
    int lx = 0, ly = 0;
    if (act) {
      const int s = base + p;
      const uint32_t cell = d.cell[s];
      x = d.x[s];
      y = d.y[s];
      u1 = d.ux[s];
      u2 = d.uy[s];
      u3 = d.uz[s];
      q = d.q[s];
      int ix, iy;
      unpack2(cell, ix, iy);
      lx = ix - cx0;
      ly = iy - cy0;
      if (lx < 0 || lx >= TX || ly < 0 || ly >= TY) {
        bad = true;
        act = false;
        lx = ly = 0;
      }
    }
    // ---- gather
    float wx[NP], wy[NP], wxh[NP], wyh[NP];
    spline<ORDER>(x, wx);
    spline<ORDER>(y, wy);
    splineh<ORDER>(x, x < 0.f ? 1.f : 0.f, wxh);
    splineh<ORDER>(y, y < 0.f ? 1.f : 0.f, wyh);
    const int gb = (ly + O::GLO + O::LP) * GX + lx + O::GLO + O::LP;
    const int gbx = gb - (x < 0.f ? 1 : 0);
    const int gby = gb - (y < 0.f ? GX : 0);
    const int gbxy = gbx - (y < 0.f ? GX : 0);
    float e1 = 0.f, e2 = 0.f, e3 = 0.f, b1 = 0.f, b2 = 0.f, b3 = 0.f;
#pragma unroll
    for (int k2 = 0; k2 < NP; k2++) {
      float l1 = 0.f, l2 = 0.f, l3 = 0.f;
#pragma unroll
      for (int k1 = 0; k1 < NP; k1++) {
        l1 = fma_(sF[0 * GN + gbx + k2 * GX + k1], wxh[k1], l1);
        l2 = fma_(sF[1 * GN + gby + k2 * GX + k1], wx[k1], l2);
        l3 = fma_(sF[2 * GN + gb + k2 * GX + k1], wx[k1], l3);
      }
      e1 = fma_(l1, wy[k2], e1);
      e2 = fma_(l2, wyh[k2], e2);
      e3 = fma_(l3, wy[k2], e3);
    }
#pragma unroll
    for (int k2 = 0; k2 < NP; k2++) {
      float l1 = 0.f, l2 = 0.f, l3 = 0.f;
#pragma unroll
      for (int k1 = 0; k1 < NP; k1++) {
        l1 = fma_(sF[3 * GN + gby + k2 * GX + k1], wx[k1], l1);
        l2 = fma_(sF[4 * GN + gbx + k2 * GX + k1], wxh[k1], l2);
        l3 = fma_(sF[5 * GN + gbxy + k2 * GX + k1], wxh[k1], l3);
      }
      b1 = fma_(l1, wyh[k2], b1);
      b2 = fma_(l2, wy[k2], b2);
      b3 = fma_(l3, wyh[k2], b3);
    }
    // ---- push
    const float ke = push<PUSHER>(tem, e1, e2, e3, b1, b2, b3, u1, u2, u3);
    if (act) ene += (double)q * (double)ke;
    // ---- position
    const float rg = rgamma(u1, u2, u3);
    const float v1 = u1 * rg, v2 = u2 * rg, v3 = u3 * rg;
    const float x1 = fma_(v1, dtdx1, x), y1 = fma_(v2, dtdx2, y);
    const float tx = trim(x1), ty = trim(y1);
    const int dix = (int)tx, diy = (int)ty;
    // ---- split; the VPs are queued one "rank" at a time (first VPs of all lanes, then
    // the second ones, ...) and full waves of 64 are deposited as soon as they are there
    VP2 vp[3];
    const int nvp = act ? split2d(x, y, x1, y1, dix, diy, q, v3, vp) : 0;
    const int nb0 = (ly + O::LP + O::JLO) * JX + lx + O::LP + O::JLO;
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int v = 0; v < 3; v++) {
      const unsigned long long m = __ballot(v < nvp);
      if (m == 0ull) break;
      if (v < nvp) {
        const int s = (qhead + qlen + __popcll(m & below)) & (QC - 1);
        Q[0][s] = vp[v].x0;
        Q[1][s] = vp[v].y0;
        Q[2][s] = vp[v].x1;
        Q[3][s] = vp[v].y1;
        Q[4][s] = vp[v].q;
        Q[5][s] = vp[v].vz;
        QN[s] = nb0 + vp[v].dj * JX + vp[v].di;
      }
      qlen += __popcll(m);
      if (qlen >= 64) {
        deposit(qhead, 64);
        qhead = (qhead + 64) & (QC - 1);
        qlen -= 64;
      }
    }
    // ---- store: stay in place, or leave through the mover buffer
    const int nix = wrapi(lx + cx0 + dix, nx), niy = wrapi(ly + cy0 + diy, ny);
    const uint32_t ncell = pack2(nix, niy);
    const bool stay = (nix / TX + g.ntx * (niy / TY)) == tile;
    const float xs = x1 - tx, ys = y1 - ty;
    if (act && stay) {
      const int s = base + p;
      d.cell[s] = ncell;
      d.x[s] = xs;
      d.y[s] = ys;
      d.ux[s] = u1;
      d.uy[s] = u2;
      d.uz[s] = u3;
    }
    const bool mv = act && !stay;
    const unsigned long long mm = __ballot(mv);
    if (mv) {
      const int slot = base + w0 + mcount + __popcll(mm & ((1ull << lane) - 1ull));
      const int s = base + p;
      d.mcell[slot] = ncell;
      d.mx[slot] = xs;
      d.my[slot] = ys;
      d.mux[slot] = u1;
      d.muy[slot] = u2;
      d.muz[slot] = u3;
      d.mq[slot] = q;
      d.mpid[slot] = d.pid[s];
      d.midx[slot] = p;
    }
    mcount += __popcll(mm);
  }
  if (qlen > 0) deposit(qhead, qlen);

  if (bad) atomicOr(&g.ctrl->error, 1);
  if (lane == 0) d.mv_cnt[tile * NW + wave] = mcount;
  if (threadIdx.x == 0) d.mv_stride[tile] = per_wave;
  ene = wave_sum(ene);
  if (lane == 0) s_ene[wave] = ene;
  __syncthreads();
  if (threadIdx.x == 0) d.epart[tile] = ((s_ene[0] + s_ene[1]) + s_ene[2]) + s_ene[3];
  // ---- fold the LDS current tile into the global fixed-point J (periodic wrap)
  for (int i = threadIdx.x; i < JN; i += NT) {
    const int gx = wrapi(cx0 + (i % JX) - O::JLO, nx), gy = wrapi(cy0 + (i / JX) - O::JLO, ny);
    const long long gi = gx + (long long)nx * gy;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const unsigned long long v = sJ[c * JN + i];
      if (v) atomicAdd(&g.J[c * nc + gi], v);
    }
  }
}

#define OSR_INST2(O, P, X, Y) \
  template __global__ void push_deposit_2d_t<O, P, X, Y>(const SpecDesc*, const int2*);
#define OSR_INST2T(X, Y)                                                      \
  OSR_INST2(1, 0, X, Y) OSR_INST2(2, 0, X, Y) OSR_INST2(3, 0, X, Y)           \
  OSR_INST2(1, 1, X, Y) OSR_INST2(2, 1, X, Y) OSR_INST2(3, 1, X, Y)
OSR_INST2T(8, 8)
OSR_INST2T(16, 16)
OSR_INST2T(32, 16)

}  // namespace osr
