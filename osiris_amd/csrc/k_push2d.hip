// k_push2d.hip -- fused 2D push + deposit with compile-time tile shape.
//
// One workgroup of NW waves (4 or 8) per (species, tile); by default 8 for tiles of 512
// cells or more, whose LDS field and current tiles are then shared by twice as many
// waves.  Compared with the generic kernel in k_tiles.hip:
//  * TX/TY are template parameters, so every LDS stencil offset is an immediate;
//  * particle, mover and field arrays are accessed through global-address-space
//    pointers: uniform base + 32-bit offset, no per-array 64-bit address arithmetic, and
//    the loads do not share the LDS wait counter as flat accesses would;
//  * the first virtual particle (VP) of every particle is deposited straight from
//    registers, with the start-point splines of the gather (VP 0 starts at the old
//    position); a trajectory that crosses a cell face has 1 or 2 more VPs, which each
//    wave appends to a private LDS queue and deposits 64 at a time, so the deposit runs
//    on full waves instead of looping max(#VP) times over partly idle lanes;
//  * the current of all waves is accumulated into one int64 fixed-point LDS tile; integer
//    adds are exact and order independent, so the deposit is bit-reproducible run to run
//    (FP32 LDS atomics were measured to be far slower here: a 40x longer LDS wait).
// The arithmetic of gather, push, position advance and split is the one of the generic
// kernel (the reference's AVX2 kernel, simd/os-spec-push-avx2.c:471-760).
#include "osr_device.hpp"
#include "osr_types.hpp"

namespace osr {

typedef __attribute__((address_space(1))) float gf32;
typedef __attribute__((address_space(1))) uint32_t gu32;
typedef __attribute__((address_space(1))) int gi32;
typedef __attribute__((address_space(1))) unsigned long long gu64;
typedef __attribute__((address_space(1))) char gch;

// &b[i] through a 32-bit byte offset: base in scalar registers + one 32-bit VGPR, the
// form the saddr global loads and stores take directly (no 64-bit address arithmetic per
// array).  i * sizeof(*b) < 2^32 holds for every per-tile particle index.
template <class P>
OSR_DEV P el(P b, uint32_t i) {
  return (P)((gch*)b + i * (uint32_t)sizeof(*b));
}

// extra row pitch of the LDS field and current tiles (bank-conflict tuning)
#ifndef OSR_FIELD_PAD
#define OSR_FIELD_PAD 0
#endif
#ifndef OSR_CUR_PAD
#define OSR_CUR_PAD 0
#endif
constexpr int kFieldPad = OSR_FIELD_PAD;
constexpr int kCurPad = OSR_CUR_PAD;

// periodic wrap of a cell index at most one period outside [0, n)
OSR_DEV int wrap1(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

template <int ORDER, int PUSHER, int TX, int TY, int NW>
__global__ __launch_bounds__(64 * NW, 4) void push_deposit_2d_t(const SpecDesc* __restrict__ descs,
                                                                const int2* __restrict__ work) {
  using O = Ord<ORDER>;
  constexpr int NP = O::NP;
  constexpr int NT = 64 * NW;
  // staged field tile GX x GY and current tile JX x JY, stored with row pitches GS / JS
  constexpr int GX = TX + O::GLO + O::GHI, GY = TY + O::GLO + O::GHI;
  constexpr int JX = TX + O::JLO + O::JHI, JY = TY + O::JLO + O::JHI;
  constexpr int GS = GX + kFieldPad, GN = GS * GY;
  constexpr int JS = JX + kCurPad, JN = JS * JY;
  constexpr int QC = 128;  // VP queue capacity per wave (< 64 pending + 64 appended)
  // staged fields: components that share a stencil are paired so that one 64-bit LDS
  // read serves both: (E1, B2) at (half x, full y), (E2, B1) at (full x, half y);
  // E3 (full, full) and B3 (half, half) are single
  __shared__ float2 sEB12[GN], sEB21[GN];
  __shared__ float sE3[GN], sB3[GN];
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

  {
    const gf32* __restrict__ E = (const gf32*)g.E;
    const gf32* __restrict__ B = (const gf32*)g.B;
    for (int i = threadIdx.x; i < GX * GY; i += NT) {
      const int col = i % GX, row = i / GX, li = row * GS + col;
      const int gx = wrapi(cx0 + col - O::GLO, nx), gy = wrapi(cy0 + row - O::GLO, ny);
      const long long gi = gx + (long long)nx * gy;
      sEB12[li] = make_float2(E[gi], B[nc + gi]);
      sEB21[li] = make_float2(E[nc + gi], B[gi]);
      sE3[li] = E[2 * nc + gi];
      sB3[li] = B[2 * nc + gi];
    }
  }
  for (int i = threadIdx.x; i < 3 * JN; i += NT) sJ[i] = 0ull;
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // uniform per workgroup: keep them (and the array bases below) in scalar registers
  const int base = __builtin_amdgcn_readfirstlane(d.off[tile]);
  const int n = __builtin_amdgcn_readfirstlane(d.cnt[tile]);
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
  // particle and mover arrays of this tile, global address space
  const gu32* __restrict__ Pcell = (const gu32*)d.cell + base;
  const gf32* __restrict__ Px = (const gf32*)d.x + base;
  const gf32* __restrict__ Py = (const gf32*)d.y + base;
  const gf32* __restrict__ Pux = (const gf32*)d.ux + base;
  const gf32* __restrict__ Puy = (const gf32*)d.uy + base;
  const gf32* __restrict__ Puz = (const gf32*)d.uz + base;
  const gf32* __restrict__ Pq = (const gf32*)d.q + base;
  const gu32* __restrict__ Ppid = (const gu32*)d.pid + base;
  gu32* __restrict__ Wcell = (gu32*)d.cell + base;
  gf32* __restrict__ Wx = (gf32*)d.x + base;
  gf32* __restrict__ Wy = (gf32*)d.y + base;
  gf32* __restrict__ Wux = (gf32*)d.ux + base;
  gf32* __restrict__ Wuy = (gf32*)d.uy + base;
  gf32* __restrict__ Wuz = (gf32*)d.uz + base;
  gu32* __restrict__ Mcell = (gu32*)d.mcell + base;
  gf32* __restrict__ Mx = (gf32*)d.mx + base;
  gf32* __restrict__ My = (gf32*)d.my + base;
  gf32* __restrict__ Mux = (gf32*)d.mux + base;
  gf32* __restrict__ Muy = (gf32*)d.muy + base;
  gf32* __restrict__ Muz = (gf32*)d.muz + base;
  gf32* __restrict__ Mq = (gf32*)d.mq + base;
  gu32* __restrict__ Mpid = (gu32*)d.mpid + base;
  gi32* __restrict__ Midx = (gi32*)d.midx + base;

  // deposit of the queued VPs in slots [head, head + cnt) (cnt <= 64, one per lane)
  auto deposit_queue = [&](int head, int cnt) {
    if (lane < cnt) {
      const int s = (head + lane) & (QC - 1);
      const float x0 = Q[0][s], y0 = Q[1][s];
      float S0x[NP], S0y[NP];
      spline<ORDER>(x0, S0x);
      spline<ORDER>(y0, S0y);
      dep_vp<ORDER>(J1, J2, J3, S0x, S0y, x0, y0, Q[2][s], Q[3][s], Q[4][s], Q[5][s], QN[s], JS,
                     jn1, jn2, jscale);
    }
  };

  double ene = 0.0;
  int mcount = 0;
  int qhead = 0, qlen = 0;
  bool bad = false;
  // The particle loads of the next iteration are issued before the current one is
  // processed (software pipelining): a wave otherwise waits a full memory latency per
  // iteration, and the bytes in flight per CU would cap the bandwidth well below peak.
  uint32_t f_cell = 0, f_pid = 0;
  float f_x = 0.f, f_y = 0.f, f_u1 = 0.f, f_u2 = 0.f, f_u3 = 0.f, f_q = 0.f;
  auto prefetch = [&](int pf) {
    if (pf < w1) {
      const unsigned upf = (unsigned)pf;
      f_cell = *el(Pcell, upf);
      f_x = *el(Px, upf);
      f_y = *el(Py, upf);
      f_u1 = *el(Pux, upf);
      f_u2 = *el(Puy, upf);
      f_u3 = *el(Puz, upf);
      f_q = *el(Pq, upf);
      f_pid = *el(Ppid, upf);
    }
  };
  prefetch(w0 + lane);
  for (int p0 = w0; p0 < w1; p0 += 64) {
    const int p = p0 + lane;
    const unsigned up = (unsigned)p;  // 32-bit unsigned offsets from the scalar bases
    bool act = p < w1;
    float x = 0.f, y = 0.f, u1 = 0.f, u2 = 0.f, u3 = 0.f, q = 0.f;
    int lx = 0, ly = 0;
    const uint32_t cell = f_cell, pid = f_pid;
    if (act) {
      x = f_x;
      y = f_y;
      u1 = f_u1;
      u2 = f_u2;
      u3 = f_u3;
      q = f_q;
    }
    prefetch(p + 64);
    if (act) {
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
    const int gb = (ly + O::GLO + O::LP) * GS + lx + O::GLO + O::LP;
    const int gbx = gb - (x < 0.f ? 1 : 0);
    const int gby = gb - (y < 0.f ? GS : 0);
    const int gbxy = gbx - (y < 0.f ? GS : 0);
    float e1 = 0.f, e2 = 0.f, e3 = 0.f, b1 = 0.f, b2 = 0.f, b3 = 0.f;
#pragma unroll
    for (int k2 = 0; k2 < NP; k2++) {
      float le1 = 0.f, le2 = 0.f, le3 = 0.f, lb1 = 0.f, lb2 = 0.f, lb3 = 0.f;
#pragma unroll
      for (int k1 = 0; k1 < NP; k1++) {
        const float2 p12 = sEB12[gbx + k2 * GS + k1];
        const float2 p21 = sEB21[gby + k2 * GS + k1];
        le1 = fma_(p12.x, wxh[k1], le1);
        le2 = fma_(p21.x, wx[k1], le2);
        le3 = fma_(sE3[gb + k2 * GS + k1], wx[k1], le3);
        lb1 = fma_(p21.y, wx[k1], lb1);
        lb2 = fma_(p12.y, wxh[k1], lb2);
        lb3 = fma_(sB3[gbxy + k2 * GS + k1], wxh[k1], lb3);
      }
      e1 = fma_(le1, wy[k2], e1);
      e2 = fma_(le2, wyh[k2], e2);
      e3 = fma_(le3, wy[k2], e3);
      b1 = fma_(lb1, wyh[k2], b1);
      b2 = fma_(lb2, wy[k2], b2);
      b3 = fma_(lb3, wyh[k2], b3);
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
    // ---- split; VP 0 is deposited in place, VPs 1 and 2 go through the queue
    VP2 vp[3];
    const int nvp = act ? split2d_fast(x, y, x1, y1, dix, diy, q, v3, vp) : 0;
    const int nb0 = (ly + O::LP + O::JLO) * JS + lx + O::LP + O::JLO;
    if (act)
      dep_vp<ORDER>(J1, J2, J3, wx, wy, vp[0].x0, vp[0].y0, vp[0].x1, vp[0].y1, vp[0].q,
                    vp[0].vz, nb0, JS, jn1, jn2, jscale);
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int v = 1; v < 3; v++) {
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
        QN[s] = nb0 + vp[v].dj * JS + vp[v].di;
      }
      qlen += __popcll(m);
      if (qlen >= 64) {
        deposit_queue(qhead, 64);
        qhead = (qhead + 64) & (QC - 1);
        qlen -= 64;
      }
    }
    // ---- store: stay in place, or leave through the mover buffer
    const int nix = wrap1(lx + cx0 + dix, nx), niy = wrap1(ly + cy0 + diy, ny);
    const uint32_t ncell = pack2(nix, niy);
    const bool stay = (nix / TX + g.ntx * (niy / TY)) == tile;
    const float xs = x1 - tx, ys = y1 - ty;
    if (act && stay) {
      *el(Wcell, up) = ncell;
      *el(Wx, up) = xs;
      *el(Wy, up) = ys;
      *el(Wux, up) = u1;
      *el(Wuy, up) = u2;
      *el(Wuz, up) = u3;
    }
    const bool mv = act && !stay;
    const unsigned long long mm = __ballot(mv);
    if (mv) {
      const unsigned slot = (unsigned)(w0 + mcount + __popcll(mm & below));
      *el(Mcell, slot) = ncell;
      *el(Mx, slot) = xs;
      *el(My, slot) = ys;
      *el(Mux, slot) = u1;
      *el(Muy, slot) = u2;
      *el(Muz, slot) = u3;
      *el(Mq, slot) = q;
      *el(Mpid, slot) = pid;
      *el(Midx, slot) = p;
    }
    mcount += __popcll(mm);
  }
  if (qlen > 0) deposit_queue(qhead, qlen);

  if (bad) atomicOr(&g.ctrl->error, 1);
  if (lane == 0) d.mv_cnt[tile * NW + wave] = mcount;
  if (threadIdx.x == 0) d.mv_stride[tile] = per_wave;
  ene = wave_sum(ene);
  if (lane == 0) s_ene[wave] = ene;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = s_ene[0];
#pragma unroll
    for (int w = 1; w < NW; w++) t += s_ene[w];
    d.epart[tile] = t;
  }
  // ---- fold the LDS current tile into the global fixed-point J (periodic wrap)
  gu64* __restrict__ GJ = (gu64*)g.J;
  for (int i = threadIdx.x; i < JX * JY; i += NT) {
    const int col = i % JX, row = i / JX, li = row * JS + col;
    const int gx = wrapi(cx0 + col - O::JLO, nx), gy = wrapi(cy0 + row - O::JLO, ny);
    const long long gi = gx + (long long)nx * gy;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const unsigned long long v = sJ[c * JN + li];
      if (v) __hip_atomic_fetch_add(&GJ[c * nc + gi], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

#define OSR_INST2(O, P, X, Y, W) \
  template __global__ void push_deposit_2d_t<O, P, X, Y, W>(const SpecDesc*, const int2*);
#define OSR_INST2T(X, Y, W)                                                            \
  OSR_INST2(1, 0, X, Y, W) OSR_INST2(2, 0, X, Y, W) OSR_INST2(3, 0, X, Y, W)           \
  OSR_INST2(1, 1, X, Y, W) OSR_INST2(2, 1, X, Y, W) OSR_INST2(3, 1, X, Y, W)
OSR_INST2T(8, 8, 4)
OSR_INST2T(16, 16, 4)
OSR_INST2T(16, 16, 8)
OSR_INST2T(32, 16, 4)
OSR_INST2T(32, 16, 8)

}  // namespace osr
