// k_tiles.hip -- the fused tile kernels of the resident fast path.
//
//   push_deposit_{2d,3d}<ORDER,PUSHER>: one workgroup per (species, tile).  E and B of
//     the tile plus the stencil halo are staged in LDS; every wave walks a contiguous
//     slice of the tile's particles (coalesced SoA loads), gathers from LDS, pushes
//     (Boris or Vay), advances the position, splits the trajectory into virtual
//     particles and deposits them into an LDS current tile with int64 fixed-point
//     atomics.  Integer addition is associative, so the deposited current is bit-for-bit
//     independent of the order in which lanes, waves and workgroups add their share:
//     the reduction is reproducible run to run.  The LDS tile is then added to the global
//     fixed-point J with the periodic wrap (the guard-cell fold of the reference).
//     Particles that stay in the tile are written back in place; leavers go to the
//     wave's region of the mover buffer (ballot-ranked, so the order is deterministic).
//   tile_fill: moves the leavers into their destination tiles, filling the holes they
//     left; a deterministic stable compaction.  A tile that would overflow its capacity
//     is left unchanged and flags the step (Ctrl::halt); the host re-lays the particles
//     out before the next push.
//
// Reference: spec/os-spec-push.f03:1438-1605 (advance_deposit_2d/3d),
// emf/os-emf-interpolate.f03:554-779 (GET_EMF_2D/3D), spec/os-spec-current.f03:659-905,
// 2691-2905 (split + DEP_CURRENT), simd/os-spec-push-avx2.c:471-760 (fused SIMD kernel),
// particles/os-particles.f03:1416-1443 (private J + reduce), spec/os-species.f03:269-441
// (counting sort of particles by cell).
#include "osr_device.hpp"
#include "osr_types.hpp"

namespace osr {

OSR_DEV int wrapi(int i, int n) {
  while (i < 0) i += n;
  while (i >= n) i -= n;
  return i;
}

// Exclusive rank of `flag` among the threads of the block in thread order, and the total.
// Uses wave ballots and one LDS exchange; all threads must call it.
template <int NT>
OSR_DEV int block_rank(bool flag, int* s_wtot, int& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long m = __ballot(flag);
  int r = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) s_wtot[wave] = __popcll(m);
  __syncthreads();
  int before = 0;
  total = 0;
  for (int w = 0; w < NT / 64; w++) {
    int c = s_wtot[w];
    if (w < wave) before += c;
    total += c;
  }
  __syncthreads();
  return before + r;
}

// block_rank for PF flags per thread at once, ranked in the order (p, thread): item p of
// thread t comes after every item p' < p and after item p of the threads below t.  One
// LDS exchange (two barriers) for all PF; all threads must call it.
template <int NT, int PF>
OSR_DEV void block_rank_multi(const bool (&flag)[PF], int (*s_cnt)[NT / 64], int (&rank)[PF],
                              int& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  int own[PF];
#pragma unroll
  for (int p = 0; p < PF; p++) {
    const unsigned long long m = __ballot(flag[p]);
    own[p] = __popcll(m & below);
    if (lane == 0) s_cnt[p][wave] = __popcll(m);
  }
  __syncthreads();
  int base = 0;
#pragma unroll
  for (int p = 0; p < PF; p++) {
    int before = 0, tp = 0;
    for (int w = 0; w < NT / 64; w++) {
      const int c = s_cnt[p][w];
      if (w < wave) before += c;
      tp += c;
    }
    rank[p] = base + before + own[p];
    base += tp;
  }
  total = base;
  __syncthreads();
}

// ------------------------------------------------------------------------------------
// 2D push + deposit
// ------------------------------------------------------------------------------------
template <int ORDER, int PUSHER, int NT>
__global__ __launch_bounds__(NT) void push_deposit_2d(const SpecDesc* __restrict__ descs,
                                                      const int2* __restrict__ work) {
  using O = Ord<ORDER>;
  constexpr int NP = O::NP;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double s_ene[NT / 64];

  const int2 wk = work[blockIdx.x];
  const SpecDesc& d = descs[wk.x];
  const GridDesc& g = d.g;
  if (g.ctrl->halt != 0) return;
  const int tile = wk.y;
  const int TX = g.tx, TY = g.ty;
  const int GX = TX + O::GLO + O::GHI, GY = TY + O::GLO + O::GHI;
  const int JX = TX + O::JLO + O::JHI, JY = TY + O::JLO + O::JHI;
  const int GN = GX * GY, JN = JX * JY;
  unsigned long long* sJ = (unsigned long long*)smem;      // [3][JN]
  float* sF = (float*)(smem + sizeof(unsigned long long) * 3 * JN);  // E[3][GN], B[3][GN]
  const int cx0 = (tile % g.ntx) * TX, cy0 = (tile / g.ntx) * TY;
  const int nx = g.nx, ny = g.ny;
  const long long nc = g.ncell;

  for (int i = threadIdx.x; i < GN; i += NT) {
    int gx = wrapi(cx0 + (i % GX) - O::GLO, nx), gy = wrapi(cy0 + (i / GX) - O::GLO, ny);
    long long gi = gx + (long long)nx * gy;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      sF[c * GN + i] = g.E[c * nc + gi];
      sF[(3 + c) * GN + i] = g.B[c * nc + gi];
    }
  }
  for (int i = threadIdx.x; i < 3 * JN; i += NT) sJ[i] = 0ull;
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int NW = NT / 64;
  const int base = d.off[tile];
  const int n = d.cnt[tile];
  const int per_wave = ((n + NT - 1) / NT) * 64;
  const int w0 = wave * per_wave;
  const int w1 = min(w0 + per_wave, n);
  const float tem = d.tem;
  const float dtdx1 = g.dt_dx[0], dtdx2 = g.dt_dx[1];
  const float jn1 = g.jnorm[0], jn2 = g.jnorm[1];
  const double jscale = g.jscale;
  const float* sE1 = sF;
  const float* sE2 = sF + GN;
  const float* sE3 = sF + 2 * GN;
  const float* sB1 = sF + 3 * GN;
  const float* sB2 = sF + 4 * GN;
  const float* sB3 = sF + 5 * GN;
  unsigned long long* sJ1 = sJ;
  unsigned long long* sJ2 = sJ + JN;
  unsigned long long* sJ3 = sJ + 2 * JN;

  double ene = 0.0;
  int mcount = 0;
  bool bad = false;
  for (int p0 = w0; p0 < w1; p0 += 64) {
    const int p = p0 + lane;
    bool act = p < w1;
    uint32_t cell = 0;
    float x = 0.f, y = 0.f, u1 = 0.f, u2 = 0.f, u3 = 0.f, q = 0.f;
    int lx = 0, ly = 0;
    if (act) {
      const int s = base + p;
      cell = d.cell[s];
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
    // ---- gather (GET_EMF_2D, same FMA order as the AVX2 kernel)
    float wx[NP], wy[NP], wxh[NP], wyh[NP];
    spline<ORDER>(x, wx);
    spline<ORDER>(y, wy);
    const float hx = x < 0.f ? 1.f : 0.f, hy = y < 0.f ? 1.f : 0.f;
    splineh<ORDER>(x, hx, wxh);
    splineh<ORDER>(y, hy, wyh);
    const int gx = lx + O::GLO + O::LP, gy = ly + O::GLO + O::LP;
    const int gxh = gx - (x < 0.f), gyh = gy - (y < 0.f);
    float e1 = 0.f, e2 = 0.f, e3 = 0.f, b1 = 0.f, b2 = 0.f, b3 = 0.f;
#pragma unroll
    for (int k2 = 0; k2 < NP; k2++) {
      float l1 = 0.f, l2 = 0.f, l3 = 0.f;
#pragma unroll
      for (int k1 = 0; k1 < NP; k1++) {
        l1 = fma_(sE1[(gy + k2) * GX + gxh + k1], wxh[k1], l1);
        l2 = fma_(sE2[(gyh + k2) * GX + gx + k1], wx[k1], l2);
        l3 = fma_(sE3[(gy + k2) * GX + gx + k1], wx[k1], l3);
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
        l1 = fma_(sB1[(gyh + k2) * GX + gx + k1], wx[k1], l1);
        l2 = fma_(sB2[(gy + k2) * GX + gxh + k1], wxh[k1], l2);
        l3 = fma_(sB3[(gyh + k2) * GX + gxh + k1], wxh[k1], l3);
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
    // ---- split + deposit
    VP2 vp[3];
    const int nvp = act ? split2d(x, y, x1, y1, dix, diy, q, v3, vp) : 0;
    for (int v = 0; v < nvp; v++) {
      const VP2 V = pick(vp, v);
      float S0x[NP], S0y[NP];
      spline<ORDER>(V.x0, S0x);
      spline<ORDER>(V.y0, S0y);
      const int jx0 = lx + V.di + O::LP + O::JLO, jy0 = ly + V.dj + O::LP + O::JLO;
      dep_vp<ORDER>(sJ1, sJ2, sJ3, S0x, S0y, V.x0, V.y0, V.x1, V.y1, V.q, V.vz, jy0 * JX + jx0,
                    JX, jn1, jn2, jscale);
    }
    // ---- store: stay in place, or leave through the mover buffer
    int nix = wrapi(lx + cx0 + dix, nx), niy = wrapi(ly + cy0 + diy, ny);
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
  if (bad) atomicOr(&g.ctrl->error, 1);
  if (lane == 0) d.mv_cnt[tile * NW + wave] = mcount;
  if (threadIdx.x == 0) d.mv_stride[tile] = per_wave;
  ene = wave_sum(ene);
  if (lane == 0) s_ene[wave] = ene;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < NW; w++) t += s_ene[w];
    d.epart[tile] = t;
  }
  // ---- fold the LDS current tile into global J with the periodic wrap
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

// ------------------------------------------------------------------------------------
// 3D push + deposit
// ------------------------------------------------------------------------------------
template <int ORDER, int PUSHER, int NT>
__global__ __launch_bounds__(NT) void push_deposit_3d(const SpecDesc* __restrict__ descs,
                                                      const int2* __restrict__ work) {
  using O = Ord<ORDER>;
  constexpr int NP = O::NP;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double s_ene[NT / 64];

  const int2 wk = work[blockIdx.x];
  const SpecDesc& d = descs[wk.x];
  const GridDesc& g = d.g;
  if (g.ctrl->halt != 0) return;
  const int tile = wk.y;
  const int TX = g.tx, TY = g.ty, TZ = g.tz;
  const int GX = TX + O::GLO + O::GHI, GY = TY + O::GLO + O::GHI, GZ = TZ + O::GLO + O::GHI;
  const int JX = TX + O::JLO + O::JHI, JY = TY + O::JLO + O::JHI, JZ = TZ + O::JLO + O::JHI;
  const int GN = GX * GY * GZ, JN = JX * JY * JZ;
  unsigned long long* sJ = (unsigned long long*)smem;
  float* sF = (float*)(smem + sizeof(unsigned long long) * 3 * JN);
  // per-wave queues of the virtual particles 1..3 (x0, y0, z0, x1, y1, z1, q; node)
  constexpr int QC = kVpQueue3;
  float* sQall = sF + 6 * GN;
  int* sQnall = (int*)(sQall + (NT / 64) * 7 * QC);
  const int ttx = tile % g.ntx, tty = (tile / g.ntx) % g.nty, ttz = tile / (g.ntx * g.nty);
  const int cx0 = ttx * TX, cy0 = tty * TY, cz0 = ttz * TZ;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const long long nc = g.ncell;

  for (int i = threadIdx.x; i < GN; i += NT) {
    const int lx = i % GX, ly = (i / GX) % GY, lz = i / (GX * GY);
    const int gx = wrapi(cx0 + lx - O::GLO, nx), gy = wrapi(cy0 + ly - O::GLO, ny),
              gz = wrapi(cz0 + lz - O::GLO, nz);
    const long long gi = gx + (long long)nx * (gy + (long long)ny * gz);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      sF[c * GN + i] = g.E[c * nc + gi];
      sF[(3 + c) * GN + i] = g.B[c * nc + gi];
    }
  }
  for (int i = threadIdx.x; i < 3 * JN; i += NT) sJ[i] = 0ull;
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int NW = NT / 64;
  const int base = d.off[tile];
  const int n = d.cnt[tile];
  const int per_wave = ((n + NT - 1) / NT) * 64;
  const int w0 = wave * per_wave;
  const int w1 = min(w0 + per_wave, n);
  const float tem = d.tem;
  const float dtdx1 = g.dt_dx[0], dtdx2 = g.dt_dx[1], dtdx3 = g.dt_dx[2];
  const float jn1 = g.jnorm[0], jn2 = g.jnorm[1], jn3 = g.jnorm[2];
  const double jscale = g.jscale;
  const int GXY = GX * GY, JXY = JX * JY;

  float* const Q = sQall + wave * 7 * QC;
  int* const QN = sQnall + wave * QC;
  int qhead = 0, qlen = 0;
  // deposit of the queued VPs in slots [head, head + cnt) (cnt <= 64, one per lane)
  auto deposit_queue = [&](int head, int cnt) {
    if (lane < cnt) {
      const int sl = (head + lane) & (QC - 1);
      const float x0 = Q[sl], y0 = Q[QC + sl], z0 = Q[2 * QC + sl];
      float S0x[NP], S0y[NP], S0z[NP];
      spline<ORDER>(x0, S0x);
      spline<ORDER>(y0, S0y);
      spline<ORDER>(z0, S0z);
      dep_vp3<ORDER>(sJ, JN, S0x, S0y, S0z, x0, y0, z0, Q[3 * QC + sl], Q[4 * QC + sl],
                     Q[5 * QC + sl], Q[6 * QC + sl], QN[sl], JX, JXY, jn1, jn2, jn3, jscale);
    }
  };

  double ene = 0.0;
  int mcount = 0;
  bool bad = false;
  for (int p0 = w0; p0 < w1; p0 += 64) {
    const int p = p0 + lane;
    bool act = p < w1;
    uint32_t cell = 0;
    float x = 0.f, y = 0.f, z = 0.f, u1 = 0.f, u2 = 0.f, u3 = 0.f, q = 0.f;
    int lx = 0, ly = 0, lz = 0;
    if (act) {
      const int s = base + p;
      cell = d.cell[s];
      x = d.x[s];
      y = d.y[s];
      z = d.z[s];
      u1 = d.ux[s];
      u2 = d.uy[s];
      u3 = d.uz[s];
      q = d.q[s];
      int ix, iy, iz;
      unpack3(cell, ix, iy, iz);
      lx = ix - cx0;
      ly = iy - cy0;
      lz = iz - cz0;
      if (lx < 0 || lx >= TX || ly < 0 || ly >= TY || lz < 0 || lz >= TZ) {
        bad = true;
        act = false;
        lx = ly = lz = 0;
      }
    }
    float wx[NP], wy[NP], wz[NP], wxh[NP], wyh[NP], wzh[NP];
    spline<ORDER>(x, wx);
    spline<ORDER>(y, wy);
    spline<ORDER>(z, wz);
    splineh<ORDER>(x, x < 0.f ? 1.f : 0.f, wxh);
    splineh<ORDER>(y, y < 0.f ? 1.f : 0.f, wyh);
    splineh<ORDER>(z, z < 0.f ? 1.f : 0.f, wzh);
    const int gx = lx + O::GLO + O::LP, gy = ly + O::GLO + O::LP, gz = lz + O::GLO + O::LP;
    const int gxh = gx - (x < 0.f), gyh = gy - (y < 0.f), gzh = gz - (z < 0.f);
    const float* sE1 = sF;
    const float* sE2 = sF + GN;
    const float* sE3 = sF + 2 * GN;
    const float* sB1 = sF + 3 * GN;
    const float* sB2 = sF + 4 * GN;
    const float* sB3 = sF + 5 * GN;
    // GET_EMF_3D line/plane order (os-emf-interpolate.f03:697-773); the AVX2 3D kernel
    // uses the same fma order along k1, then k2, then k3.
    float e1 = 0.f, e2 = 0.f, e3 = 0.f;
#pragma unroll
    for (int k3 = 0; k3 < NP; k3++) {
      float p1 = 0.f, p2 = 0.f, p3 = 0.f;
#pragma unroll
      for (int k2 = 0; k2 < NP; k2++) {
        float l1 = 0.f, l2 = 0.f, l3 = 0.f;
#pragma unroll
        for (int k1 = 0; k1 < NP; k1++) {
          l1 = fma_(sE1[(gz + k3) * GXY + (gy + k2) * GX + gxh + k1], wxh[k1], l1);
          l2 = fma_(sE2[(gz + k3) * GXY + (gyh + k2) * GX + gx + k1], wx[k1], l2);
          l3 = fma_(sE3[(gzh + k3) * GXY + (gy + k2) * GX + gx + k1], wx[k1], l3);
        }
        p1 = fma_(l1, wy[k2], p1);
        p2 = fma_(l2, wyh[k2], p2);
        p3 = fma_(l3, wy[k2], p3);
      }
      e1 = fma_(p1, wz[k3], e1);
      e2 = fma_(p2, wz[k3], e2);
      e3 = fma_(p3, wzh[k3], e3);
    }
    float b1 = 0.f, b2 = 0.f, b3 = 0.f;
#pragma unroll
    for (int k3 = 0; k3 < NP; k3++) {
      float p1 = 0.f, p2 = 0.f, p3 = 0.f;
#pragma unroll
      for (int k2 = 0; k2 < NP; k2++) {
        float l1 = 0.f, l2 = 0.f, l3 = 0.f;
#pragma unroll
        for (int k1 = 0; k1 < NP; k1++) {
          l1 = fma_(sB1[(gzh + k3) * GXY + (gyh + k2) * GX + gx + k1], wx[k1], l1);
          l2 = fma_(sB2[(gzh + k3) * GXY + (gy + k2) * GX + gxh + k1], wxh[k1], l2);
          l3 = fma_(sB3[(gz + k3) * GXY + (gyh + k2) * GX + gxh + k1], wxh[k1], l3);
        }
        p1 = fma_(l1, wyh[k2], p1);
        p2 = fma_(l2, wy[k2], p2);
        p3 = fma_(l3, wyh[k2], p3);
      }
      b1 = fma_(p1, wzh[k3], b1);
      b2 = fma_(p2, wzh[k3], b2);
      b3 = fma_(p3, wz[k3], b3);
    }
    const float ke = push<PUSHER>(tem, e1, e2, e3, b1, b2, b3, u1, u2, u3);
    if (act) ene += (double)q * (double)ke;
    const float rg = rgamma(u1, u2, u3);
    const float x1 = fma_(u1 * rg, dtdx1, x), y1 = fma_(u2 * rg, dtdx2, y),
                z1 = fma_(u3 * rg, dtdx3, z);
    const float tx = trim(x1), ty = trim(y1), tz = trim(z1);
    const int dix = (int)tx, diy = (int)ty, diz = (int)tz;
    // split; VP 0 is deposited in place with the gather's splines (it starts at the old
    // position), VPs 1..3 of the lanes whose trajectory crosses faces go through the queue
    VP3 vp[4];
    const int nvp = act ? split3d(x, y, z, x1, y1, z1, dix, diy, diz, q, vp) : 0;
    const int nb0 = (lz + O::LP + O::JLO) * JXY + (ly + O::LP + O::JLO) * JX + lx + O::LP + O::JLO;
    if (act)
      dep_vp3<ORDER>(sJ, JN, wx, wy, wz, vp[0].x0, vp[0].y0, vp[0].z0, vp[0].x1, vp[0].y1,
                     vp[0].z1, vp[0].q, nb0, JX, JXY, jn1, jn2, jn3, jscale);
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int v = 1; v < 4; v++) {
      const unsigned long long m = __ballot(v < nvp);
      if (m == 0ull) break;
      if (v < nvp) {
        const VP3 V = pick(vp, v);
        const int sl = (qhead + qlen + __popcll(m & below)) & (QC - 1);
        Q[0 * QC + sl] = V.x0;
        Q[1 * QC + sl] = V.y0;
        Q[2 * QC + sl] = V.z0;
        Q[3 * QC + sl] = V.x1;
        Q[4 * QC + sl] = V.y1;
        Q[5 * QC + sl] = V.z1;
        Q[6 * QC + sl] = V.q;
        QN[sl] = nb0 + V.dk * JXY + V.dj * JX + V.di;
      }
      qlen += __popcll(m);
      if (qlen >= 64) {
        deposit_queue(qhead, 64);
        qhead = (qhead + 64) & (QC - 1);
        qlen -= 64;
      }
    }
    const int nix = wrapi(lx + cx0 + dix, nx), niy = wrapi(ly + cy0 + diy, ny),
              niz = wrapi(lz + cz0 + diz, nz);
    const uint32_t ncell = pack3(nix, niy, niz);
    const bool stay = (nix / TX + g.ntx * (niy / TY + g.nty * (niz / TZ))) == tile;
    const float xs = x1 - tx, ys = y1 - ty, zs = z1 - tz;
    if (act && stay) {
      const int s = base + p;
      d.cell[s] = ncell;
      d.x[s] = xs;
      d.y[s] = ys;
      d.z[s] = zs;
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
      d.mz[slot] = zs;
      d.mux[slot] = u1;
      d.muy[slot] = u2;
      d.muz[slot] = u3;
      d.mq[slot] = q;
      d.mpid[slot] = d.pid[s];
      d.midx[slot] = p;
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
    double t = 0.0;
    for (int w = 0; w < NW; w++) t += s_ene[w];
    d.epart[tile] = t;
  }
  for (int i = threadIdx.x; i < JN; i += NT) {
    const int lx = i % JX, ly = (i / JX) % JY, lz = i / JXY;
    const int gx = wrapi(cx0 + lx - O::JLO, nx), gy = wrapi(cy0 + ly - O::JLO, ny),
              gz = wrapi(cz0 + lz - O::JLO, nz);
    const long long gi = gx + (long long)nx * (gy + (long long)ny * gz);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const unsigned long long v = sJ[c * JN + i];
      if (v) atomicAdd(&g.J[c * nc + gi], v);
    }
  }
}

// ------------------------------------------------------------------------------------
// Tile maintenance
// ------------------------------------------------------------------------------------
// Distinct neighbour tiles of `tile` (excluding itself) in a fixed order.
OSR_DEV int neighbours(const GridDesc& g, int tile, int* nb) {
  int tx, ty, tz;
  if (g.dims == 2) {
    tx = tile % g.ntx;
    ty = tile / g.ntx;
    tz = 0;
  } else {
    tx = tile % g.ntx;
    ty = (tile / g.ntx) % g.nty;
    tz = tile / (g.ntx * g.nty);
  }
  const int zr = g.dims == 2 ? 0 : 1;
  int k = 0;
  for (int dz = -zr; dz <= zr; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        if (dx == 0 && dy == 0 && dz == 0) continue;
        const int nxt = wrapi(tx + dx, g.ntx), nyt = wrapi(ty + dy, g.nty);
        const int nzt = g.dims == 2 ? 0 : wrapi(tz + dz, g.ntz);
        const int t = nxt + g.ntx * (nyt + g.nty * nzt);
        if (t == tile) continue;
        bool dup = false;
        for (int j = 0; j < k; j++) dup |= nb[j] == t;
        if (!dup) nb[k++] = t;
      }
  return k;
}

OSR_DEV void copy_from_mover(const SpecDesc& d, int src, int dst, int dims) {
  d.cell[dst] = d.mcell[src];
  d.x[dst] = d.mx[src];
  d.y[dst] = d.my[src];
  if (dims == 3) d.z[dst] = d.mz[src];
  d.ux[dst] = d.mux[src];
  d.uy[dst] = d.muy[src];
  d.uz[dst] = d.muz[src];
  d.q[dst] = d.mq[src];
  d.pid[dst] = d.mpid[src];
}

OSR_DEV void copy_particle(const SpecDesc& d, int src, int dst, int dims) {
  d.cell[dst] = d.cell[src];
  d.x[dst] = d.x[src];
  d.y[dst] = d.y[src];
  if (dims == 3) d.z[dst] = d.z[src];
  d.ux[dst] = d.ux[src];
  d.uy[dst] = d.uy[src];
  d.uz[dst] = d.uz[src];
  d.q[dst] = d.q[src];
  d.pid[dst] = d.pid[src];
}


// Move the movers destined to this tile into it and close the holes (deterministic:
// sources = incoming movers by (neighbour, wave, slot), then the non-hole particles of
// the tail [newcnt, cnt) ascending; targets = holes below newcnt ascending, then the
// appended slots [cnt, newcnt)).  A tile that cannot take its incoming particles (its
// capacity, or the LDS lists of this kernel, would overflow) changes nothing, marks
// itself with inc = -1 and raises Ctrl::halt; the host then re-lays the species out from
// the tiles that did fill plus the mover buffers.
// NT threads per tile; INC_MAX incoming movers and TAIL_BITS holes left in the tail are the
// most one fill can handle (LDS lists); beyond them the tile flags an overflow.
// Segment of item i in a concatenation of n lists with start offsets ofs[0..n]
// (ofs[0] = 0, non-decreasing, 0 <= i < ofs[n]): the last l with ofs[l] <= i, which skips
// empty lists.  Binary search: the fill walks up to 8 x 8 (2D) or 26 x 4 (3D) lists.
OSR_DEV int seg_find(const int* ofs, int n, int i) {
  int lo = 0, hi = n;  // ofs[lo] <= i < ofs[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (ofs[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

constexpr int kFillPF = 8;  // candidate movers per thread whose cells are loaded together

template <int NT, int INC_MAX, int TAIL_BITS>
__global__ __launch_bounds__(NT) void tile_fill(const SpecDesc* __restrict__ descs,
                                                const int2* __restrict__ work) {
  constexpr int MAXL = 26 * 4;  // candidate lists: neighbours x waves
  __shared__ int s_nb[26];
  __shared__ int s_nnb;
  __shared__ int s_lofs[MAXL + 1];
  __shared__ int s_lbase[MAXL];
  __shared__ int s_wtot[NT / 64];
  __shared__ int s_pcnt[kFillPF][NT / 64];
  __shared__ int s_hpre[17];
  __shared__ int s_inc[INC_MAX];
  __shared__ unsigned int s_bits[TAIL_BITS / 32];
  const int2 wk = work[blockIdx.x];
  const SpecDesc& d = descs[wk.x];
  const GridDesc& g = d.g;
  if (g.ctrl->halt >= 2) return;
  const int tile = wk.y;
  const int NW = d.waves;
  const int dims = g.dims;
  const int base = d.off[tile];
  const int n = d.cnt[tile];
  const int stride = d.mv_stride[tile];
  if (threadIdx.x == 0) {
    s_nnb = neighbours(g, tile, s_nb);
    s_hpre[0] = 0;
    for (int w = 0; w < NW; w++) s_hpre[w + 1] = s_hpre[w] + d.mv_cnt[tile * NW + w];
  }
  __syncthreads();
  const int L = s_nnb * NW;
  if (L > MAXL) {  // more mover lists than the LDS tables hold: never with 26 x 4 or 8 x 8
    if (threadIdx.x == 0) atomicOr(&g.ctrl->error, 2);
    return;
  }
  if (threadIdx.x < 64) {  // wave 0: list bases and the prefix of the list lengths
    const int lane = threadIdx.x;
    int carry = 0;
    for (int c0 = 0; c0 < L; c0 += 64) {
      const int l = c0 + lane;
      int v = 0;
      if (l < L) {
        const int nt = s_nb[l / NW], w = l % NW;
        v = d.mv_cnt[nt * NW + w];
        s_lbase[l] = d.off[nt] + w * d.mv_stride[nt];
      }
      for (int o = 1; o < 64; o <<= 1) {  // inclusive scan over the wave
        const int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
      }
      if (l < L) s_lofs[l + 1] = carry + v;
      carry += __shfl(v, 63, 64);
    }
    if (lane == 0) s_lofs[0] = 0;
  }
  __syncthreads();
  // a tile of the guard band of open x walls takes no movers: the particles that left
  // through a wall (ix < 1 or ix > nx) are dropped (spec/os-spec-boundary.f03:476-513)
  const bool absorb = (tile % g.ntx) * g.tx >= interior_nx(g);
  const int T = absorb ? 0 : s_lofs[L];
  const int H = s_hpre[NW];
  // pass 1: collect the incoming movers, in candidate order.  Candidates are taken PF NT
  // at a time, candidate c0 + p NT + thread as item p of the thread, and the cells of all
  // of them are loaded before any is ranked: one memory round trip per PF NT candidates
  // (a tile's neighbours hold some 8 x 200 movers on the 2D two-stream workload).
  int I = 0;
  for (int c0 = 0; c0 < T; c0 += NT * kFillPF) {
    uint32_t cell[kFillPF];
    int src[kFillPF];
#pragma unroll
    for (int p = 0; p < kFillPF; p++) {
      const int i = c0 + p * NT + threadIdx.x;
      src[p] = -1;
      cell[p] = 0u;
      if (i < T) {
        const int l = seg_find(s_lofs, L, i);
        src[p] = s_lbase[l] + (i - s_lofs[l]);
        cell[p] = d.mcell[src[p]];
      }
    }
    bool f[kFillPF];
#pragma unroll
    for (int p = 0; p < kFillPF; p++) f[p] = src[p] >= 0 && tile_of(g, cell[p]) == tile;
    int r[kFillPF], tot;
    block_rank_multi<NT, kFillPF>(f, s_pcnt, r, tot);
#pragma unroll
    for (int p = 0; p < kFillPF; p++)
      if (f[p] && I + r[p] < INC_MAX) s_inc[I + r[p]] = src[p];
    I += tot;
  }
  // every wave's entries of s_inc are in LDS before any wave reads them in pass 2 (the
  // hole count below has barriers only when the tile loses more movers than it gains)
  __syncthreads();
  const int nn = n - H + I;
  if (I > INC_MAX || nn > d.cap[tile] - 64 * NW || H - I > TAIL_BITS) {
    if (threadIdx.x == 0) {
      d.inc[tile] = -1;
      // the count the tile needs after this step, or -1 if more capacity would not help
      // (the incoming or tail lists exceed this kernel's LDS tables)
      d.newcnt[tile] = (I > INC_MAX || H - I > TAIL_BITS) ? -1 : nn;
      atomicOr(&g.ctrl->halt, 1);
    }
    return;
  }
  auto hole = [&](int k) -> int {
    const int w = seg_find(s_hpre, NW, k);
    return d.midx[base + w * stride + (k - s_hpre[w])];
  };
  // number of holes below nn (holes are ascending)
  int hlo = H;
  if (I < H) {
    int c = 0;
    for (int k = threadIdx.x; k < H; k += NT) c += hole(k) < nn;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0) s_wtot[threadIdx.x >> 6] = c;
    __syncthreads();
    hlo = 0;
    for (int w = 0; w < NT / 64; w++) hlo += s_wtot[w];
    __syncthreads();
  }
  auto target = [&](int k) -> int { return k < hlo ? hole(k) : n + (k - H); };
  // pass 2: incoming movers into the holes / appended slots
  for (int r = threadIdx.x; r < I; r += NT) copy_from_mover(d, s_inc[r], base + target(r), dims);
  // pass 3: tail particles into the remaining holes
  if (I < H) {
    const int tail = n - nn;  // = H - I
    for (int i = threadIdx.x; i < (tail + 31) / 32; i += NT) s_bits[i] = 0u;
    __syncthreads();
    for (int k = hlo + threadIdx.x; k < H; k += NT) {
      const int t = hole(k) - nn;
      atomicOr(&s_bits[t >> 5], 1u << (t & 31));
    }
    __syncthreads();
    int r1 = 0;
    for (int i0 = 0; i0 < tail; i0 += NT) {
      const int i = i0 + threadIdx.x;
      const bool f = i < tail && !((s_bits[i >> 5] >> (i & 31)) & 1u);
      int tot;
      const int r = block_rank<NT>(f, s_wtot, tot);
      if (f) copy_particle(d, base + nn + i, base + target(I + r1 + r), dims);
      r1 += tot;
    }
  }
  if (threadIdx.x == 0) {
    d.cnt[tile] = nn;
    d.inc[tile] = I;
  }
}

// ------------------------------------------------------------------------------------
// Device re-layout after a step whose fill overflowed a tile: one workgroup per tile
// copies the tile's particle slots [0, cnt) and its pending mover lists (the valid
// entries of each wave's region) from the old arrays to the same relative slots of the
// tile in the new, larger arrays.  Counts, mover counts and strides are per tile and stay
// valid; the fill of the overflowed tiles then runs again on the new layout.
// ------------------------------------------------------------------------------------
struct RegrowArrays {
  uint32_t* a[19];  // particle arrays (cell x y z ux uy uz q pid), mover arrays (mcell mx
                    // my mz mux muy muz mq mpid midx)
};
__global__ __launch_bounds__(256) void regrow_move(const SpecDesc* __restrict__ descs, int sid,
                                                   RegrowArrays src, RegrowArrays dst,
                                                   const int* __restrict__ new_off) {
  const SpecDesc& d = descs[sid];
  const int tile = blockIdx.x;
  const int o0 = d.off[tile], o1 = new_off[tile];
  const int n = d.cnt[tile];
  const int NW = d.waves, stride = d.mv_stride[tile];
  const bool three = d.g.dims == 3;
  for (int a = 0; a < 9; a++) {
    if (a == 3 && !three) continue;
    const uint32_t* S = src.a[a] + o0;
    uint32_t* D = dst.a[a] + o1;
    for (int i = threadIdx.x; i < n; i += 256) D[i] = S[i];
  }
  for (int w = 0; w < NW; w++) {
    const int m = d.mv_cnt[tile * NW + w];
    for (int a = 9; a < 19; a++) {
      if (a == 12 && !three) continue;
      const uint32_t* S = src.a[a] + o0 + w * stride;
      uint32_t* D = dst.a[a] + o1 + w * stride;
      for (int i = threadIdx.x; i < m; i += 256) D[i] = S[i];
    }
  }
}

// ------------------------------------------------------------------------------------
// Re-interleave the particles of each tile: slot order (rank within the cell, cell), the
// order the loader creates, so that the 64 lanes of a wave hold 64 different, consecutive
// cells and their LDS field reads and current updates spread over the banks.  The fill
// perturbs that order step by step; this restores it every `period` steps (after the
// fill of steps period-1, 2 period-1, ...).  Ranks follow the slot order, so the result
// is deterministic; the current is an integer sum and no arithmetic depends on the order
// (only the FP64 energy partials, at round-off).  Each particle array of the tile is
// permuted through LDS (coalesced read, LDS scatter, coalesced write back).  A step whose
// fill overflowed is skipped (the host re-lays those particles out), as are tiles with
// more than MAXN particles or RMAX particles in one cell.
// ------------------------------------------------------------------------------------
template <int NT, int RMAX, int MAXC, int MAXN>
__global__ __launch_bounds__(NT) void tile_reorder(const SpecDesc* __restrict__ descs,
                                                   const int2* __restrict__ work, int period) {
  constexpr int MAXW = RMAX * MAXC / 32;
  __shared__ int s_cnt[MAXC];
  __shared__ unsigned int s_bits[MAXW];
  __shared__ int s_wpre[MAXW];
  __shared__ int s_key[MAXN];
  __shared__ uint32_t s_tmp[MAXN];
  __shared__ int s_tot[NT / 64];
  __shared__ int s_abort;
  const int2 wk = work[blockIdx.x];
  const SpecDesc& d = descs[wk.x];
  const GridDesc& g = d.g;
  if (period <= 0 || g.ctrl->halt != 0 || g.ctrl->steps % period != period - 1) return;
  const int tile = wk.y;
  const int base = d.off[tile], n = d.cnt[tile];
  const int dims = g.dims;
  const int C = g.tx * g.ty * (dims == 3 ? g.tz : 1);
  if (n < 2 || n > MAXN || C > MAXC) return;
  const int cx0 = (tile % g.ntx) * g.tx, cy0 = ((tile / g.ntx) % g.nty) * g.ty;
  const int cz0 = dims == 3 ? (tile / (g.ntx * g.nty)) * g.tz : 0;
  for (int c = threadIdx.x; c < C; c += NT) s_cnt[c] = 0;
  if (threadIdx.x == 0) s_abort = 0;
  // the tile's packed cells, coalesced, into s_key (each entry is replaced by its key)
  for (int p = threadIdx.x; p < n; p += NT) s_key[p] = (int)d.cell[base + p];
  __syncthreads();
  int nbits = 0;  // bits of a local cell index
  while ((1 << nbits) < C) nbits++;
  if (threadIdx.x < 64) {  // one wave: ranks within the cell in slot order
    const int lane = threadIdx.x;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int c0 = 0; c0 < n; c0 += 64) {
      const int p = c0 + lane;
      int lc = -1 - lane;  // distinct for the lanes past the end
      if (p < n) {
        int ix, iy, iz = 0;
        if (dims == 2)
          unpack2((uint32_t)s_key[p], ix, iy);
        else
          unpack3((uint32_t)s_key[p], ix, iy, iz);
        const int lx = ix - cx0, ly = iy - cy0, lz = iz - cz0;
        if (lx < 0 || lx >= g.tx || ly < 0 || ly >= g.ty || lz < 0 ||
            lz >= (dims == 3 ? g.tz : 1))
          s_abort = 1;
        else
          lc = cell_key(lx, ly, lz, dims, g.tx, g.ty, dims == 3 ? g.tz : 1);
      }
      // lanes holding the same cell: one ballot per bit of the cell index
      unsigned long long same = __ballot(lc >= 0);
      for (int b = 0; b < nbits; b++) {
        const unsigned long long m = __ballot(lc >= 0 && ((lc >> b) & 1));
        same &= ((lc >> b) & 1) ? m : ~m;
      }
      const int before = __popcll(same & below);
      const int after = __popcll(same & ~below) - 1;
      if (lc >= 0) {
        const int r = s_cnt[lc] + before;
        if (r >= RMAX) s_abort = 1;
        s_key[p] = r * C + lc;
      }
      __builtin_amdgcn_wave_barrier();
      if (lc >= 0 && after == 0) s_cnt[lc] += before + 1;
      __builtin_amdgcn_wave_barrier();
    }
  }
  __syncthreads();
  if (s_abort) return;
  // occupancy of the keys (rank, cell): occupied iff the cell holds more than rank
  const int nw = (RMAX * C + 31) / 32;
  for (int w = threadIdx.x; w < nw; w += NT) {
    unsigned int m = 0u;
    int r = (32 * w) / C, c = (32 * w) % C;  // key 32 w = r C + c, walked bit by bit
    for (int b = 0; b < 32; b++) {
      if (r < RMAX && s_cnt[c] > r) m |= 1u << b;
      if (++c == C) {
        c = 0;
        r++;
      }
    }
    s_bits[w] = m;
  }
  __syncthreads();
  // exclusive prefix of the word populations: each thread owns consecutive words
  constexpr int PER = (MAXW + NT - 1) / NT;
  const int w0 = threadIdx.x * PER;
  int own = 0;
  for (int k = 0; k < PER; k++)
    if (w0 + k < nw) own += __popc(s_bits[w0 + k]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int inc = own;
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) s_tot[wave] = inc;
  __syncthreads();
  int run = inc - own;
  for (int w = 0; w < wave; w++) run += s_tot[w];
  for (int k = 0; k < PER; k++)
    if (w0 + k < nw) {
      s_wpre[w0 + k] = run;
      run += __popc(s_bits[w0 + k]);
    }
  __syncthreads();
  // new slot of every particle, in place of its key
  for (int p = threadIdx.x; p < n; p += NT) {
    const int k = s_key[p];
    s_key[p] = s_wpre[k >> 5] + __popc(s_bits[k >> 5] & ((1u << (k & 31)) - 1u));
  }
  __syncthreads();
  // permute each array through LDS: coalesced 16-byte reads of 4 consecutive slots per
  // thread (tile bases are multiples of 64 slots), LDS scatter, coalesced writes back
  constexpr int PT = MAXN / (4 * NT);  // 4-slot groups per thread
  int dst[PT][4];
#pragma unroll
  for (int j = 0; j < PT; j++)
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int p = 4 * (threadIdx.x + NT * j) + e;
      dst[j][e] = p < n ? s_key[p] : -1;
    }
  uint32_t* arr[9] = {d.cell, (uint32_t*)d.x, (uint32_t*)d.y, (uint32_t*)d.z, (uint32_t*)d.ux,
                      (uint32_t*)d.uy, (uint32_t*)d.uz, (uint32_t*)d.q, d.pid};
#pragma unroll
  for (int a = 0; a < 9; a++) {
    if (a == 3 && dims != 3) continue;
    uint32_t* A = arr[a] + base;
#pragma unroll
    for (int j = 0; j < PT; j++) {
      const int q = 4 * (threadIdx.x + NT * j);
      if (q + 3 < n) {
        const uint4 v = *(const uint4*)(A + q);
        s_tmp[dst[j][0]] = v.x;
        s_tmp[dst[j][1]] = v.y;
        s_tmp[dst[j][2]] = v.z;
        s_tmp[dst[j][3]] = v.w;
      } else {
#pragma unroll
        for (int e = 0; e < 4; e++)
          if (q + e < n) s_tmp[dst[j][e]] = A[q + e];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PT; j++) {
      const int q = 4 * (threadIdx.x + NT * j);
      if (q + 3 < n) {
        *(uint4*)(A + q) = make_uint4(s_tmp[q], s_tmp[q + 1], s_tmp[q + 2], s_tmp[q + 3]);
      } else {
#pragma unroll
        for (int e = 0; e < 4; e++)
          if (q + e < n) A[q + e] = s_tmp[q + e];
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------
// explicit instantiations
// ------------------------------------------------------------------------------------
#define OSR_INST(O, P)                                                                    \
  template __global__ void push_deposit_2d<O, P, 256>(const SpecDesc*, const int2*);      \
  template __global__ void push_deposit_3d<O, P, 256>(const SpecDesc*, const int2*);
OSR_INST(1, 0)
OSR_INST(2, 0)
OSR_INST(3, 0)
OSR_INST(1, 1)
OSR_INST(2, 1)
OSR_INST(3, 1)
template __global__ void tile_fill<256, 4096, 32768>(const SpecDesc*, const int2*);
template __global__ void tile_reorder<256, 64, 512, 4096>(const SpecDesc*, const int2*, int);

}  // namespace osr
