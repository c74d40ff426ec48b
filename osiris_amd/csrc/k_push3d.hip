// k_push3d.hip -- fused 3D push + deposit with compile-time tile shape.
//
// One workgroup of 4 waves per (species, tile).  Compared with the generic 3D kernel in
// k_tiles.hip:
//  * TX/TY/TZ are template parameters, so the LDS pitches of the staged field and current
//    tiles are constants and every stencil offset is an immediate (the generic kernel
//    adds runtime row and plane pitches for each of the 4 x 4 lines of a cubic stencil,
//    about 400 VALU instructions per 64 particles in the gather alone);
//  * particle and mover arrays are addressed through global-address-space pointers with
//    scalar bases and 32-bit offsets;
//  * the particle loads of the next iteration are issued before the current particle is
//    processed.
// The arithmetic (gather line/plane order, pusher, position advance, split, deposit) is
// the generic kernel's, operation for operation, so the results are bit-identical.
#include "osr_device.hpp"
#include "osr_types.hpp"

namespace osr {

typedef __attribute__((address_space(3))) const float lcf32;

// LDS pointer to element `elem` of the LDS array `base`, made opaque to the compiler: the
// loads of one stencil line then share this one base register and carry their small
// constant offsets in the instruction (ds_read2_b32 offsets are 8 bits of dwords).  Left
// visible, the array's absolute LDS address is folded into every constant, which no
// longer fits, and a paired load costs one more address add.  The empty asm emits no
// instruction.
OSR_DEV lcf32* lds_line(const float* base, int elem) {
  unsigned a = (unsigned)(size_t)(lcf32*)base + 4u * (unsigned)elem;
  asm("" : "+v"(a));
  return (lcf32*)(size_t)a;
}

template <int ORDER, int TX, int TY, int TZ>
struct Tile3 {
  using O = Ord<ORDER>;
  static constexpr int GX = TX + O::GLO + O::GHI, GY = TY + O::GLO + O::GHI,
                       GZ = TZ + O::GLO + O::GHI;
  static constexpr int JX = TX + O::JLO + O::JHI, JY = TY + O::JLO + O::JHI,
                       JZ = TZ + O::JLO + O::JHI;
  // LDS pitches of the staged fields: rows of RS floats, planes of GXY floats.  For 8x8
  // tiles 12 and 144 (padded for orders 1 and 2): the 4x4x2 cell block of a half wave
  // (cell_key) then starts its stencils in 32 distinct banks.
  static constexpr int RS = (TX == 8 && GX <= 12) ? 12 : GX;
  static constexpr int GYP = (TY == 8 && GY <= 12) ? 12 : GY;
  static constexpr int GXY = RS * GYP, GN = GXY * GZ;
  static constexpr int JXY = JX * JY, JN = JXY * JZ;
  static constexpr int NT = 256, NW = NT / 64, QC = kVpQueue3;
  static constexpr size_t lds = 6 * GN * 4 + 3 * JN * 8 + NW * QC * 8 * 4;
};

template <int ORDER, int PUSHER, int TX, int TY, int TZ>
__global__ __launch_bounds__(256, 2) void push_deposit_3d_t(const SpecDesc* __restrict__ descs,
                                                            const int2* __restrict__ work) {
  using O = Ord<ORDER>;
  using T = Tile3<ORDER, TX, TY, TZ>;
  constexpr int NP = O::NP;
  constexpr int GX = T::GX, GY = T::GY, GZ = T::GZ, RS = T::RS, GXY = T::GXY, GN = T::GN;
  constexpr int JX = T::JX, JY = T::JY, JXY = T::JXY, JN = T::JN;
  constexpr int NT = T::NT, NW = T::NW, QC = T::QC;
  __shared__ float sF[6 * GN];
  __shared__ unsigned long long sJ[3 * JN];
  __shared__ float sQall[NW * 7 * QC];
  __shared__ int sQnall[NW * QC];
  __shared__ double s_ene[NW];

  const int2 wk = work[blockIdx.x];
  const SpecDesc& d = descs[wk.x];
  const GridDesc& g = d.g;
  if (g.ctrl->halt != 0) return;
  const int tile = wk.y;
  const int ttx = tile % g.ntx, tty = (tile / g.ntx) % g.nty, ttz = tile / (g.ntx * g.nty);
  const int cx0 = ttx * TX, cy0 = tty * TY, cz0 = ttz * TZ;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const long long nc = g.ncell;

  {
    const gf32* __restrict__ E = (const gf32*)g.E;
    const gf32* __restrict__ B = (const gf32*)g.B;
    for (int i = threadIdx.x; i < GX * GY * GZ; i += NT) {
      const int lx = i % GX, ly = (i / GX) % GY, lz = i / (GX * GY);
      const int gx = wrapi(cx0 + lx - O::GLO, nx), gy = wrapi(cy0 + ly - O::GLO, ny),
                gz = wrapi(cz0 + lz - O::GLO, nz);
      const long long gi = gx + (long long)nx * (gy + (long long)ny * gz);
      const int li = lz * GXY + ly * RS + lx;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        sF[c * GN + li] = E[c * nc + gi];
        sF[(3 + c) * GN + li] = B[c * nc + gi];
      }
    }
  }
  for (int i = threadIdx.x; i < 3 * JN; i += NT) sJ[i] = 0ull;
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int base = __builtin_amdgcn_readfirstlane(d.off[tile]);
  const int n = __builtin_amdgcn_readfirstlane(d.cnt[tile]);
  const int per_wave = ((n + NT - 1) / NT) * 64;
  const int w0 = wave * per_wave;
  const int w1 = min(w0 + per_wave, n);
  const float tem = d.tem;
  const float dtdx1 = g.dt_dx[0], dtdx2 = g.dt_dx[1], dtdx3 = g.dt_dx[2];
  const float jn1 = g.jnorm[0], jn2 = g.jnorm[1], jn3 = g.jnorm[2];
  const double jscale = g.jscale;
  float* const Q = sQall + wave * 7 * QC;
  int* const QN = sQnall + wave * QC;
  int qhead = 0, qlen = 0;

  const gu32* __restrict__ Pcell = (const gu32*)d.cell + base;
  const gf32* __restrict__ Px = (const gf32*)d.x + base;
  const gf32* __restrict__ Py = (const gf32*)d.y + base;
  const gf32* __restrict__ Pz = (const gf32*)d.z + base;
  const gf32* __restrict__ Pux = (const gf32*)d.ux + base;
  const gf32* __restrict__ Puy = (const gf32*)d.uy + base;
  const gf32* __restrict__ Puz = (const gf32*)d.uz + base;
  const gf32* __restrict__ Pq = (const gf32*)d.q + base;
  const gu32* __restrict__ Ppid = (const gu32*)d.pid + base;
  gu32* __restrict__ Wcell = (gu32*)d.cell + base;
  gf32* __restrict__ Wx = (gf32*)d.x + base;
  gf32* __restrict__ Wy = (gf32*)d.y + base;
  gf32* __restrict__ Wz = (gf32*)d.z + base;
  gf32* __restrict__ Wux = (gf32*)d.ux + base;
  gf32* __restrict__ Wuy = (gf32*)d.uy + base;
  gf32* __restrict__ Wuz = (gf32*)d.uz + base;
  gu32* __restrict__ Mcell = (gu32*)d.mcell + base;
  gf32* __restrict__ Mx = (gf32*)d.mx + base;
  gf32* __restrict__ My = (gf32*)d.my + base;
  gf32* __restrict__ Mz = (gf32*)d.mz + base;
  gf32* __restrict__ Mux = (gf32*)d.mux + base;
  gf32* __restrict__ Muy = (gf32*)d.muy + base;
  gf32* __restrict__ Muz = (gf32*)d.muz + base;
  gf32* __restrict__ Mq = (gf32*)d.mq + base;
  gu32* __restrict__ Mpid = (gu32*)d.mpid + base;
  gi32* __restrict__ Midx = (gi32*)d.midx + base;

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
  uint32_t f_cell = 0, f_pid = 0;
  float f_x = 0.f, f_y = 0.f, f_z = 0.f, f_u1 = 0.f, f_u2 = 0.f, f_u3 = 0.f, f_q = 0.f;
  auto prefetch = [&](int pf) {
    if (pf < w1) {
      const unsigned upf = (unsigned)pf;
      f_cell = *el(Pcell, upf);
      f_x = *el(Px, upf);
      f_y = *el(Py, upf);
      f_z = *el(Pz, upf);
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
    const unsigned up = (unsigned)p;
    bool act = p < w1;
    float x = 0.f, y = 0.f, z = 0.f, u1 = 0.f, u2 = 0.f, u3 = 0.f, q = 0.f;
    const uint32_t cell = f_cell, pid = f_pid;
    if (act) {
      x = f_x;
      y = f_y;
      z = f_z;
      u1 = f_u1;
      u2 = f_u2;
      u3 = f_u3;
      q = f_q;
    }
    prefetch(p + 64);
    int lx = 0, ly = 0, lz = 0;
    if (act) {
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
    // stencil origins (full and half-shifted) as LDS element offsets
    const int g0 = (lz + O::GLO + O::LP) * GXY + (ly + O::GLO + O::LP) * RS + lx + O::GLO + O::LP;
    const int hx = x < 0.f ? 1 : 0, hy = y < 0.f ? RS : 0, hz = z < 0.f ? GXY : 0;
    const int oE1 = g0 - hx, oE2 = GN + g0 - hy, oE3 = 2 * GN + g0 - hz;
    const int oB1 = 3 * GN + g0 - hy - hz, oB2 = 4 * GN + g0 - hx - hz, oB3 = 5 * GN + g0 - hx - hy;
    // GET_EMF_3D line/plane order (os-emf-interpolate.f03:697-773): fma along k1, then k2,
    // then k3
    float e1 = 0.f, e2 = 0.f, e3 = 0.f;
#pragma unroll
    for (int k3 = 0; k3 < NP; k3++) {
      float p1 = 0.f, p2 = 0.f, p3 = 0.f;
      lcf32* L1 = lds_line(sF, oE1 + k3 * GXY);
      lcf32* L2 = lds_line(sF, oE2 + k3 * GXY);
      lcf32* L3 = lds_line(sF, oE3 + k3 * GXY);
#pragma unroll
      for (int k2 = 0; k2 < NP; k2++) {
        float l1 = 0.f, l2 = 0.f, l3 = 0.f;
#pragma unroll
        for (int k1 = 0; k1 < NP; k1++) {
          const int o = k2 * RS + k1;
          l1 = fma_(L1[o], wxh[k1], l1);
          l2 = fma_(L2[o], wx[k1], l2);
          l3 = fma_(L3[o], wx[k1], l3);
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
      lcf32* L1 = lds_line(sF, oB1 + k3 * GXY);
      lcf32* L2 = lds_line(sF, oB2 + k3 * GXY);
      lcf32* L3 = lds_line(sF, oB3 + k3 * GXY);
#pragma unroll
      for (int k2 = 0; k2 < NP; k2++) {
        float l1 = 0.f, l2 = 0.f, l3 = 0.f;
#pragma unroll
        for (int k1 = 0; k1 < NP; k1++) {
          const int o = k2 * RS + k1;
          l1 = fma_(L1[o], wx[k1], l1);
          l2 = fma_(L2[o], wxh[k1], l2);
          l3 = fma_(L3[o], wxh[k1], l3);
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
    VP3 vp[4];
    const int nvp = act ? split3d(x, y, z, x1, y1, z1, dix, diy, diz, q, vp) : 0;
    const int nb0 = (lz + O::LP + O::JLO) * JXY + (ly + O::LP + O::JLO) * JX + lx + O::LP + O::JLO;
    if (act)
      dep_vp3<ORDER>(sJ, JN, wx, wy, wz, vp[0].x0, vp[0].y0, vp[0].z0, vp[0].x1, vp[0].y1,
                     vp[0].z1, vp[0].q, nb0, JX, JXY, jn1, jn2, jn3, jscale);
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll 1
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
      *el(Wcell, up) = ncell;
      *el(Wx, up) = xs;
      *el(Wy, up) = ys;
      *el(Wz, up) = zs;
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
      *el(Mz, slot) = zs;
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
    double t = 0.0;
    for (int w = 0; w < NW; w++) t += s_ene[w];
    d.epart[tile] = t;
  }
  gu64* __restrict__ GJ = (gu64*)g.J;
  for (int i = threadIdx.x; i < JN; i += NT) {
    const int lx = i % JX, ly = (i / JX) % JY, lz = i / JXY;
    const int gx = wrapi(cx0 + lx - O::JLO, nx), gy = wrapi(cy0 + ly - O::JLO, ny),
              gz = wrapi(cz0 + lz - O::JLO, nz);
    const long long gi = gx + (long long)nx * (gy + (long long)ny * gz);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const unsigned long long v = sJ[c * JN + i];
      if (v) __hip_atomic_fetch_add(&GJ[c * nc + gi], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

#define OSR_INST3(O, P, X, Y, Z) \
  template __global__ void push_deposit_3d_t<O, P, X, Y, Z>(const SpecDesc*, const int2*);
#define OSR_INST3T(X, Y, Z)                                                   \
  OSR_INST3(1, 0, X, Y, Z) OSR_INST3(2, 0, X, Y, Z) OSR_INST3(3, 0, X, Y, Z) \
  OSR_INST3(1, 1, X, Y, Z) OSR_INST3(2, 1, X, Y, Z) OSR_INST3(3, 1, X, Y, Z)
OSR_INST3T(8, 8, 4)
#undef OSR_INST3T
#undef OSR_INST3

}  // namespace osr
