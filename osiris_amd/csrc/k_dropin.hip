// k_dropin.hip -- kernels behind the stateless drop-in entry points.
//
// They operate directly on the reference's array layout: particles ix(D,N) (1-based
// cells), x(D,N), u(3,N), q(N) interleaved as in t_species (spec/os-spec-define.f03:
// 391-412), fields f(3, 1-gc:nx+gc, ...) with the guard cells of the caller
// (vdf/os-vdf-define.f03:35-116).  No tiles, no sorting: one thread per particle,
// gather from global memory, deposit with global int64 fixed-point atomics (reproducible
// run to run), then one pass adds the converted current to the caller's float array.
// Same arithmetic as FNAME(vadvance_deposit_{2d,3d}_sN) in simd/os-spec-push-avx2.c.
#include "osr_device.hpp"
#include "osr_types.hpp"

namespace osr {

struct DropinArgs {
  int* ix;
  float* x;
  float* u;
  const float* q;
  int np;
  const float* e;
  const float* b;
  int emf_ld[3];   // array extents in cells (guards included)
  int emf_off[3];  // lower guard cells
  unsigned long long* jfix;
  int j_ld[3];
  int j_off[3];
  float tem;
  float dt_dx[3];
  float jnorm[3];
  double jscale;
  double* epart;  // per block
};

template <int ORDER, int PUSHER, int DIMS>
__global__ __launch_bounds__(256) void dropin_push(DropinArgs a) {
  using O = Ord<ORDER>;
  constexpr int NP = O::NP;
  __shared__ double s_ene[4];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double ene = 0.0;
  if (i < a.np) {
    // field index of Fortran cell (fi,fj,fk) component c: ((fi-1+off0) + ld0*(...))*3 + c
    auto fidx = [&](int fi, int fj, int fk) -> long long {
      long long o = (long long)(fi - 1 + a.emf_off[0]);
      o += (long long)a.emf_ld[0] * (fj - 1 + a.emf_off[1]);
      if (DIMS == 3) o += (long long)a.emf_ld[0] * a.emf_ld[1] * (fk - 1 + a.emf_off[2]);
      return 3 * o;
    };
    auto jidx = [&](int fi, int fj, int fk) -> long long {
      long long o = (long long)(fi - 1 + a.j_off[0]);
      o += (long long)a.j_ld[0] * (fj - 1 + a.j_off[1]);
      if (DIMS == 3) o += (long long)a.j_ld[0] * a.j_ld[1] * (fk - 1 + a.j_off[2]);
      return 3 * o;
    };
    const int i1 = a.ix[DIMS * i], i2 = a.ix[DIMS * i + 1];
    const int i3 = DIMS == 3 ? a.ix[DIMS * i + 2] : 1;
    const float x = a.x[DIMS * i], y = a.x[DIMS * i + 1];
    const float z = DIMS == 3 ? a.x[DIMS * i + 2] : 0.f;
    float u1 = a.u[3 * i], u2 = a.u[3 * i + 1], u3 = a.u[3 * i + 2];
    const float q = a.q[i];
    float wx[NP], wy[NP], wz[NP], wxh[NP], wyh[NP], wzh[NP];
    spline<ORDER>(x, wx);
    spline<ORDER>(y, wy);
    splineh<ORDER>(x, x < 0.f ? 1.f : 0.f, wxh);
    splineh<ORDER>(y, y < 0.f ? 1.f : 0.f, wyh);
    if (DIMS == 3) {
      spline<ORDER>(z, wz);
      splineh<ORDER>(z, z < 0.f ? 1.f : 0.f, wzh);
    }
    const int a1 = i1 + O::LP, a2 = i2 + O::LP, a3 = i3 + O::LP;
    const int h1 = a1 - (x < 0.f), h2 = a2 - (y < 0.f), h3 = a3 - (z < 0.f);
    float e1 = 0.f, e2 = 0.f, e3 = 0.f, b1 = 0.f, b2 = 0.f, b3 = 0.f;
    if (DIMS == 2) {
#pragma unroll
      for (int k2 = 0; k2 < NP; k2++) {
        float l1 = 0.f, l2 = 0.f, l3 = 0.f;
#pragma unroll
        for (int k1 = 0; k1 < NP; k1++) {
          l1 = fma_(a.e[fidx(h1 + k1, a2 + k2, 1) + 0], wxh[k1], l1);
          l2 = fma_(a.e[fidx(a1 + k1, h2 + k2, 1) + 1], wx[k1], l2);
          l3 = fma_(a.e[fidx(a1 + k1, a2 + k2, 1) + 2], wx[k1], l3);
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
          l1 = fma_(a.b[fidx(a1 + k1, h2 + k2, 1) + 0], wx[k1], l1);
          l2 = fma_(a.b[fidx(h1 + k1, a2 + k2, 1) + 1], wxh[k1], l2);
          l3 = fma_(a.b[fidx(h1 + k1, h2 + k2, 1) + 2], wxh[k1], l3);
        }
        b1 = fma_(l1, wyh[k2], b1);
        b2 = fma_(l2, wy[k2], b2);
        b3 = fma_(l3, wyh[k2], b3);
      }
    } else {
#pragma unroll
      for (int k3 = 0; k3 < NP; k3++) {
        float p1 = 0.f, p2 = 0.f, p3 = 0.f, r1 = 0.f, r2 = 0.f, r3 = 0.f;
#pragma unroll
        for (int k2 = 0; k2 < NP; k2++) {
          float l1 = 0.f, l2 = 0.f, l3 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f;
#pragma unroll
          for (int k1 = 0; k1 < NP; k1++) {
            l1 = fma_(a.e[fidx(h1 + k1, a2 + k2, a3 + k3) + 0], wxh[k1], l1);
            l2 = fma_(a.e[fidx(a1 + k1, h2 + k2, a3 + k3) + 1], wx[k1], l2);
            l3 = fma_(a.e[fidx(a1 + k1, a2 + k2, h3 + k3) + 2], wx[k1], l3);
            m1 = fma_(a.b[fidx(a1 + k1, h2 + k2, h3 + k3) + 0], wx[k1], m1);
            m2 = fma_(a.b[fidx(h1 + k1, a2 + k2, h3 + k3) + 1], wxh[k1], m2);
            m3 = fma_(a.b[fidx(h1 + k1, h2 + k2, a3 + k3) + 2], wxh[k1], m3);
          }
          p1 = fma_(l1, wy[k2], p1);
          p2 = fma_(l2, wyh[k2], p2);
          p3 = fma_(l3, wy[k2], p3);
          r1 = fma_(m1, wyh[k2], r1);
          r2 = fma_(m2, wy[k2], r2);
          r3 = fma_(m3, wyh[k2], r3);
        }
        e1 = fma_(p1, wz[k3], e1);
        e2 = fma_(p2, wz[k3], e2);
        e3 = fma_(p3, wzh[k3], e3);
        b1 = fma_(r1, wzh[k3], b1);
        b2 = fma_(r2, wzh[k3], b2);
        b3 = fma_(r3, wz[k3], b3);
      }
    }
    const float ke = push<PUSHER>(a.tem, e1, e2, e3, b1, b2, b3, u1, u2, u3);
    ene = (double)q * (double)ke;
    a.u[3 * i] = u1;
    a.u[3 * i + 1] = u2;
    a.u[3 * i + 2] = u3;
    const float rg = rgamma(u1, u2, u3);
    const float v1 = u1 * rg, v2 = u2 * rg, v3 = u3 * rg;
    const float x1 = fma_(v1, a.dt_dx[0], x), y1 = fma_(v2, a.dt_dx[1], y);
    const float z1 = DIMS == 3 ? fma_(v3, a.dt_dx[2], z) : 0.f;
    const float t1 = trim(x1), t2 = trim(y1), t3 = DIMS == 3 ? trim(z1) : 0.f;
    a.x[DIMS * i] = x1 - t1;
    a.x[DIMS * i + 1] = y1 - t2;
    a.ix[DIMS * i] = i1 + (int)t1;
    a.ix[DIMS * i + 1] = i2 + (int)t2;
    if (DIMS == 3) {
      a.x[DIMS * i + 2] = z1 - t3;
      a.ix[DIMS * i + 2] = i3 + (int)t3;
    }
    const double js = a.jscale;
    if (DIMS == 2) {
      VP2 vp[3];
      const int nvp = split2d(x, y, x1, y1, (int)t1, (int)t2, q, v3, vp);
      for (int v = 0; v < nvp; v++) {
        const VP2 V = pick(vp, v);
        float S0x[NP], S1x[NP], S0y[NP], S1y[NP], wl1[NP], wl2[NP];
        spline<ORDER>(V.x0, S0x);
        spline<ORDER>(V.x1, S1x);
        spline<ORDER>(V.y0, S0y);
        spline<ORDER>(V.y1, S1y);
        wl<ORDER>(V.q * a.jnorm[0], V.x0, V.x1, wl1);
        wl<ORDER>(V.q * a.jnorm[1], V.y0, V.y1, wl2);
        const float qvz = (V.q * V.vz) * (1.0f / 3.0f);
#pragma unroll
        for (int k2 = 0; k2 < NP; k2++) {
          const float wp1 = S0y[k2] + S1y[k2];
#pragma unroll
          for (int k1 = 0; k1 < NP; k1++) {
            const long long o = jidx(i1 + V.di + O::LP + k1, i2 + V.dj + O::LP + k2, 1);
            const float wp2 = S0x[k1] + S1x[k1];
            float tt = fma_(S1y[k2], S1x[k1], S0x[k1] * S0y[k2]);
            const float t2b = fma_(S0y[k2], S1x[k1], S0x[k1] * S1y[k2]);
            tt = fma_(t2b, 0.5f, tt);
            if (k1 < NP - 1) atomicAdd(&a.jfix[o], (unsigned long long)to_fix(wl1[k1] * wp1, js));
            if (k2 < NP - 1) atomicAdd(&a.jfix[o + 1], (unsigned long long)to_fix(wl2[k2] * wp2, js));
            atomicAdd(&a.jfix[o + 2], (unsigned long long)to_fix(qvz * tt, js));
          }
        }
      }
    } else {
      VP3 vp[4];
      const int nvp = split3d(x, y, z, x1, y1, z1, (int)t1, (int)t2, (int)t3, q, vp);
      for (int v = 0; v < nvp; v++) {
        const VP3 V = pick(vp, v);
        float S0x[NP], S1x[NP], S0y[NP], S1y[NP], S0z[NP], S1z[NP], wl1[NP], wl2[NP], wl3[NP];
        spline<ORDER>(V.x0, S0x);
        spline<ORDER>(V.x1, S1x);
        spline<ORDER>(V.y0, S0y);
        spline<ORDER>(V.y1, S1y);
        spline<ORDER>(V.z0, S0z);
        spline<ORDER>(V.z1, S1z);
        wl<ORDER>(V.q * a.jnorm[0], V.x0, V.x1, wl1);
        wl<ORDER>(V.q * a.jnorm[1], V.y0, V.y1, wl2);
        wl<ORDER>(V.q * a.jnorm[2], V.z0, V.z1, wl3);
#pragma unroll
        for (int k3 = 0; k3 < NP; k3++)
#pragma unroll
          for (int k2 = 0; k2 < NP; k2++) {
            const float wp1 = fma_(0.5f, fma_(S1y[k2], S0z[k3], S0y[k2] * S1z[k3]),
                                   fma_(S1y[k2], S1z[k3], S0y[k2] * S0z[k3]));
#pragma unroll
            for (int k1 = 0; k1 < NP; k1++) {
              const long long o = jidx(i1 + V.di + O::LP + k1, i2 + V.dj + O::LP + k2,
                                       i3 + V.dk + O::LP + k3);
              const float wp2 = fma_(0.5f, fma_(S1x[k1], S0z[k3], S0x[k1] * S1z[k3]),
                                     fma_(S1x[k1], S1z[k3], S0x[k1] * S0z[k3]));
              const float wp3 = fma_(0.5f, fma_(S1x[k1], S0y[k2], S0x[k1] * S1y[k2]),
                                     fma_(S1x[k1], S1y[k2], S0x[k1] * S0y[k2]));
              if (k1 < NP - 1) atomicAdd(&a.jfix[o], (unsigned long long)to_fix(wl1[k1] * wp1, js));
              if (k2 < NP - 1) atomicAdd(&a.jfix[o + 1], (unsigned long long)to_fix(wl2[k2] * wp2, js));
              if (k3 < NP - 1) atomicAdd(&a.jfix[o + 2], (unsigned long long)to_fix(wl3[k3] * wp3, js));
            }
          }
      }
    }
  }
  ene = wave_sum(ene);
  if ((threadIdx.x & 63) == 0) s_ene[threadIdx.x >> 6] = ene;
  __syncthreads();
  if (threadIdx.x == 0) a.epart[blockIdx.x] = ((s_ene[0] + s_ene[1]) + s_ene[2]) + s_ene[3];
}

// j += J_fixed * 2^-shift over the whole guarded array
__global__ __launch_bounds__(256) void dropin_jadd(float* j, const unsigned long long* jfix,
                                                   long long n, double jinv) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) j[i] = j[i] + (float)((double)(long long)jfix[i] * jinv);
}

#define OSR_DINST(O, P)                                       \
  template __global__ void dropin_push<O, P, 2>(DropinArgs); \
  template __global__ void dropin_push<O, P, 3>(DropinArgs);
OSR_DINST(1, 0)
OSR_DINST(2, 0)
OSR_DINST(3, 0)
OSR_DINST(1, 1)
OSR_DINST(2, 1)
OSR_DINST(3, 1)

}  // namespace osr
