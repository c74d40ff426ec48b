// osr_device.hpp -- device-side building blocks of the PIC hot path for gfx950.
//
// Arithmetic follows the reference's AVX2 single precision kernels operation by
// operation (the same fused multiply-adds in the same places, IEEE division and square
// root), so that gather, push and position advance give bit-identical results to
// FNAME(vadvance_deposit_*) of simd/os-spec-push-avx2.c on the same inputs.  The file is
// compiled with -ffp-contract=off: every FMA below is explicit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define OSR_DEV __device__ __forceinline__

namespace osr {

OSR_DEV float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// ------------------------------------------------------------------------------------
// Correctly rounded FP32 reciprocal, quotient and square root for the operands of the
// push path.  Each starts
// from the hardware estimate (v_rcp_f32, v_rsq_f32, accurate to about 1 ulp) and
// corrects it with fused multiply-adds, after Markstein (P. Markstein, "Computation of
// elementary functions on the IBM RISC System/6000 processor", IBM J. Res. Dev. 34,
// 1990; and "IA-64 and Elementary Functions", 2000, ch. 8):
//  * rcp_rn: one Newton step y = r + r (1 - b r) from the estimate r;
//  * div_rn: q = a y, the exact remainder a - b q, and q + (a - b q) y.  With y the
//    correctly rounded 1/b this is the correctly rounded a/b;
//  * sqrt_rn: g = x y and h = y/2 from the estimate y of 1/sqrt(x), one coupled Newton
//    step on both, then g + (x - g g) h.
// The sequences are correctly rounded on the ranges below, checked bit for bit against
// the IEEE operations on the device: rcp_rn for every normal b with a normal reciprocal;
// sqrt_rn for every x >= 2^-100 (below about 2^-102 the remainder underflows); div_rn
// for every pair of mantissas, which stands for every pair of normal operands with
// |a| >= 2^-79 (the remainder then stays normal) whose quotient is normal, and for +0
// numerators (a -0 numerator gives +0).  The push path
// uses them only there: sqrt_rn on 1 + u^2, div_rn on numerators that are tem, 1, 2 or
// +-1/2 - x0 (at least 2^-25 in magnitude, or exactly +0).  The kinetic energy
// u^2 / (gamma + 1), whose numerator can be denormal, and Vay's inner roots keep the
// IEEE operations.  The compiler's IEEE sequences (-fhip-fp32-correctly-rounded-divide-
// sqrt) also scale operands for the denormal and overflow ranges; the shorter forms
// leave that out.
// ------------------------------------------------------------------------------------
OSR_DEV float rcp_rn(float b) {
  const float r = __builtin_amdgcn_rcpf(b);
  return fma_(fma_(-b, r, 1.0f), r, r);
}
OSR_DEV float div_rn(float a, float b) {
  const float y = rcp_rn(b);
  const float q = a * y;
  return fma_(fma_(-b, q, a), y, q);
}
OSR_DEV float sqrt_rn(float x) {
  const float y = __builtin_amdgcn_rsqf(x);
  float g = x * y, h = 0.5f * y;
  const float r = fma_(-g, h, 0.5f);
  g = fma_(g, r, g);
  h = fma_(h, r, h);
  return fma_(fma_(-g, g, x), h, g);
}

// Stencil parameters per interpolation order (emf/os-emf-interpolate.f03:385-461,
// spec/os-spec-current.f03:2495-2533).  LP: lower stencil point of spline/wl.
template <int ORDER> struct Ord {
  static constexpr int NP = ORDER + 1;
  static constexpr int LP = ORDER == 1 ? 0 : -1;
  // field halo needed by the gather (half-shifted stencil reaches one lower)
  static constexpr int GLO = 1 - LP;                // 1, 2, 2
  static constexpr int GHI = LP + ORDER;            // 1, 1, 2
  // current halo touched by the deposit (virtual particle cells are ix-1..ix+1)
  static constexpr int JLO = 1 - LP;                // 1, 2, 2
  static constexpr int JHI = 1 + LP + ORDER;        // 2, 2, 3
};

// ------------------------------------------------------------------------------------
// Splines: simd/splines-avx2.h:26-123 (vspline_s1..s3, vsplineh_s1..s3).
// ------------------------------------------------------------------------------------
template <int ORDER> OSR_DEV void spline(float x, float* w);
template <> OSR_DEV void spline<1>(float x, float* w) {
  w[0] = 0.5f - x;
  w[1] = 0.5f + x;
}
template <> OSR_DEV void spline<2>(float x, float* w) {
  float t0 = 0.5f - x, t1 = 0.5f + x;
  w[0] = (t0 * t0) * 0.5f;
  w[1] = fma_(t0, t1, 0.5f);
  w[2] = (t1 * t1) * 0.5f;
}
template <> OSR_DEV void spline<3>(float x, float* w) {
  const float c1_6 = 1.0f / 6.0f, c2_3 = 2.0f / 3.0f;
  float t0 = 0.5f - x, t1 = 0.5f + x;
  float t2 = t0 * t0, t3 = t1 * t1;
  t0 = t0 * t2;
  t1 = t1 * t3;
  w[0] = c1_6 * t0;
  w[1] = fma_(0.5f, t1, c2_3 - t3);
  w[2] = fma_(0.5f, t0, c2_3 - t2);
  w[3] = c1_6 * t1;
}
// h is 1.0f when x < 0, else 0.0f
template <int ORDER> OSR_DEV void splineh(float x, float h, float* w);
template <> OSR_DEV void splineh<1>(float x, float h, float* w) {
  w[0] = (1.0f - h) - x;
  w[1] = h + x;
}
template <> OSR_DEV void splineh<2>(float x, float h, float* w) {
  float t0 = (1.0f - h) - x, t1 = h + x;
  w[0] = (t0 * t0) * 0.5f;
  w[1] = fma_(t0, t1, 0.5f);
  w[2] = (t1 * t1) * 0.5f;
}
template <> OSR_DEV void splineh<3>(float x, float h, float* w) {
  const float c1_6 = 1.0f / 6.0f, c2_3 = 2.0f / 3.0f;
  float t0 = (1.0f - h) - x, t1 = h + x;
  float t2 = t0 * t0, t3 = t1 * t1;
  t0 = t0 * t2;
  t1 = t1 * t3;
  w[0] = c1_6 * t0;
  w[1] = fma_(0.5f, t1, c2_3 - t3);
  w[2] = fma_(0.5f, t0, c2_3 - t2);
  w[3] = c1_6 * t1;
}

// Longitudinal weights wl[0..NP-1] with wl[NP-1] = 0: simd/os-spec-current-avx2.c:20-64
template <int ORDER> OSR_DEV void wl(float qn, float x0, float x1, float* w);
template <> OSR_DEV void wl<1>(float qn, float x0, float x1, float* w) {
  w[0] = qn * (x1 - x0);
  w[1] = 0.0f;
}
template <> OSR_DEV void wl<2>(float qn, float x0, float x1, float* w) {
  float d = x1 - x0, s1_2 = 0.5f - x0, p1_2 = 0.5f + x0;
  float n = qn * d;
  w[0] = n * fma_(-0.5f, d, s1_2);
  w[1] = n * fma_(0.5f, d, p1_2);
  w[2] = 0.0f;
}
template <> OSR_DEV void wl<3>(float qn, float x0, float x1, float* w) {
  const float c1_3 = 1.0f / 3.0f;
  float d = x1 - x0, s = 0.5f - x0, p = 0.5f + x0, d_3 = c1_3 * d;
  float a = 0.5f * fma_(s, s, -(d * (s - d_3)));
  float b = fma_(-d, x0 + d_3, fma_(-x0, x0, 0.75f));
  float c = 0.5f * fma_(p, p, d * (p + d_3));
  float n = qn * d;
  w[0] = n * a;
  w[1] = n * b;
  w[2] = n * c;
  w[3] = 0.0f;
}

// ------------------------------------------------------------------------------------
// Boris push: simd/os-spec-push-avx2.c:90-146 (VDUDT_BORIS).  e/b are the gathered
// fields; returns the time centred u^2/(gamma+1) (per particle, float, as vene).
// ------------------------------------------------------------------------------------
OSR_DEV float boris(float tem, float e1, float e2, float e3, float b1, float b2, float b3,
                    float& u1, float& u2, float& u3) {
  e1 = e1 * tem;
  e2 = e2 * tem;
  e3 = e3 * tem;
  float ut1 = u1 + e1, ut2 = u2 + e2, ut3 = u3 + e3;
  float utsq = fma_(ut3, ut3, fma_(ut2, ut2, ut1 * ut1));
  float gamma = sqrt_rn(utsq + 1.0f);
  float ene = utsq / (gamma + 1.0f);  // IEEE: utsq can be denormal
  float tg = div_rn(tem, gamma);
  b1 = b1 * tg;
  b2 = b2 * tg;
  b3 = b3 * tg;
  float v1 = fma_(b3, ut2, ut1);
  float v2 = fma_(b1, ut3, ut2);
  float v3 = fma_(b2, ut1, ut3);
  v1 = fma_(-b2, ut3, v1);
  v2 = fma_(-b3, ut1, v2);
  v3 = fma_(-b1, ut2, v3);
  float otsq = fma_(b1, b1, 1.0f);
  otsq = fma_(b2, b2, otsq);
  otsq = fma_(b3, b3, otsq);
  otsq = div_rn(2.0f, otsq);
  b1 = b1 * otsq;
  b2 = b2 * otsq;
  b3 = b3 * otsq;
  ut1 = fma_(b3, v2, ut1);
  ut2 = fma_(b1, v3, ut2);
  ut3 = fma_(b2, v1, ut3);
  ut1 = fma_(-b2, v3, ut1);
  ut2 = fma_(-b3, v1, ut2);
  ut3 = fma_(-b1, v2, ut3);
  u1 = ut1 + e1;
  u2 = ut2 + e2;
  u3 = ut3 + e3;
  return ene;
}

// ------------------------------------------------------------------------------------
// Vay push: spec/os-spec-push.f03:636-796, in the Fortran operation order (no FMA).
// Returns uc^2/(gamma_c+1) of the time centred momentum (:772-792).
// ------------------------------------------------------------------------------------
OSR_DEV float vay(float tem, float e1, float e2, float e3, float b1, float b2, float b3,
                  float& u1, float& u2, float& u3) {
  e1 *= tem; e2 *= tem; e3 *= tem;
  b1 *= tem; b2 *= tem; b3 *= tem;
  float rgam = rcp_rn(sqrt_rn(1.0f + u1 * u1 + u2 * u2 + u3 * u3));
  float v1 = u1 * rgam, v2 = u2 * rgam, v3 = u3 * rgam;
  float bpsq = b1 * b1 + b2 * b2 + b3 * b3;
  float ut1 = u1 + e1 + v2 * b3 - v3 * b2;
  float ut2 = u2 + e2 + v3 * b1 - v1 * b3;
  float ut3 = u3 + e3 + v1 * b2 - v2 * b1;
  float w1 = ut1 + e1, w2 = ut2 + e2, w3 = ut3 + e3;
  float ustar = w1 * b1 + w2 * b2 + w3 * b3;
  float sigma = 1.0f + w1 * w1 + w2 * w2 + w3 * w3 - bpsq;
  rgam = 1.0f / __builtin_sqrtf(0.5f * (sigma + __builtin_sqrtf(sigma * sigma +
                                                                 4.0f * (bpsq + ustar * ustar))));
  b1 *= rgam; b2 *= rgam; b3 *= rgam;
  float spar = rcp_rn(1.0f + b1 * b1 + b2 * b2 + b3 * b3);
  float uptp = w1 * b1 + w2 * b2 + w3 * b3;
  float o1 = u1, o2 = u2, o3 = u3;
  u1 = spar * (w1 + uptp * b1 + w2 * b3 - w3 * b2);
  u2 = spar * (w2 + uptp * b2 + w3 * b1 - w1 * b3);
  u3 = spar * (w3 + uptp * b3 + w1 * b2 - w2 * b1);
  float uc1 = 0.5f * (o1 + u1), uc2 = 0.5f * (o2 + u2), uc3 = 0.5f * (o3 + u3);
  float uc2s = uc1 * uc1 + uc2 * uc2 + uc3 * uc3;
  float gc = sqrt_rn(uc2s + 1.0f);
  return uc2s / (gc + 1.0f);  // IEEE: uc2s can be denormal
}

template <int PUSHER>
OSR_DEV float push(float tem, float e1, float e2, float e3, float b1, float b2, float b3,
                   float& u1, float& u2, float& u3) {
  if (PUSHER == 0) return boris(tem, e1, e2, e3, b1, b2, b3, u1, u2, u3);
  return vay(tem, e1, e2, e3, b1, b2, b3, u1, u2, u3);
}

// 1/gamma: simd/os-spec-push-avx2.c:156-161 (VRGAMMA)
OSR_DEV float rgamma(float u1, float u2, float u3) {
  return rcp_rn(sqrt_rn(fma_(u3, u3, fma_(u2, u2, fma_(u1, u1, 1.0f)))));
}

// Cell shift keeping x in [-0.5, 0.5): spec/os-spec-push.f03:1332-1359, vmntrim
OSR_DEV float trim(float x) { return (x >= 0.5f ? 1.0f : 0.0f) - (x < -0.5f ? 1.0f : 0.0f); }

// ------------------------------------------------------------------------------------
// Fixed point conversion for the deterministic current accumulation.  For |v*scale|
// < 2^51, fma(v, scale, 1.5*2^52) rounds v*scale to the nearest integer in the low
// mantissa bits; subtracting the bit pattern of 1.5*2^52 yields it as int64.
// ------------------------------------------------------------------------------------
OSR_DEV long long to_fix(float v, double scale) {
  double t = __builtin_fma((double)v, scale, 6755399441055744.0);
  return __double_as_longlong(t) - 0x4338000000000000LL;
}
// the integer held by t = fma(a, b, 1.5*2^52), as the 64-bit pattern of the sum
OSR_DEV unsigned long long fix_bits(double t) {
  return (unsigned long long)(__double_as_longlong(t) - 0x4338000000000000LL);
}

// Charge-conserving deposit of one VP into the LDS current tile (DEP_CURRENT_2D,
// spec/os-spec-current.f03:2691-2796); S0x/S0y are the splines of its start point and
// nb the LDS node of its first stencil point, js the row pitch of the tile (a compile-time
// constant after inlining in the fixed-tile kernels).
//
// The weights are those of the reference: j1 += wl1(k1) (S0y + S1y)(k2),
// j2 += (S0x + S1x)(k1) wl2(k2), j3 += (q vz / 3) [S0x S0y + S1x S1y + (S0x S1y + S1x S0y)/2].
// Each product is formed exactly in FP64 from FP32 factors that already carry the fixed-
// point scale 2^s (a power of two, so the scaling is exact) and rounded once to the int64
// grid by fma(a, b, 1.5*2^52): one FMA and one integer add per contribution.  The j3 bracket
// is evaluated as A S0y + B S1y with A = S0x + S1x/2, B = S1x + S0x/2 (two FP32 operations
// per node).  |a b 2^s| < 2^51 for every contribution (bound of the scale choice).
template <int ORDER>
OSR_DEV void dep_vp(unsigned long long* __restrict__ J1, unsigned long long* __restrict__ J2,
                    unsigned long long* __restrict__ J3, const float* S0x, const float* S0y,
                    float x0, float y0, float x1, float y1, float q, float vz, int nb, int js,
                    float jn1, float jn2, double jscale) {
  constexpr int NP = ORDER + 1;
  constexpr double M = 6755399441055744.0;  // 1.5 * 2^52
  float S1x[NP], S1y[NP], wl1[NP], wl2[NP];
  spline<ORDER>(x1, S1x);
  spline<ORDER>(y1, S1y);
  wl<ORDER>(q * jn1, x0, x1, wl1);
  wl<ORDER>(q * jn2, y0, y1, wl2);
  double a1[NP - 1], a2[NP - 1], b1[NP], b2[NP];
  float A[NP], B[NP];
#pragma unroll
  for (int k = 0; k < NP - 1; k++) {
    a1[k] = (double)wl1[k] * jscale;
    a2[k] = (double)wl2[k] * jscale;
  }
#pragma unroll
  for (int k = 0; k < NP; k++) {
    b1[k] = (double)(S0y[k] + S1y[k]);
    b2[k] = (double)(S0x[k] + S1x[k]);
    A[k] = fma_(S1x[k], 0.5f, S0x[k]);
    B[k] = fma_(S0x[k], 0.5f, S1x[k]);
  }
  const double c3 = (double)((q * vz) * (1.0f / 3.0f)) * jscale;
#pragma unroll
  for (int k2 = 0; k2 < NP; k2++) {
#pragma unroll
    for (int k1 = 0; k1 < NP; k1++) {
      const int idx = nb + k2 * js + k1;
      const float t1 = fma_(B[k1], S1y[k2], A[k1] * S0y[k2]);
      if (k1 < NP - 1) atomicAdd(&J1[idx], fix_bits(__builtin_fma(a1[k1], b1[k2], M)));
      if (k2 < NP - 1) atomicAdd(&J2[idx], fix_bits(__builtin_fma(a2[k2], b2[k1], M)));
      atomicAdd(&J3[idx], fix_bits(__builtin_fma(c3, (double)t1, M)));
    }
  }
}

// ------------------------------------------------------------------------------------
// Virtual particles.  split_2d: spec/os-spec-current.f03:659-888, vsplit2D
// simd/split-vec.h:105-222 (scalar IEEE arithmetic, as compiled by the reference).
// ------------------------------------------------------------------------------------
struct VP2 {
  float x0, y0, x1, y1, q, vz;
  int di, dj;  // cell offset relative to the particle's old cell
};

OSR_DEV int split2d(float xo, float yo, float xn, float yn, int d1, int d2, float q, float vz,
                    VP2* vp) {
  // All three records are written unconditionally at the end from scalar variables;
  // conditional stores into different array elements would be merged by the compiler
  // into stores through a selected pointer, which forces the array into scratch memory.
  const int cross = (d1 != 0) + 2 * (d2 != 0);
  const float f1 = (float)d1, f2 = (float)d2;
  VP2 a = {xo, yo, xn, yn, q, vz, 0, 0};
  VP2 b = a, c = a;
  int n = 1;
  if (cross == 1) {
    const float xint = 0.5f * f1;
    const float delta = div_rn(xint - xo, xn - xo);
    const float yint = yo + (yn - yo) * delta;
    a = {xo, yo, xint, yint, q, vz * delta, 0, 0};
    b = {-xint, yint, xn - f1, yn, q, vz * (1.0f - delta), d1, 0};
    n = 2;
  } else if (cross == 2) {
    const float yint = 0.5f * f2;
    const float delta = div_rn(yint - yo, yn - yo);
    const float xint = xo + (xn - xo) * delta;
    a = {xo, yo, xint, yint, q, vz * delta, 0, 0};
    b = {xint, -yint, xn, yn - f2, q, vz * (1.0f - delta), 0, d2};
    n = 2;
  } else if (cross == 3) {
    const float xint = 0.5f * f1;
    float delta = div_rn(xint - xo, xn - xo);
    const float yint = yo + (yn - yo) * delta;
    if (yint >= -0.5f && yint < 0.5f) {
      const float vzint = vz * delta;
      const float vzint2 = vz * (1.0f - delta);
      const float yint2 = 0.5f * f2;
      delta = div_rn(yint2 - yint, yn - yint);
      const float xint2 = -xint + (xn - xint) * delta;
      a = {xo, yo, xint, yint, q, vzint, 0, 0};
      b = {-xint, yint, xint2, yint2, q, vzint2 * delta, d1, 0};
      c = {xint2, -yint2, xn - f1, yn - f2, q, vzint2 * (1.0f - delta), d1, d2};
    } else {
      const float vzint = vz * delta;
      const float yint2 = 0.5f * f2;
      delta = div_rn(yint2 - yo, yint - yo);
      const float xint2 = xo + (xint - xo) * delta;
      a = {xo, yo, xint2, yint2, q, vzint * delta, 0, 0};
      b = {xint2, -yint2, xint, yint - f2, q, vzint * (1.0f - delta), 0, d2};
      c = {-xint, yint - f2, xn - f1, yn - f2, q, vz - vzint, d1, d2};
    }
    n = 3;
  }
  vp[0] = a;
  vp[1] = b;
  vp[2] = c;
  return n;
}

// split2d with one division shared by the single-crossing cases (x face or y face)
// instead of one divergent branch per case.  Every VP field is computed by the same
// expression as in split2d, so the results are bit-identical; the double crossing
// (a rare case) keeps its own branch.
OSR_DEV int split2d_fast(float xo, float yo, float xn, float yn, int d1, int d2, float q,
                         float vz, VP2* vp) {
  const bool cx = d1 != 0, cy = d2 != 0;
  const float f1 = (float)d1, f2 = (float)d2;
  VP2 a = {xo, yo, xn, yn, q, vz, 0, 0};
  VP2 b = a, c = a;
  if (cx || cy) {
    // first crossing: the x face when the x cell changes, else the y face
    const float xint = 0.5f * f1, yface = 0.5f * f2;
    const float num = cx ? xint - xo : yface - yo;
    const float den = cx ? xn - xo : yn - yo;
    const float delta = div_rn(num, den);
    const float o0 = cx ? yo : xo, o1 = cx ? yn : xn;
    const float oint = o0 + (o1 - o0) * delta;  // yint (x face) or xint (y face)
    a.x1 = cx ? xint : oint;
    a.y1 = cx ? oint : yface;
    a.vz = vz * delta;
    b.x0 = cx ? -xint : oint;
    b.y0 = cx ? oint : -yface;
    b.x1 = cx ? xn - f1 : xn;
    b.y1 = cx ? yn : yn - f2;
    b.vz = vz * (1.0f - delta);
    b.di = d1;
    b.dj = d2;
    if (cx && cy) {
      const float yint = oint;
      const float yint2 = 0.5f * f2;
      if (yint >= -0.5f && yint < 0.5f) {
        const float vzint = vz * delta;
        const float vzint2 = vz * (1.0f - delta);
        const float delta2 = div_rn(yint2 - yint, yn - yint);
        const float xint2 = -xint + (xn - xint) * delta2;
        a = {xo, yo, xint, yint, q, vzint, 0, 0};
        b = {-xint, yint, xint2, yint2, q, vzint2 * delta2, d1, 0};
        c = {xint2, -yint2, xn - f1, yn - f2, q, vzint2 * (1.0f - delta2), d1, d2};
      } else {
        const float vzint = vz * delta;
        const float delta2 = div_rn(yint2 - yo, yint - yo);
        const float xint2 = xo + (xint - xo) * delta2;
        a = {xo, yo, xint2, yint2, q, vzint * delta2, 0, 0};
        b = {xint2, -yint2, xint, yint - f2, q, vzint * (1.0f - delta2), 0, d2};
        c = {-xint, yint - f2, xn - f1, yn - f2, q, vz - vzint, d1, d2};
      }
    }
  }
  vp[0] = a;
  vp[1] = b;
  vp[2] = c;
  return 1 + (cx || cy) + (cx && cy);
}

struct VP3 {
  float x0, y0, z0, x1, y1, z1, q;
  int di, dj, dk;
};

// Select virtual particle v of a register array field by field, so that the array
// stays in registers (no dynamic indexing of private memory).
// (values are passed by value: a conditional over lvalues would select addresses)
template <class T>
OSR_DEV T sel3(int v, T a, T b, T c) { return v == 0 ? a : (v == 1 ? b : c); }
template <class T>
OSR_DEV T sel4(int v, T a, T b, T c, T d) { return v == 0 ? a : (v == 1 ? b : (v == 2 ? c : d)); }
#define OSR_SEL3(F) r.F = sel3(v, vp[0].F, vp[1].F, vp[2].F)
#define OSR_SEL4(F) r.F = sel4(v, vp[0].F, vp[1].F, vp[2].F, vp[3].F)
OSR_DEV VP2 pick(const VP2* vp, int v) {
  VP2 r;
  OSR_SEL3(x0); OSR_SEL3(y0); OSR_SEL3(x1); OSR_SEL3(y1);
  OSR_SEL3(q); OSR_SEL3(vz); OSR_SEL3(di); OSR_SEL3(dj);
  return r;
}
OSR_DEV VP3 pick(const VP3* vp, int v) {
  VP3 r;
  OSR_SEL4(x0); OSR_SEL4(y0); OSR_SEL4(z0); OSR_SEL4(x1); OSR_SEL4(y1); OSR_SEL4(z1);
  OSR_SEL4(q); OSR_SEL4(di); OSR_SEL4(dj); OSR_SEL4(dk);
  return r;
}
#undef OSR_SEL3
#undef OSR_SEL4

// Charge-conserving deposit of one 3D VP into the LDS current tile (DEP_CURRENT_3D,
// spec/os-spec-current.f03:2800-2905): j1 += wl1(k1) wp1(k2,k3), j2 += wl2(k2) wp2(k1,k3),
// j3 += wl3(k3) wp3(k1,k2), each wp = S0 S0 + S1 S1 + (S0 S1 + S1 S0)/2 over the two other
// directions, evaluated in FP32 in the reference's operation order (:2876-2890).  As in
// 2D, every contribution is the exact FP64 product of wl 2^s and wp, rounded once to the
// int64 grid: one FMA and one integer add per node.  S0x/S0y/S0z are the splines of the
// VP's start point, nb its first stencil node, jy/jz the row and plane pitches of the
// tile, JN the size of one component plane.
template <int ORDER>
OSR_DEV void dep_vp3(unsigned long long* __restrict__ J, int JN, const float* S0x,
                     const float* S0y, const float* S0z, float x0, float y0, float z0, float x1,
                     float y1, float z1, float q, int nb, int jy, int jz, float jn1, float jn2,
                     float jn3, double jscale) {
  constexpr int NP = ORDER + 1;
  constexpr double M = 6755399441055744.0;  // 1.5 * 2^52
  float S1x[NP], S1y[NP], S1z[NP], wl1[NP], wl2[NP], wl3[NP];
  spline<ORDER>(x1, S1x);
  spline<ORDER>(y1, S1y);
  spline<ORDER>(z1, S1z);
  wl<ORDER>(q * jn1, x0, x1, wl1);
  wl<ORDER>(q * jn2, y0, y1, wl2);
  wl<ORDER>(q * jn3, z0, z1, wl3);
  double a1[NP - 1], a2[NP - 1], a3[NP - 1];
#pragma unroll
  for (int k = 0; k < NP - 1; k++) {
    a1[k] = (double)wl1[k] * jscale;
    a2[k] = (double)wl2[k] * jscale;
    a3[k] = (double)wl3[k] * jscale;
  }
  // wp(a, b) over the pair of directions other than the current's
  auto wp = [](float s0a, float s1a, float s0b, float s1b) {
    return fma_(0.5f, fma_(s1a, s0b, s0a * s1b), fma_(s1a, s1b, s0a * s0b));
  };
#pragma unroll
  for (int k3 = 0; k3 < NP; k3++) {
#pragma unroll
    for (int k2 = 0; k2 < NP; k2++) {
      const double b1 = (double)wp(S0y[k2], S1y[k2], S0z[k3], S1z[k3]);
#pragma unroll
      for (int k1 = 0; k1 < NP; k1++) {
        const int idx = nb + k3 * jz + k2 * jy + k1;
        if (k1 < NP - 1) atomicAdd(&J[idx], fix_bits(__builtin_fma(a1[k1], b1, M)));
        if (k2 < NP - 1) {
          const double b2 = (double)wp(S0x[k1], S1x[k1], S0z[k3], S1z[k3]);
          atomicAdd(&J[JN + idx], fix_bits(__builtin_fma(a2[k2], b2, M)));
        }
        if (k3 < NP - 1) {
          const double b3 = (double)wp(S0x[k1], S1x[k1], S0y[k2], S1y[k2]);
          atomicAdd(&J[2 * JN + idx], fix_bits(__builtin_fma(a3[k3], b3, M)));
        }
      }
    }
  }
}

// split_3d: spec/os-spec-current.f03:895-1601 (vsplit3D, simd/split-vec.h:231-).
OSR_DEV VP3 mk3(float x0, float y0, float z0, float x1, float y1, float z1, float q, int i, int j,
                int k) {
  VP3 v;
  v.x0 = x0; v.y0 = y0; v.z0 = z0; v.x1 = x1; v.y1 = y1; v.z1 = z1; v.q = q;
  v.di = i; v.dj = j; v.dk = k;
  return v;
}

OSR_DEV int split3d(float xo, float yo, float zo, float xn, float yn, float zn, int d1, int d2,
                    int d3, float q, VP3* vp) {
  const bool cx = d1 != 0, cy = d2 != 0, cz = d3 != 0;
  const int cross = cx + 2 * cy + 4 * cz;
  float xint, yint, zint, xint2, yint2, zint2, delta;
  const float f1 = (float)d1, f2 = (float)d2, f3 = (float)d3;
  if (cross == 0) {
    vp[0] = mk3(xo, yo, zo, xn, yn, zn, q, 0, 0, 0);
    return 1;
  }
  if (cross == 1 || cross == 2 || cross == 4) {
    // One face crossed (cases 1, 2 and 4 of the reference): one division on the crossed
    // axis and the same interpolation of the other two, so a wave whose lanes cross
    // different faces runs one path instead of three.  Each field is the expression of
    // its case (a shift of 0 leaves a coordinate exactly as the case leaves it).
    const float fk = cx ? f1 : (cy ? f2 : f3);
    const float ok = cx ? xo : (cy ? yo : zo), nk = cx ? xn : (cy ? yn : zn);
    const float tint = 0.5f * fk;
    delta = div_rn(tint - ok, nk - ok);
    xint = cx ? tint : xo + (xn - xo) * delta;
    yint = cy ? tint : yo + (yn - yo) * delta;
    zint = cz ? tint : zo + (zn - zo) * delta;
    vp[0] = mk3(xo, yo, zo, xint, yint, zint, q, 0, 0, 0);
    vp[1] = mk3(cx ? -xint : xint, cy ? -yint : yint, cz ? -zint : zint, xn - f1, yn - f2,
                zn - f3, q, d1, d2, d3);
    return 2;
  }
  switch (cross) {
  case 3:
    xint = 0.5f * f1;
    delta = div_rn(xint - xo, xn - xo);
    yint = yo + (yn - yo) * delta;
    zint = zo + (zn - zo) * delta;
    if (yint >= -0.5f && yint < 0.5f) {
      yint2 = 0.5f * f2;
      delta = div_rn(yint2 - yint, yn - yint);
      xint2 = -xint + (xn - xint) * delta;
      zint2 = zint + (zn - zint) * delta;
      vp[0] = mk3(xo, yo, zo, xint, yint, zint, q, 0, 0, 0);
      vp[1] = mk3(-xint, yint, zint, xint2, yint2, zint2, q, d1, 0, 0);
      vp[2] = mk3(xint2, -yint2, zint2, xn - f1, yn - f2, zn, q, d1, d2, 0);
    } else {
      yint2 = 0.5f * f2;
      delta = div_rn(yint2 - yo, yint - yo);
      xint2 = xo + (xint - xo) * delta;
      zint2 = zo + (zint - zo) * delta;
      vp[0] = mk3(xo, yo, zo, xint2, yint2, zint2, q, 0, 0, 0);
      vp[1] = mk3(xint2, -yint2, zint2, xint, yint - f2, zint, q, 0, d2, 0);
      vp[2] = mk3(-xint, yint - f2, zint, xn - f1, yn - f2, zn, q, d1, d2, 0);
    }
    return 3;
  case 5:
    xint = 0.5f * f1;
    delta = div_rn(xint - xo, xn - xo);
    yint = yo + (yn - yo) * delta;
    zint = zo + (zn - zo) * delta;
    zint2 = 0.5f * f3;
    if (zint >= -0.5f && zint < 0.5f) {
      delta = div_rn(zint2 - zint, zn - zint);
      xint2 = -xint + (xn - xint) * delta;
      yint2 = yint + (yn - yint) * delta;
      vp[0] = mk3(xo, yo, zo, xint, yint, zint, q, 0, 0, 0);
      vp[1] = mk3(-xint, yint, zint, xint2, yint2, zint2, q, d1, 0, 0);
      vp[2] = mk3(xint2, yint2, -zint2, xn - f1, yn, zn - f3, q, d1, 0, d3);
    } else {
      delta = div_rn(zint2 - zo, zint - zo);
      xint2 = xo + (xint - xo) * delta;
      yint2 = yo + (yint - yo) * delta;
      vp[0] = mk3(xo, yo, zo, xint2, yint2, zint2, q, 0, 0, 0);
      vp[1] = mk3(xint2, yint2, -zint2, xint, yint, zint - f3, q, 0, 0, d3);
      vp[2] = mk3(-xint, yint, zint - f3, xn - f1, yn, zn - f3, q, d1, 0, d3);
    }
    return 3;
  case 6:
    yint = 0.5f * f2;
    delta = div_rn(yint - yo, yn - yo);
    xint = xo + (xn - xo) * delta;
    zint = zo + (zn - zo) * delta;
    zint2 = 0.5f * f3;
    if (zint >= -0.5f && zint < 0.5f) {
      delta = div_rn(zint2 - zint, zn - zint);
      xint2 = xint + (xn - xint) * delta;
      yint2 = -yint + (yn - yint) * delta;
      vp[0] = mk3(xo, yo, zo, xint, yint, zint, q, 0, 0, 0);
      vp[1] = mk3(xint, -yint, zint, xint2, yint2, zint2, q, 0, d2, 0);
      vp[2] = mk3(xint2, yint2, -zint2, xn, yn - f2, zn - f3, q, 0, d2, d3);
    } else {
      delta = div_rn(zint2 - zo, zint - zo);
      xint2 = xo + (xint - xo) * delta;
      yint2 = yo + (yint - yo) * delta;
      vp[0] = mk3(xo, yo, zo, xint2, yint2, zint2, q, 0, 0, 0);
      vp[1] = mk3(xint2, yint2, -zint2, xint, yint, zint - f3, q, 0, 0, d3);
      vp[2] = mk3(xint, -yint, zint - f3, xn, yn - f2, zn - f3, q, 0, d2, d3);
    }
    return 3;
  default: {  // 7: x, y and z crossings
    xint = 0.5f * f1;
    delta = div_rn(xint - xo, xn - xo);
    yint = yo + (yn - yo) * delta;
    zint = zo + (zn - zo) * delta;
    if (yint >= -0.5f && yint < 0.5f) {
      yint2 = 0.5f * f2;
      delta = div_rn(yint2 - yint, yn - yint);
      xint2 = -xint + (xn - xint) * delta;
      zint2 = zint + (zn - zint) * delta;
      vp[0] = mk3(xo, yo, zo, xint, yint, zint, q, 0, 0, 0);
      vp[1] = mk3(-xint, yint, zint, xint2, yint2, zint2, q, d1, 0, 0);
      vp[2] = mk3(xint2, -yint2, zint2, xn - f1, yn - f2, zn, q, d1, d2, 0);
    } else {
      yint2 = 0.5f * f2;
      delta = div_rn(yint2 - yo, yint - yo);
      xint2 = xo + (xint - xo) * delta;
      zint2 = zo + (zint - zo) * delta;
      vp[0] = mk3(xo, yo, zo, xint2, yint2, zint2, q, 0, 0, 0);
      vp[1] = mk3(xint2, -yint2, zint2, xint, yint - f2, zint, q, 0, d2, 0);
      vp[2] = mk3(-xint, yint - f2, zint, xn - f1, yn - f2, zn, q, d1, d2, 0);
    }
    // one of the three needs an extra z split (os-spec-current.f03:1561-1590); the new
    // virtual particle is appended last and the ones after the split one move one cell
    // in z.  Static indices only (no dynamic register-array indexing).
    zint = 0.5f * f3;
    const bool c0 = vp[0].z1 < -0.5f || vp[0].z1 >= 0.5f;
    const bool c1 = vp[1].z1 < -0.5f || vp[1].z1 >= 0.5f;
    const bool c2 = vp[2].z1 < -0.5f || vp[2].z1 >= 0.5f;
    if (!(c0 || c1 || c2)) return 3;
    const int s = c0 ? 0 : (c1 ? 1 : 2);
    const VP3 o = pick(vp, s);
    delta = div_rn(zint - o.z0, o.z1 - o.z0);
    xint = o.x0 + (o.x1 - o.x0) * delta;
    yint = o.y0 + (o.y1 - o.y0) * delta;
    vp[3] = mk3(xint, yint, -zint, o.x1, o.y1, o.z1 - f3, o.q, o.di, o.dj, o.dk + d3);
#pragma unroll
    for (int jj = 0; jj < 3; jj++) {
      if (jj == s) {
        vp[jj].x1 = xint;
        vp[jj].y1 = yint;
        vp[jj].z1 = zint;
      } else if (jj > s) {
        vp[jj].z0 = vp[jj].z0 - f3;
        vp[jj].z1 = vp[jj].z1 - f3;
        vp[jj].dk = vp[jj].dk + d3;
      }
    }
    return 4;
  }
  }
}

// Wave-level helpers (64-wide wavefronts)
OSR_DEV int lane_id() { return __lane_id(); }

OSR_DEV double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

}  // namespace osr
