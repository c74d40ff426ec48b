/*
 * pic_oracle.c -- CPU restatement of the OSIRIS per-timestep PIC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product
 * library never links it and never falls back to it.
 *
 * It restates the reference Fortran algorithms in plain C, operation by operation, so
 * that a build with REAL=double reproduces the FP64 reference build and a build with
 * REAL=float (and -ffp-contract=off) reproduces the Fortran single precision build.
 * Every function cites the reference file:line it follows (paths relative to
 * source/ of osiris-code/osiris).
 *
 * Conventions (identical to the reference):
 *   - grids are stored like t_vdf%f2 / f3: component fastest, then x, y, z, with
 *     Fortran indices running from 1-gcl to nx+gcu (vdf/os-vdf-define.f03:35-116);
 *   - particle cell indices ix(d,N) are 1-based Fortran indices, positions x(d,N) lie in
 *     [-0.5,0.5) relative to the cell, momenta u(3,N) are proper velocities, q(N) charges
 *     (spec/os-spec-define.f03:391-412).
 *
 * Pinning (see the tests): splines/wl against closed forms and partition of unity; the
 * fused Boris advance/deposit against the reference's own AVX2 kernels compiled from
 * simd/ (oracle/_ref) on identical inputs; deposit against discrete continuity with the
 * independent rho deposit; Vay against the exact ExB drift, the exact force-free motion
 * with a momentum component along B (the gamma solve's (u'.tau)^2 term) and the exact
 * acceleration along B; the step order (push, energy rows, fold, Yee) against the energy
 * conservation of a cold Langmuir oscillation, O((w_p dt)^2) only when the field and
 * kinetic rows are time-aligned; Yee against the closed-form
 * numerical dispersion of a vacuum plane wave and energy conservation; Lindman walls
 * against the closed-form Mur reflection; VPML walls against the Yee lattice continued
 * through the wall (attenuation off) and their measured reflections.  Parity unpinned:
 * the last bit of the VPML attenuation coefficients against a reference binary, whose
 * expf and FMA contraction (-Ofast -march=native) the restatement cannot see.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL double
#endif
typedef REAL real;
#define R(v) ((real)(v))

/* os-param.f90:227 */
#define P_CACHE_SIZE 1024

typedef struct {
  int dims;   /* 2 or 3 */
  int nx[3];  /* interior cells */
  int gcl[3]; /* lower guard cells */
  int gcu[3]; /* upper guard cells */
} orc_grid;

static inline long ld_(const orc_grid *g, int d) {
  return (long)g->nx[d] + g->gcl[d] + g->gcu[d];
}

/* linear cell offset of Fortran index (i,j,k) in a t_vdf with 1-gcl lower bound */
static inline long cidx(const orc_grid *g, int i, int j, int k) {
  long o = (long)(i - 1 + g->gcl[0]);
  if (g->dims > 1) o += ld_(g, 0) * (long)(j - 1 + g->gcl[1]);
  if (g->dims > 2) o += ld_(g, 0) * ld_(g, 1) * (long)(k - 1 + g->gcl[2]);
  return o;
}
#define FLD(f, g, c, i, j, k) ((f)[3 * cidx((g), (i), (j), (k)) + (c)])

long orc_grid_cells(const orc_grid *g) {
  long n = ld_(g, 0);
  if (g->dims > 1) n *= ld_(g, 1);
  if (g->dims > 2) n *= ld_(g, 2);
  return n;
}

int orc_real_size(void) { return (int)sizeof(real); }

/* ------------------------------------------------------------------------------------
 * Splines.  emf/os-emf-interpolate.f03:171-293 (spline_s1..s3, splineh_s1..s3) and
 * spec/os-spec-current.f03:488-550 (same formulas).  s[k] holds the weight of stencil
 * point LP+k, LP = 0/-1/-1 for orders 1/2/3 (os-emf-interpolate.f03:385-461).
 * ------------------------------------------------------------------------------------ */
static inline int lp_of(int order) { return order == 1 ? 0 : -1; }

void orc_spline(int order, real x, real *s) {
  switch (order) {
  case 1:
    s[0] = R(0.5) - x;
    s[1] = R(0.5) + x;
    break;
  case 2: {
    real t0 = R(0.5) - x, t1 = R(0.5) + x;
    s[0] = R(0.5) * (t0 * t0);
    s[1] = R(0.5) + t0 * t1;
    s[2] = R(0.5) * (t1 * t1);
  } break;
  case 3: {
    const real c1_6 = R(1.0) / R(6.0), c2_3 = R(2.0) / R(3.0);
    real t0 = R(0.5) - x, t1 = R(0.5) + x;
    real t2 = t0 * t0, t3 = t1 * t1;
    t0 = t0 * t2;
    t1 = t1 * t3;
    s[0] = c1_6 * t0;
    s[1] = (c2_3 - t3) + R(0.5) * t1;
    s[2] = (c2_3 - t2) + R(0.5) * t0;
    s[3] = c1_6 * t1;
  } break;
  }
}

/* half-cell shifted spline, os-emf-interpolate.f03:183-193, 215-232, 264-292 */
void orc_splineh(int order, real x, int h, real *s) {
  real t0 = (real)(1 - h) - x, t1 = (real)h + x;
  switch (order) {
  case 1:
    s[0] = t0;
    s[1] = t1;
    break;
  case 2:
    s[0] = R(0.5) * (t0 * t0);
    s[1] = R(0.5) + t0 * t1;
    s[2] = R(0.5) * (t1 * t1);
    break;
  case 3: {
    const real c1_6 = R(1.0) / R(6.0), c2_3 = R(2.0) / R(3.0);
    real t2 = t0 * t0, t3 = t1 * t1;
    t0 = t0 * t2;
    t1 = t1 * t3;
    s[0] = c1_6 * t0;
    s[1] = (c2_3 - t3) + R(0.5) * t1;
    s[2] = (c2_3 - t2) + R(0.5) * t0;
    s[3] = c1_6 * t1;
  } break;
  }
}

/* Longitudinal current weights, spec/os-spec-current.f03:383-441 (wl_s1..wl_s3).
 * wl[k] is the weight of point P0+k; the entry P1 is zeroed as in DEP_CURRENT
 * (os-spec-current.f03:2752-2753). */
void orc_wl(int order, real qn, real x0, real x1, real *wl) {
  switch (order) {
  case 1:
    wl[0] = qn * (x1 - x0);
    wl[1] = 0;
    break;
  case 2: {
    real d = x1 - x0, s1_2 = R(0.5) - x0, p1_2 = R(0.5) + x0;
    wl[0] = s1_2 - R(0.5) * d;
    wl[1] = p1_2 + R(0.5) * d;
    real n = qn * d;
    wl[0] = n * wl[0];
    wl[1] = n * wl[1];
    wl[2] = 0;
  } break;
  case 3: {
    const real c1_3 = R(1.0) / R(3.0);
    real d = x1 - x0, s1_2 = R(0.5) - x0, p1_2 = R(0.5) + x0, d_3 = c1_3 * d;
    wl[0] = R(0.5) * (s1_2 * s1_2 - d * (s1_2 - d_3));
    wl[1] = (R(0.75) - x0 * x0) - d * (x0 + d_3);
    wl[2] = R(0.5) * (p1_2 * p1_2 + d * (p1_2 + d_3));
    real n = qn * d;
    wl[0] = n * wl[0];
    wl[1] = n * wl[1];
    wl[2] = n * wl[2];
    wl[3] = 0;
  } break;
  }
}

/* ------------------------------------------------------------------------------------
 * Field gather.  emf/os-emf-interpolate.f03:554-648 (GET_EMF_2D), 652-779 (GET_EMF_3D);
 * signbit :358-371.
 * ------------------------------------------------------------------------------------ */
static void get_emf_2d(const orc_grid *g, int order, const real *e, const real *b,
                       const int *ix, const real *x, int np, real *ep, real *bp) {
  const int n = order + 1, lp = lp_of(order);
  real w1[4], w1h[4], w2[4], w2h[4];
  for (int l = 0; l < np; l++) {
    int i1 = ix[2 * l], i2 = ix[2 * l + 1];
    real dx1 = x[2 * l], dx2 = x[2 * l + 1];
    int h1 = dx1 < 0 ? 1 : 0, h2 = dx2 < 0 ? 1 : 0;
    int i1h = i1 - h1, i2h = i2 - h2;
    orc_spline(order, dx1, w1);
    orc_splineh(order, dx1, h1, w1h);
    orc_spline(order, dx2, w2);
    orc_splineh(order, dx2, h2, w2h);
    real f1 = 0, f2 = 0, f3 = 0;
    for (int k2 = 0; k2 < n; k2++) {
      real l1 = 0, l2 = 0, l3 = 0;
      for (int k1 = 0; k1 < n; k1++) {
        l1 = l1 + FLD(e, g, 0, i1h + lp + k1, i2 + lp + k2, 1) * w1h[k1];
        l2 = l2 + FLD(e, g, 1, i1 + lp + k1, i2h + lp + k2, 1) * w1[k1];
        l3 = l3 + FLD(e, g, 2, i1 + lp + k1, i2 + lp + k2, 1) * w1[k1];
      }
      f1 = f1 + l1 * w2[k2];
      f2 = f2 + l2 * w2h[k2];
      f3 = f3 + l3 * w2[k2];
    }
    ep[3 * l] = f1;
    ep[3 * l + 1] = f2;
    ep[3 * l + 2] = f3;
    f1 = f2 = f3 = 0;
    for (int k2 = 0; k2 < n; k2++) {
      real l1 = 0, l2 = 0, l3 = 0;
      for (int k1 = 0; k1 < n; k1++) {
        l1 = l1 + FLD(b, g, 0, i1 + lp + k1, i2h + lp + k2, 1) * w1[k1];
        l2 = l2 + FLD(b, g, 1, i1h + lp + k1, i2 + lp + k2, 1) * w1h[k1];
        l3 = l3 + FLD(b, g, 2, i1h + lp + k1, i2h + lp + k2, 1) * w1h[k1];
      }
      f1 = f1 + l1 * w2h[k2];
      f2 = f2 + l2 * w2[k2];
      f3 = f3 + l3 * w2h[k2];
    }
    bp[3 * l] = f1;
    bp[3 * l + 1] = f2;
    bp[3 * l + 2] = f3;
  }
}

static void get_emf_3d(const orc_grid *g, int order, const real *e, const real *b,
                       const int *ix, const real *x, int np, real *ep, real *bp) {
  const int n = order + 1, lp = lp_of(order);
  real w1[4], w1h[4], w2[4], w2h[4], w3[4], w3h[4];
  for (int l = 0; l < np; l++) {
    int i1 = ix[3 * l], i2 = ix[3 * l + 1], i3 = ix[3 * l + 2];
    real dx1 = x[3 * l], dx2 = x[3 * l + 1], dx3 = x[3 * l + 2];
    int h1 = dx1 < 0, h2 = dx2 < 0, h3 = dx3 < 0;
    int i1h = i1 - h1, i2h = i2 - h2, i3h = i3 - h3;
    orc_spline(order, dx1, w1);
    orc_splineh(order, dx1, h1, w1h);
    orc_spline(order, dx2, w2);
    orc_splineh(order, dx2, h2, w2h);
    orc_spline(order, dx3, w3);
    orc_splineh(order, dx3, h3, w3h);
    real f1 = 0, f2 = 0, f3 = 0;
    for (int k3 = 0; k3 < n; k3++) {
      real p1 = 0, p2 = 0, p3 = 0;
      for (int k2 = 0; k2 < n; k2++) {
        real l1 = 0, l2 = 0, l3 = 0;
        for (int k1 = 0; k1 < n; k1++) {
          l1 = l1 + FLD(e, g, 0, i1h + lp + k1, i2 + lp + k2, i3 + lp + k3) * w1h[k1];
          l2 = l2 + FLD(e, g, 1, i1 + lp + k1, i2h + lp + k2, i3 + lp + k3) * w1[k1];
          l3 = l3 + FLD(e, g, 2, i1 + lp + k1, i2 + lp + k2, i3h + lp + k3) * w1[k1];
        }
        p1 = p1 + l1 * w2[k2];
        p2 = p2 + l2 * w2h[k2];
        p3 = p3 + l3 * w2[k2];
      }
      f1 = f1 + p1 * w3[k3];
      f2 = f2 + p2 * w3[k3];
      f3 = f3 + p3 * w3h[k3];
    }
    ep[3 * l] = f1;
    ep[3 * l + 1] = f2;
    ep[3 * l + 2] = f3;
    f1 = f2 = f3 = 0;
    for (int k3 = 0; k3 < n; k3++) {
      real p1 = 0, p2 = 0, p3 = 0;
      for (int k2 = 0; k2 < n; k2++) {
        real l1 = 0, l2 = 0, l3 = 0;
        for (int k1 = 0; k1 < n; k1++) {
          l1 = l1 + FLD(b, g, 0, i1 + lp + k1, i2h + lp + k2, i3h + lp + k3) * w1[k1];
          l2 = l2 + FLD(b, g, 1, i1h + lp + k1, i2 + lp + k2, i3h + lp + k3) * w1h[k1];
          l3 = l3 + FLD(b, g, 2, i1h + lp + k1, i2h + lp + k2, i3 + lp + k3) * w1h[k1];
        }
        p1 = p1 + l1 * w2h[k2];
        p2 = p2 + l2 * w2[k2];
        p3 = p3 + l3 * w2h[k2];
      }
      f1 = f1 + p1 * w3h[k3];
      f2 = f2 + p2 * w3h[k3];
      f3 = f3 + p3 * w3[k3];
    }
    bp[3 * l] = f1;
    bp[3 * l + 1] = f2;
    bp[3 * l + 2] = f3;
  }
}

void orc_get_emf(const orc_grid *g, int order, const real *e, const real *b, const int *ix,
                 const real *x, int np, real *ep, real *bp) {
  if (g->dims == 2)
    get_emf_2d(g, order, e, b, ix, x, np, ep, bp);
  else
    get_emf_3d(g, order, e, b, ix, x, np, ep, bp);
}

/* ------------------------------------------------------------------------------------
 * Momentum pushers on pre-gathered fields.
 * Boris: spec/os-spec-push.f03:356-632, non radiation-reaction path 466-574.
 *   tem = real(0.5*(dt/rqm)) :396; energy accumulated per p_cache_size chunk :488-510.
 * Vay:   spec/os-spec-push.f03:636-796.
 * Both return the chunk-accumulated sum of q*u^2/(gamma+1) in double.
 * ------------------------------------------------------------------------------------ */
double orc_dudt_boris(real rqm, double dt, real *u, const real *q, real *ep, real *bp,
                      int np) {
  real tem = (real)(0.5 * (dt / (double)rqm));
  double energy = 0;
  for (int c0 = 0; c0 < np; c0 += P_CACHE_SIZE) {
    int c1 = c0 + P_CACHE_SIZE < np ? c0 + P_CACHE_SIZE : np;
    double loc_ene = 0;
    for (int i = c0; i < c1; i++) {
      real *p = u + 3 * i, *e = ep + 3 * i, *b = bp + 3 * i;
      e[0] = e[0] * tem;
      e[1] = e[1] * tem;
      e[2] = e[2] * tem;
      real ut1 = p[0] + e[0], ut2 = p[1] + e[1], ut3 = p[2] + e[2];
      real u2 = (ut1 * ut1 + ut2 * ut2) + ut3 * ut3;
      real gamma = (real)sqrt(u2 + 1);
      loc_ene = loc_ene + (double)(q[i] * u2) / ((double)gamma + 1.0);
      real gam_tem = tem / gamma;
      b[0] = b[0] * gam_tem;
      b[1] = b[1] * gam_tem;
      b[2] = b[2] * gam_tem;
      real p1 = ut1 + ut2 * b[2], p2 = ut2 + ut3 * b[0], p3 = ut3 + ut1 * b[1];
      p1 = p1 - ut3 * b[1];
      p2 = p2 - ut1 * b[2];
      p3 = p3 - ut2 * b[0];
      real otsq = R(2.0) / (((R(1.0) + b[0] * b[0]) + b[1] * b[1]) + b[2] * b[2]);
      b[0] = b[0] * otsq;
      b[1] = b[1] * otsq;
      b[2] = b[2] * otsq;
      ut1 = ut1 + p2 * b[2];
      ut2 = ut2 + p3 * b[0];
      ut3 = ut3 + p1 * b[1];
      ut1 = ut1 - p3 * b[1];
      ut2 = ut2 - p1 * b[2];
      ut3 = ut3 - p2 * b[0];
      p[0] = ut1 + e[0];
      p[1] = ut2 + e[1];
      p[2] = ut3 + e[2];
    }
    energy = energy + loc_ene;
  }
  return energy;
}

double orc_dudt_vay(real rqm, double dt, real *u, const real *q, real *ep, real *bp, int np) {
  real tem = (real)(0.5 * dt / (double)rqm);
  double energy = 0;
  for (int c0 = 0; c0 < np; c0 += P_CACHE_SIZE) {
    int c1 = c0 + P_CACHE_SIZE < np ? c0 + P_CACHE_SIZE : np;
    double loc_ene = 0;
    for (int i = c0; i < c1; i++) {
      real *p = u + 3 * i, *e = ep + 3 * i, *b = bp + 3 * i;
      e[0] *= tem;
      e[1] *= tem;
      e[2] *= tem;
      b[0] *= tem;
      b[1] *= tem;
      b[2] *= tem;
      real rgam = R(1.0) / (real)sqrt(R(1.0) + p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
      real v1 = p[0] * rgam, v2 = p[1] * rgam, v3 = p[2] * rgam;
      real bpsq = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
      real ut1 = p[0] + e[0] + v2 * b[2] - v3 * b[1];
      real ut2 = p[1] + e[1] + v3 * b[0] - v1 * b[2];
      real ut3 = p[2] + e[2] + v1 * b[1] - v2 * b[0];
      /* temp_vec = u' */
      real w1 = ut1 + e[0], w2 = ut2 + e[1], w3 = ut3 + e[2];
      real ustar = w1 * b[0] + w2 * b[1] + w3 * b[2];
      real sigma = R(1.0) + w1 * w1 + w2 * w2 + w3 * w3 - bpsq;
      rgam = R(1.0) /
             (real)sqrt(R(0.5) * (sigma + (real)sqrt(sigma * sigma +
                                                      R(4.0) * (bpsq + ustar * ustar))));
      b[0] *= rgam;
      b[1] *= rgam;
      b[2] *= rgam;
      real spar = R(1.0) / (R(1.0) + b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
      real uptp = w1 * b[0] + w2 * b[1] + w3 * b[2];
      real o1 = p[0], o2 = p[1], o3 = p[2];
      p[0] = spar * (w1 + uptp * b[0] + w2 * b[2] - w3 * b[1]);
      p[1] = spar * (w2 + uptp * b[1] + w3 * b[0] - w1 * b[2]);
      p[2] = spar * (w3 + uptp * b[2] + w1 * b[1] - w2 * b[0]);
      real uc1 = R(0.5) * (o1 + p[0]), uc2 = R(0.5) * (o2 + p[1]), uc3 = R(0.5) * (o3 + p[2]);
      real uc2s = uc1 * uc1 + uc2 * uc2 + uc3 * uc3;
      real gamma_c = (real)sqrt(uc2s + 1);
      loc_ene = loc_ene + (double)(q[i] * uc2s / (gamma_c + 1));
    }
    energy = energy + loc_ene;
  }
  return energy;
}

/* ------------------------------------------------------------------------------------
 * ntrim, spec/os-spec-push.f03:1332-1359
 * ------------------------------------------------------------------------------------ */
static inline int ntrim(real x) { return (x < R(-0.5) ? -1 : 0) + (x >= R(0.5) ? 1 : 0); }
int orc_ntrim(real x) { return ntrim(x); }

/* ------------------------------------------------------------------------------------
 * Virtual particles and trajectory splitting.
 * t_vp2D / t_vp3D: spec/os-spec-current.f03:52-66
 * split_2d: os-spec-current.f03:659-888;  split_3d: :895-1601
 * ------------------------------------------------------------------------------------ */
typedef struct {
  real x0, y0, x1, y1, q, vz;
  int i, j;
} vp2d;
typedef struct {
  real x0, y0, z0, x1, y1, z1, q;
  int i, j, k;
} vp3d;

static int split_2d(const int *dxi, const real *xnew, const int *ixold, const real *xold,
                    const real *q, const real *rgamma, const real *u, int np, vp2d *vp) {
  int k = 0;
  for (int i = 0; i < np; i++) {
    const int d1 = dxi[2 * i], d2 = dxi[2 * i + 1];
    const real xo = xold[2 * i], yo = xold[2 * i + 1];
    const real xn = xnew[2 * i], yn = xnew[2 * i + 1];
    const int io = ixold[2 * i], jo = ixold[2 * i + 1];
    const int cross = abs(d1) + 2 * abs(d2);
    real xint, yint, xint2, yint2, delta, vz, vzint;
    switch (cross) {
    case 0:
      vp[k] = (vp2d){xo, yo, xn, yn, q[i], u[3 * i + 2] * rgamma[i], io, jo};
      k++;
      break;
    case 1:
      xint = R(0.5) * d1;
      delta = (xint - xo) / (xn - xo);
      yint = yo + (yn - yo) * delta;
      vz = u[3 * i + 2] * rgamma[i];
      vp[k++] = (vp2d){xo, yo, xint, yint, q[i], vz * delta, io, jo};
      vp[k++] = (vp2d){-xint, yint, xn - d1, yn, q[i], vz * (1 - delta), io + d1, jo};
      break;
    case 2:
      yint = R(0.5) * d2;
      delta = (yint - yo) / (yn - yo);
      xint = xo + (xn - xo) * delta;
      vz = u[3 * i + 2] * rgamma[i];
      vp[k++] = (vp2d){xo, yo, xint, yint, q[i], vz * delta, io, jo};
      vp[k++] = (vp2d){xint, -yint, xn, yn - d2, q[i], vz * (1 - delta), io, jo + d2};
      break;
    case 3:
      xint = R(0.5) * d1;
      delta = (xint - xo) / (xn - xo);
      yint = yo + (yn - yo) * delta;
      vz = u[3 * i + 2] * rgamma[i];
      if ((yint >= R(-0.5)) && (yint < R(0.5))) {
        vp[k++] = (vp2d){xo, yo, xint, yint, q[i], vz * delta, io, jo};
        vzint = vz * (1 - delta);
        yint2 = R(0.5) * d2;
        delta = (yint2 - yint) / (yn - yint);
        xint2 = -xint + (xn - xint) * delta;
        vp[k++] = (vp2d){-xint, yint, xint2, yint2, q[i], vzint * delta, io + d1, jo};
        vp[k++] = (vp2d){xint2, -yint2, xn - d1, yn - d2, q[i], vzint * (1 - delta), io + d1,
                         jo + d2};
      } else {
        vzint = vz * delta;
        yint2 = R(0.5) * d2;
        delta = (yint2 - yo) / (yint - yo);
        xint2 = xo + (xint - xo) * delta;
        vp[k++] = (vp2d){xo, yo, xint2, yint2, q[i], vzint * delta, io, jo};
        vp[k++] = (vp2d){xint2, -yint2, xint, yint - d2, q[i], vzint * (1 - delta), io, jo + d2};
        vp[k++] = (vp2d){-xint, yint - d2, xn - d1, yn - d2, q[i], vz - vzint, io + d1, jo + d2};
      }
      break;
    }
  }
  return k;
}

static inline vp3d mk3(real x0, real y0, real z0, real x1, real y1, real z1, real q, int i,
                       int j, int k) {
  vp3d v = {x0, y0, z0, x1, y1, z1, q, i, j, k};
  return v;
}

static int split_3d(const int *dxi, const real *xnew, const int *ixold, const real *xold,
                    const real *q, int np, vp3d *vp) {
  int k = 0;
  for (int i = 0; i < np; i++) {
    const int d1 = dxi[3 * i], d2 = dxi[3 * i + 1], d3 = dxi[3 * i + 2];
    const real xo = xold[3 * i], yo = xold[3 * i + 1], zo = xold[3 * i + 2];
    const real xn = xnew[3 * i], yn = xnew[3 * i + 1], zn = xnew[3 * i + 2];
    const int io = ixold[3 * i], jo = ixold[3 * i + 1], ko = ixold[3 * i + 2];
    const real qq = q[i];
    const int cross = abs(d1) + 2 * abs(d2) + 4 * abs(d3);
    real xint, yint, zint, xint2, yint2, zint2, delta;
    switch (cross) {
    case 0:
      vp[k++] = mk3(xo, yo, zo, xn, yn, zn, qq, io, jo, ko);
      break;
    case 1: /* :940-975 */
      xint = R(0.5) * d1;
      delta = (xint - xo) / (xn - xo);
      yint = yo + (yn - yo) * delta;
      zint = zo + (zn - zo) * delta;
      vp[k++] = mk3(xo, yo, zo, xint, yint, zint, qq, io, jo, ko);
      vp[k++] = mk3(-xint, yint, zint, xn - d1, yn, zn, qq, io + d1, jo, ko);
      break;
    case 2: /* :977-1012 */
      yint = R(0.5) * d2;
      delta = (yint - yo) / (yn - yo);
      xint = xo + (xn - xo) * delta;
      zint = zo + (zn - zo) * delta;
      vp[k++] = mk3(xo, yo, zo, xint, yint, zint, qq, io, jo, ko);
      vp[k++] = mk3(xint, -yint, zint, xn, yn - d2, zn, qq, io, jo + d2, ko);
      break;
    case 3: /* :1015-1133 */
      xint = R(0.5) * d1;
      delta = (xint - xo) / (xn - xo);
      yint = yo + (yn - yo) * delta;
      zint = zo + (zn - zo) * delta;
      if ((yint >= R(-0.5)) && (yint < R(0.5))) {
        vp[k++] = mk3(xo, yo, zo, xint, yint, zint, qq, io, jo, ko);
        yint2 = R(0.5) * d2;
        delta = (yint2 - yint) / (yn - yint);
        xint2 = -xint + (xn - xint) * delta;
        zint2 = zint + (zn - zint) * delta;
        vp[k++] = mk3(-xint, yint, zint, xint2, yint2, zint2, qq, io + d1, jo, ko);
        vp[k++] = mk3(xint2, -yint2, zint2, xn - d1, yn - d2, zn, qq, io + d1, jo + d2, ko);
      } else {
        yint2 = R(0.5) * d2;
        delta = (yint2 - yo) / (yint - yo);
        xint2 = xo + (xint - xo) * delta;
        zint2 = zo + (zint - zo) * delta;
        vp[k++] = mk3(xo, yo, zo, xint2, yint2, zint2, qq, io, jo, ko);
        vp[k++] = mk3(xint2, -yint2, zint2, xint, yint - d2, zint, qq, io, jo + d2, ko);
        vp[k++] = mk3(-xint, yint - d2, zint, xn - d1, yn - d2, zn, qq, io + d1, jo + d2, ko);
      }
      break;
    case 4: /* :1135-1170 */
      zint = R(0.5) * d3;
      delta = (zint - zo) / (zn - zo);
      xint = xo + (xn - xo) * delta;
      yint = yo + (yn - yo) * delta;
      vp[k++] = mk3(xo, yo, zo, xint, yint, zint, qq, io, jo, ko);
      vp[k++] = mk3(xint, yint, -zint, xn, yn, zn - d3, qq, io, jo, ko + d3);
      break;
    case 5: /* :1173-1300 */
      xint = R(0.5) * d1;
      delta = (xint - xo) / (xn - xo);
      yint = yo + (yn - yo) * delta;
      zint = zo + (zn - zo) * delta;
      zint2 = R(0.5) * d3;
      if ((zint >= R(-0.5)) && (zint < R(0.5))) {
        vp[k++] = mk3(xo, yo, zo, xint, yint, zint, qq, io, jo, ko);
        delta = (zint2 - zint) / (zn - zint);
        xint2 = -xint + (xn - xint) * delta;
        yint2 = yint + (yn - yint) * delta;
        vp[k++] = mk3(-xint, yint, zint, xint2, yint2, zint2, qq, io + d1, jo, ko);
        vp[k++] = mk3(xint2, yint2, -zint2, xn - d1, yn, zn - d3, qq, io + d1, jo, ko + d3);
      } else {
        delta = (zint2 - zo) / (zint - zo);
        xint2 = xo + (xint - xo) * delta;
        yint2 = yo + (yint - yo) * delta;
        vp[k++] = mk3(xo, yo, zo, xint2, yint2, zint2, qq, io, jo, ko);
        vp[k++] = mk3(xint2, yint2, -zint2, xint, yint, zint - d3, qq, io, jo, ko + d3);
        vp[k++] = mk3(-xint, yint, zint - d3, xn - d1, yn, zn - d3, qq, io + d1, jo, ko + d3);
      }
      break;
    case 6: /* :1302-1428 */
      yint = R(0.5) * d2;
      delta = (yint - yo) / (yn - yo);
      xint = xo + (xn - xo) * delta;
      zint = zo + (zn - zo) * delta;
      if ((zint >= R(-0.5)) && (zint < R(0.5))) {
        vp[k++] = mk3(xo, yo, zo, xint, yint, zint, qq, io, jo, ko);
        zint2 = R(0.5) * d3;
        delta = (zint2 - zint) / (zn - zint);
        xint2 = xint + (xn - xint) * delta;
        yint2 = -yint + (yn - yint) * delta;
        vp[k++] = mk3(xint, -yint, zint, xint2, yint2, zint2, qq, io, jo + d2, ko);
        vp[k++] = mk3(xint2, yint2, -zint2, xn, yn - d2, zn - d3, qq, io, jo + d2, ko + d3);
      } else {
        zint2 = R(0.5) * d3;
        delta = (zint2 - zo) / (zint - zo);
        xint2 = xo + (xint - xo) * delta;
        yint2 = yo + (yint - yo) * delta;
        vp[k++] = mk3(xo, yo, zo, xint2, yint2, zint2, qq, io, jo, ko);
        vp[k++] = mk3(xint2, yint2, -zint2, xint, yint, zint - d3, qq, io, jo, ko + d3);
        vp[k++] = mk3(xint, -yint, zint - d3, xn, yn - d2, zn - d3, qq, io, jo + d2, ko + d3);
      }
      break;
    case 7: { /* :1431-1592 */
      xint = R(0.5) * d1;
      delta = (xint - xo) / (xn - xo);
      yint = yo + (yn - yo) * delta;
      zint = zo + (zn - zo) * delta;
      if ((yint >= R(-0.5)) && (yint < R(0.5))) {
        vp[k++] = mk3(xo, yo, zo, xint, yint, zint, qq, io, jo, ko);
        yint2 = R(0.5) * d2;
        delta = (yint2 - yint) / (yn - yint);
        xint2 = -xint + (xn - xint) * delta;
        zint2 = zint + (zn - zint) * delta;
        vp[k++] = mk3(-xint, yint, zint, xint2, yint2, zint2, qq, io + d1, jo, ko);
        vp[k++] = mk3(xint2, -yint2, zint2, xn - d1, yn - d2, zn, qq, io + d1, jo + d2, ko);
      } else {
        yint2 = R(0.5) * d2;
        delta = (yint2 - yo) / (yint - yo);
        xint2 = xo + (xint - xo) * delta;
        zint2 = zo + (zint - zo) * delta;
        vp[k++] = mk3(xo, yo, zo, xint2, yint2, zint2, qq, io, jo, ko);
        vp[k++] = mk3(xint2, -yint2, zint2, xint, yint - d2, zint, qq, io, jo + d2, ko);
        vp[k++] = mk3(-xint, yint - d2, zint, xn - d1, yn - d2, zn, qq, io + d1, jo + d2, ko);
      }
      /* one of the 3 vp requires an additional z split, :1561-1590 */
      zint = R(0.5) * d3;
      for (int vpidx = k - 3; vpidx < k; vpidx++) {
        if ((vp[vpidx].z1 < R(-0.5)) || (vp[vpidx].z1 >= R(0.5))) {
          vp3d *o = &vp[vpidx];
          delta = (zint - o->z0) / (o->z1 - o->z0);
          xint = o->x0 + (o->x1 - o->x0) * delta;
          yint = o->y0 + (o->y1 - o->y0) * delta;
          vp[k] = mk3(xint, yint, -zint, o->x1, o->y1, o->z1 - d3, o->q, o->i, o->j, o->k + d3);
          k++;
          o->x1 = xint;
          o->y1 = yint;
          o->z1 = zint;
          for (int jj = vpidx + 1; jj < k - 1; jj++) {
            vp[jj].z0 = vp[jj].z0 - d3;
            vp[jj].z1 = vp[jj].z1 - d3;
            vp[jj].k = vp[jj].k + d3;
          }
          break;
        }
      }
    } break;
    }
  }
  return k;
}

/* ------------------------------------------------------------------------------------
 * Current deposition from virtual particles.
 * DEP_CURRENT_2D: spec/os-spec-current.f03:2691-2796; DEP_CURRENT_3D: :2800-2905
 * P0/P1 template params :2495-2533.  jnorm = real(dx/dt/2) (2D), real(dx/dt/3) (3D).
 * ------------------------------------------------------------------------------------ */
static void dep_vp_2d(const orc_grid *g, int order, real *j, const vp2d *vp, int nsplit,
                      real jn1, real jn2) {
  const int n = order + 1, p0 = lp_of(order);
  const real c1_3 = R(1.0) / R(3.0);
  real S0x[4], S1x[4], S0y[4], S1y[4], wl1[4], wl2[4], wp1[4], wp2[4];
  for (int i = 0; i < nsplit; i++) {
    const vp2d *v = &vp[i];
    real qnx = v->q * jn1, qny = v->q * jn2, qvz = c1_3 * (v->q * v->vz);
    orc_spline(order, v->x0, S0x);
    orc_spline(order, v->x1, S1x);
    orc_spline(order, v->y0, S0y);
    orc_spline(order, v->y1, S1y);
    orc_wl(order, qnx, v->x0, v->x1, wl1);
    orc_wl(order, qny, v->y0, v->y1, wl2);
    for (int k = 0; k < n; k++) {
      wp1[k] = S0y[k] + S1y[k];
      wp2[k] = S0x[k] + S1x[k];
    }
    for (int k2 = 0; k2 < n; k2++)
      for (int k1 = 0; k1 < n; k1++) {
        real *jj = &FLD(j, g, 0, v->i + p0 + k1, v->j + p0 + k2, 1);
        real j1 = jj[0], j2 = jj[1], j3 = jj[2];
        j1 = j1 + wl1[k1] * wp1[k2];
        j2 = j2 + wp2[k1] * wl2[k2];
        real tmp1 = S0x[k1] * S0y[k2] + S1x[k1] * S1y[k2];
        real tmp2 = S0x[k1] * S1y[k2] + S1x[k1] * S0y[k2];
        j3 = j3 + qvz * (tmp1 + R(0.5) * tmp2);
        jj[0] = j1;
        jj[1] = j2;
        jj[2] = j3;
      }
  }
}

static void dep_vp_3d(const orc_grid *g, int order, real *j, const vp3d *vp, int nsplit,
                      real jn1, real jn2, real jn3) {
  const int n = order + 1, p0 = lp_of(order);
  real S0x[4], S1x[4], S0y[4], S1y[4], S0z[4], S1z[4], wl1[4], wl2[4], wl3[4];
  real wp1[4][4], wp2[4][4], wp3[4][4];
  for (int i = 0; i < nsplit; i++) {
    const vp3d *v = &vp[i];
    real qnx = v->q * jn1, qny = v->q * jn2, qnz = v->q * jn3;
    orc_spline(order, v->x0, S0x);
    orc_spline(order, v->x1, S1x);
    orc_spline(order, v->y0, S0y);
    orc_spline(order, v->y1, S1y);
    orc_spline(order, v->z0, S0z);
    orc_spline(order, v->z1, S1z);
    orc_wl(order, qnx, v->x0, v->x1, wl1);
    orc_wl(order, qny, v->y0, v->y1, wl2);
    orc_wl(order, qnz, v->z0, v->z1, wl3);
    for (int k2 = 0; k2 < n; k2++)
      for (int k1 = 0; k1 < n; k1++) {
        wp1[k1][k2] = (S0y[k1] * S0z[k2] + S1y[k1] * S1z[k2]) +
                      R(0.5) * (S0y[k1] * S1z[k2] + S1y[k1] * S0z[k2]);
        wp2[k1][k2] = (S0x[k1] * S0z[k2] + S1x[k1] * S1z[k2]) +
                      R(0.5) * (S0x[k1] * S1z[k2] + S1x[k1] * S0z[k2]);
        wp3[k1][k2] = (S0x[k1] * S0y[k2] + S1x[k1] * S1y[k2]) +
                      R(0.5) * (S0x[k1] * S1y[k2] + S1x[k1] * S0y[k2]);
      }
    for (int k3 = 0; k3 < n; k3++)
      for (int k2 = 0; k2 < n; k2++)
        for (int k1 = 0; k1 < n; k1++) {
          real *jj = &FLD(j, g, 0, v->i + p0 + k1, v->j + p0 + k2, v->k + p0 + k3);
          real j1 = jj[0], j2 = jj[1], j3 = jj[2];
          j1 = j1 + wl1[k1] * wp1[k2][k3];
          j2 = j2 + wl2[k2] * wp2[k1][k3];
          j3 = j3 + wl3[k3] * wp3[k1][k2];
          jj[0] = j1;
          jj[1] = j2;
          jj[2] = j3;
        }
  }
}

/* ------------------------------------------------------------------------------------
 * advance_deposit_2d / 3d: spec/os-spec-push.f03:1438-1518 / 1522-1605, called from the
 * dispatcher advance_deposit :1748-1810 (dtcycle = real(dt, p_k_part) :1780).
 * pusher: 0 = Boris, 1 = Vay.  Updates ix, x, u in place, accumulates into j (caller
 * zeroes it), returns the un-normalised time-centred energy sum (Sum q*u^2/(gamma+1)).
 * ------------------------------------------------------------------------------------ */
double orc_advance_deposit(const orc_grid *g, int order, int pusher, real rqm,
                           const double *dx, double dt_in, const real *e, const real *b,
                           real *j, int *ix, real *x, real *u, const real *q, int np) {
  const int D = g->dims;
  const double gdt = (double)(real)dt_in;
  real *ep = malloc(sizeof(real) * 3 * P_CACHE_SIZE);
  real *bp = malloc(sizeof(real) * 3 * P_CACHE_SIZE);
  double energy = 0;

  /* dudt over the whole range in cache-size chunks (get_emf per chunk) */
  for (int c0 = 0; c0 < np; c0 += P_CACHE_SIZE) {
    int nc = np - c0 < P_CACHE_SIZE ? np - c0 : P_CACHE_SIZE;
    orc_get_emf(g, order, e, b, ix + D * c0, x + D * c0, nc, ep, bp);
    if (pusher == 0)
      energy += orc_dudt_boris(rqm, gdt, u + 3 * c0, q + c0, ep, bp, nc);
    else
      energy += orc_dudt_vay(rqm, gdt, u + 3 * c0, q + c0, ep, bp, nc);
  }
  free(ep);
  free(bp);

  real dt_dx[3];
  for (int d = 0; d < D; d++) dt_dx[d] = (real)(gdt / dx[d]);
  real *xbuf = malloc(sizeof(real) * D * P_CACHE_SIZE);
  int *dxi = malloc(sizeof(int) * D * P_CACHE_SIZE);
  real *rgamma = malloc(sizeof(real) * P_CACHE_SIZE);
  void *vpb = malloc(D == 2 ? sizeof(vp2d) * 3 * P_CACHE_SIZE : sizeof(vp3d) * 4 * P_CACHE_SIZE);
  for (int c0 = 0; c0 < np; c0 += P_CACHE_SIZE) {
    int nc = np - c0 < P_CACHE_SIZE ? np - c0 : P_CACHE_SIZE;
    for (int i = 0; i < nc; i++) {
      const real *p = u + 3 * (c0 + i);
      rgamma[i] = R(1.0) / (real)sqrt(((R(1.0) + p[0] * p[0]) + p[1] * p[1]) + p[2] * p[2]);
    }
    for (int i = 0; i < nc; i++) {
      const real *p = u + 3 * (c0 + i);
      for (int d = 0; d < D; d++) {
        xbuf[D * i + d] = x[D * (c0 + i) + d] + (p[d] * rgamma[i]) * dt_dx[d];
        dxi[D * i + d] = ntrim(xbuf[D * i + d]);
      }
    }
    if (D == 2) {
      int ns = split_2d(dxi, xbuf, ix + 2 * c0, x + 2 * c0, q + c0, rgamma, u + 3 * c0, nc,
                        (vp2d *)vpb);
      dep_vp_2d(g, order, j, (vp2d *)vpb, ns, (real)(dx[0] / gdt / 2), (real)(dx[1] / gdt / 2));
    } else {
      int ns = split_3d(dxi, xbuf, ix + 3 * c0, x + 3 * c0, q + c0, nc, (vp3d *)vpb);
      dep_vp_3d(g, order, j, (vp3d *)vpb, ns, (real)(dx[0] / gdt / 3), (real)(dx[1] / gdt / 3),
                (real)(dx[2] / gdt / 3));
    }
    for (int i = 0; i < nc; i++)
      for (int d = 0; d < D; d++) {
        x[D * (c0 + i) + d] = xbuf[D * i + d] - dxi[D * i + d];
        ix[D * (c0 + i) + d] = ix[D * (c0 + i) + d] + dxi[D * i + d];
      }
  }
  free(xbuf);
  free(dxi);
  free(rgamma);
  free(vpb);
  return energy;
}

/* Splitting alone, for tests: returns VP count and writes VP records
 * (x0,y0,[z0],x1,y1,[z1],q,[vz]) and cells. */
int orc_split(int dims, int np, const int *dxi, const real *xnew, const int *ixold,
              const real *xold, const real *q, const real *rgamma, const real *u, real *vpf,
              int *vpi) {
  if (dims == 2) {
    vp2d *vp = malloc(sizeof(vp2d) * 3 * (size_t)(np > 0 ? np : 1));
    int ns = split_2d(dxi, xnew, ixold, xold, q, rgamma, u, np, vp);
    for (int k = 0; k < ns; k++) {
      real *f = vpf + 6 * k;
      f[0] = vp[k].x0; f[1] = vp[k].y0; f[2] = vp[k].x1; f[3] = vp[k].y1;
      f[4] = vp[k].q; f[5] = vp[k].vz;
      vpi[2 * k] = vp[k].i; vpi[2 * k + 1] = vp[k].j;
    }
    free(vp);
    return ns;
  } else {
    vp3d *vp = malloc(sizeof(vp3d) * 4 * (size_t)(np > 0 ? np : 1));
    int ns = split_3d(dxi, xnew, ixold, xold, q, np, vp);
    for (int k = 0; k < ns; k++) {
      real *f = vpf + 7 * k;
      f[0] = vp[k].x0; f[1] = vp[k].y0; f[2] = vp[k].z0; f[3] = vp[k].x1;
      f[4] = vp[k].y1; f[5] = vp[k].z1; f[6] = vp[k].q;
      vpi[3 * k] = vp[k].i; vpi[3 * k + 1] = vp[k].j; vpi[3 * k + 2] = vp[k].k;
    }
    free(vp);
    return ns;
  }
}

/* ------------------------------------------------------------------------------------
 * Periodic boundaries, single node.
 * Grids: vdf/os-vdf-comm.f90:142-330 update_periodic_vdf, dims in order 1..x_dim
 * (:685-692).  mode 0 = p_vdf_replace (E,B), 1 = p_vdf_add (current fold).
 * Particles: spec/os-spec-boundary.f03:1280-1290 (ix += nx below 1, ix -= nx above nx).
 * ------------------------------------------------------------------------------------ */
/* directions d0..x_dim-1 only (d0 = 1 when x is not periodic) */
static void update_periodic_from(const orc_grid *g, real *f, int nc, int mode, int d0) {
  const int D = g->dims;
  for (int dim = d0; dim < D; dim++) {
    int shift[3] = {0, 0, 0};
    shift[dim] = g->nx[dim];
    int llo[3], lhi[3], ulo[3], uhi[3];
    for (int d = 0; d < 3; d++) {
      if (d >= D) {
        llo[d] = lhi[d] = ulo[d] = uhi[d] = 1;
      } else if (d == dim) {
        llo[d] = 1 - g->gcl[d];
        lhi[d] = 0;
        ulo[d] = g->nx[d] + 1;
        uhi[d] = g->nx[d] + g->gcu[d];
      } else {
        llo[d] = ulo[d] = 1 - g->gcl[d];
        lhi[d] = uhi[d] = g->nx[d] + g->gcu[d];
      }
    }
#define IDX(i1, i2, i3) (nc * cidx(g, (i1), (i2), (i3)))
    for (int i3 = llo[2]; i3 <= lhi[2]; i3++)
      for (int i2 = llo[1]; i2 <= lhi[1]; i2++)
        for (int i1 = llo[0]; i1 <= lhi[0]; i1++)
          for (int c = 0; c < nc; c++) {
            real *a = f + IDX(i1, i2, i3) + c;
            real *s = f + IDX(i1 + shift[0], i2 + shift[1], i3 + shift[2]) + c;
            if (mode == 1) {
              real val = *a + *s;
              *a = val;
              *s = val;
            } else
              *a = *s;
          }
    for (int i3 = ulo[2]; i3 <= uhi[2]; i3++)
      for (int i2 = ulo[1]; i2 <= uhi[1]; i2++)
        for (int i1 = ulo[0]; i1 <= uhi[0]; i1++)
          for (int c = 0; c < nc; c++) {
            real *a = f + IDX(i1, i2, i3) + c;
            real *s = f + IDX(i1 - shift[0], i2 - shift[1], i3 - shift[2]) + c;
            if (mode == 1) {
              real val = *a + *s;
              *s = val;
              *a = val;
            } else
              *a = *s;
          }
#undef IDX
  }
}

void orc_update_periodic(const orc_grid *g, real *f, int nc, int mode) {
  update_periodic_from(g, f, nc, mode, 0);
}

/* the same for directions d0.. only (d0 = 1: x walls that are not periodic) */
void orc_update_periodic_from(const orc_grid *g, real *f, int nc, int mode, int d0) {
  update_periodic_from(g, f, nc, mode, d0);
}

void orc_particles_periodic(const orc_grid *g, int *ix, int np) {
  for (int i = 0; i < np; i++)
    for (int d = 0; d < g->dims; d++) {
      int *v = &ix[g->dims * i + d];
      if (*v < 1)
        *v += g->nx[d];
      else if (*v > g->nx[d])
        *v -= g->nx[d];
    }
}

/* ------------------------------------------------------------------------------------
 * Yee solver.  emf/solver/os-emf-solver-yee.f03: dbdt_2d :102-138, dbdt_3d :211-251,
 * dedt_2d :287-326, dedt_3d :385-428, advance_yee :542-568 (B dt/2, E dt, B dt/2 with a
 * periodic refresh after each sub-step).
 * ------------------------------------------------------------------------------------ */
static void dbdt(const orc_grid *g, const double *dx, double dt, real *b, const real *e) {
  real d1 = (real)(dt / dx[0]), d2 = (real)(dt / dx[1]);
  if (g->dims == 2) {
    for (int i2 = 0; i2 <= g->nx[1] + 1; i2++)
      for (int i1 = 0; i1 <= g->nx[0] + 1; i1++) {
        real *bb = &FLD(b, g, 0, i1, i2, 1);
        bb[0] = bb[0] - d2 * (FLD(e, g, 2, i1, i2 + 1, 1) - FLD(e, g, 2, i1, i2, 1));
        bb[1] = bb[1] + d1 * (FLD(e, g, 2, i1 + 1, i2, 1) - FLD(e, g, 2, i1, i2, 1));
        bb[2] = bb[2] - d1 * (FLD(e, g, 1, i1 + 1, i2, 1) - FLD(e, g, 1, i1, i2, 1)) +
                d2 * (FLD(e, g, 0, i1, i2 + 1, 1) - FLD(e, g, 0, i1, i2, 1));
      }
  } else {
    real d3 = (real)(dt / dx[2]);
    for (int i3 = 0; i3 <= g->nx[2] + 1; i3++)
      for (int i2 = 0; i2 <= g->nx[1] + 1; i2++)
        for (int i1 = 0; i1 <= g->nx[0] + 1; i1++) {
          real *bb = &FLD(b, g, 0, i1, i2, i3);
          bb[0] = bb[0] - d2 * (FLD(e, g, 2, i1, i2 + 1, i3) - FLD(e, g, 2, i1, i2, i3)) +
                  d3 * (FLD(e, g, 1, i1, i2, i3 + 1) - FLD(e, g, 1, i1, i2, i3));
          bb[1] = bb[1] + d1 * (FLD(e, g, 2, i1 + 1, i2, i3) - FLD(e, g, 2, i1, i2, i3)) -
                  d3 * (FLD(e, g, 0, i1, i2, i3 + 1) - FLD(e, g, 0, i1, i2, i3));
          bb[2] = bb[2] - d1 * (FLD(e, g, 1, i1 + 1, i2, i3) - FLD(e, g, 1, i1, i2, i3)) +
                  d2 * (FLD(e, g, 0, i1, i2 + 1, i3) - FLD(e, g, 0, i1, i2, i3));
        }
  }
}

static void dedt(const orc_grid *g, const double *dx, double dt, real *e, const real *b,
                 const real *j) {
  real d1 = (real)(dt / dx[0]), d2 = (real)(dt / dx[1]), dtif = (real)dt;
  if (g->dims == 2) {
    for (int i2 = 1; i2 <= g->nx[1] + 1; i2++)
      for (int i1 = 1; i1 <= g->nx[0] + 1; i1++) {
        real *ee = &FLD(e, g, 0, i1, i2, 1);
        const real *jj = &FLD(j, g, 0, i1, i2, 1);
        ee[0] = ee[0] - dtif * jj[0] + d2 * (FLD(b, g, 2, i1, i2, 1) - FLD(b, g, 2, i1, i2 - 1, 1));
        ee[1] = ee[1] - dtif * jj[1] - d1 * (FLD(b, g, 2, i1, i2, 1) - FLD(b, g, 2, i1 - 1, i2, 1));
        ee[2] = ee[2] - dtif * jj[2] + d1 * (FLD(b, g, 1, i1, i2, 1) - FLD(b, g, 1, i1 - 1, i2, 1)) -
                d2 * (FLD(b, g, 0, i1, i2, 1) - FLD(b, g, 0, i1, i2 - 1, 1));
      }
  } else {
    real d3 = (real)(dt / dx[2]);
    for (int i3 = 1; i3 <= g->nx[2] + 1; i3++)
      for (int i2 = 1; i2 <= g->nx[1] + 1; i2++)
        for (int i1 = 1; i1 <= g->nx[0] + 1; i1++) {
          real *ee = &FLD(e, g, 0, i1, i2, i3);
          const real *jj = &FLD(j, g, 0, i1, i2, i3);
          ee[0] = ee[0] - dtif * jj[0] +
                  d2 * (FLD(b, g, 2, i1, i2, i3) - FLD(b, g, 2, i1, i2 - 1, i3)) -
                  d3 * (FLD(b, g, 1, i1, i2, i3) - FLD(b, g, 1, i1, i2, i3 - 1));
          ee[1] = ee[1] - dtif * jj[1] +
                  d3 * (FLD(b, g, 0, i1, i2, i3) - FLD(b, g, 0, i1, i2, i3 - 1)) -
                  d1 * (FLD(b, g, 2, i1, i2, i3) - FLD(b, g, 2, i1 - 1, i2, i3));
          ee[2] = ee[2] - dtif * jj[2] +
                  d1 * (FLD(b, g, 1, i1, i2, i3) - FLD(b, g, 1, i1 - 1, i2, i3)) -
                  d2 * (FLD(b, g, 0, i1, i2, i3) - FLD(b, g, 0, i1, i2 - 1, i3));
        }
  }
}

void orc_yee_advance(const orc_grid *g, const double *dx, double dt, real *e, real *b,
                     const real *j) {
  dbdt(g, dx, dt / 2.0, b, e);
  orc_update_periodic(g, b, 3, 0);
  dedt(g, dx, dt, e, b, j);
  orc_update_periodic(g, e, 3, 0);
  dbdt(g, dx, dt / 2.0, b, e);
  orc_update_periodic(g, b, 3, 0);
}

/* ------------------------------------------------------------------------------------
 * Diagnostics.
 * total_vdf(pow=2) over interior cells, vdf/os-vdf-math.f90:625-660, in double;
 * report_energy_emf: emf/diagnostics/os-emf-diagnostics.f03:44-55 (x 0.5*dvol),
 * out = (B1,B2,B3,E1,E2,E3).
 * ------------------------------------------------------------------------------------ */
void orc_field_energy(const orc_grid *g, const real *e, const real *b, double dvol,
                      double *out) {
  double t[6] = {0, 0, 0, 0, 0, 0};
  int n3 = g->dims > 2 ? g->nx[2] : 1;
  for (int i3 = 1; i3 <= n3; i3++)
    for (int i2 = 1; i2 <= g->nx[1]; i2++)
      for (int i1 = 1; i1 <= g->nx[0]; i1++)
        for (int c = 0; c < 3; c++) {
          real vb = FLD(b, g, c, i1, i2, i3), ve = FLD(e, g, c, i1, i2, i3);
          t[c] = t[c] + (double)(vb * vb);
          t[3 + c] = t[3 + c] + (double)(ve * ve);
        }
  for (int c = 0; c < 6; c++) out[c] = t[c] * dvol * 0.5;
}

/* get_energy recalculation branch, spec/os-spec-define.f03:1041-1052 (un-normalised) */
double orc_kinetic_energy(const real *u, const real *q, int np) {
  double energy = 0;
  for (int i = 0; i < np; i++) {
    double u2 = (double)u[3 * i] * u[3 * i] + (double)u[3 * i + 1] * u[3 * i + 1] +
                (double)u[3 * i + 2] * u[3 * i + 2];
    double gamma = sqrt(u2 + 1);
    energy = energy + (double)q[i] * (u2 / (gamma + 1));
  }
  return energy;
}

/* DEPOSIT_RHO_2D/3D, spec/os-spec-charge.f90:500-580: rho on the lower cell corner with
 * spline weights of stencil LP..UP.  rho is a one-component grid (nc = 1). */
void orc_deposit_rho(const orc_grid *g, int order, const int *ix, const real *x,
                     const real *q, int np, real *rho) {
  const int n = order + 1, lp = lp_of(order), D = g->dims;
  real w1[4], w2[4], w3[4];
  for (int l = 0; l < np; l++) {
    orc_spline(order, x[D * l], w1);
    orc_spline(order, x[D * l + 1], w2);
    if (D == 3) orc_spline(order, x[D * l + 2], w3);
    int n3 = D == 3 ? n : 1;
    for (int k3 = 0; k3 < n3; k3++)
      for (int k2 = 0; k2 < n; k2++)
        for (int k1 = 0; k1 < n; k1++) {
          long c = cidx(g, ix[D * l] + lp + k1, ix[D * l + 1] + lp + k2,
                        D == 3 ? ix[D * l + 2] + lp + k3 : 1);
          if (D == 2)
            rho[c] = rho[c] + q[l] * w1[k1] * w2[k2];
          else
            rho[c] = rho[c] + q[l] * w1[k1] * w2[k2] * w3[k3];
        }
  }
}

/* div_vdf with offset 0, vdf/os-vdf-math.f90:1013-1073; div is a one-component grid */
void orc_div(const orc_grid *g, const double *dx, const real *f, real *div) {
  real r1 = (real)(1.0 / dx[0]), r2 = (real)(1.0 / dx[1]);
  real r3 = g->dims > 2 ? (real)(1.0 / dx[2]) : 0;
  int n3 = g->dims > 2 ? g->nx[2] : 1;
  for (int i3 = 1; i3 <= n3; i3++)
    for (int i2 = 1; i2 <= g->nx[1]; i2++)
      for (int i1 = 1; i1 <= g->nx[0]; i1++) {
        real v = (FLD(f, g, 0, i1, i2, i3) - FLD(f, g, 0, i1 - 1, i2, i3)) * r1 +
                 (FLD(f, g, 1, i1, i2, i3) - FLD(f, g, 1, i1, i2 - 1, i3)) * r2;
        if (g->dims > 2) v = v + (FLD(f, g, 2, i1, i2, i3) - FLD(f, g, 2, i1, i2, i3 - 1)) * r3;
        div[cidx(g, i1, i2, i3)] = v;
      }
}

/* ------------------------------------------------------------------------------------
 * One full time step of the per-step order of os-simulation.f03:498-666 for periodic
 * boundaries, no smoothing, no external fields: E/B guard refresh (:561), push+deposit
 * into a zeroed J (:588, os-particles.f03:1453-1466), particle wrap (:593), energy report
 * of E^n/B^n and the time-centred kinetic energy (:596-597), J fold (:652), Yee (:661).
 *
 * Species are passed as arrays of pointers.  fld_ene[6] and par_ene[nspec] receive the
 * step's HIST rows (normalised as report_energy_* does).
 * ------------------------------------------------------------------------------------ */
/* Current smoothing, one call = smooth_vdf (vdf/os-vdf-smooth.f90:515-540) on a grid array
 * with nc components: filter_1d/2d/3d (:545-1100).  Each direction d whose order
 * n = order[d] is positive is filtered in turn (x, then y, then z) along every line of the
 * whole array, guard cells included (ext_lb..ext_ub in the other directions).  Along the
 * line, cells int_lb..int_ub = ext_lb+n .. ext_ub-n are replaced by
 *   s(-n) o(i-n) + s(-n+1) o(i-n+1) + ... + s(n) o(i+n)
 * summed left to right in the working precision, o being the values of the line before
 * the pass (the reference's rolling buffer o(:,-n:n) holds exactly those); the other cells
 * keep their values.  coef[31 d + 15 + k] = s(k) of direction d (scoef(:, d) centred). */
void orc_smooth(const orc_grid *g, int nc, real *f, const int *order, const real *coef) {
  const int D = g->dims;
  long ld[3] = {1, 1, 1};
  for (int d = 0; d < D; d++) ld[d] = ld_(g, d);
  const long stride[3] = {nc, nc * ld[0], nc * ld[0] * ld[1]};
  long lmax = ld[0];
  if (ld[1] > lmax) lmax = ld[1];
  if (ld[2] > lmax) lmax = ld[2];
  real *line = (real *)malloc(sizeof(real) * (size_t)lmax);
  for (int d = 0; d < D; d++) {
    const int n = order[d];
    if (n <= 0) continue;
    const real *sc = coef + 31 * d + 15;
    const int a = d == 0 ? 1 : 0, b = d == 2 ? 1 : 2;  /* the two other directions */
    for (long jb = 0; jb < ld[b]; jb++)
      for (long ja = 0; ja < ld[a]; ja++)
        for (int c = 0; c < nc; c++) {
          real *base = f + c + ja * stride[a] + jb * stride[b];  /* ld[b] = 1 beyond D */
          for (long i = 0; i < ld[d]; i++) line[i] = base[i * stride[d]];
          for (long i = n; i <= ld[d] - 1 - n; i++) {
            real acc = line[i - n] * sc[-n];
            for (int k = -n + 1; k <= n; k++) acc = acc + line[i + k] * sc[k];
            base[i * stride[d]] = acc;
          }
        }
  }
  free(line);
}

void orc_step(const orc_grid *g, int order, const double *dx, double dt, real *e, real *b,
              real *j, int nspec, const int *pusher, const real *rqm, int *const *ix,
              real *const *x, real *const *u, real *const *q, const int *np, double *fld_ene,
              double *par_ene, const int *sm_order, const real *sm_coef) {
  double dvol = dx[0] * dx[1] * (g->dims > 2 ? dx[2] : 1.0);
  orc_update_periodic(g, e, 3, 0);
  orc_update_periodic(g, b, 3, 0);
  memset(j, 0, sizeof(real) * 3 * orc_grid_cells(g));
  for (int s = 0; s < nspec; s++) {
    double ene = orc_advance_deposit(g, order, pusher[s], rqm[s], dx, dt, e, b, j, ix[s], x[s],
                                     u[s], q[s], np[s]);
    orc_particles_periodic(g, ix[s], np[s]);
    if (par_ene) par_ene[s] = ene * (double)rqm[s] * dvol;
  }
  if (fld_ene) orc_field_energy(g, e, b, dvol, fld_ene);
  orc_update_periodic(g, j, 3, 1);
  /* os-simulation.f03:650-657: the folded current is smoothed before the field advance */
  if (sm_order) orc_smooth(g, 3, j, sm_order, sm_coef);
  orc_yee_advance(g, dx, dt, e, b, j);
}

/* ------------------------------------------------------------------------------------
 * Open x walls, 2D and 3D: Lindman absorbing field boundaries on both x walls
 * (emf/boundary/os-emf-lindman.f90) and absorbing particles
 * (spec/os-spec-boundary.f03:150-151, 476-513: particles with ix < 1 or ix > nx after
 * the push are removed); y (and z) stay periodic.
 *
 * The wall buffers are t_wall objects (vdf/os-wall.f90:111-215): zero at creation,
 * x range (0,1) at the lower wall and (nx, nx+2) at the upper one
 * (os-emf-lindman.f90:308-324), y index 1-gcl..ny+gcu (z likewise).  They are stored here
 * as (3, ncol, ld1[, ld2]) with ncol = 2 (lower) or 3 (upper); the upper wall's B buffer
 * is never read (os-emf-lindman.f90:572-573, 994-1011) and is not kept.
 * ------------------------------------------------------------------------------------ */
typedef struct {
  real disp_corr;  /* (1 - dt/dx1) / (1 + dt/dx1), os-emf-lindman.f90:237-238 */
  real dtdxi[3];   /* real(dt/dx, p_k_fld) * 0.5, os-emf-lindman.f90:236, 241 */
  real *ebl, *bbl; /* lower wall E and B buffers, x = 0, 1 */
  real *ebu;       /* upper wall E buffer, x = nx, nx+1, nx+2 */
} orc_lindman;

/* wall buffer entry: component c, x column col, y index j, z index k (k = 1 in 2D) */
static inline long wb_idx(const orc_grid *g, int ncol, int c, int col, int j, int k) {
  long ld1 = ld_(g, 1);
  long kk = g->dims > 2 ? (long)(k - 1 + g->gcl[2]) : 0;
  return 3 * ((long)col + ncol * ((long)(j - 1 + g->gcl[1]) + ld1 * kk)) + c;
}
#define WB(w, ncol, c, col, j, k) ((w)[wb_idx(g, ncol, c, col, j, k)])
#define Z3(g) ((g)->dims > 2)

/* update_periodic_wall with p_vdf_replace along y, then z (vdf/os-wall-comm.f90:182-344):
 * guard entries of the buffer take the periodic images, over the full range of the
 * other direction */
static void wall_periodic(const orc_grid *g, real *w, int ncol) {
  const int ny = g->nx[1], nz = Z3(g) ? g->nx[2] : 1;
  const int klo = Z3(g) ? 1 - g->gcl[2] : 1, khi = Z3(g) ? nz + g->gcu[2] : 1;
  for (int k = klo; k <= khi; k++)
    for (int j = 1 - g->gcl[1]; j <= ny + g->gcu[1]; j++) {
      int s = j < 1 ? j + ny : (j > ny ? j - ny : j);
      if (s == j) continue;
      for (int col = 0; col < ncol; col++)
        for (int c = 0; c < 3; c++) WB(w, ncol, c, col, j, k) = WB(w, ncol, c, col, s, k);
    }
  if (!Z3(g)) return;
  for (int k = klo; k <= khi; k++) {
    int s = k < 1 ? k + nz : (k > nz ? k - nz : k);
    if (s == k) continue;
    for (int j = 1 - g->gcl[1]; j <= ny + g->gcu[1]; j++)
      for (int col = 0; col < ncol; col++)
        for (int c = 0; c < 3; c++) WB(w, ncol, c, col, j, k) = WB(w, ncol, c, col, j, s);
  }
}

/* update_lindman_2d_e / update_lindman_3d_e, idir = 1 (os-emf-lindman.f90:454-535,
 * 604-750): perp = components 2, 3, over y = 1..ny (and z = 1..nz) */
static void lindman_e(const orc_grid *g, orc_lindman *w, real *e) {
  const int nx = g->nx[0], nz = Z3(g) ? g->nx[2] : 1;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= g->nx[1]; j++)
      for (int p = 1; p <= 2; p++) {
        /* lower wall, i_range = (0, 1) */
        FLD(e, g, p, 0, j, k) = WB(w->ebl, 2, p, 1, j, k) +
                                w->disp_corr * (WB(w->ebl, 2, p, 0, j, k) - FLD(e, g, p, 1, j, k));
        WB(w->ebl, 2, p, 0, j, k) = FLD(e, g, p, 0, j, k);
        WB(w->ebl, 2, p, 1, j, k) = FLD(e, g, p, 1, j, k);
      }
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= g->nx[1]; j++)
      for (int p = 1; p <= 2; p++) {
        /* upper wall, i_range = (nx, nx+2): buffer columns 1, 2 are x = nx+1, nx+2 */
        FLD(e, g, p, nx + 2, j, k) = WB(w->ebu, 3, p, 1, j, k) +
            w->disp_corr * (WB(w->ebu, 3, p, 2, j, k) - FLD(e, g, p, nx + 1, j, k));
        WB(w->ebu, 3, p, 1, j, k) = FLD(e, g, p, nx + 1, j, k);
        WB(w->ebu, 3, p, 2, j, k) = FLD(e, g, p, nx + 2, j, k);
      }
}

/* update_lindman_2d_b (os-emf-lindman.f90:539-600) / update_lindman_3d_b (:755-857),
 * idir = 1, the lower wall's buffer (the 3D routine also advances the upper wall's,
 * which is never read) */
static void lindman_b(const orc_grid *g, orc_lindman *w, const real *e) {
  const real d1 = w->dtdxi[0], d2 = w->dtdxi[1], d3 = w->dtdxi[2];
  const int nz = Z3(g) ? g->nx[2] : 1;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= g->nx[1]; j++) {
      real *bb = &WB(w->bbl, 2, 0, 0, j, k);
      if (!Z3(g)) {
        bb[0] = bb[0] - d2 * (FLD(e, g, 2, 0, j + 1, 1) - FLD(e, g, 2, 0, j, 1));
        bb[1] = bb[1] + d1 * (FLD(e, g, 2, 1, j, 1) - FLD(e, g, 2, 0, j, 1));
        bb[2] = bb[2] - d1 * (FLD(e, g, 1, 1, j, 1) - FLD(e, g, 1, 0, j, 1)) +
                d2 * (FLD(e, g, 0, 0, j + 1, 1) - FLD(e, g, 0, 0, j, 1));
      } else {
        bb[0] = bb[0] - d2 * (FLD(e, g, 2, 0, j + 1, k) - FLD(e, g, 2, 0, j, k)) +
                d3 * (FLD(e, g, 1, 0, j, k + 1) - FLD(e, g, 1, 0, j, k));
        bb[1] = bb[1] - d3 * (FLD(e, g, 0, 0, j, k + 1) - FLD(e, g, 0, 0, j, k)) +
                d1 * (FLD(e, g, 2, 1, j, k) - FLD(e, g, 2, 0, j, k));
        bb[2] = bb[2] - d1 * (FLD(e, g, 1, 1, j, k) - FLD(e, g, 1, 0, j, k)) +
                d2 * (FLD(e, g, 0, 0, j + 1, k) - FLD(e, g, 0, 0, j, k));
      }
    }
}

/* update_boundary_emf_bound for open x, periodic y (and z), one node,
 * emf/boundary/os-emf-boundary.f90:723-830: wall buffers refreshed along y, z
 * (:768-769, os-emf-lindman.f90:352-368), then the x walls restore the wall values
 * (update_emfbound_lindman, os-emf-lindman.f90:861-1011: B and E at x = 0 from the
 * lower buffers, E at x = nx+1 from the upper one, every y, z), then the periodic
 * refresh of y (and z). */
void orc_open_boundary(const orc_grid *g, orc_lindman *w, real *e, real *b) {
  wall_periodic(g, w->bbl, 2);
  wall_periodic(g, w->ebl, 2);
  wall_periodic(g, w->ebu, 3);
  const int nx = g->nx[0];
  const int klo = Z3(g) ? 1 - g->gcl[2] : 1, khi = Z3(g) ? g->nx[2] + g->gcu[2] : 1;
  for (int k = klo; k <= khi; k++)
    for (int j = 1 - g->gcl[1]; j <= g->nx[1] + g->gcu[1]; j++)
      for (int c = 0; c < 3; c++) {
        FLD(b, g, c, 0, j, k) = WB(w->bbl, 2, c, 0, j, k);
        FLD(e, g, c, 0, j, k) = WB(w->ebl, 2, c, 0, j, k);
        FLD(e, g, c, nx + 1, j, k) = WB(w->ebu, 3, c, 1, j, k);
      }
  update_periodic_from(g, e, 3, 0, 1);
  update_periodic_from(g, b, 3, 0, 1);
}
#undef WB
#undef Z3

/* advance_yee with Lindman x walls, os-emf-solver-yee.f03:542-568 with
 * update_boundary_bfld/efld (os-emf-boundary.f90:840-1112): no guard refresh between the
 * sub-steps. */
void orc_yee_advance_open(const orc_grid *g, const double *dx, double dt, real *e, real *b,
                          const real *j, orc_lindman *w) {
  dbdt(g, dx, dt / 2.0, b, e);
  lindman_b(g, w, e);
  dedt(g, dx, dt, e, b, j);
  lindman_e(g, w, e);
  dbdt(g, dx, dt / 2.0, b, e);
  lindman_b(g, w, e);
}

/* Particle boundaries: y, z periodic (as orc_particles_periodic), x absorbing: particles with
 * ix < 1 or ix > nx are removed, the others keep their order.  Returns the new count. */
int orc_particles_open(const orc_grid *g, int *ix, real *x, real *u, real *q, int np) {
  const int D = g->dims;
  int k = 0;
  for (int i = 0; i < np; i++) {
    for (int d = 1; d < D; d++) {
      int *v = &ix[D * i + d];
      if (*v < 1)
        *v += g->nx[d];
      else if (*v > g->nx[d])
        *v -= g->nx[d];
    }
    if (ix[D * i] < 1 || ix[D * i] > g->nx[0]) continue;
    if (k != i) {
      for (int d = 0; d < D; d++) {
        ix[D * k + d] = ix[D * i + d];
        x[D * k + d] = x[D * i + d];
      }
      for (int c = 0; c < 3; c++) u[3 * k + c] = u[3 * i + c];
      q[k] = q[i];
    }
    k++;
  }
  return k;
}

/* ------------------------------------------------------------------------------------
 * VPML absorbing x walls, 2D and 3D (emf/boundary/os-emf-vpml.f03; emf_bound type "vpml"
 * or "open", os-emf-boundary.f90:262-263), y (and z) periodic, no corners; in 2D with the
 * optional diffusion of the split B3 (vpml_diffuse, p_difrate = 0.1, :51, 977-1006).
 *
 * Each wall holds split E and B fields (os-emf-vpml.f03:440-452): in 2D the components
 * 1 = X, 2 = Y, 3 = ZX, 4 = ZY; in 3D 1 = XY, 2 = XZ, 3 = YX, 4 = YZ, 5 = ZX, 6 = ZY.  The
 * walls cover n = vpml_bnd_size x cells: x = 2-n .. 1 at the lower wall, nx+1 .. nx+n at
 * the upper one (setup_vpml :356-366); y index 1-gc1 .. ny+gc2 and z index 1-gc3 ..
 * nz+gc4 with lower guards max(field gcl, 1) and upper guards max(field gcu, 2) (t_wall
 * new with corner = 2, vdf/os-wall.f90:189-191).  Stored as (z row, y row, l, component)
 * with l = x - x0.  All zero at creation.  The attenuation coefficients
 * (setup_att_coeff :3586-3689) are indexed [wall][E or B][1..3][cell], cell 0-based.
 * ------------------------------------------------------------------------------------ */
#define ORC_VPML_MAX 64
typedef struct {
  int n;                              /* vpml_bnd_size */
  int gc1, gc2, gc3, gc4;             /* wall guard rows: y lower, upper, z lower, upper */
  int diffuse[2];                     /* vpml_diffuse of the lower, upper wall (2D only) */
  real dtdx2, dtdx3;                  /* real(dt/dx2), real(dt/dx3) */
  real coef[2][2][3][ORC_VPML_MAX];   /* [lower, upper][E, B][coefficient][cell] */
  real *w[2][2];                      /* [lower, upper][E, B] wall fields */
} orc_vpml;

static real rexp(real v) { return sizeof(real) == sizeof(float) ? (real)expf((float)v) : (real)exp((double)v); }

/* setup_att_coeff (os-emf-vpml.f03:3586-3689) for direction 1: p_att_nexp = 4 (:48),
 * s1 = -(0.8 (nexp+1)) / 2, s2 = 1 / (n-3); the power is the integer power t^4 = (t t)(t t)
 * and every operation is in the field precision.  dtdx1 = real(dt/dx1). */
void orc_vpml_coefs(orc_vpml *v, real dtdx1) {
  const int n = v->n;
  const real s1 = -(R(0.8) * (R(4) + R(1))) / R(2);
  const real s2 = R(1) / (real)(n - 3);
  const real dxdtdif = (R(1) - dtdx1) / (R(1) + dtdx1);
#define P4(t) (((t) * (t)) * ((t) * (t)))
  for (int j = 0; j < n; j++) {
    const real rj = (real)j;
    real t, tj, tj12, tj1, b1e, b1b;
    /* lower wall: B is the outside point */
    t = s2 * (rj - R(0.5)); tj = rexp(s1 * P4(t));
    t = s2 * rj;            tj12 = rexp(s1 * P4(t));
    t = s2 * (rj + R(0.5)); tj1 = rexp(s1 * P4(t));
    b1e = dtdx1 * (R(1) + dxdtdif * (R(1) - tj12));
    b1b = dtdx1 * (R(1) + dxdtdif * (R(1) - tj1));
    const int k = n - 1 - j;
    v->coef[0][0][0][k] = R(1) - b1e + tj12 * dtdx1;
    v->coef[0][0][1][k] = tj * b1e;
    v->coef[0][0][2][k] = dtdx1;
    v->coef[0][1][0][k] = R(1) - b1b + tj1 * dtdx1;
    v->coef[0][1][1][k] = tj12 * b1b;
    v->coef[0][1][2][k] = dtdx1;
    /* upper wall: E is the outside point */
    t = s2 * rj;            tj = rexp(s1 * P4(t));
    t = s2 * (rj + R(0.5)); tj12 = rexp(s1 * P4(t));
    t = s2 * (rj + R(1.0)); tj1 = rexp(s1 * P4(t));
    b1e = dtdx1 * (R(1) + dxdtdif * (R(1) - tj12));
    b1b = dtdx1 * (R(1) + dxdtdif * (R(1) - tj1));
    v->coef[1][0][0][j] = R(1) - b1e + tj12 * dtdx1;
    v->coef[1][0][1][j] = dtdx1;
    v->coef[1][0][2][j] = tj * b1e;
    v->coef[1][1][0][j] = R(1) - b1b + tj1 * dtdx1;
    v->coef[1][1][1][j] = dtdx1;
    v->coef[1][1][2][j] = tj12 * b1b;
  }
#undef P4
}

static inline int vp_x0(const orc_grid *g, const orc_vpml *v, int wl) {
  return wl == 0 ? 2 - v->n : g->nx[0] + 1;
}
static inline long vp_idx(const orc_grid *g, const orc_vpml *v, int wl, int c, int x, int i2,
                          int i3) {
  const int NC = g->dims > 2 ? 6 : 4;
  const long ld2 = (long)g->nx[1] + v->gc1 + v->gc2;
  const long r3 = g->dims > 2 ? (long)(i3 - 1 + v->gc3) : 0;
  return ((r3 * ld2 + (i2 - 1 + v->gc1)) * v->n + (x - vp_x0(g, v, wl))) * NC + c;
}
/* wall entry: wall wl, E (0) or B (1), component c (0-based), Fortran x, y, z indices */
#define VW(v, g, wl, eb, c, x, i2, i3) ((v)->w[wl][eb][vp_idx(g, v, wl, c, x, i2, i3)])

/* update_periodic_wall along y, then z, with p_vdf_replace (vdf/os-wall-comm.f90:182-344) */
static void vpml_periodic(const orc_grid *g, orc_vpml *v) {
  const int ny = g->nx[1], NC = g->dims > 2 ? 6 : 4;
  const int nz = g->dims > 2 ? g->nx[2] : 1;
  const int klo = g->dims > 2 ? 1 - v->gc3 : 1, khi = g->dims > 2 ? nz + v->gc4 : 1;
  for (int wl = 0; wl < 2; wl++)
    for (int eb = 0; eb < 2; eb++) {
      for (int i3 = klo; i3 <= khi; i3++)
        for (int i2 = 1 - v->gc1; i2 <= ny + v->gc2; i2++) {
          const int s = i2 < 1 ? i2 + ny : (i2 > ny ? i2 - ny : i2);
          if (s == i2) continue;
          for (int l = 0; l < v->n; l++) {
            const int x = vp_x0(g, v, wl) + l;
            for (int c = 0; c < NC; c++)
              VW(v, g, wl, eb, c, x, i2, i3) = VW(v, g, wl, eb, c, x, s, i3);
          }
        }
      if (g->dims < 3) continue;
      for (int i3 = klo; i3 <= khi; i3++) {
        const int s = i3 < 1 ? i3 + nz : (i3 > nz ? i3 - nz : i3);
        if (s == i3) continue;
        for (int i2 = 1 - v->gc1; i2 <= ny + v->gc2; i2++)
          for (int l = 0; l < v->n; l++) {
            const int x = vp_x0(g, v, wl) + l;
            for (int c = 0; c < NC; c++)
              VW(v, g, wl, eb, c, x, i2, i3) = VW(v, g, wl, eb, c, x, i2, s);
          }
      }
    }
}

/* update_interface_x1_2d (os-emf-vpml.f03:2880-2997) / update_interface_x1_3d
 * (:2997-3186) without corners.  The box value at box_to_vpml enters the wall: in 2D over
 * y = 0 .. ny+2 with ZX taking the whole Z and ZY zeroed; in 3D over y = 1 .. ny+1,
 * z = 1 .. nz+1 with XY, YX, ZX taking the whole components and XZ, YZ, ZY zeroed.  The
 * wall value at vpml_to_box enters the box as the sum of its split pair, over y (and z)
 * = 0 .. n+2. */
static void vpml_interface(const orc_grid *g, orc_vpml *v, int eb, real *f) {
  const int n = v->n, ny = g->nx[1], D3 = g->dims > 2;
  const int nz = D3 ? g->nx[2] : 1;
  for (int wl = 0; wl < 2; wl++) {
    const int x0 = vp_x0(g, v, wl);
    const int btv = wl == 0 ? x0 + n - 1 : x0, vtb = wl == 0 ? x0 + n - 2 : x0 + 1;
    if (!D3) {
      for (int i2 = 0; i2 <= ny + 2; i2++) {
        VW(v, g, wl, eb, 0, btv, i2, 1) = FLD(f, g, 0, btv, i2, 1);
        VW(v, g, wl, eb, 1, btv, i2, 1) = FLD(f, g, 1, btv, i2, 1);
        VW(v, g, wl, eb, 2, btv, i2, 1) = FLD(f, g, 2, btv, i2, 1);
        VW(v, g, wl, eb, 3, btv, i2, 1) = R(0);
      }
      for (int i2 = 0; i2 <= ny + 2; i2++) {
        FLD(f, g, 0, vtb, i2, 1) = VW(v, g, wl, eb, 0, vtb, i2, 1);
        FLD(f, g, 1, vtb, i2, 1) = VW(v, g, wl, eb, 1, vtb, i2, 1);
        FLD(f, g, 2, vtb, i2, 1) = VW(v, g, wl, eb, 2, vtb, i2, 1) + VW(v, g, wl, eb, 3, vtb, i2, 1);
      }
      continue;
    }
    for (int i3 = 1; i3 <= nz + 1; i3++)
      for (int i2 = 1; i2 <= ny + 1; i2++)
        for (int c = 0; c < 3; c++) {
          VW(v, g, wl, eb, 2 * c, btv, i2, i3) = FLD(f, g, c, btv, i2, i3);
          VW(v, g, wl, eb, 2 * c + 1, btv, i2, i3) = R(0);
        }
    for (int i3 = 0; i3 <= nz + 2; i3++)
      for (int i2 = 0; i2 <= ny + 2; i2++)
        for (int c = 0; c < 3; c++)
          FLD(f, g, c, vtb, i2, i3) =
              VW(v, g, wl, eb, 2 * c, vtb, i2, i3) + VW(v, g, wl, eb, 2 * c + 1, vtb, i2, i3);
  }
}

#define EW(c, xx, yy, zz) VW(v, g, wl, 0, c, xx, yy, zz)
#define BW(c, xx, yy, zz) VW(v, g, wl, 1, c, xx, yy, zz)
/* update_e_vpml_2d / 3d (os-emf-vpml.f03:598-722, 726-856) for the two x walls, after the
 * first B half step: interface of B, then update_e_x1_2d (:925-1048) / update_e_x1_3d
 * (:1172-1365), standard (non-corner) solver, x over (x0+1 .. x0+n-2) at the lower wall
 * and (x0+1 .. x0+n-1) at the upper one */
static void vpml_update_e(const orc_grid *g, orc_vpml *v, real *b) {
  vpml_interface(g, v, 1, b);
  const int n = v->n, ny = g->nx[1], D3 = g->dims > 2;
  const int klo = D3 ? 2 - v->gc3 : 1, khi = D3 ? g->nx[2] + v->gc4 : 1;
  const real d2 = v->dtdx2, d3 = v->dtdx3;
  for (int wl = 0; wl < 2; wl++) {
    const int x0 = vp_x0(g, v, wl);
    const int lo = x0 + 1, hi = wl == 0 ? x0 + n - 2 : x0 + n - 1;
    const real *ce = v->coef[wl][0][0];
    for (int k = klo; k <= khi; k++)
      for (int j = 2 - v->gc1; j <= ny + v->gc2; j++)
        for (int x = lo; x <= hi; x++) {
          /* coefficient cell: coef_i = x - coef_std, coef_std = x0 - 1 (lower), x0 (upper) */
          const int ci = wl == 0 ? x - x0 : x - x0 - 1;
          const real c1 = ce[ci], c2 = ce[ORC_VPML_MAX + ci], c3 = ce[2 * ORC_VPML_MAX + ci];
          if (!D3) {
            if (v->diffuse[wl]) {  /* before the cell's update, in loop order (:1001-1006) */
              const real rate = R(0.1), keep = R(1) - rate, zx = BW(2, x, j, 1);
              BW(2, x, j, 1) = keep * BW(2, x, j, 1) + rate * BW(3, x, j, 1);
              BW(3, x, j, 1) = keep * BW(3, x, j, 1) + rate * zx;
            }
            EW(0, x, j, 1) = EW(0, x, j, 1) + d2 * (BW(2, x, j, 1) + BW(3, x, j, 1) -
                                                    BW(2, x, j - 1, 1) - BW(3, x, j - 1, 1));
            EW(3, x, j, 1) = EW(3, x, j, 1) - d2 * (BW(0, x, j, 1) - BW(0, x, j - 1, 1));
            EW(1, x, j, 1) = c1 * EW(1, x, j, 1) - c2 * (BW(2, x, j, 1) + BW(3, x, j, 1)) +
                             c3 * (BW(2, x - 1, j, 1) + BW(3, x - 1, j, 1));
            EW(2, x, j, 1) = c1 * EW(2, x, j, 1) + c2 * BW(1, x, j, 1) - c3 * BW(1, x - 1, j, 1);
            continue;
          }
          EW(0, x, j, k) = EW(0, x, j, k) + d2 * (BW(4, x, j, k) + BW(5, x, j, k) -
                                                  BW(4, x, j - 1, k) - BW(5, x, j - 1, k));
          EW(5, x, j, k) = EW(5, x, j, k) - d2 * (BW(0, x, j, k) + BW(1, x, j, k) -
                                                  BW(0, x, j - 1, k) - BW(1, x, j - 1, k));
          EW(1, x, j, k) = EW(1, x, j, k) - d3 * (BW(2, x, j, k) + BW(3, x, j, k) -
                                                  BW(2, x, j, k - 1) - BW(3, x, j, k - 1));
          EW(3, x, j, k) = EW(3, x, j, k) + d3 * (BW(0, x, j, k) + BW(1, x, j, k) -
                                                  BW(0, x, j, k - 1) - BW(1, x, j, k - 1));
          EW(2, x, j, k) = c1 * EW(2, x, j, k) - c2 * (BW(4, x, j, k) + BW(5, x, j, k)) +
                           c3 * (BW(4, x - 1, j, k) + BW(5, x - 1, j, k));
          EW(4, x, j, k) = c1 * EW(4, x, j, k) + c2 * (BW(2, x, j, k) + BW(3, x, j, k)) -
                           c3 * (BW(2, x - 1, j, k) + BW(3, x - 1, j, k));
        }
  }
}

/* update_b_vpml_2d / 3d (os-emf-vpml.f03:1742-1861, 1865-...) after the E step: interface
 * of E, then update_b_x1_2d (:2066-2181) / update_b_x1_3d (:2311-2508), x over
 * (x0 .. x0+n-2) at the lower wall and (x0+1 .. x0+n-2) at the upper one */
static void vpml_update_b(const orc_grid *g, orc_vpml *v, real *e) {
  vpml_interface(g, v, 0, e);
  const int n = v->n, ny = g->nx[1], D3 = g->dims > 2;
  const int klo = D3 ? 1 - v->gc3 : 1, khi = D3 ? g->nx[2] + v->gc4 - 1 : 1;
  const real d2 = v->dtdx2, d3 = v->dtdx3;
  for (int wl = 0; wl < 2; wl++) {
    const int x0 = vp_x0(g, v, wl);
    const int lo = wl == 0 ? x0 : x0 + 1, hi = x0 + n - 2;
    const real *cb = v->coef[wl][1][0];
    for (int k = klo; k <= khi; k++)
      for (int j = 1 - v->gc1; j <= ny + v->gc2 - 1; j++)
        for (int x = lo; x <= hi; x++) {
          /* coef_std = x0 - 2 (lower), x0 (upper) */
          const int ci = wl == 0 ? x - x0 + 1 : x - x0 - 1;
          const real c1 = cb[ci], c2 = cb[ORC_VPML_MAX + ci], c3 = cb[2 * ORC_VPML_MAX + ci];
          if (!D3) {
            BW(0, x, j, 1) = BW(0, x, j, 1) - d2 * (EW(2, x, j + 1, 1) + EW(3, x, j + 1, 1) -
                                                    EW(2, x, j, 1) - EW(3, x, j, 1));
            BW(3, x, j, 1) = BW(3, x, j, 1) + d2 * (EW(0, x, j + 1, 1) - EW(0, x, j, 1));
            BW(1, x, j, 1) = c1 * BW(1, x, j, 1) + c2 * (EW(2, x + 1, j, 1) + EW(3, x + 1, j, 1)) -
                             c3 * (EW(2, x, j, 1) + EW(3, x, j, 1));
            BW(2, x, j, 1) = c1 * BW(2, x, j, 1) - c2 * EW(1, x + 1, j, 1) + c3 * EW(1, x, j, 1);
            continue;
          }
          BW(0, x, j, k) = BW(0, x, j, k) - d2 * (EW(4, x, j + 1, k) + EW(5, x, j + 1, k) -
                                                  EW(4, x, j, k) - EW(5, x, j, k));
          BW(5, x, j, k) = BW(5, x, j, k) + d2 * (EW(0, x, j + 1, k) + EW(1, x, j + 1, k) -
                                                  EW(0, x, j, k) - EW(1, x, j, k));
          BW(1, x, j, k) = BW(1, x, j, k) + d3 * (EW(2, x, j, k + 1) + EW(3, x, j, k + 1) -
                                                  EW(2, x, j, k) - EW(3, x, j, k));
          BW(3, x, j, k) = BW(3, x, j, k) - d3 * (EW(0, x, j, k + 1) + EW(1, x, j, k + 1) -
                                                  EW(0, x, j, k) - EW(1, x, j, k));
          BW(2, x, j, k) = c1 * BW(2, x, j, k) + c2 * (EW(4, x + 1, j, k) + EW(5, x + 1, j, k)) -
                           c3 * (EW(4, x, j, k) + EW(5, x, j, k));
          BW(4, x, j, k) = c1 * BW(4, x, j, k) - c2 * (EW(2, x + 1, j, k) + EW(3, x + 1, j, k)) +
                           c3 * (EW(2, x, j, k) + EW(3, x, j, k));
        }
  }
}
#undef EW
#undef BW

/* the boundary update every step starts with, VPML walls: wall guards refreshed along y, z
 * (update_boundary_vpml, os-emf-vpml.f03:506-524), then the periodic refresh of y (and z) */
void orc_open_boundary_vpml(const orc_grid *g, orc_vpml *v, real *e, real *b) {
  vpml_periodic(g, v);
  update_periodic_from(g, e, 3, 0, 1);
  update_periodic_from(g, b, 3, 0, 1);
}

/* advance_yee with VPML x walls: the wall E after the first B half step
 * (update_boundary_bfld step 1, os-emf-boundary.f90:930-935, 961-965), the wall B after the
 * E step (update_boundary_efld, :1036-1041) */
void orc_yee_advance_vpml(const orc_grid *g, const double *dx, double dt, real *e, real *b,
                          const real *j, orc_vpml *v) {
  dbdt(g, dx, dt / 2.0, b, e);
  vpml_update_e(g, v, b);
  dedt(g, dx, dt, e, b, j);
  vpml_update_b(g, v, e);
  dbdt(g, dx, dt / 2.0, b, e);
}
#undef VW

/* orc_step with open x walls: boundary update of E/B (Lindman restore + periodic y, z),
 * push+deposit, particle boundaries, energies, J folded along y, z only, Yee with the
 * Lindman wall updates; with v non-null the walls are VPML (2D) instead.  np[s] is
 * updated to the particles left. */
void orc_step_open(const orc_grid *g, int order, const double *dx, double dt, real *e, real *b,
                   real *j, orc_lindman *w, orc_vpml *v, int nspec, const int *pusher,
                   const real *rqm, int *const *ix, real *const *x, real *const *u,
                   real *const *q, int *np, double *fld_ene, double *par_ene) {
  double dvol = dx[0] * dx[1] * (g->dims > 2 ? dx[2] : 1.0);
  if (v)
    orc_open_boundary_vpml(g, v, e, b);
  else
    orc_open_boundary(g, w, e, b);
  memset(j, 0, sizeof(real) * 3 * orc_grid_cells(g));
  for (int s = 0; s < nspec; s++) {
    double ene = orc_advance_deposit(g, order, pusher[s], rqm[s], dx, dt, e, b, j, ix[s], x[s],
                                     u[s], q[s], np[s]);
    np[s] = orc_particles_open(g, ix[s], x[s], u[s], q[s], np[s]);
    if (par_ene) par_ene[s] = ene * (double)rqm[s] * dvol;
  }
  if (fld_ene) orc_field_energy(g, e, b, dvol, fld_ene);
  update_periodic_from(g, j, 3, 1, 1);
  if (v)
    orc_yee_advance_vpml(g, dx, dt, e, b, j, v);
  else
    orc_yee_advance_open(g, dx, dt, e, b, j, w);
}

/* ------------------------------------------------------------------------------------
 * Laser injection from an x wall, 2D and 3D (zpulse_wall, zpulse/os-zpulse-wall.f03).
 * orc_zpulse holds the parameters after read_input_zpulse_wall (:90-628): angles in
 * radians, envelope lengths resolved from fwhm, per_focus negated for backward pulses.
 * orc_zpulse_wall_2d is launch_zpulse_wall (:633-700) at time t: nothing outside
 * launch_time .. launch_time + duration; otherwise launch_em_wall_2d (:865-960) adds
 * real(b12), real(b3) to B2, B3 of column x = 1 (forward) or nx (backward) for
 * ip = 0 .. ny+2, twice with the second launch's phase and polarisation for pol_type != 0.
 * ------------------------------------------------------------------------------------ */
typedef struct {
  double a0, omega0, phase0, pol;
  int pol_type, backward, tenv_type, per_type, tem_mode[2];
  double rise, flat, fall, duration, range;
  double center[2], w0, focus, launch_time;
  double xmin[3], xmax[3];
} orc_zpulse;

static const double ORC_PI_2 = 1.57079632679489661923132; /* os-math.f90:14 */

/* fenv_poly, os-math.f90:83-128 */
static double zp_fenv(double tt) {
  return 10 * (tt * tt * tt) - 15 * (tt * tt * tt * tt) + 6 * (tt * tt * tt * tt * tt);
}
static double zp_poly(double x, double rise, double flat, double fall) {
  double length = rise + flat + fall;
  if (x < 0.0) return 0.0;
  if (x < rise) return zp_fenv(x / rise);
  if (x < rise + flat) return 1.0;
  if (x < length) return zp_fenv((length - x) / fall);
  return 0.0;
}
/* fenv_sin2, os-math.f90:132-165 */
static double zp_sin2(double x, double rise, double flat, double fall) {
  double length = rise + flat + fall, v;
  if (x < 0.0) return 0.0;
  if (x < rise) {
    v = sin(ORC_PI_2 * x / rise);
    return v * v;
  }
  if (x < rise + flat) return 1.0;
  if (x < length) {
    v = sin(ORC_PI_2 * (length - x) / fall);
    return v * v;
  }
  return 0.0;
}
/* hermite_double, os-math.f90:478-509 */
static double zp_hermite(int m, double u) {
  double u2 = u * u;
  switch (m) {
    case 0: return 1.0;
    case 1: return 2.0 * u;
    case 2: return 4.0 * u2 - 2.0;
    case 3: return 8.0 * (u2 * u) - 12.0 * u;
    case 4: return 16.0 * (u2 * u * u) - 48.0 * u2 + 12.0;
    case 5: return 32.0 * (u2 * u * u * u) - 160.0 * (u2 * u) + 120.0 * u;
    case 6: return 64.0 * (u2 * u * u * u * u) - 480.0 * (u2 * u * u) + 720.0 * u2 - 120.0;
    case 7:
      return 128.0 * (u2 * u * u * u * u * u) - 1344.0 * (u2 * u * u * u) + 3360.0 * (u2 * u) -
             1680.0 * u;
  }
  return 0.0;
}
/* t_env_zpulse_wall :705-746 (types 0 polynomial, 1 sin2, 2 gaussian) */
static double zp_tenv(const orc_zpulse *p, double t) {
  if (p->tenv_type == 0) return zp_poly(t, p->rise, p->flat, p->fall);
  if (p->tenv_type == 1) return zp_sin2(t, p->rise, p->flat, p->fall);
  double a = (t - 0.5 * p->range) / p->duration;
  return exp(-2 * (a * a));
}
/* per_envelope_2d_zpulse (types 0 plane, 1 hermite_gaussian) with tin_sim absent (t = 0)
 * and lon_center_zpulse_wall :781-805 (moves with the injection time) */
static double zp_per(const orc_zpulse *p, double phase0, double x1, double x2, double inj) {
  double zc = (p->tenv_type == 2 ? 0.5 * p->range : p->rise + p->flat / 2) + inj;
  double k = p->omega0, ps = p->backward ? -1.0 : 1.0, t = 0.0;
  if (p->per_type == 0) return cos(k * (x1 - zc - ps * t) + phase0);
  double z0 = k * (p->w0 * p->w0) * 0.5;
  double rho = x2 - p->center[0], rho2 = rho * rho, z = x1 - p->focus;
  double rwl2 = 1, curv = 0, gouy = 0;
  if (z != 0) {
    rwl2 = (z0 * z0) / ((z0 * z0) + (z * z));
    curv = 0.5 * rho2 * z / ((z * z) + (z0 * z0));
    gouy = (p->tem_mode[0] + 1) * atan2(z, z0);
  }
  double rwu = sqrt(2 * rwl2) / p->w0;
  return sqrt(sqrt(rwl2)) * zp_hermite(p->tem_mode[0], rho * rwu) *
         exp(-(rho2 * rwl2 / (p->w0 * p->w0))) * cos(k * (x1 - zc - ps * t) + k * curv - gouy + phase0);
}

/* per_envelope_3d_zpulse, os-zpulse-std.f03:2760-3049 (plane, hermite_gaussian) */
static double zp_per3(const orc_zpulse *p, double phase0, double x1, double x2, double x3,
                      double inj) {
  double zc = (p->tenv_type == 2 ? 0.5 * p->range : p->rise + p->flat / 2) + inj;
  double ox = x2 - p->center[0], oy = x3 - p->center[1];
  double k = p->omega0, ps = p->backward ? -1.0 : 1.0, t = 0.0;
  if (p->per_type == 0) return cos(k * (x1 - zc - ps * t) + phase0);
  double rho2 = ox * ox + oy * oy, z = x1 - p->focus, z0 = k * (p->w0 * p->w0) / 2;
  double rwl2 = 1, curv = 0, gouy = 0;
  if (z != 0) {
    rwl2 = (z0 * z0) / ((z0 * z0) + (z * z));
    curv = 0.5 * rho2 * z / ((z * z) + (z0 * z0));
    gouy = (p->tem_mode[0] + p->tem_mode[1] + 1) * atan2(z, z0);
  }
  double rwu = sqrt(2 * rwl2) / p->w0;
  return sqrt(rwl2) * zp_hermite(p->tem_mode[0], ox * rwu) * zp_hermite(p->tem_mode[1], oy * rwu) *
         exp(-(rho2 * rwl2 / (p->w0 * p->w0))) * cos(k * (x1 - zc - ps * t) + k * curv - gouy + phase0);
}

static void zp_launch(const orc_grid *g, const orc_zpulse *p, int order, const double *dx,
                      double dt, double t, double phase0, double pol, real *b) {
  int inpos = p->backward ? g->nx[0] : 1;
  double ps = p->backward ? -1.0 : 1.0;
  double z = p->backward ? -p->xmax[0] : p->xmin[0];
  double inj = t - p->launch_time;
  double rdtdx = (double)(real)(dt / dx[0]);
  double rmin = p->xmin[1];
  double cos_pol = cos(pol) * ps, sin_pol = -sin(pol) * ps;
  double amp = 2.0 * p->omega0 * p->a0, dr_2 = dx[1] / 2.0;
  if (order == 2) {
    z = z + dx[0] * 0.5;
    rmin = rmin + dr_2;
  }
  double lenv = amp * zp_tenv(p, inj);
  if (g->dims == 3) { /* launch_em_wall_3d :965-1082, corner rounded to p_k_fld */
    double rmin1 = (double)(real)p->xmin[1], rmin2 = (double)(real)p->xmin[2];
    double dr2_2 = dx[2] / 2.0;
    if (order == 2) {
      rmin1 = rmin1 + dr_2;
      rmin2 = rmin2 + dr2_2;
    }
    for (int ip2 = 0; ip2 <= g->nx[2] + 2; ip2++)
      for (int ip1 = 0; ip1 <= g->nx[1] + 2; ip1++) {
        double r1 = rmin1 + (ip1 - 1) * dx[1], r2 = rmin2 + (ip2 - 1) * dx[2];
        double r1_2 = r1 + dr_2, r2_2 = r2 + dr2_2;
        double b0 = lenv * rdtdx * zp_per3(p, phase0, z, r1, r2_2, inj) * sin_pol;
        double b90 = lenv * rdtdx * zp_per3(p, phase0, z, r1_2, r2, inj) * cos_pol;
        FLD(b, g, 1, inpos, ip1, ip2) = FLD(b, g, 1, inpos, ip1, ip2) + (real)b0;
        FLD(b, g, 2, inpos, ip1, ip2) = FLD(b, g, 2, inpos, ip1, ip2) + (real)b90;
      }
    return;
  }
  for (int ip = 0; ip <= g->nx[1] + 2; ip++) {
    double r = rmin + (ip - 1) * dx[1], r_2 = r + dr_2;
    double b12 = lenv * rdtdx * zp_per(p, phase0, z, r, inj) * sin_pol;
    double b3 = lenv * rdtdx * zp_per(p, phase0, z, r_2, inj) * cos_pol;
    FLD(b, g, 1, inpos, ip, 1) = FLD(b, g, 1, inpos, ip, 1) + (real)b12;
    FLD(b, g, 2, inpos, ip, 1) = FLD(b, g, 2, inpos, ip, 1) + (real)b3;
  }
}

void orc_zpulse_wall_2d(const orc_grid *g, const orc_zpulse *p, int order, const double *dx,
                        double dt, double t, real *b) {
  double dur = p->tenv_type == 2 ? p->range : p->rise + p->flat + p->fall;
  if (t > p->launch_time + dur || !(t >= p->launch_time)) return;
  zp_launch(g, p, order, dx, dt, t, p->phase0, p->pol, b);
  if (p->pol_type != 0)
    zp_launch(g, p, order, dx, dt, t, p->phase0 + (p->pol_type > 0 ? ORC_PI_2 : -ORC_PI_2),
              p->pol + ORC_PI_2, b);
}

/* ------------------------------------------------------------------------------------
 * Initial electrostatic field of a species (init_fields, spec/os-spec-initfields.f90:
 * 38-100, non-beam branch): rho deposited and folded, then es_solver_2d
 * (emf/os-emf-es-solver.f03:125-236): the free-space direct sum over every source cell
 * with nonzero charge, in source order (y outer, x inner), of charge * d / r^2 at the
 * Yee points of E1 (x + 1/2, y) and E2 (x, y + 1/2), each term in FP64 and accumulated
 * into the working-precision field; then times the cell volume, E3 = 0.  rho is a
 * one-component grid, ef a three-component grid (guards untouched, interior written).
 * ------------------------------------------------------------------------------------ */
void orc_es_solver_2d(const orc_grid *g, const double *dx, const real *rho, real *ef) {
  const int nx = g->nx[0], ny = g->nx[1];
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++)
      for (int c = 0; c < 3; c++) FLD(ef, g, c, i, j, 1) = 0;
  for (int ri2 = 1; ri2 <= ny; ri2++)
    for (int ri1 = 1; ri1 <= nx; ri1++) {
      double charge = (double)rho[cidx(g, ri1, ri2, 1)];
      if (charge == 0) continue;
      for (int ei2 = 1; ei2 <= ny; ei2++) {
        double dx2_e1 = (double)(ei2 - ri2);
        double dx2_e2 = dx2_e1 + 0.5;
        dx2_e1 = dx2_e1 * dx[1];
        dx2_e2 = dx2_e2 * dx[1];
        for (int ei1 = 1; ei1 <= nx; ei1++) {
          double dx1_e2 = (double)(ei1 - ri1);
          double dx1_e1 = dx1_e2 + 0.5;
          dx1_e1 = dx1_e1 * dx[0];
          dx1_e2 = dx1_e2 * dx[0];
          double r_2 = dx1_e1 * dx1_e1 + dx2_e1 * dx2_e1;
          real *e = &FLD(ef, g, 0, ei1, ei2, 1);
          e[0] = (real)((double)e[0] + charge * dx1_e1 / r_2);
          r_2 = dx1_e2 * dx1_e2 + dx2_e2 * dx2_e2;
          e[1] = (real)((double)e[1] + charge * dx2_e2 / r_2);
        }
      }
    }
  double cell_vol = dx[0] * dx[1];
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      real *e = &FLD(ef, g, 0, i, j, 1);
      e[0] = (real)((double)e[0] * cell_vol);
      e[1] = (real)((double)e[1] * cell_vol);
      e[2] = 0;
    }
}

/* es_solver_2d_beam (emf/os-emf-es-solver.f03:457-616), one node, the branch init_fields
 * takes for a species drifting along x (ufl(1) != 0, spec/os-spec-initfields.f90:77-92):
 * every source cell with nonzero charge adds rho * sign(dy / gamma / 2, 0.5 + (j - rj))
 * to E2 of its own column, for every y of the array (guards included, source order y
 * outer, x inner); after the E guard refresh B3(i) = 0.5 u1 (E2(i) + E2(i+1)) (the sum in
 * the working precision), then E2 *= gamma, E1 = E3 = 0.  gamma = sqrt(u1^2 + 1) with u1
 * squared in the particle precision (float).  open_x: E2 guards along x are not refreshed.
 * ef, bf are three-component grids (overwritten). */
void orc_es_solver_2d_beam(const orc_grid *g, const double *dx, const real *rho, float u1,
                           int open_x, real *ef, real *bf) {
  const int nx = g->nx[0], ny = g->nx[1];
  const long nc = orc_grid_cells(g);
  memset(ef, 0, sizeof(real) * 3 * nc);
  memset(bf, 0, sizeof(real) * 3 * nc);
  float u1sq = u1 * u1;
  double gamma = sqrt((double)u1sq + 1.0);
  double charge_norm = dx[1] / gamma / 2;
  for (int ri2 = 1; ri2 <= ny; ri2++)
    for (int ri1 = 1; ri1 <= nx; ri1++) {
      real r = rho[cidx(g, ri1, ri2, 1)];
      if (r == 0) continue;
      for (int ei2 = 1 - g->gcl[1]; ei2 <= ny + g->gcu[1]; ei2++) {
        double s = 0.5 + (double)(ei2 - ri2) >= 0 ? charge_norm : -charge_norm;
        real *e = &FLD(ef, g, 1, ri1, ei2, 1);
        *e = (real)((double)*e + (double)r * s);
      }
    }
  update_periodic_from(g, ef, 3, 0, open_x ? 1 : 0);
  for (int ei2 = 1 - g->gcl[1]; ei2 <= ny + g->gcu[1]; ei2++)
    for (int ei1 = 1; ei1 <= nx; ei1++) {
      real sum = FLD(ef, g, 1, ei1, ei2, 1) + FLD(ef, g, 1, ei1 + 1, ei2, 1);
      FLD(bf, g, 2, ei1, ei2, 1) = (real)(0.5 * (double)u1 * (double)sum);
    }
  update_periodic_from(g, bf, 3, 0, open_x ? 1 : 0);
  for (long c = 0; c < nc; c++) {
    ef[3 * c] = 0;
    ef[3 * c + 1] = (real)((double)ef[3 * c + 1] * gamma);
    ef[3 * c + 2] = 0;
  }
}

/* es_solver_3d_beam (emf/os-emf-es-solver.f03:795-1003), one node, the 3D branch of
 * es_solver_beam (:240-287) that init_fields takes for a species drifting along x
 * (spec/os-spec-initfields.f90:77-92): each source cell (ri1, ri2, ri3) with nonzero
 * charge adds, on its own x slice ei1 = ri1 and for every (ei2, ei3) of the array
 * (guards included; sources in the order ri3, ri2, ri1 outer to inner), the field of an
 * infinite wire, rho * charge_norm * d / r^2 with charge_norm = dx2 dx3 / gamma / 2 / pi:
 * E2 at (y + 1/2, z) with d = (ei2 - ri2 + 1/2) dx2, r^2 = d^2 + ((ei3 - ri3) dx3)^2, and
 * E3 at (y, z + 1/2) symmetrically; each term in FP64, accumulated in the working
 * precision.  After the E guard refresh: B1 = 0, B2 = -0.5 u1 (E3(i) + E3(i+1)),
 * B3 = 0.5 u1 (E2(i) + E2(i+1)) (the sums in the working precision); then E2, E3 *= gamma,
 * E1 = 0.  gamma = sqrt(u1^2 + 1) with u1 squared in the particle precision (float).
 * open_x: the guards along x are not refreshed (E beyond the upper wall stays zero).
 * ef, bf are three-component grids (overwritten). */
void orc_es_solver_3d_beam(const orc_grid *g, const double *dx, const real *rho, float u1,
                           int open_x, real *ef, real *bf) {
  const int nx = g->nx[0], ny = g->nx[1], nz = g->nx[2];
  const long nc = orc_grid_cells(g);
  const double pi = 3.14159265358979323846;
  memset(ef, 0, sizeof(real) * 3 * nc);
  memset(bf, 0, sizeof(real) * 3 * nc);
  float u1sq = u1 * u1;
  double gamma = sqrt((double)u1sq + 1.0);
  double charge_norm = dx[1] * dx[2] / gamma / 2 / pi;
  for (int ri3 = 1; ri3 <= nz; ri3++)
    for (int ri2 = 1; ri2 <= ny; ri2++)
      for (int ri1 = 1; ri1 <= nx; ri1++) {
        real r = rho[cidx(g, ri1, ri2, ri3)];
        if (r == 0) continue;
        for (int ei3 = 1 - g->gcl[2]; ei3 <= nz + g->gcu[2]; ei3++) {
          double dx3c = (double)(ei3 - ri3);
          double dx3s = dx3c * dx[2];
          double dx3s_h = (dx3c + 0.5) * dx[2];
          for (int ei2 = 1 - g->gcl[1]; ei2 <= ny + g->gcu[1]; ei2++) {
            double dx2c = (double)(ei2 - ri2);
            double dx2s = dx2c * dx[1];
            double dx2s_h = (dx2c + 0.5) * dx[1];
            real *e = &FLD(ef, g, 0, ri1, ei2, ei3);
            double r_2 = dx2s_h * dx2s_h + dx3s * dx3s;
            e[1] = (real)((double)e[1] + (double)r * charge_norm * dx2s_h / r_2);
            r_2 = dx2s * dx2s + dx3s_h * dx3s_h;
            e[2] = (real)((double)e[2] + (double)r * charge_norm * dx3s_h / r_2);
          }
        }
      }
  update_periodic_from(g, ef, 3, 0, open_x ? 1 : 0);
  for (int ei3 = 1 - g->gcl[2]; ei3 <= nz + g->gcu[2]; ei3++)
    for (int ei2 = 1 - g->gcl[1]; ei2 <= ny + g->gcu[1]; ei2++)
      for (int ei1 = 1; ei1 <= nx; ei1++) {
        real s3 = FLD(ef, g, 2, ei1, ei2, ei3) + FLD(ef, g, 2, ei1 + 1, ei2, ei3);
        real s2 = FLD(ef, g, 1, ei1, ei2, ei3) + FLD(ef, g, 1, ei1 + 1, ei2, ei3);
        FLD(bf, g, 0, ei1, ei2, ei3) = 0;
        FLD(bf, g, 1, ei1, ei2, ei3) = (real)(-(0.5 * (double)u1 * (double)s3));
        FLD(bf, g, 2, ei1, ei2, ei3) = (real)(0.5 * (double)u1 * (double)s2);
      }
  update_periodic_from(g, bf, 3, 0, open_x ? 1 : 0);
  for (long c = 0; c < nc; c++) {
    ef[3 * c] = 0;
    ef[3 * c + 1] = (real)((double)ef[3 * c + 1] * gamma);
    ef[3 * c + 2] = (real)((double)ef[3 * c + 2] * gamma);
  }
}

/* ------------------------------------------------------------------------------------
 * In-box laser pulse (zpulse, zpulse/os-zpulse-std.f03), launched once into the fields
 * along x1 (direction 1), without boost, tilt or chirp, one node.  orc_zpulse_box holds
 * the parameters after read_input_zpulse (:157-805): angles in radians, envelope lengths
 * resolved from fwhm, per_center defaulted to the box centre; lon_type 0 polynomial,
 * 1 sin2, 2 gaussian; per_type 0 plane, 1 hermite_gaussian; div_corr = 0 for plane
 * pulses or no_div_corr (init_zpulse :867-929).
 * orc_zpulse_box_add is launch_zpulse (:2350-2469): the pulse field is built in zeroed
 * grids (launch_zpulse_2d :1072-1212 / 3d :1218-1387), its divergence corrected
 * (div_corr_zpulse_2d :1681-1835 / 3d :2010-2205) and added to E and B (every cell);
 * for pol_type != 0 a second pulse with phase0 + sign(pi/2, pol_type) and pol + pi/2 is
 * built and added the same way.  order: the interpolation order (2 shifts the grid
 * origin by half a cell, :1107-1113).  e, b: three-component grids.
 * ------------------------------------------------------------------------------------ */
typedef struct {
  double a0, omega0, phase0, pol;
  int pol_type, backward, lon_type, per_type, tem_mode[2], div_corr;
  double lon_start, rise, flat, fall, lon_x0, duration, range;
  double center[2], w0, focus;
  double xmin[3];
} orc_zpulse_box;

/* lon_envelope_zpulse :2480-2561 (t ignored without boost) */
static double zb_lon(const orc_zpulse_box *p, double z) {
  if (p->lon_type == 0)
    return p->backward ? zp_poly(z - p->lon_start, p->rise, p->flat, p->fall)
                       : zp_poly(p->lon_start - z, p->rise, p->flat, p->fall);
  if (p->lon_type == 1)
    return p->backward ? zp_sin2(z - p->lon_start, p->rise, p->flat, p->fall)
                       : zp_sin2(p->lon_start - z, p->rise, p->flat, p->fall);
  if (fabs(z - p->lon_x0) > p->range / 2) return 0.0;
  double a = (z - p->lon_x0) / p->duration;
  return exp(-2 * (a * a));
}
/* lon_center_zpulse :3055-3097 */
static double zb_center(const orc_zpulse_box *p) {
  if (p->lon_type == 2) return p->lon_x0;
  return p->backward ? p->lon_start + p->rise + p->flat / 2.0
                     : p->lon_start - p->rise - p->flat / 2.0;
}
/* per_envelope_2d_zpulse :2599-2753 at t = 0 (plane, hermite_gaussian) */
static double zb_per2(const orc_zpulse_box *p, double phase0, double x1, double x2) {
  double zc = zb_center(p), k = p->omega0, ps = p->backward ? -1.0 : 1.0, t = 0.0;
  if (p->per_type == 0) return cos(k * (x1 - zc - ps * t) + phase0);
  double z0 = k * (p->w0 * p->w0) * 0.5;
  double rho = x2 - p->center[0], rho2 = rho * rho, z = x1 - p->focus;
  double rwl2 = 1, curv = 0, gouy = 0;
  if (z != 0) {
    rwl2 = (z0 * z0) / ((z0 * z0) + (z * z));
    curv = 0.5 * rho2 * z / ((z * z) + (z0 * z0));
    gouy = (p->tem_mode[0] + 1) * atan2(z, z0);
  }
  double rwu = sqrt(2 * rwl2) / p->w0;
  return sqrt(sqrt(rwl2)) * zp_hermite(p->tem_mode[0], rho * rwu) *
         exp(-(rho2 * rwl2 / (p->w0 * p->w0))) *
         cos(k * (x1 - zc - ps * t) + k * curv - gouy + phase0);
}
/* per_envelope_3d_zpulse :2760-3049 at t = 0 (plane, hermite_gaussian) */
static double zb_per3(const orc_zpulse_box *p, double phase0, double x1, double x2, double x3) {
  double zc = zb_center(p), k = p->omega0, ps = p->backward ? -1.0 : 1.0, t = 0.0;
  double ox = x2 - p->center[0], oy = x3 - p->center[1];
  if (p->per_type == 0) return cos(k * (x1 - zc - ps * t) + phase0);
  double rho2 = ox * ox + oy * oy, z = x1 - p->focus, z0 = k * (p->w0 * p->w0) / 2;
  double rwl2 = 1, curv = 0, gouy = 0;
  if (z != 0) {
    rwl2 = (z0 * z0) / ((z0 * z0) + (z * z));
    curv = 0.5 * rho2 * z / ((z * z) + (z0 * z0));
    gouy = (p->tem_mode[0] + p->tem_mode[1] + 1) * atan2(z, z0);
  }
  double rwu = sqrt(2 * rwl2) / p->w0;
  return sqrt(rwl2) * zp_hermite(p->tem_mode[0], ox * rwu) * zp_hermite(p->tem_mode[1], oy * rwu) *
         exp(-(rho2 * rwl2 / (p->w0 * p->w0))) *
         cos(k * (x1 - zc - ps * t) + k * curv - gouy + phase0);
}

/* one pulse (phase0, pol) into zeroed grids eb (E) and bb (B) */
static void zb_build(const orc_grid *g, const orc_zpulse_box *p, int order, const double *dx,
                     double phase0, double pol, real *eb, real *bb) {
  const int D = g->dims;
  int lo[3] = {1, 1, 1}, hi[3] = {1, 1, 1};
  for (int d = 0; d < D; d++) {
    lo[d] = 1 - g->gcl[d];
    hi[d] = g->nx[d] + g->gcu[d];
  }
  double cos_pol = cos(pol), sin_pol = sin(pol), amp = p->omega0 * p->a0;
  double ps = p->backward ? -1.0 : 1.0;
  double zmin = p->xmin[0], rmin = p->xmin[1], r2min = D == 3 ? p->xmin[2] : 0.0;
  double dz_2 = dx[0] / 2.0, dr_2 = dx[1] / 2.0, dr2_2 = D == 3 ? dx[2] / 2.0 : 0.0;
  if (order == 2) {
    zmin = zmin + dz_2;
    rmin = rmin + dr_2;
    r2min = r2min + dr2_2;
  }
  for (int i1 = lo[0]; i1 <= hi[0]; i1++) {
    double z = zmin + (double)(i1 - 1) * dx[0], z_2 = z + dz_2;
    double lenv = amp * zb_lon(p, z), lenv_2 = amp * zb_lon(p, z_2);
    if (D == 2) {
      for (int i2 = lo[1]; i2 <= hi[1]; i2++) {
        double r = rmin + (double)(i2 - 1) * dx[1], r_2 = r + dr_2;
        FLD(bb, g, 0, i1, i2, 1) = 0;
        FLD(bb, g, 1, i1, i2, 1) = (real)(-(lenv_2 * zb_per2(p, phase0, z_2, r) * sin_pol * ps));
        FLD(bb, g, 2, i1, i2, 1) = (real)(+(lenv_2 * zb_per2(p, phase0, z_2, r_2) * cos_pol * ps));
        FLD(eb, g, 0, i1, i2, 1) = 0;
        FLD(eb, g, 1, i1, i2, 1) = (real)(lenv * zb_per2(p, phase0, z, r_2) * cos_pol);
        FLD(eb, g, 2, i1, i2, 1) = (real)(lenv * zb_per2(p, phase0, z, r) * sin_pol);
      }
      continue;
    }
    for (int i2 = lo[1]; i2 <= hi[1]; i2++) {
      double r1 = rmin + (double)(i2 - 1) * dx[1], r1_2 = r1 + dr_2;
      for (int i3 = lo[2]; i3 <= hi[2]; i3++) {
        double r2 = r2min + (double)(i3 - 1) * dx[2], r2_2 = r2 + dr2_2;
        FLD(bb, g, 1, i1, i2, i3) = FLD(bb, g, 1, i1, i2, i3) -
                                    (real)(lenv_2 * zb_per3(p, phase0, z_2, r1, r2_2) * sin_pol * ps);
        FLD(bb, g, 2, i1, i2, i3) = FLD(bb, g, 2, i1, i2, i3) +
                                    (real)(lenv_2 * zb_per3(p, phase0, z_2, r1_2, r2) * cos_pol * ps);
        FLD(eb, g, 1, i1, i2, i3) = FLD(eb, g, 1, i1, i2, i3) +
                                    (real)(lenv * zb_per3(p, phase0, z, r1_2, r2) * cos_pol);
        FLD(eb, g, 2, i1, i2, i3) = FLD(eb, g, 2, i1, i2, i3) +
                                    (real)(lenv * zb_per3(p, phase0, z, r1, r2_2) * sin_pol);
      }
    }
  }
  if (!p->div_corr) return;
  /* div_corr_zpulse_2d / 3d: E1 and B1 integrated from the upper x edge (one node: the
   * loop over the other nodes' x range is empty and e1 = b1 = 0 at x = nx + 1) */
  double x1min = p->xmin[0], x2min = p->xmin[1], x3min = D == 3 ? p->xmin[2] : 0.0;
  (void)x2min;
  (void)x3min;
  if (order == 2) {
    x1min = x1min + dx[0] * 0.5;
    x2min = x2min + dx[1] * 0.5;
    if (D == 3) x3min = x3min + dx[2] * 0.5;
  }
  const int n3 = D == 3 ? g->nx[2] : 1;
  for (int i3 = 1; i3 <= n3; i3++)
    for (int i2 = 1; i2 <= g->nx[1]; i2++) {
      double e1 = 0.0, b1 = 0.0;
      FLD(eb, g, 0, g->nx[0] + 1, i2, i3) = (real)e1;
      FLD(bb, g, 0, g->nx[0] + 1, i2, i3) = (real)b1;
      for (int i1 = g->nx[0]; i1 >= 1 - g->gcl[0]; i1--) {
        if (zb_lon(p, x1min + i1 * dx[0]) > 0.0) {
          double e2p = FLD(eb, g, 1, i1 + 1, i2, i3), e2m = FLD(eb, g, 1, i1 + 1, i2 - 1, i3);
          if (D == 2) {
            e1 = e1 + dx[0] * (e2p - e2m) / dx[1];
          } else {
            double e3p = FLD(eb, g, 2, i1 + 1, i2, i3), e3m = FLD(eb, g, 2, i1 + 1, i2, i3 - 1);
            e1 = e1 + dx[0] * ((e2p - e2m) / dx[1] + (e3p - e3m) / dx[2]);
          }
          FLD(eb, g, 0, i1, i2, i3) = (real)e1;
        } else {
          e1 = 0.0;
        }
        if (zb_lon(p, x1min + (i1 - 0.5) * dx[0]) > 0.0) {
          double b2p = FLD(bb, g, 1, i1, i2 + 1, i3), b2m = FLD(bb, g, 1, i1, i2, i3);
          if (D == 2) {
            b1 = b1 + dx[0] * (b2p - b2m) / dx[1];
          } else {
            double b3p = FLD(bb, g, 2, i1, i2, i3 + 1), b3m = FLD(bb, g, 2, i1, i2, i3);
            b1 = b1 + dx[0] * ((b2p - b2m) / dx[1] + (b3p - b3m) / dx[2]);
          }
          FLD(bb, g, 0, i1, i2, i3) = (real)b1;
        } else {
          b1 = 0.0;
        }
      }
    }
}

void orc_zpulse_box_add(const orc_grid *g, const orc_zpulse_box *p, int order, const double *dx,
                        real *e, real *b) {
  const long nc = orc_grid_cells(g);
  real *eb = (real *)calloc((size_t)(3 * nc), sizeof(real));
  real *bb = (real *)calloc((size_t)(3 * nc), sizeof(real));
  double phase0 = p->phase0, pol = p->pol;
  for (int k = 0; k < (p->pol_type != 0 ? 2 : 1); k++) {
    if (k == 1) {
      memset(eb, 0, sizeof(real) * 3 * nc);
      memset(bb, 0, sizeof(real) * 3 * nc);
      phase0 = phase0 + (p->pol_type > 0 ? ORC_PI_2 : -ORC_PI_2);
      pol = pol + ORC_PI_2;
    }
    zb_build(g, p, order, dx, phase0, pol, eb, bb);
    for (long c = 0; c < 3 * nc; c++) {
      b[c] = b[c] + bb[c];
      e[c] = e[c] + eb[c];
    }
  }
  free(eb);
  free(bb);
}
