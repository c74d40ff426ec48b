// k_field.hip -- Yee field advance, current conversion, per-step diagnostics.
//
// Yee: emf/solver/os-emf-solver-yee.f03 dbdt_2d :102-138, dbdt_3d :211-251,
// dedt_2d :287-326, dedt_3d :385-428, advance_yee :542-568.  Periodic boundaries are
// index wraps on the interior arrays, which reproduces the reference's guard-cell
// refresh (update_periodic_vdf, vdf/os-vdf-comm.f90:142-330) on every interior cell.
// The arithmetic keeps the reference's operation order in FP32 (no contraction).
//
// The three sub-steps are memory-bound stencils over planar component arrays: one thread
// per cell, consecutive threads on consecutive x (coalesced).  The E step also converts
// the fixed-point current to float and clears it for the next deposit, and the first B
// step accumulates the field energy of E^n, B^n (the fields the particles were pushed
// with) into per-block FP64 partial sums.
#include "osr_device.hpp"
#include "osr_types.hpp"

namespace osr {

struct FieldDesc {
  GridDesc g;
  double* fe_part;   // [nblk][6] B1,B2,B3,E1,E2,E3 partial sums of squares
  int nblk;          // energy partials: one per kEneBlk blocks of 256 cells
  // finalisation
  int nspec;
  double* ep[kMaxSpecies];  // per-species tile energy partials
  int ep_n[kMaxSpecies];
  double ep_scale[kMaxSpecies];  // rqm * dvol
  double fe_scale;               // 0.5 * dvol
  double* hist;                  // rows of (6 + nspec) doubles
  long long hist_cap;
  // Yee coefficients: real(dt/dx) for the half B steps and the full E step, real(dt)
  float bdt[3];
  float edt[3];
  float dtif;
  // when set, the E step reads a float current from here instead of the fixed-point J
  float* Jf;
  int jf_zero;  // clear Jf after reading it (the plain path accumulates into it)
  // current smoothing: passes (one per filtered direction, x before y before z) from the
  // source current into the ping-pong buffers; the last output is Jf, and the E step then
  // clears the source (1: fixed-point J, 2: the plain path's float J)
  int sm_np;
  int sm_dir[3];
  int sm_n[3];
  int sm_src;  // 0 fixed-point J, 1 float J (plain path)
  int jsrc_clear;
  float sm_coef[3][31];  // per pass, s(k) at [15 + k]
  float* sm_buf[2];
  // open x walls (g.open_nx > 0): Lindman buffers, planar, one value per row (y, z) of the
  // wall, R = ny nz values per entry: ebl[3][2] (x = 0, 1), bbl[3] (x = 0), ebu[3][2]
  // (x = nx+1, nx+2); then E3(x = 0, y = 1, z) [nz] and E2(x = 0, y, z = 1) [ny] as they
  // were when the step began
  float* wall;
  float wall_dc;     // real((1 - dt/dx1) / (1 + dt/dx1))
  float wall_dt[3];  // real(dt/dx) * 0.5
  // VPML x walls instead (vp_n > 0): split-field walls vp_n cells thick; wall holds
  // [lower, upper][E, B][component][cell l][row y + ny z] (components X, Y, ZX, ZY in 2D,
  // XY, XZ, YX, YZ, ZX, ZY in 3D; cell l is x = 2 - vp_n + l at the lower wall,
  // x = nx + 1 + l at the upper one), then the attenuation coefficients
  // [lower, upper][E, B][1..3][cell]
  int vp_n;
  float vp_dt2, vp_dt3;  // real(dt/dx2), real(dt/dx3)
  int vp_diff[2];  // 2D vpml_diffuse per wall (lower, upper)
  // laser walls: las_k launches per step, into ring columns las_col[k]; for table row
  // r = step - *las_first: las_mask[r][k] (0: the launch adds nothing) and the values
  // las_tab[r][k][2][ny] (B2, B3 increments of rows y = 1..ny)
  int las_k;
  int las_rows;      // steps held by the table
  int las_col[16];
  const float* las_tab;
  const int* las_mask;
  const long long* las_first;
};

// entries of FieldDesc::wall before the saved values (units of ny nz values)
constexpr int kWallEntries = 15;

OSR_DEV int wrp(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// coordinates of cell c of the grid (c < ncell < 2^32): x, and the row y + ny z split into
// y and z, with the grid's multiply-high divisors
OSR_DEV void cell_xyz(const GridDesc& g, long long c, int& i, int& j, int& k) {
  const uint32_t u = (uint32_t)c;
  const uint32_t row = g.div_nx.div(u);
  const uint32_t z = g.div_ny.div(row);
  i = (int)(u - row * (uint32_t)g.nx);
  j = (int)(row - z * (uint32_t)g.ny);
  k = (int)z;
}
// a wall row r = y + ny z (r < ny nz) split into y and z
OSR_DEV void row_yz(const GridDesc& g, long long r, int& j, int& k) {
  const uint32_t u = (uint32_t)r;
  const uint32_t z = g.div_ny.div(u);
  j = (int)(u - z * (uint32_t)g.ny);
  k = (int)z;
}

// 256-cell blocks per workgroup of the B half step that also sums the field energy: the
// FP64 block reduction is paid once per 1024 cells
constexpr int kEneBlk = 4;

// one row of the Lindman wall B update (defined with the wall kernels below)
template <int STEP, int DIMS, bool RESTORE = false>
OSR_DEV void wall_b_row(const FieldDesc& fd, long long r);

template <int DIMS, bool ENERGY>
__global__ __launch_bounds__(256) void yee_dbdt(const FieldDesc* __restrict__ fds,
                                                const int2* __restrict__ work) {
  __shared__ double s_red[4][6];
  constexpr int NB = ENERGY ? kEneBlk : 1;
  const int2 wk = work[blockIdx.x];
  const FieldDesc& fd = fds[wk.x];
  const GridDesc& g = fd.g;
  if (ENERGY && wk.y < 0) {
    // a block of rows of the lower Lindman wall: its B update after the first half step
    // (wall_b<1>) reads E and its own buffer only, neither of which this kernel writes
    wall_b_row<1, DIMS>(fd, (long long)(-1 - wk.y) * 256 + threadIdx.x);
    return;
  }
  if (g.ctrl->halt >= 2) return;
  const long long nc = g.ncell;
  const float d1 = fd.bdt[0], d2 = fd.bdt[1], d3 = fd.bdt[2];
  const int nx = g.nx, ny = g.ny;
  // open x walls: the reference advances B on x = 0 .. nx+1 only (dbdt_2d loops), so the
  // guard band keeps its values elsewhere; the energy sums the interior
  const int nxi = interior_nx(g);
  const float* __restrict__ E1 = g.E;
  const float* __restrict__ E2 = g.E + nc;
  const float* __restrict__ E3 = g.E + 2 * nc;
  float* __restrict__ B1 = g.B;
  float* __restrict__ B2 = g.B + nc;
  float* __restrict__ B3 = g.B + 2 * nc;
  // every load of the workgroup's NB cells per thread is issued before the first store
  // (E and B are distinct arrays)
  long long idx[NB];
  int ii[NB];
  float e1[NB], e2[NB], e3[NB], b1[NB], b2[NB], b3[NB];
  float e3j[NB], e3i[NB], e2i[NB], e1j[NB], e2k[NB], e1k[NB];
#pragma unroll
  for (int qb = 0; qb < NB; qb++) {
    idx[qb] = ((long long)wk.y * NB + qb) * 256 + threadIdx.x;
    ii[qb] = 0;
    e1[qb] = e2[qb] = e3[qb] = b1[qb] = b2[qb] = b3[qb] = 0.f;
    e3j[qb] = e3i[qb] = e2i[qb] = e1j[qb] = e2k[qb] = e1k[qb] = 0.f;
    if (idx[qb] >= nc) continue;
    const long long id = idx[qb];
    int i, j, k;
    cell_xyz(g, id, i, j, k);
    const long long ip = id - i + wrp(i + 1, nx);
    const long long jp = id + (long long)(wrp(j + 1, ny) - j) * nx;
    ii[qb] = i;
    e1[qb] = E1[id];
    e2[qb] = E2[id];
    e3[qb] = E3[id];
    b1[qb] = B1[id];
    b2[qb] = B2[id];
    b3[qb] = B3[id];
    e3j[qb] = E3[jp];
    e3i[qb] = E3[ip];
    e2i[qb] = E2[ip];
    e1j[qb] = E1[jp];
    if (DIMS == 3) {
      const int nz = g.nz;
      const long long kp = id + (long long)(wrp(k + 1, nz) - k) * nx * ny;
      e2k[qb] = E2[kp];
      e1k[qb] = E1[kp];
    }
  }
  double acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int qb = 0; qb < NB; qb++) {
    if (idx[qb] >= nc) break;
    const int i = ii[qb];
    const bool live = i <= nxi || i == nx - 1;
    if (ENERGY && i < nxi) {
      acc[0] += (double)b1[qb] * b1[qb];
      acc[1] += (double)b2[qb] * b2[qb];
      acc[2] += (double)b3[qb] * b3[qb];
      acc[3] += (double)e1[qb] * e1[qb];
      acc[4] += (double)e2[qb] * e2[qb];
      acc[5] += (double)e3[qb] * e3[qb];
    }
    float n1, n2, n3;
    if (DIMS == 2) {
      n1 = b1[qb] - d2 * (e3j[qb] - e3[qb]);
      n2 = b2[qb] + d1 * (e3i[qb] - e3[qb]);
      n3 = b3[qb] - d1 * (e2i[qb] - e2[qb]) + d2 * (e1j[qb] - e1[qb]);
    } else {
      n1 = b1[qb] - d2 * (e3j[qb] - e3[qb]) + d3 * (e2k[qb] - e2[qb]);
      n2 = b2[qb] + d1 * (e3i[qb] - e3[qb]) - d3 * (e1k[qb] - e1[qb]);
      n3 = b3[qb] - d1 * (e2i[qb] - e2[qb]) + d2 * (e1j[qb] - e1[qb]);
    }
    if (live) {
      const long long id = idx[qb];
      B1[id] = n1;
      B2[id] = n2;
      B3[id] = n3;
    }
  }
  if (ENERGY) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 6; c++) {
      double v = wave_sum(acc[c]);
      if (lane == 0) s_red[wave][c] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
      const int c = threadIdx.x;
      fd.fe_part[(long long)wk.y * 6 + c] =
          ((s_red[0][c] + s_red[1][c]) + s_red[2][c]) + s_red[3][c];
    }
  }
}

template <int DIMS>
__global__ __launch_bounds__(256) void yee_dedt(const FieldDesc* __restrict__ fds,
                                                const int2* __restrict__ work) {
  const int2 wk = work[blockIdx.x];
  const FieldDesc& fd = fds[wk.x];
  const GridDesc& g = fd.g;
  if (g.ctrl->halt >= 2) return;
  const long long nc = g.ncell;
  const long long idx = (long long)wk.y * 256 + threadIdx.x;
  if (idx >= nc) return;
  const float d1 = fd.edt[0], d2 = fd.edt[1], d3 = fd.edt[2], dtif = fd.dtif;
  const int nx = g.nx, ny = g.ny;
  int i, j, k;
  cell_xyz(g, idx, i, j, k);
  const long long im = idx - i + wrp(i - 1, nx);
  const long long jm = idx + (long long)(wrp(j - 1, ny) - j) * nx;
  float* E1 = g.E;
  float* E2 = g.E + nc;
  float* E3 = g.E + 2 * nc;
  const float* B1 = g.B;
  const float* B2 = g.B + nc;
  const float* B3 = g.B + 2 * nc;
  unsigned long long* J = g.J;
  float j1, j2, j3;
  if (fd.Jf) {
    j1 = fd.Jf[idx];
    j2 = fd.Jf[nc + idx];
    j3 = fd.Jf[2 * nc + idx];
    if (fd.jf_zero) {
      fd.Jf[idx] = 0.f;
      fd.Jf[nc + idx] = 0.f;
      fd.Jf[2 * nc + idx] = 0.f;
    }
    if (fd.jsrc_clear == 1) {
      J[idx] = 0ull;
      J[nc + idx] = 0ull;
      J[2 * nc + idx] = 0ull;
    } else if (fd.jsrc_clear == 2) {
      float* Jp = (float*)J;
      Jp[idx] = 0.f;
      Jp[nc + idx] = 0.f;
      Jp[2 * nc + idx] = 0.f;
    }
  } else {
    const double jinv = g.jinv;
    j1 = (float)((double)(long long)J[idx] * jinv);
    j2 = (float)((double)(long long)J[nc + idx] * jinv);
    j3 = (float)((double)(long long)J[2 * nc + idx] * jinv);
    J[idx] = 0ull;
    J[nc + idx] = 0ull;
    J[2 * nc + idx] = 0ull;
  }
  const float b1 = B1[idx], b2 = B2[idx], b3 = B3[idx];
  float e1 = E1[idx], e2 = E2[idx], e3 = E3[idx];
  if (DIMS == 2) {
    e1 = e1 - dtif * j1 + d2 * (b3 - B3[jm]);
    e2 = e2 - dtif * j2 - d1 * (b3 - B3[im]);
    e3 = e3 - dtif * j3 + d1 * (b2 - B2[im]) - d2 * (b1 - B1[jm]);
  } else {
    const int nz = g.nz;
    const long long km = idx + (long long)(wrp(k - 1, nz) - k) * nx * ny;
    e1 = e1 - dtif * j1 + d2 * (b3 - B3[jm]) - d3 * (b2 - B2[km]);
    e2 = e2 - dtif * j2 + d3 * (b1 - B1[km]) - d1 * (b3 - B3[im]);
    e3 = e3 - dtif * j3 + d1 * (b2 - B2[im]) - d2 * (b1 - B1[jm]);
  }
  // open x walls: E is advanced on x = 1 .. nx+1 only (dedt_2d loops); the current of
  // every column is still converted and cleared above
  if (i > interior_nx(g)) return;
  E1[idx] = e1;
  E2[idx] = e2;
  E3[idx] = e3;
}

// ------------------------------------------------------------------------------------
// Lindman absorbing walls at x = 0 (ring column nx-1) and x = nx+1, nx+2 (columns
// open_nx, open_nx+1), one thread per row r = y + ny z of the wall.
// emf/boundary/os-emf-lindman.f90: update_lindman_2d_e / 3d_e :454-535, 604-750 (after the
// E step), update_lindman_2d_b / 3d_b :539-600, 755-857 (after each B half step),
// update_emfbound_lindman :861-1011 (wall values restored before the next push).  The
// wall buffers start at zero (vdf/os-wall.f90:189-205).
// ------------------------------------------------------------------------------------
OSR_DEV float* wall_at(const FieldDesc& fd, int k, long long R) { return fd.wall + (size_t)k * R; }

// B buffer of the lower wall.  STEP 2: the reference reads E3(0, ny+1, z) and E2(0, y,
// nz+1), guard cells that nothing refreshes between the start of the step and the second
// B half step, so they still hold E3(0, 1, z) and E2(0, y, 1) of the start of the step
// (saved by wall_e).  E1 on the wall is always zero.
// RESTORE (after the second B half step): the same thread then puts its row's wall values
// back into the box, as wall_restore does.  The B update reads E only at x = 0 and x = 1;
// there the restore rewrites the values the box already holds (wall_e left E2, E3 at x = 0
// equal to their buffers, E1 there is never advanced and stays zero, and no kernel since
// wall_e writes E).  The writes that change the box (B at x = 0, E1 at x = nx+1) touch
// cells no B update reads, so the rows need no ordering and one launch does both.
template <int STEP, int DIMS, bool RESTORE>
OSR_DEV void wall_b_row(const FieldDesc& fd, long long r) {
  const GridDesc& g = fd.g;
  const int halt = g.ctrl->halt;  // checked before the first store, the loads in flight
  const int ny = g.ny, nz = g.nz, nx = g.nx;
  const long long R = (long long)ny * nz;
  if (r >= R) return;
  int j, k;
  row_yz(g, r, j, k);
  const long long nc = g.ncell;
  const float* E1 = g.E;
  const float* E2 = g.E + nc;
  const float* E3 = g.E + 2 * nc;
  const long long rj = (j + 1 == ny ? 0 : j + 1) + (long long)k * ny;  // row y+1
  const long long c0 = r * nx + nx - 1, c1 = r * nx;
  const long long cj = rj * nx + nx - 1;
  const float* stale3 = wall_at(fd, 15, R);
  const float* stale2 = stale3 + nz;
  const float e3j = (STEP == 2 && j == ny - 1) ? stale3[k] : E3[cj];
  float* bb = wall_at(fd, 6, R);
  const float d1 = fd.wall_dt[0], d2 = fd.wall_dt[1], d3 = fd.wall_dt[2];
  float n0, n1, n2;
  if (DIMS == 2) {
    n0 = bb[r] - d2 * (e3j - E3[c0]);
    n1 = bb[R + r] + d1 * (E3[c1] - E3[c0]);
    n2 = bb[2 * R + r] - d1 * (E2[c1] - E2[c0]) + d2 * (E1[cj] - E1[c0]);
  } else {
    const long long rk = j + (long long)(k + 1 == nz ? 0 : k + 1) * ny;  // row z+1
    const long long ck = rk * nx + nx - 1;
    const float e2k = (STEP == 2 && k == nz - 1) ? stale2[j] : E2[ck];
    n0 = bb[r] - d2 * (e3j - E3[c0]) + d3 * (e2k - E2[c0]);
    n1 = bb[R + r] - d3 * (E1[ck] - E1[c0]) + d1 * (E3[c1] - E3[c0]);
    n2 = bb[2 * R + r] - d1 * (E2[c1] - E2[c0]) + d2 * (E1[cj] - E1[c0]);
  }
  float el[3], eu[3];
  const long long row = r * nx;
  const int nxi = g.open_nx;
  if (RESTORE) {
    const float* ebl = wall_at(fd, 0, R);
    const float* ebu = wall_at(fd, 9, R);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      el[c] = ebl[(2 * c) * R + r];
      eu[c] = ebu[(2 * c) * R + r];
    }
  }
  if (halt >= 2) return;
  bb[r] = n0;
  bb[R + r] = n1;
  bb[2 * R + r] = n2;
  if (RESTORE) {
    const float nb[3] = {n0, n1, n2};
#pragma unroll
    for (int c = 0; c < 3; c++) {
      g.E[c * nc + row + nx - 1] = el[c];
      g.B[c * nc + row + nx - 1] = nb[c];
      g.E[c * nc + row + nxi] = eu[c];
    }
  }
}

template <int STEP, int DIMS, bool RESTORE = false>
__global__ __launch_bounds__(256) void wall_b(const FieldDesc* __restrict__ fds) {
  wall_b_row<STEP, DIMS, RESTORE>(*fds, (long long)blockIdx.x * 256 + threadIdx.x);
}

__global__ __launch_bounds__(256) void wall_e(const FieldDesc* __restrict__ fds) {
  const FieldDesc& fd = *fds;
  const GridDesc& g = fd.g;
  const int halt = g.ctrl->halt;  // checked before the first store, the loads in flight
  const int ny = g.ny, nz = g.nz, nx = g.nx, nxi = g.open_nx;
  const long long R = (long long)ny * nz;
  const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
  if (r >= R) return;
  int j, k;
  row_yz(g, r, j, k);
  const long long nc = g.ncell;
  const float dc = fd.wall_dc;
  float* ebl = wall_at(fd, 0, R);
  float* ebu = wall_at(fd, 9, R);
  float* stale3 = wall_at(fd, 15, R);
  float* stale2 = stale3 + nz;
  const long long row = r * nx;
  // lower wall: E(0) = buf(1) + dc (buf(0) - E(1)); buffers <- E(0), E(1)
  // upper wall: E(nx+2) = buf(nx+1) + dc (buf(nx+2) - E(nx+1)); buffers <- E(nx+1), E(nx+2)
  float lb0[3], lb1[3], e1v[3], ub1[3], ub2[3], env[3];
#pragma unroll
  for (int p = 1; p <= 2; p++) {
    const float* E = g.E + p * nc + row;
    lb0[p] = ebl[(2 * p + 0) * R + r];
    lb1[p] = ebl[(2 * p + 1) * R + r];
    e1v[p] = E[0];
    ub1[p] = ebu[(2 * p + 0) * R + r];
    ub2[p] = ebu[(2 * p + 1) * R + r];
    env[p] = E[nxi];
  }
  if (halt >= 2) return;
  if (j == 0) stale3[k] = lb0[2];  // E3(0, 1, z) before this update
  if (k == 0) stale2[j] = lb0[1];  // E2(0, y, 1)
#pragma unroll
  for (int p = 1; p <= 2; p++) {
    float* E = g.E + p * nc + row;
    const float v = lb1[p] + dc * (lb0[p] - e1v[p]);
    E[nx - 1] = v;
    ebl[(2 * p + 0) * R + r] = v;
    ebl[(2 * p + 1) * R + r] = e1v[p];
    const float w = ub1[p] + dc * (ub2[p] - env[p]);
    E[nxi + 1] = w;
    ebu[(2 * p + 0) * R + r] = env[p];
    ebu[(2 * p + 1) * R + r] = w;
  }
}

// the wall values the next push reads: E and B at x = 0 and E at x = nx+1 from the buffers
// (the E1 entries are never written and stay zero)
__global__ __launch_bounds__(256) void wall_restore(const FieldDesc* __restrict__ fds) {
  const FieldDesc& fd = *fds;
  const GridDesc& g = fd.g;
  if (g.ctrl->halt >= 2) return;
  const int nx = g.nx, nxi = g.open_nx;
  const long long R = (long long)g.ny * g.nz;
  const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
  if (r >= R) return;
  const long long nc = g.ncell, row = r * nx;
  const float* ebl = wall_at(fd, 0, R);
  const float* bbl = wall_at(fd, 6, R);
  const float* ebu = wall_at(fd, 9, R);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    g.E[c * nc + row + nx - 1] = ebl[(2 * c) * R + r];
    g.B[c * nc + row + nx - 1] = bbl[c * R + r];
    g.E[c * nc + row + nxi] = ebu[(2 * c) * R + r];
  }
}

// ------------------------------------------------------------------------------------
// VPML absorbing x walls, 2D and 3D (emf/boundary/os-emf-vpml.f03), one thread per (wall,
// cell, row).  EB = 0 runs after the first B half step (update_e_vpml_2d/3d :598-856):
// the B interface (update_interface_x1_2d/3d :2880-3186: box B at box_to_vpml into the
// wall, the whole component in the first of its split pair and zero in the second; wall
// B at vpml_to_box into the box as the sum of each pair), then the wall E update
// (update_e_x1_2d :925-1048, update_e_x1_3d :1172-1365).  EB = 1 runs after the E step
// (update_b_vpml_2d/3d :1742-...): the E interface, then the wall B update
// (update_b_x1_2d :2066-2181, update_b_x1_3d :2311-2508).
// The rows are periodic: the reference's guard rows of the wall are refreshed from the
// interior rows before every step, so its interior rows evolve as on a periodic wall.
// The one wall cell a thread would read while the interface writes it (box_to_vpml,
// the upper wall's first cell for E, the lower wall's last cell for B) is read from the
// box field it is copied from.
// ------------------------------------------------------------------------------------
// split components per wall field: X, Y, ZX, ZY in 2D; XY, XZ, YX, YZ, ZX, ZY in 3D
template <int DIMS>
OSR_DEV float* vp_arr(const FieldDesc& fd, int wl, int eb, int c) {
  constexpr int NC = DIMS == 3 ? 6 : 4;
  return fd.wall + (size_t)((wl * 2 + eb) * NC + c) * fd.vp_n * fd.g.ny * fd.g.nz;
}

// vpml_diffuse (2D, update_e_x1_2d :977-1006): before the wall E update each cell of its
// range mixes the split B3 pair, ZX' = 0.9 ZX + 0.1 ZY, ZY' = 0.9 ZY + 0.1 ZX; a separate
// pass, as every read of the E update sees mixed values (cells earlier in the reference's
// loop order, or the cell itself).  This holds for interpolation orders 2 and 3, whose
// walls have at least two lower guard rows; with one (order 1) the reference leaves row 0
// unmixed and its rows 1 and ny+1 part, which the library refuses.  The B interface runs
// before the mixing in the reference, so this pass also copies the vpml_to_box cell
// (inside the mixed range) into the box, from the values before mixing.
__global__ __launch_bounds__(256) void vpml_diffuse(const FieldDesc* __restrict__ fds) {
  const FieldDesc& fd = *fds;
  const GridDesc& g = fd.g;
  if (g.ctrl->halt >= 2) return;
  const int n = fd.vp_n, ny = g.ny;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= 2LL * n * ny) return;
  const int j = (int)(t % ny);
  const int l = (int)((t / ny) % n);
  const int wl = (int)(t / ((long long)ny * n));
  if (!fd.vp_diff[wl] || l < 1 || l > (wl == 0 ? n - 2 : n - 1)) return;
  const long long o = (long long)l * ny + j;
  float* zx = vp_arr<2>(fd, wl, 1, 2);
  float* zy = vp_arr<2>(fd, wl, 1, 3);
  const float a = zx[o], b = zy[o];
  if (l == (wl == 0 ? n - 2 : 1)) {  // vpml_to_box: x = 0 or nx+2
    const long long nc = g.ncell;
    const long long cv = (long long)j * g.nx + (wl == 0 ? g.nx - 1 : g.open_nx + 1);
    g.B[cv] = vp_arr<2>(fd, wl, 1, 0)[o];
    g.B[nc + cv] = vp_arr<2>(fd, wl, 1, 1)[o];
    g.B[2 * nc + cv] = a + b;
  }
  const float rate = 0.1f, keep = 1.0f - rate;
  zx[o] = keep * a + rate * b;
  zy[o] = keep * b + rate * a;
}

template <int EB, int DIMS>
__global__ __launch_bounds__(256) void vpml_update(const FieldDesc* __restrict__ fds) {
  constexpr int NC = DIMS == 3 ? 6 : 4;
  const FieldDesc& fd = *fds;
  const GridDesc& g = fd.g;
  if (g.ctrl->halt >= 2) return;
  const int n = fd.vp_n, ny = g.ny, nz = g.nz, nx = g.nx;
  const long long R = (long long)ny * nz;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= 2 * n * R) return;
  const long long r = t % R;
  const int l = (int)((t / R) % n);
  const int wl = (int)(t / (R * n));
  int j, k;
  row_yz(g, r, j, k);
  const long long nc = g.ncell;
  // ring columns of the wall's interface cells: x = 1 and 0 (lower), nx+1 and nx+2 (upper)
  const int btv = wl == 0 ? n - 1 : 0, vtb = wl == 0 ? n - 2 : 1;
  const long long cb = r * nx + (wl == 0 ? 0 : g.open_nx);
  const long long cv = r * nx + (wl == 0 ? nx - 1 : g.open_nx + 1);
  const long long o = (long long)l * R + r;
  float* const F = EB == 0 ? g.B : g.E;  // the box field of the interface
  float* W[NC];                          // its wall counterpart
  float* U[NC];                          // the wall field updated
#pragma unroll
  for (int c = 0; c < NC; c++) {
    W[c] = vp_arr<DIMS>(fd, wl, EB == 0 ? 1 : 0, c);
    U[c] = vp_arr<DIMS>(fd, wl, EB == 0 ? 0 : 1, c);
  }
  if (l == btv) {
    if (DIMS == 2) {
      W[0][o] = F[cb];
      W[1][o] = F[nc + cb];
      W[2][o] = F[2 * nc + cb];
      W[3][o] = 0.0f;
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        W[2 * c][o] = F[c * nc + cb];
        W[2 * c + 1][o] = 0.0f;
      }
    }
  } else if (l == vtb && !(DIMS == 2 && EB == 0 && fd.vp_diff[wl])) {  // else vpml_diffuse
    if (DIMS == 2) {
      F[cv] = W[0][o];
      F[nc + cv] = W[1][o];
      F[2 * nc + cv] = W[2][o] + W[3][o];
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) F[c * nc + cv] = W[2 * c][o] + W[2 * c + 1][o];
    }
  }
  const float d2 = fd.vp_dt2, d3 = fd.vp_dt3;
  const float* cf = fd.wall + (size_t)4 * NC * n * R + (size_t)(wl * 2 + EB) * 3 * n;
  // rows y-1, y+1 (and z-1, z+1) of the same cell
  const long long ojm = o + (j == 0 ? ny - 1 : -1), ojp = o + (j + 1 == ny ? 1 - ny : 1);
  const long long okm = o + (k == 0 ? (long long)(nz - 1) * ny : -(long long)ny);
  const long long okp = o + (k + 1 == nz ? -(long long)(nz - 1) * ny : (long long)ny);
  if (EB == 0) {
    if (l < 1 || l > (wl == 0 ? n - 2 : n - 1)) return;
    const int ci = wl == 0 ? l : l - 1;
    const float c1 = cf[ci], c2 = cf[n + ci], c3 = cf[2 * n + ci];
    const long long ol = o - R;  // cell x-1
    const bool from_box = l - 1 == btv;  // upper wall, first updated cell: x-1 = nx+1
    if (DIMS == 2) {
      const float bl1 = from_box ? F[nc + cb] : W[1][ol];
      const float bl23 = from_box ? F[2 * nc + cb] + 0.0f : W[2][ol] + W[3][ol];
      const float b0 = W[0][o], b1 = W[1][o], b2 = W[2][o], b3 = W[3][o];
      U[0][o] = U[0][o] + d2 * (b2 + b3 - W[2][ojm] - W[3][ojm]);
      U[3][o] = U[3][o] - d2 * (b0 - W[0][ojm]);
      U[1][o] = c1 * U[1][o] - c2 * (b2 + b3) + c3 * bl23;
      U[2][o] = c1 * U[2][o] + c2 * b1 - c3 * bl1;
    } else {
      const float blY = from_box ? F[nc + cb] + 0.0f : W[2][ol] + W[3][ol];
      const float blZ = from_box ? F[2 * nc + cb] + 0.0f : W[4][ol] + W[5][ol];
      const float bX = W[0][o] + W[1][o], bY = W[2][o] + W[3][o], bZ = W[4][o] + W[5][o];
      U[0][o] = U[0][o] + d2 * (bZ - W[4][ojm] - W[5][ojm]);
      U[5][o] = U[5][o] - d2 * (bX - W[0][ojm] - W[1][ojm]);
      U[1][o] = U[1][o] - d3 * (bY - W[2][okm] - W[3][okm]);
      U[3][o] = U[3][o] + d3 * (bX - W[0][okm] - W[1][okm]);
      U[2][o] = c1 * U[2][o] - c2 * bZ + c3 * blZ;
      U[4][o] = c1 * U[4][o] + c2 * bY - c3 * blY;
    }
  } else {
    if (l < (wl == 0 ? 0 : 1) || l > n - 2) return;
    const int ci = wl == 0 ? l + 1 : l - 1;
    const float c1 = cf[ci], c2 = cf[n + ci], c3 = cf[2 * n + ci];
    const long long or1 = o + R;  // cell x+1
    const bool from_box = l + 1 == btv;  // lower wall, last updated cell: x+1 = 1
    if (DIMS == 2) {
      const float el1 = from_box ? F[nc + cb] : W[1][or1];
      const float el23 = from_box ? F[2 * nc + cb] + 0.0f : W[2][or1] + W[3][or1];
      const float e0 = W[0][o], e1 = W[1][o], e2 = W[2][o], e3 = W[3][o];
      U[0][o] = U[0][o] - d2 * (W[2][ojp] + W[3][ojp] - e2 - e3);
      U[3][o] = U[3][o] + d2 * (W[0][ojp] - e0);
      U[1][o] = c1 * U[1][o] + c2 * el23 - c3 * (e2 + e3);
      U[2][o] = c1 * U[2][o] - c2 * el1 + c3 * e1;
    } else {
      const float elY = from_box ? F[nc + cb] + 0.0f : W[2][or1] + W[3][or1];
      const float elZ = from_box ? F[2 * nc + cb] + 0.0f : W[4][or1] + W[5][or1];
      const float e0 = W[0][o], e1 = W[1][o], e2 = W[2][o], e3 = W[3][o];
      const float e4 = W[4][o], e5 = W[5][o];
      U[0][o] = U[0][o] - d2 * (W[4][ojp] + W[5][ojp] - e4 - e5);
      U[5][o] = U[5][o] + d2 * (W[0][ojp] + W[1][ojp] - e0 - e1);
      U[1][o] = U[1][o] + d3 * (W[2][okp] + W[3][okp] - e2 - e3);
      U[3][o] = U[3][o] - d3 * (W[0][okp] + W[1][okp] - e0 - e1);
      U[2][o] = c1 * U[2][o] + c2 * elZ - c3 * (e4 + e5);
      U[4][o] = c1 * U[4][o] - c2 * elY + c3 * (e2 + e3);
    }
  }
}

// zpulse_wall launches (launch_em_wall_2d, zpulse/os-zpulse-wall.f03:865-960): B2, B3 of
// one x column += the FP32 values evaluated on the host, launches in list order; before
// the push of the step.
__global__ __launch_bounds__(256) void laser_wall(const FieldDesc* __restrict__ fds) {
  const FieldDesc& fd = *fds;
  const GridDesc& g = fd.g;
  const int halt = g.ctrl->halt;  // loaded together with the step and the table start
  const long long step = g.ctrl->steps, first = *fd.las_first;
  if (halt >= 2) return;
  const int nx = g.nx;
  const long long R = (long long)g.ny * g.nz;  // rows (y, z) of the wall column
  const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
  if (r >= R) return;
  const long long row = step - first;
  if (row < 0 || row >= fd.las_rows) return;
  const long long nc = g.ncell;
  for (int k = 0; k < fd.las_k; k++) {
    const long long e = row * fd.las_k + k;
    if (!fd.las_mask[e]) continue;
    const float* v = fd.las_tab + e * 2 * R;
    const long long c = r * nx + fd.las_col[k];
    g.B[nc + c] = g.B[nc + c] + v[r];
    g.B[2 * nc + c] = g.B[2 * nc + c] + v[R + r];
  }
}


// Neumaier compensated accumulation
OSR_DEV void kadd(double& s, double& c, double v) {
  const double t = s + v;
  if (fabs(s) >= fabs(v))
    c += (s - t) + v;
  else
    c += (v - t) + s;
  s = t;
}

// Merge of two compensated sums (s, c) and (s2, c2): error-free TwoSum of the leading
// parts, the rounding error joins the compensation.
OSR_DEV void kmerge(double& s, double& c, double s2, double c2) {
  const double t = s + s2;
  const double bp = t - s;
  const double err = (s - (t - bp)) + (s2 - bp);
  c = (c + c2) + err;
  s = t;
}

// One block per simulation: compensated reduction of the step's partial sums into one
// history row, then the step counter and the halt protocol.  The order of every addition
// is fixed (strided per-thread sums, then a fixed shuffle tree per wave and the waves in
// order), so the row is reproducible bit for bit.  All 6 + nspec columns go through the
// reduction together: the six field columns in one pass over the blocks (each thread
// reads the six contiguous partials of its blocks, several blocks' loads issued before
// they are accumulated), the shuffle trees of all columns interleaved, and the final
// merge of the 16 wave sums done by one thread per column, so the dependent FP64 chains
// of the columns overlap instead of following one another.
constexpr int kFinNT = 1024;
constexpr int kFinCols = 6 + kMaxSpecies;

// the reduction and the step-end protocol, run by one workgroup of kFinNT threads
OSR_DEV void finalize_body(const FieldDesc& fd) {
  __shared__ double w_s[kFinNT / 64][kFinCols], w_c[kFinNT / 64][kFinCols];
  const GridDesc& g = fd.g;
  Ctrl* ctrl = g.ctrl;
  // read with the partial sums; a halted simulation writes nothing (the reduction itself
  // only reads, so it need not wait for the flag)
  const int halt = ctrl->halt;
  const long long step = ctrl->steps;
  const int nspec = fd.nspec;
  const int ncol = 6 + nspec;
  constexpr int U = 2;  // items per column loaded ahead per thread (the order of the adds
                       // is the same for any U)
  double s[kFinCols], c[kFinCols];
#pragma unroll
  for (int col = 0; col < kFinCols; col++) s[col] = c[col] = 0.0;
  const int tid = threadIdx.x;
  const int nblk = fd.nblk;
  // first pass: a thread's first U items of the field columns and of the first kPre
  // species, all loads issued before the first addition (one memory round trip instead
  // of one per column group); kPre bounds the registers the loads ahead take
  constexpr int kPre = 2;
  {
    double vf[U][6], vs[kPre][U];
    int epn[kPre];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int b = tid + u * kFinNT;
#pragma unroll
      for (int col = 0; col < 6; col++)
        vf[u][col] = b < nblk ? fd.fe_part[(long long)b * 6 + col] : 0.0;
    }
#pragma unroll
    for (int sp = 0; sp < kPre; sp++) {
      epn[sp] = sp < nspec ? fd.ep_n[sp] : 0;
      const double* ep = fd.ep[sp];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int t = tid + u * kFinNT;
        vs[sp][u] = t < epn[sp] ? ep[t] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (tid + u * kFinNT < nblk)
#pragma unroll
        for (int col = 0; col < 6; col++) kadd(s[col], c[col], vf[u][col]);
#pragma unroll
    for (int sp = 0; sp < kPre; sp++)
#pragma unroll
      for (int u = 0; u < U; u++)
        if (tid + u * kFinNT < epn[sp]) kadd(s[6 + sp], c[6 + sp], vs[sp][u]);
  }
  // the remaining items of every column (many cell blocks or tiles, more species), in
  // order
  for (int b0 = tid + U * kFinNT; b0 < nblk; b0 += kFinNT * U) {
    double v[U][6];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int b = b0 + u * kFinNT;
#pragma unroll
      for (int col = 0; col < 6; col++)
        v[u][col] = b < nblk ? fd.fe_part[(long long)b * 6 + col] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (b0 + u * kFinNT < nblk)
#pragma unroll
        for (int col = 0; col < 6; col++) kadd(s[col], c[col], v[u][col]);
  }
#pragma unroll
  for (int sp = 0; sp < kMaxSpecies; sp++) {
    if (sp >= nspec) break;
    const double* ep = fd.ep[sp];
    const int n = fd.ep_n[sp];
    for (int t0 = tid + (sp < kPre ? U * kFinNT : 0); t0 < n; t0 += kFinNT * U) {
      double v[U];
#pragma unroll
      for (int u = 0; u < U; u++) v[u] = t0 + u * kFinNT < n ? ep[t0 + u * kFinNT] : 0.0;
#pragma unroll
      for (int u = 0; u < U; u++)
        if (t0 + u * kFinNT < n) kadd(s[6 + sp], c[6 + sp], v[u]);
    }
  }
  // per wave: the fixed shuffle tree, all columns at each level
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int col = 0; col < kFinCols; col++) {
      if (col >= ncol) break;
      const double s2 = __shfl_down(s[col], off), c2 = __shfl_down(c[col], off);
      kmerge(s[col], c[col], s2, c2);
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int col = 0; col < kFinCols; col++) {
      w_s[wave][col] = s[col];
      w_c[wave][col] = c[col];
    }
  }
  __syncthreads();
  if (halt >= 2) return;  // uniform over the block, after its last barrier
  // the waves in order, one thread per column
  if (threadIdx.x < ncol) {
    const int col = threadIdx.x;
    double ss = w_s[0][col], cc = w_c[0][col];
    for (int w = 1; w < kFinNT / 64; w++) kmerge(ss, cc, w_s[w][col], w_c[w][col]);
    if (step < fd.hist_cap)
      fd.hist[step * ncol + col] = (ss + cc) * (col < 6 ? fd.fe_scale : fd.ep_scale[col - 6]);
  }
  __syncthreads();  // every read of the control block above precedes its update
  if (threadIdx.x == 0) {
    ctrl->steps = step + 1;
    if (halt == 1)
      ctrl->halt = 2;
    else if (ctrl->target > 0 && step + 1 >= ctrl->target)
      ctrl->halt = 3;
  }
}

__global__ __launch_bounds__(kFinNT) void finalize_step(const FieldDesc* __restrict__ fds,
                                                        const int* __restrict__ work) {
  finalize_body(fds[work[blockIdx.x]]);
}

// Lindman walls: the lower wall's second B update with the restore of the wall values
// (wall_b<2, DIMS, true>), kFinNT rows per workgroup, and the end of the step in the
// workgroup that finishes last (a ticket on Ctrl::done), so no kernel launch follows.
// Every workgroup has read the halt flag before it takes its ticket, so the flag changes
// only once none of them looks at it any more; the reduction reads partial sums of
// earlier kernels only.
template <int DIMS>
__global__ __launch_bounds__(kFinNT) void wall_b2_end(const FieldDesc* __restrict__ fds) {
  __shared__ int s_last;
  const FieldDesc& fd = *fds;
  wall_b_row<2, DIMS, true>(fd, (long long)blockIdx.x * kFinNT + threadIdx.x);
  __syncthreads();
  if (threadIdx.x == 0) s_last = atomicAdd(&fd.g.ctrl->done, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!s_last) return;
  if (threadIdx.x == 0) fd.g.ctrl->done = 0;
  finalize_body(fd);
}

// ------------------------------------------------------------------------------------
// Gauss-law diagnostic (FP64): rho on cell corners with the same splines as
// DEPOSIT_RHO_2D/3D (spec/os-spec-charge.f90:500-580), accumulated in int64 fixed point,
// and div E (vdf/os-vdf-math.f90:1013-1073, offset 0).
// ------------------------------------------------------------------------------------
// charge-density splines in FP64 (spec/os-spec-charge.f90 uses the interpolation
// splines, emf/os-emf-interpolate.f03:171-293)
template <int ORDER>
OSR_DEV void dspline(double x, double* w) {
  const double t0 = 0.5 - x, t1 = 0.5 + x;
  if (ORDER == 1) {
    w[0] = t0;
    w[1] = t1;
  } else if (ORDER == 2) {
    w[0] = 0.5 * (t0 * t0);
    w[1] = 0.5 + t0 * t1;
    w[2] = 0.5 * (t1 * t1);
  } else {
    const double t2 = t0 * t0, t3 = t1 * t1;
    w[0] = (t0 * t2) / 6.0;
    w[1] = (2.0 / 3.0 - t3) + 0.5 * (t1 * t3);
    w[2] = (2.0 / 3.0 - t2) + 0.5 * (t0 * t2);
    w[3] = (t1 * t3) / 6.0;
  }
}

template <int ORDER>
__global__ __launch_bounds__(256) void rho_deposit(const SpecDesc* __restrict__ descs,
                                                   const int2* __restrict__ work,
                                                   unsigned long long* rho, double scale) {
  using O = Ord<ORDER>;
  constexpr int NP = O::NP;
  const int2 wk = work[blockIdx.x];
  const SpecDesc& d = descs[wk.x];
  const GridDesc& g = d.g;
  const int tile = wk.y;
  const int base = d.off[tile], n = d.cnt[tile];
  const int dims = g.dims;
  for (int p = threadIdx.x; p < n; p += 256) {
    const int s = base + p;
    int ix, iy, iz = 0;
    if (dims == 2)
      unpack2(d.cell[s], ix, iy);
    else
      unpack3(d.cell[s], ix, iy, iz);
    double w1[NP], w2[NP], w3[NP];
    dspline<ORDER>((double)d.x[s], w1);
    dspline<ORDER>((double)d.y[s], w2);
    if (dims == 3)
      dspline<ORDER>((double)d.z[s], w3);
    else
#pragma unroll
      for (int k = 0; k < NP; k++) w3[k] = k == 0 ? 1.0 : 0.0;
    const double q = d.q[s];
#pragma unroll
    for (int k3 = 0; k3 < NP; k3++) {
      if (dims == 2 && k3 > 0) break;
      const long long zoff =
          dims == 3 ? (long long)g.nx * g.ny * wrp(iz + O::LP + k3, g.nz) : 0;
#pragma unroll
      for (int k2 = 0; k2 < NP; k2++) {
        const long long row = zoff + (long long)g.nx * wrp(iy + O::LP + k2, g.ny);
#pragma unroll
        for (int k1 = 0; k1 < NP; k1++) {
          const double w = ((q * w1[k1]) * w2[k2]) * w3[k3];
          const long long v = __double2ll_rn(w * scale);
          atomicAdd(&rho[row + wrp(ix + O::LP + k1, g.nx)], (unsigned long long)v);
        }
      }
    }
  }
}

__global__ __launch_bounds__(256) void gauss_residual(GridDesc g, const unsigned long long* rho,
                                                      double rscale, double r1, double r2,
                                                      double r3, double* res) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= g.ncell) return;
  const long long nc = g.ncell;
  const int nx = g.nx, ny = g.ny;
  int i, j, k;
  cell_xyz(g, idx, i, j, k);
  const long long im = idx - i + wrp(i - 1, nx);
  const long long jm = idx + (long long)(wrp(j - 1, ny) - j) * nx;
  double dv = ((double)g.E[idx] - (double)g.E[im]) * r1 +
              ((double)g.E[nc + idx] - (double)g.E[nc + jm]) * r2;
  if (g.dims == 3) {
    const long long km = idx + (long long)(wrp(k - 1, g.nz) - k) * nx * ny;
    dv += ((double)g.E[2 * nc + idx] - (double)g.E[2 * nc + km]) * r3;
  }
  res[idx] = dv - (double)(long long)rho[idx] * rscale;
}

// ------------------------------------------------------------------------------------
// Initial electrostatic field (init_fields, spec/os-spec-initfields.f90:38-100, and
// es_solver_2d, emf/os-emf-es-solver.f03:125-236): the free-space direct sum of
// charge * d / r^2 over every interior source cell with nonzero charge, per target cell
// in source order (y outer, x inner), each term in FP64 and accumulated into an FP32
// field as the reference's p_k_fld grid does; then times the cell volume and added to E.
// ------------------------------------------------------------------------------------

// the fixed-point charge density of the interior, rounded to FP32 (the reference's rho
// grid) and widened to FP64, row-major
__global__ __launch_bounds__(256) void es_charge(GridDesc g, const unsigned long long* rho,
                                                 double inv, double* charge) {
  const int nxi = interior_nx(g);
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long long)nxi * g.ny * g.nz) return;
  const long long c = (t / nxi) * g.nx + t % nxi;  // rows y + ny z of nx cells
  charge[t] = (double)(float)((double)(long long)rho[c] * inv);
}

constexpr int kEsChunk = 1024;

// one thread per target cell; source charges staged through LDS (every lane reads the same
// source: broadcast reads, uniform branches)
__global__ __launch_bounds__(256) void es_direct_2d(GridDesc g, const double* __restrict__ charge,
                                                    double d0, double d1, double vol) {
  __shared__ double s_q[kEsChunk];
  const int nxi = interior_nx(g), ny = g.ny;
  const long long nt = (long long)nxi * ny;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  const bool live = t < nt;
  const int ei1 = live ? (int)(t % nxi) : 0, ei2 = live ? (int)(t / nxi) : 0;
  float a1 = 0.f, a2 = 0.f;
  for (long long c0 = 0; c0 < nt; c0 += kEsChunk) {
    const int m = (int)(nt - c0 < kEsChunk ? nt - c0 : kEsChunk);
    __syncthreads();
    for (int k = threadIdx.x; k < m; k += 256) s_q[k] = charge[c0 + k];
    __syncthreads();
    if (!live) continue;
    int ri1 = (int)(c0 % nxi), ri2 = (int)(c0 / nxi);
    for (int k = 0; k < m; k++) {
      const double q = s_q[k];
      if (q != 0) {
        double y1 = (double)(ei2 - ri2);
        double y2 = y1 + 0.5;
        y1 = y1 * d1;
        y2 = y2 * d1;
        double x2 = (double)(ei1 - ri1);
        double x1 = x2 + 0.5;
        x1 = x1 * d0;
        x2 = x2 * d0;
        double r2 = x1 * x1 + y1 * y1;
        a1 = (float)((double)a1 + q * x1 / r2);
        r2 = x2 * x2 + y2 * y2;
        a2 = (float)((double)a2 + q * y2 / r2);
      }
      if (++ri1 == nxi) {
        ri1 = 0;
        ri2++;
      }
    }
  }
  if (!live) return;
  a1 = (float)((double)a1 * vol);
  a2 = (float)((double)a2 * vol);
  const long long nc = g.ncell, c = (long long)ei2 * g.nx + ei1;
  g.E[c] = g.E[c] + a1;
  g.E[nc + c] = g.E[nc + c] + a2;
  g.E[2 * nc + c] = g.E[2 * nc + c] + 0.0f;  // E3 of the ES field is 0 (-0 becomes +0)
}

// es_solver_2d_beam (emf/os-emf-es-solver.f03:457-616), a species drifting along x: E2 of
// each cell = sum over the charge of its own column, in source order (y ascending), of
// rho * sign(dy / gamma / 2, 0.5 + (j - rj)), accumulated in FP32 through FP64 sums
// (before the gamma scaling); written to e2u [ny][nxi]
__global__ __launch_bounds__(256) void es_beam_e2(GridDesc g, const double* __restrict__ charge,
                                                  double cn, float* e2u) {
  const int nxi = interior_nx(g), ny = g.ny;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long long)nxi * ny) return;
  const int i = (int)(t % nxi), j = (int)(t / nxi);
  float a = 0.f;
  for (int rj = 0; rj < ny; rj++) {
    const double q = charge[(long long)rj * nxi + i];
    if (q != 0) a = (float)((double)a + q * (j - rj >= 0 ? cn : -cn));
  }
  e2u[t] = a;
}

// then B3 = 0.5 u1 (E2(i) + E2(i+1)) (FP32 sum; E2 beyond an open upper wall is zero),
// E2 *= gamma, E1 = E3 = 0, and the fields are added to E and B (add(emf%e, ef),
// add(emf%b, bf)); zero components are added too (-0 becomes +0 as in the reference)
__global__ __launch_bounds__(256) void es_beam_apply(GridDesc g, const float* __restrict__ e2u,
                                                     float u1, double gamma) {
  const int nxi = interior_nx(g), ny = g.ny;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long long)nxi * ny) return;
  const int i = (int)(t % nxi), j = (int)(t / nxi);
  const float e2 = e2u[t];
  float e2n;
  if (i + 1 < nxi)
    e2n = e2u[t + 1];
  else
    e2n = g.open_nx ? 0.f : e2u[(long long)j * nxi];  // periodic image of x = 1
  const float sum = e2 + e2n;
  const float b3 = (float)(0.5 * (double)u1 * (double)sum);
  const long long nc = g.ncell, c = (long long)j * g.nx + i;
  g.E[c] = g.E[c] + 0.0f;
  g.E[nc + c] = g.E[nc + c] + (float)((double)e2 * gamma);
  g.E[2 * nc + c] = g.E[2 * nc + c] + 0.0f;
  g.B[c] = g.B[c] + 0.0f;
  g.B[nc + c] = g.B[nc + c] + 0.0f;
  g.B[2 * nc + c] = g.B[2 * nc + c] + b3;
}

// es_solver_3d_beam (emf/os-emf-es-solver.f03:795-1003), a species drifting along x in
// 3D: each cell of x slice i gets the field of the infinite wires of that slice's charge,
// sources in the reference's order (z outer, y inner), each term
// q * cn * d / r^2 in FP64 and accumulated in FP32 through FP64 sums: E2 at (y + 1/2, z),
// E3 at (y, z + 1/2).  One workgroup per (slice, 256 targets); the slice's charges are
// staged through LDS (every lane reads the same source: broadcast reads).  Writes the
// unboosted E2, E3 to e23u [2][nz][ny][nxi].
__global__ __launch_bounds__(256) void es_beam3d_e(GridDesc g, const double* __restrict__ charge,
                                                   double cn, double d2, double d3, float* e23u) {
  __shared__ double s_q[kEsChunk];
  const int nxi = interior_nx(g), ny = g.ny, nz = g.nz;
  const int i = blockIdx.y;
  const long long ns = (long long)ny * nz;  // sources (and targets) of the slice
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  const bool live = t < ns;
  const int ej = live ? (int)(t % ny) : 0, ek = live ? (int)(t / ny) : 0;
  float a2 = 0.f, a3 = 0.f;
  for (long long c0 = 0; c0 < ns; c0 += kEsChunk) {
    const int m = (int)(ns - c0 < kEsChunk ? ns - c0 : kEsChunk);
    __syncthreads();
    for (int k = threadIdx.x; k < m; k += 256) s_q[k] = charge[(c0 + k) * nxi + i];
    __syncthreads();
    if (!live) continue;
    int rj = (int)(c0 % ny), rk = (int)(c0 / ny);
    for (int k = 0; k < m; k++) {
      const double q = s_q[k];
      if (q != 0) {
        const double dx3c = (double)(ek - rk);
        const double dx3s = dx3c * d3;
        const double dx3s_h = (dx3c + 0.5) * d3;
        const double dx2c = (double)(ej - rj);
        const double dx2s = dx2c * d2;
        const double dx2s_h = (dx2c + 0.5) * d2;
        double r2 = dx2s_h * dx2s_h + dx3s * dx3s;
        a2 = (float)((double)a2 + q * cn * dx2s_h / r2);
        r2 = dx2s * dx2s + dx3s_h * dx3s_h;
        a3 = (float)((double)a3 + q * cn * dx3s_h / r2);
      }
      if (++rj == ny) {
        rj = 0;
        rk++;
      }
    }
  }
  if (!live) return;
  const long long nt = (long long)nxi * ns, o = t * nxi + i;
  e23u[o] = a2;
  e23u[nt + o] = a3;
}

// then B2 = -0.5 u1 (E3(i) + E3(i+1)), B3 = 0.5 u1 (E2(i) + E2(i+1)) (FP32 sums; the field
// beyond an open upper wall is zero), E2, E3 *= gamma, E1 = B1 = 0, and the fields are
// added to E and B (add(emf%e, ef), add(emf%b, bf)); zero components are added too
__global__ __launch_bounds__(256) void es_beam3d_apply(GridDesc g, const float* __restrict__ e23u,
                                                       float u1, double gamma) {
  const int nxi = interior_nx(g);
  const long long nt = (long long)nxi * g.ny * g.nz;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= nt) return;
  const int i = (int)(t % nxi);
  const long long row = t / nxi;
  const float e2 = e23u[t], e3 = e23u[nt + t];
  float e2n, e3n;
  if (i + 1 < nxi) {
    e2n = e23u[t + 1];
    e3n = e23u[nt + t + 1];
  } else if (g.open_nx) {
    e2n = e3n = 0.f;
  } else {  // periodic image of x = 1
    e2n = e23u[row * nxi];
    e3n = e23u[nt + row * nxi];
  }
  const float s2 = e2 + e2n, s3 = e3 + e3n;
  const float b2 = (float)(-(0.5 * (double)u1 * (double)s3));
  const float b3 = (float)(0.5 * (double)u1 * (double)s2);
  const long long nc = g.ncell, c = row * g.nx + i;
  g.E[c] = g.E[c] + 0.0f;
  g.E[nc + c] = g.E[nc + c] + (float)((double)e2 * gamma);
  g.E[2 * nc + c] = g.E[2 * nc + c] + (float)((double)e3 * gamma);
  g.B[c] = g.B[c] + 0.0f;
  g.B[nc + c] = g.B[nc + c] + b2;
  g.B[2 * nc + c] = g.B[2 * nc + c] + b3;
}

template __global__ void yee_dbdt<2, true>(const FieldDesc*, const int2*);
template __global__ void yee_dbdt<2, false>(const FieldDesc*, const int2*);
template __global__ void yee_dbdt<3, true>(const FieldDesc*, const int2*);
template __global__ void yee_dbdt<3, false>(const FieldDesc*, const int2*);
template __global__ void yee_dedt<2>(const FieldDesc*, const int2*);
template __global__ void wall_b2_end<2>(const FieldDesc*);
template __global__ void wall_b2_end<3>(const FieldDesc*);
template __global__ void vpml_update<0, 2>(const FieldDesc*);
template __global__ void vpml_update<1, 2>(const FieldDesc*);
template __global__ void vpml_update<0, 3>(const FieldDesc*);
template __global__ void vpml_update<1, 3>(const FieldDesc*);
template __global__ void yee_dedt<3>(const FieldDesc*, const int2*);
template __global__ void rho_deposit<1>(const SpecDesc*, const int2*, unsigned long long*, double);
template __global__ void rho_deposit<2>(const SpecDesc*, const int2*, unsigned long long*, double);
template __global__ void rho_deposit<3>(const SpecDesc*, const int2*, unsigned long long*, double);

}  // namespace osr
