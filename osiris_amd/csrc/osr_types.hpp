// osr_types.hpp -- descriptors shared by host code and kernels.
//
// Device data layout of one simulation (all resident in HBM):
//   fields   E, B : float, planar [3][nz][ny][nx] (interior cells only; periodic
//                   boundaries are index wraps, so no guard cells are stored).  With
//                   open x walls the x extent is a ring of the open_nx interior columns
//                   followed by a guard band of whole tiles: columns open_nx, open_nx+1,
//                   ... are the upper guard cells x = nx+1, nx+2, ... and the last
//                   columns, reached by the same index wrap from column 0, the lower
//                   guard cells x = 0, -1, ... (1-based cell indices of the reference);
//   current  J    : int64 fixed point, planar [3][nz][ny][nx]; the value is J * 2^jshift;
//   particles     : structure of arrays, grouped by tile.  Tile t owns the slots
//                   [off[t], off[t]+cap[t]) of every array and its first cnt[t] slots hold
//                   particles whose cell lies inside the tile.  A particle is
//                   cell (packed uint32), x, y[, z] in [-0.5, 0.5) relative to the cell
//                   centre, ux, uy, uz (proper velocity), q, pid (upload order).
//   movers        : one buffer per species with the same tile layout; the push kernel
//                   writes particles that leave their tile there, one contiguous region
//                   per (tile, wave), and the fill kernel moves them into their new tile.
#pragma once
#include <stdint.h>

namespace osr {

constexpr int kMaxSpecies = 8;
// Push workgroups have 4 waves, or 8 for the 2D kernels with compile-time tiles of at least
// 512 cells (by default; 4 or 8 on request for 16x16 and 32x16): their field and current
// tiles in LDS are then shared by twice as many waves, so two workgroups (16 waves) fit in
// a CU's LDS where three 4-wave ones (12 waves) did.
constexpr int kMaxPushWaves = 8;
// capacity of the per-wave virtual-particle queue of the 3D push (pending < 64 plus up to
// 64 appended per VP level)
constexpr int kVpQueue3 = 128;
// default waves per workgroup of the 2D compile-time tile kernels
__host__ __device__ constexpr int push2d_default_waves(int tx, int ty) {
  return tx * ty >= 512 ? 8 : 4;
}

// Unsigned division of 32-bit values by a divisor fixed per grid, as a multiply-high and
// two shifts: the round-up method of Granlund and Montgomery ("Division by invariant
// integers using multiplication", PLDI 1994, fig. 4.1), exact for every dividend below
// 2^32 and every divisor 1 <= d <= 2^31.  With l = ceil(log2 d) and
// m = floor(2^32 (2^l - d) / d) + 1:  q = (t + ((n - t) >> 1)) >> (l - 1), t = mulhi(m, n)
// (d = 1: m = 1, both shifts 0).  The field kernels split cell indices (< 2^32 for every
// grid the library accepts) into coordinates with it instead of 64-bit divisions.
struct FastDiv {
  uint32_t m;
  uint32_t s1, s2;
  uint32_t d;
  __host__ void set(uint32_t dv) {
    uint32_t l = 0;
    while ((1ull << l) < dv) l++;
    d = dv;
    m = (uint32_t)((((1ull << 32) * ((1ull << l) - dv)) / dv) + 1);
    s1 = l > 0 ? 1 : 0;
    s2 = l > 0 ? l - 1 : 0;
  }
  __host__ __device__ uint32_t div(uint32_t n) const {
    const uint32_t t = (uint32_t)(((uint64_t)m * n) >> 32);
    return (t + ((n - t) >> s1)) >> s2;
  }
};

// Control block, one per simulation (device memory).
struct Ctrl {
  int halt;          // 0 run, 1 a tile overflowed this step, 2 stopped at a step end,
                     // 3 parked: the step target was reached
  int error;         // sticky: 1 = particle found outside its tile (layout invariant broken)
  long long steps;   // completed steps
  long long target;  // park (halt = 3) once steps reaches target; 0 = no target
  unsigned done;     // workgroups of a step-ending kernel that have finished (the last one
                     // ends the step and sets it back to 0)
  unsigned pad;
};

struct GridDesc {
  int dims;
  int nx, ny, nz;
  long long ncell;
  int tx, ty, tz;     // tile size in cells
  int ntx, nty, ntz;  // tiles per dimension
  int ntiles;
  int open_nx;  // 0: x periodic; else open x walls and this many interior columns (< nx)
  float* E;
  float* B;
  unsigned long long* J;
  double jscale;     // 2^jshift
  double jinv;       // 2^-jshift
  float dt_dx[3];    // (float)(dt/dx)
  float jnorm[3];    // (float)(dx/dt/2) in 2D, (float)(dx/dt/3) in 3D
  Ctrl* ctrl;
  FastDiv div_nx, div_ny;  // cell index c -> x = c mod nx, row c / nx; row -> y, z
};

// the divisors of a grid whose extents are set (host)
__host__ inline void set_grid_divisors(GridDesc& g) {
  g.div_nx.set((uint32_t)g.nx);
  g.div_ny.set((uint32_t)g.ny);
}

struct SpecDesc {
  GridDesc g;
  // particles
  uint32_t* cell;
  float* x;
  float* y;
  float* z;
  float* ux;
  float* uy;
  float* uz;
  float* q;
  uint32_t* pid;
  // movers
  uint32_t* mcell;
  float* mx;
  float* my;
  float* mz;
  float* mux;
  float* muy;
  float* muz;
  float* mq;
  uint32_t* mpid;
  int* midx;       // slot the mover left (relative to the tile offset)
  // tiles
  int* off;
  int* cnt;
  int* cap;
  int* mv_cnt;     // [ntiles][waves]
  int* mv_stride;  // [ntiles] per-wave region size used by the last push
  int* newcnt;     // [ntiles]
  int* inc;        // [ntiles] incoming count
  double* epart;   // [ntiles] time-centred energy partial sums (un-normalised)
  float tem;       // (float)(0.5*dt/rqm)
  int waves;       // waves per workgroup in the push kernel
};

// Cell packing.  2D: ix | iy<<16.  3D: ix | iy<<11 | iz<<22.
__host__ __device__ inline uint32_t pack2(int ix, int iy) {
  return (uint32_t)ix | ((uint32_t)iy << 16);
}
__host__ __device__ inline uint32_t pack3(int ix, int iy, int iz) {
  return (uint32_t)ix | ((uint32_t)iy << 11) | ((uint32_t)iz << 22);
}
__host__ __device__ inline void unpack2(uint32_t c, int& ix, int& iy) {
  ix = (int)(c & 0xFFFFu);
  iy = (int)(c >> 16);
}
__host__ __device__ inline void unpack3(uint32_t c, int& ix, int& iy, int& iz) {
  ix = (int)(c & 0x7FFu);
  iy = (int)((c >> 11) & 0x7FFu);
  iz = (int)(c >> 22);
}

// interior cells along x (the ring without the guard band of open x walls)
__host__ __device__ inline int interior_nx(const GridDesc& g) {
  return g.open_nx ? g.open_nx : g.nx;
}

__host__ __device__ inline int tile_of(const GridDesc& g, uint32_t c) {
  if (g.dims == 2) {
    int ix, iy;
    unpack2(c, ix, iy);
    return ix / g.tx + g.ntx * (iy / g.ty);
  }
  int ix, iy, iz;
  unpack3(c, ix, iy, iz);
  return ix / g.tx + g.ntx * (iy / g.ty + g.nty * (iz / g.tz));
}

// Order of the cells of a tile in the particle layout: the loader places particle k of
// a cell at slot k * ncell + key, the in-tile reorder restores (rank within the cell,
// key) order, and the host layout uses the same keys.  Consecutive lanes of a wave
// therefore hold consecutive keys.  3D tiles whose sides are multiples of (4, 4, 2) are
// walked in blocks of 4 x 4 x 2 cells (x fastest inside a block, blocks x fastest), so
// that the 32 lanes of a half wave hold one compact block and their LDS stencil reads
// spread over the banks (with the 12-float row and 144-float plane pitches of the 8x8x4
// kernel, lx + 12 ly + 144 lz covers 32 distinct banks for the cells of one block).
// Other tiles are walked x fastest.  The key only orders particles: any order gives the
// same results.
__host__ __device__ inline bool blocked_tile(int dims, int tx, int ty, int tz) {
  return dims == 3 && tx % 4 == 0 && ty % 4 == 0 && tz % 2 == 0;
}
__host__ __device__ inline int cell_key(int lx, int ly, int lz, int dims, int tx, int ty, int tz) {
  if (!blocked_tile(dims, tx, ty, tz)) return lx + tx * (ly + ty * lz);
  const int blk = (lx >> 2) + (tx >> 2) * ((ly >> 2) + (ty >> 2) * (lz >> 1));
  return 32 * blk + (lx & 3) + 4 * (ly & 3) + 16 * (lz & 1);
}
// inverse of cell_key over a full tile
__host__ __device__ inline void key_cell(int key, int dims, int tx, int ty, int tz, int& lx, int& ly,
                                         int& lz) {
  if (!blocked_tile(dims, tx, ty, tz)) {
    lx = key % tx;
    ly = (key / tx) % ty;
    lz = key / (tx * ty);
    return;
  }
  const int blk = key >> 5, w = key & 31;
  const int bx = blk % (tx >> 2), by = (blk / (tx >> 2)) % (ty >> 2), bz = blk / ((tx >> 2) * (ty >> 2));
  lx = 4 * bx + (w & 3);
  ly = 4 * by + ((w >> 2) & 3);
  lz = 2 * bz + (w >> 4);
}

}  // namespace osr
