// k_loader.hip -- synthetic particle generation directly in the tile layout.
//
// Positions: regular lattice inside the cell (spec/psource/os-psource-std.f03:822-869:
// (2*i - 1 - n) / (2n) along each dimension) or uniform random in [-0.5, 0.5).
// Momenta: u = ufl + uth * N(0,1) per component (spec/os-spec-udist.f03:577-579 plus the
// fluid momentum).  Random numbers come from a counter-based generator (SplitMix64 of
// seed, particle id and draw index), so the result does not depend on the launch shape.
// Charges: q = sign(rqm) * density / ppc (os-psource-std.f03:846).
//
// Slot p of a tile holds the (p / ncell_tile)-th particle of the cell of key
// p % ncell_tile (cell_key): consecutive lanes of a wave work on different cells, which
// spreads their LDS current atomics over different banks and addresses.
#include "osr_device.hpp"
#include "osr_types.hpp"

namespace osr {

OSR_DEV uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// uniform double in (0, 1]
OSR_DEV double urand(uint64_t seed, uint64_t id, uint32_t k) {
  const uint64_t h = splitmix64(seed ^ splitmix64(id * 0x100000001B3ull + k));
  return ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

OSR_DEV double gauss(uint64_t seed, uint64_t id, uint32_t k) {
  const double u1 = urand(seed, id, 2 * k), u2 = urand(seed, id, 2 * k + 1);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

struct GenArgs {
  int ppc[3];
  int ppc_tot;
  float q;
  double ufl[3];
  double uth[3];
  uint64_t seed;
  int random_pos;
};

__global__ __launch_bounds__(256) void generate_particles(const SpecDesc* __restrict__ descs,
                                                          int sid, GenArgs a) {
  const SpecDesc& d = descs[sid];
  const GridDesc& g = d.g;
  const int tile = blockIdx.x;
  int tx0, ty0, tz0, wx, wy, wz;
  if (g.dims == 2) {
    tx0 = (tile % g.ntx) * g.tx;
    ty0 = (tile / g.ntx) * g.ty;
    tz0 = 0;
    wz = 1;
  } else {
    tx0 = (tile % g.ntx) * g.tx;
    ty0 = ((tile / g.ntx) % g.nty) * g.ty;
    tz0 = (tile / (g.ntx * g.nty)) * g.tz;
    wz = min(g.tz, g.nz - tz0);
  }
  wx = min(g.tx, interior_nx(g) - tx0);
  wy = min(g.ty, g.ny - ty0);
  const int nct = wx * wy * wz;
  const int n = d.cnt[tile];
  const int base = d.off[tile];
  // full tiles follow the in-tile cell order (cell_key), partial ones are walked x first
  const bool full = wx == g.tx && wy == g.ty && wz == (g.dims == 3 ? g.tz : 1);
  for (int p = blockIdx.y * 256 + threadIdx.x; p < n; p += 256 * gridDim.y) {
    const int c = p % nct, k = p / nct;
    int lx, ly, lz;
    if (full) {
      key_cell(c, g.dims, g.tx, g.ty, g.dims == 3 ? g.tz : 1, lx, ly, lz);
    } else {
      lx = c % wx;
      ly = (c / wx) % wy;
      lz = c / (wx * wy);
    }
    const int ix = tx0 + lx, iy = ty0 + ly, iz = tz0 + lz;
    const long long gcell = ix + (long long)interior_nx(g) * (iy + (long long)g.ny * iz);
    const uint64_t id = (uint64_t)gcell * (uint64_t)a.ppc_tot + (uint64_t)k;
    float pos[3];
    if (a.random_pos) {
      for (int dd = 0; dd < g.dims; dd++) {
        // uniform in [-0.5, 0.5): 24-bit grid, exactly representable in float
        const uint64_t h = splitmix64(a.seed ^ splitmix64(id * 0x100000001B3ull + 100 + dd));
        pos[dd] = (float)((double)(h >> 40) * (1.0 / 16777216.0)) - 0.5f;
      }
    } else {
      // lattice index of particle k: dimension 1 slowest (os-psource-std.f03:846-866)
      int rem = k;
      int li[3] = {0, 0, 0};
      for (int dd = g.dims - 1; dd >= 0; dd--) {
        li[dd] = rem % a.ppc[dd];
        rem /= a.ppc[dd];
      }
      for (int dd = 0; dd < g.dims; dd++)
        pos[dd] = (float)((2 * (li[dd] + 1) - 1 - a.ppc[dd]) * (0.5 / a.ppc[dd]));
    }
    const int s = base + p;
    d.cell[s] = g.dims == 2 ? pack2(ix, iy) : pack3(ix, iy, iz);
    d.x[s] = pos[0];
    d.y[s] = pos[1];
    if (g.dims == 3) d.z[s] = pos[2];
    d.ux[s] = (float)(a.ufl[0] + a.uth[0] * gauss(a.seed, id, 0));
    d.uy[s] = (float)(a.ufl[1] + a.uth[1] * gauss(a.seed, id, 1));
    d.uz[s] = (float)(a.ufl[2] + a.uth[2] * gauss(a.seed, id, 2));
    d.q[s] = a.q;
    d.pid[s] = (uint32_t)id;
  }
}

}  // namespace osr
