// fastdiv_check.cpp -- host check of osr::FastDiv (osiris_amd/csrc/osr_types.hpp) against
// the integer division, built with g++ by tests/test_fastdiv.py (test infrastructure).
//
// Divisors: every d from 1 to 70000 (every grid extent the library accepts) and a spread
// of larger ones up to 2^31; dividends: the edges (0, 1, multiples of d and their
// neighbours, 2^32 - 1) and pseudo-random values; for a few divisors, every dividend.
// Prints the first mismatch and exits 1, or prints "ok" and exits 0.
#include <cstdint>
#include <cstdio>

#define __host__
#define __device__
#include "../../osiris_amd/csrc/osr_types.hpp"

static uint64_t rng = 88172645463325252ull;
static uint32_t next32() {
  rng ^= rng << 13;
  rng ^= rng >> 7;
  rng ^= rng << 17;
  return (uint32_t)(rng >> 16);
}

static bool check(uint32_t d, uint32_t n) {
  osr::FastDiv f;
  f.set(d);
  const uint32_t q = f.div(n);
  if (q != n / d) {
    std::printf("mismatch: d=%u n=%u got %u want %u\n", d, n, q, n / d);
    return false;
  }
  return true;
}

static bool sweep(uint32_t d) {
  const uint32_t edges[] = {0u, 1u, 2u, d - 1u, d, d + 1u, 0xFFFFFFFFu, 0xFFFFFFFEu, 0x80000000u,
                            0x7FFFFFFFu};
  for (uint32_t n : edges)
    if (!check(d, n)) return false;
  for (uint64_t k = 1; k * d <= 0xFFFFFFFFull && k < 4096; k = k * 3 + 1)
    for (int e = -1; e <= 1; e++)
      if (!check(d, (uint32_t)(k * d + e))) return false;
  const uint64_t top = 0xFFFFFFFFull / d * d;  // the largest multiple of d
  for (int e = -2; e <= 0; e++)
    if (!check(d, (uint32_t)(top + e))) return false;
  for (int r = 0; r < 64; r++)
    if (!check(d, next32())) return false;
  return true;
}

int main(int argc, char** argv) {
  for (uint32_t d = 1; d <= 70000; d++)
    if (!sweep(d)) return 1;
  for (int r = 0; r < 200000; r++) {
    const uint32_t d = (next32() & 0x7FFFFFFFu) + 1u;
    if (!sweep(d)) return 1;
  }
  if (!sweep(0x80000000u)) return 1;
  // every dividend for divisors whose multiplier needs the full 32 bits (7, 641, 65535)
  // and grid extents of the benchmark workloads
  if (argc > 1) {
    const uint32_t ds[] = {7u, 641u, 65535u, 1024u, 2048u, 2062u, 1000u};
    for (uint32_t d : ds) {
      osr::FastDiv f;
      f.set(d);
      for (uint64_t n = 0; n <= 0xFFFFFFFFull; n++)
        if (f.div((uint32_t)n) != (uint32_t)n / d) {
          std::printf("mismatch: d=%u n=%llu\n", d, (unsigned long long)n);
          return 1;
        }
    }
  }
  std::printf("ok\n");
  return 0;
}
