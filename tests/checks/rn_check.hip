// rn_check.hip -- device checks of the correctly rounded reciprocal, quotient and square
// root sequences of osr_device.hpp (rcp_rn, div_rn, sqrt_rn) against the IEEE operations
// the compiler emits under -fhip-fp32-correctly-rounded-divide-sqrt.  Test infrastructure:
// loaded only by tests/test_gpu_rounding.py.
//
// Every check counts the inputs whose result differs in any bit and keeps the smallest
// such input index, so a failing test can name one.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../osiris_amd/csrc/osr_device.hpp"

namespace {

__device__ uint32_t f2u(float f) { return __float_as_uint(f); }
__device__ float u2f(uint32_t u) { return __uint_as_float(u); }

struct Result {
  unsigned long long bad;
  unsigned long long first;  // smallest failing index (all ones: none)
};

__device__ void report(Result* res, unsigned long long idx) {
  atomicAdd(&res->bad, 1ull);
  atomicMin(&res->first, idx);
}

// bit patterns first .. first + n - 1 (both signs when signs = 2)
__global__ void k_rcp(uint32_t first, uint64_t n, int signs, Result* res) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    for (int sg = 0; sg < signs; sg++) {
      const float b = u2f((first + (uint32_t)i) | (sg ? 0x80000000u : 0u));
      const float got = osr::rcp_rn(b), want = 1.0f / b;
      if (f2u(got) != f2u(want)) report(res, i);
    }
  }
}

__global__ void k_sqrt(uint32_t first, uint64_t n, Result* res) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const float x = u2f(first + (uint32_t)i);
    const float got = osr::sqrt_rn(x), want = __builtin_sqrtf(x);
    if (f2u(got) != f2u(want)) report(res, i);
  }
}

// a = 1 + ia 2^-23 for ia in [a_first, a_first + na), b = 1 + ib 2^-23 for every ib:
// one thread per b mantissa, looping over the a mantissas
__global__ void k_div(uint32_t a_first, uint32_t na, Result* res) {
  const uint32_t ib = blockIdx.x * blockDim.x + threadIdx.x;
  if (ib >= (1u << 23)) return;
  const float b = u2f(0x3f800000u | ib);
  for (uint32_t k = 0; k < na; k++) {
    const uint32_t ia = a_first + k;
    const float a = u2f(0x3f800000u | ia);
    const float got = osr::div_rn(a, b), want = a / b;
    if (f2u(got) != f2u(want)) report(res, ((unsigned long long)ia << 23) | ib);
  }
}

// explicit operand pairs (zeros, signs, exponent spread)
__global__ void k_div_pairs(const float* a, const float* b, uint64_t n, Result* res) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float got = osr::div_rn(a[i], b[i]), want = a[i] / b[i];
  if (f2u(got) != f2u(want)) report(res, i);
}

template <class F>
int run(F launch, unsigned long long* bad, unsigned long long* first) {
  Result* d = nullptr;
  if (hipMalloc(&d, sizeof(Result)) != hipSuccess) return 1;
  Result h{0ull, ~0ull};
  if (hipMemcpy(d, &h, sizeof(Result), hipMemcpyHostToDevice) != hipSuccess) return 2;
  launch(d);
  if (hipGetLastError() != hipSuccess) return 3;
  if (hipMemcpy(&h, d, sizeof(Result), hipMemcpyDeviceToHost) != hipSuccess) return 4;
  (void)hipFree(d);
  *bad = h.bad;
  *first = h.first;
  return 0;
}

}  // namespace

extern "C" {

int rn_check_rcp(uint32_t first, uint64_t n, int signs, unsigned long long* bad,
                 unsigned long long* first_bad) {
  if (n > (1ull << 32) - first) return 5;
  return run([=](Result* d) {
    hipLaunchKernelGGL(k_rcp, dim3(8192), dim3(256), 0, 0, first, n, signs, d);
  }, bad, first_bad);
}

int rn_check_sqrt(uint32_t first, uint64_t n, unsigned long long* bad,
                  unsigned long long* first_bad) {
  if (n > (1ull << 32) - first) return 5;
  return run([=](Result* d) {
    hipLaunchKernelGGL(k_sqrt, dim3(8192), dim3(256), 0, 0, first, n, d);
  }, bad, first_bad);
}

int rn_check_div(uint32_t a_first, uint32_t na, unsigned long long* bad,
                 unsigned long long* first_bad) {
  if (a_first >= (1u << 23) || na > (1u << 23) - a_first) return 5;
  return run([=](Result* d) {
    hipLaunchKernelGGL(k_div, dim3((1u << 23) / 256), dim3(256), 0, 0, a_first, na, d);
  }, bad, first_bad);
}

int rn_check_div_pairs(const float* a, const float* b, uint64_t n, unsigned long long* bad,
                       unsigned long long* first_bad) {
  float *da = nullptr, *db = nullptr;
  if (hipMalloc(&da, 4 * n) != hipSuccess || hipMalloc(&db, 4 * n) != hipSuccess) return 1;
  if (hipMemcpy(da, a, 4 * n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(db, b, 4 * n, hipMemcpyHostToDevice) != hipSuccess)
    return 2;
  const int r = run([=](Result* d) {
    hipLaunchKernelGGL(k_div_pairs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0,
                       (const float*)da, (const float*)db, n, d);
  }, bad, first_bad);
  (void)hipFree(da);
  (void)hipFree(db);
  return r;
}

}  // extern "C"
