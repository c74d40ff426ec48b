// mock_hip.h -- a validating stand-in for the HIP runtime, for checking the library's host
// logic on a machine without a GPU (test infrastructure only; never linked into the
// product).
//
// Force-included before osiris_amd/csrc/capi.hip: every runtime call of the host code is
// redirected to a mock.  Device memory is host memory tracked in a registry (size and
// allocation kind: plain vs stream-ordered pool), copies and fills check that they stay
// inside live allocations, freeing checks that the release matches the allocation kind,
// streams and events complete immediately (hipEventQuery can be made to report "not
// ready" at random, to exercise in-flight scheduling paths), and graph capture records
// kernel launches for replay.  Kernel launches are handed to osr_mock_exec overloads
// (capi_mock.hip) that validate every descriptor pointer the kernel would dereference.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <functional>
#include <map>
#include <mutex>
#include <vector>
#include <execinfo.h>

namespace mock {
struct Alloc { size_t n; bool async; };
inline std::map<uintptr_t, Alloc>& live() { static std::map<uintptr_t, Alloc> m; return m; }
inline long long& nerr() { static long long e = 0; return e; }
inline void err(const char* what, const void* p, size_t need) {
  if (nerr()++ < 50) {
    fprintf(stderr, "MOCK ERROR: %s ptr=%p need=%zu\n", what, p, need);
    void* bt[16];
    int n = backtrace(bt, 16);
    backtrace_symbols_fd(bt, n, 2);
  }
}
// true if [p, p+need) lies inside one live allocation
inline bool valid(const void* p, size_t need) {
  if (!p) return need == 0;
  auto& m = live();
  auto it = m.upper_bound((uintptr_t)p);
  if (it == m.begin()) return false;
  --it;
  return (uintptr_t)p + need <= it->first + it->second.n;
}
inline void check(const void* p, size_t need, const char* what) { if (!valid(p, need)) err(what, p, need); }
struct Launch { const void* k; dim3 g; std::function<void()> run; };
inline bool& capturing() { static bool c = false; return c; }
inline std::vector<Launch>& cap() { static std::vector<Launch> v; return v; }
struct Graph { std::vector<Launch> l; };
}  // namespace mock

template <class... A>
void osr_mock_exec(const void*, dim3, A...) {}

template <class T> const T* mock_c(T* p) { return p; }
template <class T> const T* mock_c(const T* p) { return p; }
template <class T> T mock_c(T v) { return v; }
template <class... A>
void mock_launch(const void* k, dim3 g, dim3, size_t, hipStream_t, A... a) {
  mock::Launch L{k, g, [=]() { osr_mock_exec(k, g, mock_c(a)...); }};
  if (mock::capturing()) mock::cap().push_back(L);
  else L.run();
}
#undef hipLaunchKernelGGL
#define hipLaunchKernelGGL(K, G, B, L, S, ...) mock_launch((const void*)(K), dim3(G), dim3(B), (size_t)(L), S, __VA_ARGS__)

inline hipError_t mock_malloc(void** p, size_t n, bool async) {
  void* q = aligned_alloc(256, (n + 255) / 256 * 256);
  if (!q) return hipErrorOutOfMemory;
  mock::live()[(uintptr_t)q] = {n, async};
  *p = q;
  return hipSuccess;
}
inline hipError_t mock_free(void* p, bool async) {
  if (!p) return hipSuccess;
  auto it = mock::live().find((uintptr_t)p);
  if (it == mock::live().end()) { mock::err("free of unknown pointer", p, 0); return hipErrorInvalidValue; }
  if (it->second.async != async) mock::err("free kind mismatch", p, 0);
  mock::live().erase(it);
  if (getenv("MOCK_NOREUSE")) {  // keep freed memory mapped but dead: stale pointers fail checks
    static std::vector<void*> dead;
    dead.push_back(p);
  } else {
    free(p);
  }
  return hipSuccess;
}
inline hipError_t m_hipMalloc(void** p, size_t n) { return mock_malloc(p, n, false); }
inline hipError_t m_hipMallocAsync(void** p, size_t n, hipStream_t) { return mock_malloc(p, n, true); }
inline hipError_t m_hipFree(void* p) { return mock_free(p, false); }
inline hipError_t m_hipFreeAsync(void* p, hipStream_t) { return mock_free(p, true); }
inline bool is_dev(const void* p) { return mock::valid(p, 1); }
inline hipError_t m_hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k) {
  if (k == hipMemcpyHostToDevice || k == hipMemcpyDeviceToDevice) mock::check(d, n, "memcpy dst");
  if (k == hipMemcpyDeviceToHost || k == hipMemcpyDeviceToDevice) mock::check(s, n, "memcpy src");
  memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t m_hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) {
  if (mock::capturing()) { mock::err("memcpy in capture", d, n); }
  return m_hipMemcpy(d, s, n, k);
}
inline hipError_t m_hipMemset(void* d, int v, size_t n) { mock::check(d, n, "memset"); memset(d, v, n); return hipSuccess; }
inline hipError_t m_hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { return m_hipMemset(d, v, n); }
inline hipError_t m_hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = (hipStream_t)new int(1); return hipSuccess; }
inline hipError_t m_hipStreamDestroy(hipStream_t s) { delete (int*)s; return hipSuccess; }
inline hipError_t m_hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t m_hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t m_hipEventCreate(hipEvent_t* e) { *e = (hipEvent_t)new int(2); return hipSuccess; }
inline hipError_t m_hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return m_hipEventCreate(e); }
inline hipError_t m_hipEventDestroy(hipEvent_t e) { delete (int*)e; return hipSuccess; }
inline hipError_t m_hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
extern "C" double g_mock_notready_p;
inline hipError_t m_hipEventQuery(hipEvent_t) { return (rand() / (double)RAND_MAX) < g_mock_notready_p ? hipErrorNotReady : hipSuccess; }
inline hipError_t m_hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t m_hipEventElapsedTime(float* t, hipEvent_t, hipEvent_t) { *t = 0.f; return hipSuccess; }
inline hipError_t m_hipHostMalloc(void** p, size_t n, unsigned) { *p = malloc(n); return hipSuccess; }
inline hipError_t m_hipHostFree(void* p) { free(p); return hipSuccess; }
inline hipError_t m_hipGetLastError() { return hipSuccess; }
inline const char* m_hipGetErrorString(hipError_t) { return "mock"; }
inline hipError_t m_hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
inline hipError_t m_hipGetDevice(int* d) { *d = 0; return hipSuccess; }
inline hipError_t m_hipMemGetInfo(size_t* f, size_t* t) { *f = (size_t)8 << 30; *t = (size_t)64 << 30; return hipSuccess; }
inline hipError_t m_hipDeviceGetDefaultMemPool(hipMemPool_t* p, int) { *p = (hipMemPool_t)0x1; return hipSuccess; }
inline hipError_t m_hipMemPoolSetAttribute(hipMemPool_t, hipMemPoolAttr, void*) { return hipSuccess; }
inline hipError_t m_hipMemPoolTrimTo(hipMemPool_t, size_t) { return hipSuccess; }
template <class F> inline hipError_t m_hipFuncSetAttribute(F, hipFuncAttribute, int) { return hipSuccess; }
inline hipError_t m_hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { mock::capturing() = true; mock::cap().clear(); return hipSuccess; }
inline hipError_t m_hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
  mock::capturing() = false;
  auto* G = new mock::Graph{mock::cap()};
  *g = (hipGraph_t)G;
  return hipSuccess;
}
inline hipError_t m_hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  *e = (hipGraphExec_t) new mock::Graph(*(mock::Graph*)g);
  return hipSuccess;
}
inline hipError_t m_hipGraphDestroy(hipGraph_t g) { delete (mock::Graph*)g; return hipSuccess; }
inline hipError_t m_hipGraphExecDestroy(hipGraphExec_t e) { delete (mock::Graph*)e; return hipSuccess; }
inline hipError_t m_hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto& L : ((mock::Graph*)e)->l) L.run();
  return hipSuccess;
}
#define hipMalloc m_hipMalloc
#define hipMallocAsync m_hipMallocAsync
#define hipFree m_hipFree
#define hipFreeAsync m_hipFreeAsync
#define hipMemcpy m_hipMemcpy
#define hipMemcpyAsync m_hipMemcpyAsync
#define hipMemset m_hipMemset
#define hipMemsetAsync m_hipMemsetAsync
#define hipStreamCreateWithFlags m_hipStreamCreateWithFlags
#define hipStreamDestroy m_hipStreamDestroy
#define hipStreamSynchronize m_hipStreamSynchronize
#define hipDeviceSynchronize m_hipDeviceSynchronize
#define hipEventCreate m_hipEventCreate
#define hipEventCreateWithFlags m_hipEventCreateWithFlags
#define hipEventDestroy m_hipEventDestroy
#define hipEventRecord m_hipEventRecord
#define hipEventQuery m_hipEventQuery
#define hipEventSynchronize m_hipEventSynchronize
#define hipEventElapsedTime m_hipEventElapsedTime
#define hipHostMalloc m_hipHostMalloc
#define hipHostFree m_hipHostFree
#define hipGetLastError m_hipGetLastError
#define hipGetErrorString m_hipGetErrorString
#define hipGetDeviceCount m_hipGetDeviceCount
#define hipGetDevice m_hipGetDevice
#define hipMemGetInfo m_hipMemGetInfo
#define hipDeviceGetDefaultMemPool m_hipDeviceGetDefaultMemPool
#define hipMemPoolSetAttribute m_hipMemPoolSetAttribute
#define hipMemPoolTrimTo m_hipMemPoolTrimTo
#define hipFuncSetAttribute m_hipFuncSetAttribute
#define hipStreamBeginCapture m_hipStreamBeginCapture
#define hipStreamEndCapture m_hipStreamEndCapture
#define hipGraphInstantiate m_hipGraphInstantiate
#define hipGraphDestroy m_hipGraphDestroy
#define hipGraphExecDestroy m_hipGraphExecDestroy
#define hipGraphLaunch m_hipGraphLaunch
