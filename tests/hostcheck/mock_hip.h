// mock_hip.h -- a validating stand-in for the HIP runtime, for checking the library's host
// logic on a machine without a GPU (test infrastructure only; never linked into the
// product).
//
// Force-included before osiris_amd/csrc/capi.hip: every runtime call of the host code is
// redirected to a mock.  Device memory is host memory tracked in a registry (size and
// allocation kind: plain vs stream-ordered pool).
//
// Streams are modelled as queues that execute later, in an order the mock chooses: work
// enqueued on a stream (kernel launches, graph launches, async copies and fills, event
// records, stream-ordered allocations and frees) runs when the host synchronises on it
// (stream or event synchronisation, a successful event query, device synchronisation,
// hipFree) or when the mock advances a random stream by a random amount on any runtime
// call.  Streams carry no ordering between each other (the library creates them all
// non-blocking), so the mock checks the rules of stream-ordered memory:
//   * a kernel, copy or fill touches only live allocations when it EXECUTES (not only
//     when it is enqueued);
//   * a stream-ordered free (hipFreeAsync) executes while no other stream still holds
//     queued work that touches the block: the pool could hand the block to the next
//     allocation at that moment;
//   * a block from hipMallocAsync is touched by another stream, or by a host-synchronous
//     copy, only after its allocating stream has executed the allocation;
//   * a host-synchronous copy or fill (null stream, which does not order with
//     non-blocking streams) touches no block that a stream still has queued work on;
//   * pinned host memory is not freed while a queued copy targets it;
//   * work enqueued on a stream touches no block written by a host-synchronous copy or
//     fill (null stream) since the last device-wide synchronisation: the streams do not
//     wait for the null stream, so such work may not see the write.
// Graph capture records kernel launches for replay (each replay enqueues them again).
// Kernel launches are handed to osr_mock_exec overloads (capi_mock.hip), which validate
// every descriptor pointer the kernel would dereference and report, through mock::check,
// the allocations it touches.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <random>
#include <set>
#include <vector>
#include <execinfo.h>

namespace mock {
struct Alloc {
  size_t n;
  bool async;
  bool ready;           // stream-ordered allocation executed by its stream
  hipStream_t stream;   // allocating stream (async)
  bool pinned;          // host pinned memory (hipHostMalloc)
};
inline std::map<uintptr_t, Alloc>& live() { static std::map<uintptr_t, Alloc> m; return m; }
inline long long& nerr() { static long long e = 0; return e; }
inline long long& nordered() { static long long e = 0; return e; }  // stream-order violations
inline void err(const char* what, const void* p, size_t need) {
  if (nerr()++ < 50) {
    fprintf(stderr, "MOCK ERROR: %s ptr=%p need=%zu\n", what, p, need);
    void* bt[16];
    int n = backtrace(bt, 16);
    backtrace_symbols_fd(bt, n, 2);
  }
}
inline void order_err(const char* what, const void* p) {
  nordered()++;
  err(what, p, 0);
}
// allocation holding [p, p+need), or live().end()
inline std::map<uintptr_t, Alloc>::iterator find(const void* p, size_t need) {
  auto& m = live();
  if (!p) return m.end();
  auto it = m.upper_bound((uintptr_t)p);
  if (it == m.begin()) return m.end();
  --it;
  if ((uintptr_t)p + need > it->first + it->second.n) return m.end();
  return it;
}
inline bool valid(const void* p, size_t need) {
  if (!p) return need == 0;
  return find(p, need) != live().end();
}
// touch collection: while `collecting`, every checked pointer's allocation is recorded
inline bool& collecting() { static bool c = false; return c; }
inline std::set<uintptr_t>& touched() { static std::set<uintptr_t> t; return t; }
// the stream whose work is executing (nullptr: the host)
inline hipStream_t& cur_stream() { static hipStream_t s = nullptr; return s; }
inline void check(const void* p, size_t need, const char* what) {
  auto it = find(p, need);
  if (!p && need == 0) return;
  if (it == live().end()) {
    err(what, p, need);
    return;
  }
  if (collecting()) touched().insert(it->first);
  if (!collecting() && it->second.async && !it->second.ready && cur_stream() != it->second.stream)
    order_err("pool block touched before its allocating stream executed the allocation", p);
}

// ---- streams -------------------------------------------------------------------------
struct Op {
  std::function<void()> run;
  std::set<uintptr_t> touch;
  int event = 0;  // > 0: event record marker
};
struct Stream {
  std::deque<Op> q;
};
inline std::map<hipStream_t, Stream>& streams() { static std::map<hipStream_t, Stream> m; return m; }
struct Event {
  long long recorded = 0, done = 0;  // record sequence numbers
};
inline std::map<int, Event>& events() { static std::map<int, Event> m; return m; }
inline std::mt19937_64& rng() { static std::mt19937_64 r(777); return r; }
extern "C" double g_mock_progress_p;  // per runtime call: chance that a random stream advances

inline bool pending_touch(uintptr_t base, hipStream_t except, bool any_stream = false) {
  for (auto& kv : streams()) {
    if (!any_stream && kv.first == except) continue;
    for (const Op& o : kv.second.q)
      if (o.touch.count(base)) return true;
  }
  return false;
}
inline void exec_one(hipStream_t s) {
  auto& q = streams()[s].q;
  if (q.empty()) return;
  Op o = std::move(q.front());
  q.pop_front();
  hipStream_t prev = cur_stream();
  cur_stream() = s;
  o.run();
  cur_stream() = prev;
}
inline void drain(hipStream_t s) {
  while (!streams()[s].q.empty()) exec_one(s);
}
// blocks written by host-synchronous copies or fills on the null stream since the last
// device-wide synchronisation: the library's streams do not wait for the null stream, so
// work enqueued on one of them may not see such a write
inline std::set<uintptr_t>& null_written() { static std::set<uintptr_t> w; return w; }
inline void drain_all() {
  for (auto& kv : streams()) drain(kv.first);
  null_written().clear();
}
// work enqueued on stream s touching `touch`
inline void check_null_order(hipStream_t s, const std::set<uintptr_t>& touch) {
  if (!s) return;  // the null stream itself is ordered with its copies
  for (uintptr_t b : touch)
    if (null_written().count(b))
      order_err("stream work touches a block written by a null-stream copy it is not ordered after",
                (void*)b);
}
inline void tick() {
  std::uniform_real_distribution<double> U(0, 1);
  if (streams().empty() || U(rng()) >= g_mock_progress_p) return;
  std::vector<hipStream_t> busy;
  for (auto& kv : streams())
    if (!kv.second.q.empty()) busy.push_back(kv.first);
  if (busy.empty()) return;
  hipStream_t s = busy[rng()() % busy.size()];
  const int n = 1 + (int)(rng()() % 8);
  for (int k = 0; k < n; k++) exec_one(s);
}
// touch set of a unit of work, computed by running its validation in collecting mode
template <class F>
inline std::set<uintptr_t> touch_of(F&& validate) {
  collecting() = true;
  touched().clear();
  validate();
  collecting() = false;
  return touched();
}
inline void enqueue(hipStream_t s, Op o) { streams()[s].q.push_back(std::move(o)); }

struct Launch {
  const void* k;
  dim3 g;
  std::function<void()> validate;  // checks only (no effects)
  std::function<void()> run;       // checks + emulated effects
};
inline bool& capturing() { static bool c = false; return c; }
inline std::vector<Launch>& cap() { static std::vector<Launch> v; return v; }
struct Graph { std::vector<Launch> l; };
inline bool& effects() { static bool e = true; return e; }
inline void enqueue_launch(const Launch& L, hipStream_t s) {
  Op o;
  effects() = false;
  o.touch = touch_of(L.validate);
  effects() = true;
  check_null_order(s, o.touch);
  o.run = L.run;
  enqueue(s, std::move(o));
}
}  // namespace mock

// kernels without a dedicated check: every non-null pointer argument must point into a
// live allocation (a by-value argument is not a pointer and is skipped)
template <class T> void mock_arg_check(const T* p) {
  if (p) mock::check((const void*)p, 1, "kernel pointer argument");
}
template <class T> void mock_arg_check(const T&) {}
template <class... A>
void osr_mock_exec(const void*, dim3, A... a) {
  (mock_arg_check(a), ...);
}

template <class T> const T* mock_c(T* p) { return p; }
template <class T> const T* mock_c(const T* p) { return p; }
template <class T> T mock_c(T v) { return v; }
template <class... A>
void mock_launch(const void* k, dim3 g, dim3, size_t, hipStream_t s, A... a) {
  mock::Launch L{k, g, [=]() { osr_mock_exec(k, g, mock_c(a)...); },
                 [=]() { osr_mock_exec(k, g, mock_c(a)...); }};
  if (mock::capturing()) mock::cap().push_back(L);
  else mock::enqueue_launch(L, s);
  mock::tick();
}
#undef hipLaunchKernelGGL
#define hipLaunchKernelGGL(K, G, B, L, S, ...) mock_launch((const void*)(K), dim3(G), dim3(B), (size_t)(L), S, __VA_ARGS__)

inline hipError_t mock_malloc(void** p, size_t n, bool async, hipStream_t s) {
  mock::tick();
  void* q = aligned_alloc(256, (n + 255) / 256 * 256);
  if (!q) return hipErrorOutOfMemory;
  mock::live()[(uintptr_t)q] = {n, async, !async, s, false};
  if (async) {
    const uintptr_t base = (uintptr_t)q;
    mock::Op o;
    o.run = [base]() {
      auto it = mock::live().find(base);
      if (it != mock::live().end()) it->second.ready = true;
    };
    mock::enqueue(s, std::move(o));
  }
  *p = q;
  return hipSuccess;
}
inline void mock_release(uintptr_t base) {
  auto it = mock::live().find(base);
  if (it == mock::live().end()) return;
  mock::live().erase(it);
  if (getenv("MOCK_NOREUSE")) {  // keep freed memory mapped but dead: stale pointers fail checks
    static std::vector<void*> dead;
    dead.push_back((void*)base);
  } else {
    free((void*)base);
  }
}
inline hipError_t m_hipMalloc(void** p, size_t n) { return mock_malloc(p, n, false, nullptr); }
inline hipError_t m_hipMallocAsync(void** p, size_t n, hipStream_t s) { return mock_malloc(p, n, true, s); }
inline hipError_t m_hipFree(void* p) {
  if (!p) return hipSuccess;
  mock::drain_all();  // hipFree synchronises the device
  auto it = mock::live().find((uintptr_t)p);
  if (it == mock::live().end()) { mock::err("free of unknown pointer", p, 0); return hipErrorInvalidValue; }
  if (it->second.async) mock::err("free kind mismatch", p, 0);
  mock_release((uintptr_t)p);
  return hipSuccess;
}
inline hipError_t m_hipFreeAsync(void* p, hipStream_t s) {
  if (!p) return hipSuccess;
  mock::tick();
  auto it = mock::live().find((uintptr_t)p);
  if (it == mock::live().end()) { mock::err("free of unknown pointer", p, 0); return hipErrorInvalidValue; }
  if (!it->second.async) mock::err("free kind mismatch", p, 0);
  const uintptr_t base = (uintptr_t)p;
  mock::Op o;
  o.run = [base, s]() {
    // the pool may reuse the block from here on: no other stream may still have work on it
    if (mock::pending_touch(base, s))
      mock::order_err("stream-ordered free while another stream has queued work on the block",
                      (void*)base);
    mock_release(base);
  };
  mock::enqueue(s, std::move(o));
  return hipSuccess;
}
// host-synchronous access (null stream): no queued stream work may touch the block
inline void host_access(const void* p, size_t n, const char* what) {
  mock::check(p, n, what);
  auto it = mock::find(p, n);
  if (it == mock::live().end()) return;
  if (it->second.async && !it->second.ready)
    mock::order_err("host-synchronous access to a pool block before its allocation executed", p);
  if (mock::pending_touch(it->first, nullptr, true))
    mock::order_err("host-synchronous copy/fill races queued stream work on the block", p);
}
inline hipError_t m_hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k) {
  mock::tick();
  if (mock::capturing()) mock::err("synchronous copy during capture", d, n);
  if (k == hipMemcpyHostToDevice || k == hipMemcpyDeviceToDevice) {
    host_access(d, n, "memcpy dst");
    auto it = mock::find(d, n);
    if (it != mock::live().end()) mock::null_written().insert(it->first);
  }
  if (k == hipMemcpyDeviceToHost || k == hipMemcpyDeviceToDevice) host_access(s, n, "memcpy src");
  memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t m_hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st) {
  mock::tick();
  if (mock::capturing()) { mock::err("memcpy in capture", d, n); }
  auto validate = [=]() {
    if (k == hipMemcpyHostToDevice || k == hipMemcpyDeviceToDevice) mock::check(d, n, "memcpyAsync dst");
    if (k == hipMemcpyDeviceToHost || k == hipMemcpyDeviceToDevice) mock::check(s, n, "memcpyAsync src");
    if (k == hipMemcpyDeviceToHost) {
      auto it = mock::find(d, n);
      if (it != mock::live().end() && it->second.pinned && mock::collecting())
        mock::touched().insert(it->first);
    }
  };
  mock::Op o;
  o.touch = mock::touch_of(validate);
  mock::check_null_order(st, o.touch);
  o.run = [=]() {
    validate();
    memcpy(d, s, n);
  };
  mock::enqueue(st, std::move(o));
  return hipSuccess;
}
inline hipError_t m_hipMemset(void* d, int v, size_t n) {
  mock::tick();
  host_access(d, n, "memset");
  auto it = mock::find(d, n);
  if (it != mock::live().end()) mock::null_written().insert(it->first);
  memset(d, v, n);
  return hipSuccess;
}
inline hipError_t m_hipMemsetAsync(void* d, int v, size_t n, hipStream_t st) {
  mock::tick();
  if (mock::capturing()) mock::err("memset in capture", d, n);
  mock::Op o;
  o.touch = mock::touch_of([=]() { mock::check(d, n, "memsetAsync"); });
  mock::check_null_order(st, o.touch);
  o.run = [=]() {
    mock::check(d, n, "memsetAsync");
    if (mock::valid(d, n)) memset(d, v, n);
  };
  mock::enqueue(st, std::move(o));
  return hipSuccess;
}
inline hipError_t m_hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  *s = (hipStream_t)new int(1);
  mock::streams()[*s];
  return hipSuccess;
}
inline hipError_t m_hipStreamDestroy(hipStream_t s) {
  mock::drain(s);  // queued work still completes
  mock::streams().erase(s);
  delete (int*)s;
  return hipSuccess;
}
inline hipError_t m_hipStreamSynchronize(hipStream_t s) { mock::drain(s); return hipSuccess; }
inline hipError_t m_hipDeviceSynchronize() { mock::drain_all(); return hipSuccess; }
inline int& next_event() { static int e = 0; return e; }
inline hipError_t m_hipEventCreate(hipEvent_t* e) {
  const int id = ++next_event();
  mock::events()[id];
  *e = (hipEvent_t)(uintptr_t)id;
  return hipSuccess;
}
inline hipError_t m_hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return m_hipEventCreate(e); }
inline hipError_t m_hipEventDestroy(hipEvent_t e) { mock::events().erase((int)(uintptr_t)e); return hipSuccess; }
inline hipError_t m_hipEventRecord(hipEvent_t e, hipStream_t s) {
  const int id = (int)(uintptr_t)e;
  auto& ev = mock::events()[id];
  const long long seq = ++ev.recorded;
  mock::Op o;
  o.event = id;
  o.run = [id, seq]() {
    auto it = mock::events().find(id);
    if (it != mock::events().end()) it->second.done = std::max(it->second.done, seq);
  };
  mock::enqueue(s, std::move(o));
  mock::tick();
  return hipSuccess;
}
// the stream holding event e's latest record, or nullptr
inline hipStream_t event_stream(int id) {
  for (auto& kv : mock::streams())
    for (const mock::Op& o : kv.second.q)
      if (o.event == id) return kv.first;
  return nullptr;
}
inline hipError_t m_hipEventSynchronize(hipEvent_t e) {
  const int id = (int)(uintptr_t)e;
  while (hipStream_t s = event_stream(id)) {
    auto& q = mock::streams()[s].q;
    bool last = false;
    while (!q.empty() && !last) {
      last = q.front().event == id;
      mock::exec_one(s);
    }
  }
  return hipSuccess;
}
// a stream waiting on an event: the mock runs the event's stream up to the record (an
// order the device may also take); in capture the launches are recorded in issue order
inline hipError_t m_hipStreamWaitEvent(hipStream_t, hipEvent_t e, unsigned) {
  if (!mock::capturing()) m_hipEventSynchronize(e);
  return hipSuccess;
}
extern "C" double g_mock_notready_p;
inline hipError_t m_hipEventQuery(hipEvent_t e) {
  mock::tick();
  const int id = (int)(uintptr_t)e;
  const auto& ev = mock::events()[id];
  if (ev.done >= ev.recorded) return hipSuccess;
  // not finished: sometimes let its stream run up to the record
  if ((rand() / (double)RAND_MAX) >= g_mock_notready_p) {
    m_hipEventSynchronize(e);
    return hipSuccess;
  }
  return hipErrorNotReady;
}
inline hipError_t m_hipEventElapsedTime(float* t, hipEvent_t, hipEvent_t) { *t = 0.f; return hipSuccess; }
inline hipError_t m_hipHostMalloc(void** p, size_t n, unsigned) {
  *p = malloc(n);
  mock::live()[(uintptr_t)*p] = {n, false, true, nullptr, true};
  return hipSuccess;
}
inline hipError_t m_hipHostFree(void* p) {
  if (!p) return hipSuccess;
  auto it = mock::live().find((uintptr_t)p);
  if (it == mock::live().end() || !it->second.pinned) {
    mock::err("hipHostFree of unknown pointer", p, 0);
    return hipErrorInvalidValue;
  }
  if (mock::pending_touch((uintptr_t)p, nullptr, true))
    mock::order_err("pinned host memory freed while a queued copy targets it", p);
  mock::live().erase(it);
  free(p);
  return hipSuccess;
}
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
inline hipError_t m_hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
  for (auto& L : ((mock::Graph*)e)->l) mock::enqueue_launch(L, s);
  mock::tick();
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
#define hipStreamWaitEvent m_hipStreamWaitEvent
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
