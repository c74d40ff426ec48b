// capi_mock.hip -- the library's real host code (capi.hip, ensemble.hip) over the mock
// runtime of mock_hip.h (test infrastructure only).
//
// Kernels are not executed.  Each launch validates what the kernel would touch: every
// work item's descriptor, every array a descriptor points to (live, large enough, tile
// offsets consistent, tile counts within the push limit); kernels without a dedicated
// check have every pointer argument checked.  The checks run when the launch is enqueued
// (to record the allocations it touches) and again when its stream executes it.  The
// particle generator and the step finalisation are emulated at execution (particles
// placed on their cells, step counter, history rows = step + column, parking at the step
// target), and the fill can inject tile overflows at random so that the re-layout path
// runs.
#define OSR_ENSEMBLE_HOST_CHECK 1
#include "mock_hip.h"
#include "../../osiris_amd/csrc/capi.hip"
#include <random>
#include <set>

namespace osr {
static std::mt19937_64 g_rng(12345);
extern "C" double g_mock_overflow_p;
double g_mock_overflow_p = 0.0;
extern "C" double g_mock_notready_p;
double g_mock_notready_p = 0.0;
extern "C" double g_mock_progress_p;
double g_mock_progress_p = 0.0;
extern "C" long long mock_errors() { return mock::nerr(); }
extern "C" long long mock_order_errors() { return mock::nordered(); }
extern "C" long long mock_live() { return (long long)mock::live().size(); }

static void check_grid(const GridDesc& g) {
  mock::check(g.E, 3 * g.ncell * 4, "E");
  mock::check(g.B, 3 * g.ncell * 4, "B");
  mock::check(g.J, 3 * g.ncell * 8, "J");
  mock::check(g.ctrl, sizeof(Ctrl), "ctrl");
}
static void check_spec(const SpecDesc& d) {
  check_grid(d.g);
  const int nt = d.g.ntiles;
  bool ok = mock::valid(d.off, 4 * nt) && mock::valid(d.cnt, 4 * nt) && mock::valid(d.cap, 4 * nt);
  mock::check(d.off, 4 * nt, "off");
  mock::check(d.cnt, 4 * nt, "cnt");
  mock::check(d.cap, 4 * nt, "cap");
  mock::check(d.mv_cnt, 4 * nt * 4, "mv_cnt");
  mock::check(d.mv_stride, 4 * nt, "mv_stride");
  mock::check(d.newcnt, 4 * nt, "newcnt");
  mock::check(d.inc, 4 * nt, "inc");
  mock::check(d.epart, 8 * nt, "epart");
  if (!ok) return;
  const size_t tot = (size_t)d.off[nt - 1] + d.cap[nt - 1];
  const size_t b = 4 * tot;
  const void* arr[] = {d.cell, d.x, d.y, d.ux, d.uy, d.uz, d.q, d.pid, d.mcell, d.mx, d.my, d.mux,
                       d.muy, d.muz, d.mq, d.mpid, d.midx};
  for (const void* p : arr) mock::check(p, b, "particle array");
  if (d.g.dims == 3) { mock::check(d.z, b, "z"); mock::check(d.mz, b, "mz"); }
  for (int t = 0; t < nt; t++) {
    if (d.cnt[t] > d.cap[t] - 256) mock::err("tile count above push limit", d.cnt, (size_t)t);
    if (t && d.off[t] != d.off[t - 1] + d.cap[t - 1]) mock::err("tile offsets", d.off, (size_t)t);
  }
}
static void check_field(const FieldDesc& f) {
  check_grid(f.g);
  mock::check(f.fe_part, (size_t)f.nblk * 48, "fe_part");
  for (int i = 0; i < f.nspec; i++) mock::check(f.ep[i], 8 * (size_t)f.ep_n[i], "ep");
  mock::check(f.hist, (size_t)f.hist_cap * (6 + f.nspec) * 8, "hist");
}

void osr_mock_exec(const void* k, dim3 g, const SpecDesc* d, const int2* w) {
  mock::check(w, (size_t)g.x * 8, "work list");
  std::set<int> seen;
  for (unsigned b = 0; b < g.x; b++) {
    const int2 wk = w[b];
    if (!mock::valid(&d[wk.x], sizeof(SpecDesc))) { mock::err("desc index", &d[wk.x], sizeof(SpecDesc)); continue; }
    mock::check(&d[wk.x], sizeof(SpecDesc), "desc");
    if (wk.y < 0 || wk.y >= d[wk.x].g.ntiles) mock::err("tile index", d, (size_t)wk.y);
    if (seen.insert(wk.x).second) check_spec(d[wk.x]);
  }
  if (mock::effects() && k == (const void*)tile_fill<256, 4096, 32768>) {
    std::uniform_real_distribution<double> U(0, 1);
    for (int i : seen) {
      Ctrl* c = d[i].g.ctrl;
      if (c->halt == 0 && U(g_rng) < g_mock_overflow_p) {
        c->halt = 1;
        d[i].inc[0] = -1;
        d[i].newcnt[0] = d[i].cnt[0];  // capacity would help: the device re-layout runs
      }
    }
  }
}
void osr_mock_exec(const void*, dim3 g, const FieldDesc* f, const int2* w) {
  mock::check(w, (size_t)g.x * 8, "yee work");
  std::set<int> seen;
  for (unsigned b = 0; b < g.x; b++) {
    const int2 wk = w[b];
    if (!mock::valid(&f[wk.x], sizeof(FieldDesc))) { mock::err("fd index", &f[wk.x], 0); continue; }
    mock::check(&f[wk.x], sizeof(FieldDesc), "fd");
    if ((long long)wk.y * 256 >= f[wk.x].g.ncell + 255) mock::err("yee block", f, (size_t)wk.y);
    if (seen.insert(wk.x).second) check_field(f[wk.x]);
  }
}
// the step end (history row = step + column, step counter, halt protocol)
static void emulate_step_end(const FieldDesc& fd) {
  Ctrl* c = fd.g.ctrl;
  if (!mock::valid(c, sizeof(Ctrl)) || c->halt >= 2) return;
  const long long step = c->steps;
  const int ncol = 6 + fd.nspec;
  if (step < fd.hist_cap)
    for (int col = 0; col < ncol; col++) fd.hist[step * ncol + col] = (double)step + col;
  c->steps = step + 1;
  if (c->halt == 1) c->halt = 2;
  else if (c->target > 0 && step + 1 >= c->target) c->halt = 3;
}
void osr_mock_exec(const void*, dim3 g, const FieldDesc* f, const int* w) {
  mock::check(w, (size_t)g.x * 4, "fin work");
  for (unsigned b = 0; b < g.x; b++) {
    if (!mock::valid(&f[w[b]], sizeof(FieldDesc))) { mock::err("fin fd", &f[w[b]], 0); continue; }
    mock::check(&f[w[b]], sizeof(FieldDesc), "fin fd");
    const FieldDesc& fd = f[w[b]];
    check_field(fd);
    if (mock::effects()) emulate_step_end(fd);
  }
}
// kernels of one simulation's walls and lasers (one field descriptor): the descriptor and
// every array it names are checked; the Lindman step end (wall_b2_end) also ends the step
void osr_mock_exec(const void* k, dim3, const FieldDesc* f) {
  if (!mock::valid(f, sizeof(FieldDesc))) { mock::err("wall fd", f, 0); return; }
  mock::check(f, sizeof(FieldDesc), "wall fd");
  check_field(*f);
  if (f->wall) mock::check(f->wall, 4, "wall buffers");
  if (f->las_k) {
    mock::check(f->las_tab, 4, "laser table");
    mock::check(f->las_mask, 4, "laser mask");
    mock::check(f->las_first, 8, "laser first row");
  }
  if (mock::effects() && (k == (const void*)wall_b2_end<2> || k == (const void*)wall_b2_end<3>))
    emulate_step_end(*f);
}
void osr_mock_exec(const void*, dim3, const SpecDesc* d, int sid, GenArgs a) {
  mock::check(&d[sid], sizeof(SpecDesc), "gen desc");
  check_spec(d[sid]);
  if (!mock::effects()) return;
  const SpecDesc& D = d[sid];
  const GridDesc& g = D.g;
  for (int tile = 0; tile < g.ntiles; tile++) {
    int tx0, ty0, tz0, wz;
    if (g.dims == 2) { tx0 = (tile % g.ntx) * g.tx; ty0 = (tile / g.ntx) * g.ty; tz0 = 0; wz = 1; }
    else { tx0 = (tile % g.ntx) * g.tx; ty0 = ((tile / g.ntx) % g.nty) * g.ty; tz0 = (tile / (g.ntx * g.nty)) * g.tz; wz = std::min(g.tz, g.nz - tz0); }
    const int wx = std::min(g.tx, g.nx - tx0), wy = std::min(g.ty, g.ny - ty0);
    const int nct = wx * wy * wz;
    for (int p = 0; p < D.cnt[tile]; p++) {
      const int c = p % nct;
      const int ix = tx0 + c % wx, iy = ty0 + (c / wx) % wy, iz = tz0 + c / (wx * wy);
      const int s = D.off[tile] + p;
      D.cell[s] = g.dims == 2 ? pack2(ix, iy) : pack3(ix, iy, iz);
      D.x[s] = 0.f; D.y[s] = 0.f; if (g.dims == 3) D.z[s] = 0.f;
      D.ux[s] = D.uy[s] = D.uz[s] = 0.f;
      D.q[s] = a.q;
      D.pid[s] = (uint32_t)(s);
    }
  }
}
// regrow_move: the tile's particle slots and valid mover entries, old layout -> new
void osr_mock_exec(const void*, dim3 g, const SpecDesc* d, int sid, RegrowArrays src,
                   RegrowArrays dst, const int* new_off) {
  mock::check(&d[sid], sizeof(SpecDesc), "regrow desc");
  const SpecDesc& D = d[sid];
  check_spec(D);
  mock::check(new_off, 4 * (size_t)g.x, "regrow offsets");
  if (!mock::effects()) return;
  const bool three = D.g.dims == 3;
  for (unsigned t = 0; t < g.x; t++) {
    const int o0 = D.off[t], o1 = new_off[t], n = D.cnt[t], NW = D.waves, st = D.mv_stride[t];
    for (int a = 0; a < 9; a++) {
      if (a == 3 && !three) continue;
      mock::check(src.a[a] + o0, 4 * (size_t)n, "regrow src");
      mock::check(dst.a[a] + o1, 4 * (size_t)n, "regrow dst");
      memcpy(dst.a[a] + o1, src.a[a] + o0, 4 * (size_t)n);
    }
    for (int w = 0; w < NW; w++) {
      const int m = D.mv_cnt[t * NW + w];
      for (int a = 9; a < 19; a++) {
        if (a == 12 && !three) continue;
        if (!m) continue;
        mock::check(src.a[a] + o0 + w * st, 4 * (size_t)m, "regrow mover src");
        mock::check(dst.a[a] + o1 + w * st, 4 * (size_t)m, "regrow mover dst");
        memcpy(dst.a[a] + o1 + w * st, src.a[a] + o0 + w * st, 4 * (size_t)m);
      }
    }
  }
}
}  // namespace osr
#include <csignal>
#include <unistd.h>
static void mock_segv(int) {
  void* bt[32];
  int n = backtrace(bt, 32);
  backtrace_symbols_fd(bt, n, 2);
  _exit(99);
}
__attribute__((constructor)) static void mock_init() { signal(SIGSEGV, mock_segv); signal(SIGBUS, mock_segv); }

// Planted stream-order mistakes, to show that the checks above catch them: returns the
// number of violations reported for (1) a pool block freed on the allocation stream while
// a kernel queued on another stream still uses it, (2) a host-synchronous copy into a
// block with queued kernel work, (3) a kernel on another stream touching a pool block
// before its allocating stream executed the allocation, (4) a kernel on a stream reading
// a block that a host-synchronous copy (null stream) wrote just before.  Expected: 4.
__global__ void mock_touch_kernel(float* p) { (void)p; }
extern "C" long long mock_selftest_planted() {
  const long long before = mock::nordered();
  hipStream_t a, b;
  hipStreamCreateWithFlags(&a, hipStreamNonBlocking);
  hipStreamCreateWithFlags(&b, hipStreamNonBlocking);
  const double keep_p = g_mock_progress_p;
  g_mock_progress_p = 0.0;  // nothing advances unless synchronised
  void* p = nullptr;
  hipMallocAsync(&p, 4096, a);
  hipStreamSynchronize(a);
  hipLaunchKernelGGL(mock_touch_kernel, dim3(1), dim3(64), 0, b, (float*)p);
  hipFreeAsync(p, a);
  hipStreamSynchronize(a);  // (1)
  hipStreamSynchronize(b);
  void* q = nullptr;
  hipMalloc(&q, 4096);
  hipLaunchKernelGGL(mock_touch_kernel, dim3(1), dim3(64), 0, b, (float*)q);
  float h[4] = {0, 0, 0, 0};
  hipMemcpy(q, h, sizeof(h), hipMemcpyHostToDevice);  // (2)
  hipStreamSynchronize(b);
  hipFree(q);
  void* r = nullptr;
  hipLaunchKernelGGL(mock_touch_kernel, dim3(1), dim3(64), 0, a, (float*)nullptr);  // a busy
  hipMallocAsync(&r, 4096, a);
  hipLaunchKernelGGL(mock_touch_kernel, dim3(1), dim3(64), 0, b, (float*)r);
  hipStreamSynchronize(b);  // (3)
  hipStreamSynchronize(a);
  hipFreeAsync(r, a);
  hipStreamSynchronize(a);
  void* t = nullptr;
  hipMalloc(&t, 4096);
  hipMemcpy(t, h, sizeof(h), hipMemcpyHostToDevice);
  hipLaunchKernelGGL(mock_touch_kernel, dim3(1), dim3(64), 0, b, (float*)t);  // (4)
  hipStreamSynchronize(b);
  hipFree(t);
  hipStreamDestroy(a);
  hipStreamDestroy(b);
  g_mock_progress_p = keep_p;
  return mock::nordered() - before;
}
