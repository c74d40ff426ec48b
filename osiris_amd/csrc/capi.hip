// capi.hip -- host side of the C ABI declared in include/osiris_amd.h.
//
// One translation unit: the kernel files are included here so that every kernel
// template is instantiated next to the code that launches it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/osiris_amd.h"
#include "k_dropin.hip"
#include "k_field.hip"
#include "k_smooth.hip"
#include "k_loader.hip"
#include "k_tiles.hip"
#include "k_push2d.hip"
#include "k_push3d.hip"
#include "k_plain.hip"

using namespace osr;

// ------------------------------------------------------------------------------------
// errors
// ------------------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPC(call)                                                                       \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(OSR_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),   \
                  __FILE__, __LINE__);                                                   \
  } while (0)

#define TRY(call)              \
  do {                         \
    int r_ = (call);           \
    if (r_ != OSR_OK) return r_; \
  } while (0)

#include "output.hip"
#include "zpulse.hip"

namespace {

constexpr int kNT = 256;  // threads per workgroup of the push kernels
constexpr int kGraphSteps = 4;  // steps per replay of the captured step graph
constexpr int kWaves = kNT / 64;

// Device allocations made while g_alloc_stream is set (ensemble jobs) come from the
// stream-ordered pool: retiring a job then frees its memory without the implicit device
// synchronisation of hipFree, so the other jobs keep the GPU busy.
thread_local hipStream_t g_alloc_stream = nullptr;
thread_local bool g_alloc_async = false;

struct DBuf {
  void* p = nullptr;
  size_t n = 0;
  bool async = false;
  hipStream_t st = nullptr;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) {
      if (async)
        (void)hipFreeAsync(p, st);
      else
        (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
    async = false;
  }
  int alloc(size_t bytes) {
    release();
    if (bytes == 0) bytes = 256;
    hipError_t e;
    if (g_alloc_async) {
      e = hipMallocAsync(&p, bytes, g_alloc_stream);
      if (e == hipSuccess) e = hipStreamSynchronize(g_alloc_stream);
      async = true;
      st = g_alloc_stream;
    } else {
      e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess) {
      if (p && async) (void)hipFreeAsync(p, st);
      p = nullptr;
      async = false;
      (void)hipGetLastError();
      return fail(OSR_ERR_NOMEM, "device allocation of %zu bytes failed", bytes);
    }
    n = bytes;
    return OSR_OK;
  }
  // exchange ownership, including how the memory must be released
  void swap(DBuf& o) {
    std::swap(p, o.p);
    std::swap(n, o.n);
    std::swap(async, o.async);
    std::swap(st, o.st);
  }
  template <class T>
  T* as() const {
    return (T*)p;
  }
};

// Copy between host and device ordered on stream `st`: it runs after the work already
// queued there (a download sees the kernels launched on `st` before it) and is complete
// when the call returns, so kernels launched afterwards on any stream see an upload and
// the host buffer may be reused.  A plain hipMemcpy runs on the null stream, which the
// library's non-blocking streams do not wait for: a kernel launched on one of them right
// after it may read the device memory as it was before the copy.
int copy_on(hipStream_t st, void* dst, const void* src, size_t n, hipMemcpyKind kind) {
  if (n == 0) return OSR_OK;
  HIPC(hipMemcpyAsync(dst, src, n, kind, st));
  HIPC(hipStreamSynchronize(st));
  return OSR_OK;
}

struct Species {
  float rqm = -1.f;
  int pusher = OSR_BORIS;
  int64_t np = 0;
  int64_t cap_total = 0;
  double qabs = 0.0, qmax = 0.0;
  DBuf cell, x, y, z, ux, uy, uz, q, pid;
  DBuf mcell, mx, my, mz, mux, muy, muz, mq, mpid, midx;
  DBuf off, cnt, cap, mv_cnt, mv_stride, newcnt, inc, epart;
  std::vector<int> h_off, h_cap;
  // plain path intermediates
  DBuf pl_ep, pl_bp, pl_vp, pl_nvp;
  int epart_n = 0;
};

// host-side particle list used by uploads and re-layouts
struct HostParts {
  std::vector<uint32_t> cell, pid;
  std::vector<float> x, y, z, ux, uy, uz, q;
  size_t size() const { return cell.size(); }
  void reserve(size_t n) {
    cell.reserve(n); pid.reserve(n); x.reserve(n); y.reserve(n); z.reserve(n);
    ux.reserve(n); uy.reserve(n); uz.reserve(n); q.reserve(n);
  }
};

}  // namespace

struct osr_sim {
  osr_sim_desc desc{};
  GridDesc g{};
  int dims = 2, order = 2;
  int ntiles = 0;
  int nblk = 0;  // Yee blocks of 256 cells
  int nblk_e = 0;  // energy workgroups of the first B half step (kEneBlk blocks each)
  int nwork_e = 0;  // its work items: the nblk_e blocks, then (Lindman walls) the blocks of
                    // 256 rows of the lower wall's first B update, as (0, -1 - block)
  double dx[3] = {1, 1, 1};
  double dt = 0;
  DBuf E, B, J, ctrl, fe_part, hist;
  int64_t hist_cap = 0;
  std::vector<std::unique_ptr<Species>> sp;
  DBuf d_spec, d_work, d_fd, d_ywork, d_ework, d_fwork;
  FieldDesc fd{};
  hipStream_t stream = nullptr;
  hipGraphExec_t gexec = nullptr;    // one step
  hipGraphExec_t gexec_k = nullptr;  // kGraphSteps steps
  bool graph_ok = false;
  int64_t steps = 0;
  int64_t relayouts = 0;
  bool plain = false;
  size_t lds_push = 0;
  int push_waves = 4;  // waves per push workgroup (SpecDesc::waves)
  int reorder = 0;     // tile_reorder period in steps (0: off)
  int sm_order[3] = {0, 0, 0};  // current smoothing per direction (0: none)
  float sm_coef[3][31] = {};
  DBuf sm_buf[2];
  DBuf wall;  // open x walls: Lindman buffers or VPML walls (FieldDesc::wall)
  int vp_n = 0;  // VPML wall thickness (0: Lindman walls)
  std::vector<osr_zp::Wall> zp;  // laser walls, launch order
  DBuf zp_tab, zp_mask, zp_first;  // FieldDesc::las_* (kZpRows steps per fill)
  Ctrl* h_ctrl = nullptr;  // pinned mirror
  int64_t desc_version = 0;  // bumped whenever the device descriptors change

  osr_sim() = default;
  osr_sim(const osr_sim&) = delete;
  osr_sim& operator=(const osr_sim&) = delete;
  // Also runs when osr_sim_create fails part way: the stream, graphs and pinned mirror
  // go with the handle; the device buffers are released by their own destructors.
  ~osr_sim() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (gexec_k) (void)hipGraphExecDestroy(gexec_k);
    if (h_ctrl) (void)hipHostFree(h_ctrl);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int ceil_div(int a, int b) { return (a + b - 1) / b; }

constexpr int kZpRows = 256;  // steps of laser values per host evaluation (the longest batch)

// steps held by the laser table: kZpRows, fewer when a wall plane is large (the table holds
// at most 2^26 floats)
int zp_rows(const osr_sim* s, int K) {
  const long long per = 2ll * s->g.ny * s->g.nz * std::max(K, 1);
  return (int)std::max(1ll, std::min<long long>(kZpRows, (1ll << 26) / per));
}

// setup_att_coeff (emf/boundary/os-emf-vpml.f03:3586-3689) for direction 1, every
// operation in FP32 as the reference's p_k_fld: s1 = -(0.8 (nexp + 1)) / 2 with
// p_att_nexp = 4, s2 = 1 / (n - 3), attenuation exp(s1 (s2 p)^4) with the integer power
// (t t)(t t); out = [lower, upper][E, B][1..3][cell], the lower wall's cells reversed
// (cell n-1-j holds point j)
void vpml_coefs(int n, float dtdx1, float* out) {
  const float s1 = -(0.8f * (4.0f + 1.0f)) / 2.0f;
  const float s2 = 1.0f / (float)(n - 3);
  const float r = (1.0f - dtdx1) / (1.0f + dtdx1);
  auto att = [&](float p) {
    const float t = s2 * p;
    return expf(s1 * ((t * t) * (t * t)));
  };
  auto at = [&](int wl, int eb, int k, int cell) -> float& {
    return out[((wl * 2 + eb) * 3 + k) * n + cell];
  };
  for (int j = 0; j < n; j++) {
    const float fj = (float)j;
    // lower wall: B is the outside point
    float tj = att(fj - 0.5f), tj12 = att(fj), tj1 = att(fj + 0.5f);
    float b1e = dtdx1 * (1.0f + r * (1.0f - tj12));
    float b1b = dtdx1 * (1.0f + r * (1.0f - tj1));
    const int k = n - 1 - j;
    at(0, 0, 0, k) = 1.0f - b1e + tj12 * dtdx1;
    at(0, 0, 1, k) = tj * b1e;
    at(0, 0, 2, k) = dtdx1;
    at(0, 1, 0, k) = 1.0f - b1b + tj1 * dtdx1;
    at(0, 1, 1, k) = tj12 * b1b;
    at(0, 1, 2, k) = dtdx1;
    // upper wall: E is the outside point
    tj = att(fj);
    tj12 = att(fj + 0.5f);
    tj1 = att(fj + 1.0f);
    b1e = dtdx1 * (1.0f + r * (1.0f - tj12));
    b1b = dtdx1 * (1.0f + r * (1.0f - tj1));
    at(1, 0, 0, j) = 1.0f - b1e + tj12 * dtdx1;
    at(1, 0, 1, j) = dtdx1;
    at(1, 0, 2, j) = tj * b1e;
    at(1, 1, 0, j) = 1.0f - b1b + tj1 * dtdx1;
    at(1, 1, 1, j) = dtdx1;
    at(1, 1, 2, j) = tj12 * b1b;
  }
}

// launches per step of the laser walls (two for a circularly polarised pulse)
int zp_launches(const osr_sim* s) {
  int k = 0;
  for (auto& w : s->zp) k += w.pol_type != 0 ? 2 : 1;
  return k;
}

// LDS bytes of the push kernel for a tile shape
size_t push_lds_bytes(int dims, int order, const int* t) {
  const int glo = order == 1 ? 1 : 2, ghi = order == 3 ? 2 : 1;
  const int jlo = order == 1 ? 1 : 2, jhi = order == 1 ? 2 : (order == 2 ? 2 : 3);
  size_t gn = 1, jn = 1;
  for (int d = 0; d < dims; d++) {
    gn *= (size_t)(t[d] + glo + ghi);
    jn *= (size_t)(t[d] + jlo + jhi);
  }
  size_t q = 0;  // 3D: per-wave VP queues of kVpQueue3 entries of 8 words
  if (dims == 3) q = (size_t)kWaves * kVpQueue3 * 8 * sizeof(float);
  return 3 * jn * sizeof(unsigned long long) + 6 * gn * sizeof(float) + q;
}

// fixed-point shift of the current: the largest one for which neither the sum of all
// contributions that can reach one node nor a single contribution can leave the
// representable range (see the bound in the design notes of the deposit kernel).
void update_jscale(osr_sim* s) {
  double qabs = 0.0, qmax = 0.0;
  for (auto& p : s->sp) {
    qabs += p->qabs;
    qmax = std::max(qmax, p->qmax);
  }
  const double wsum = s->dims == 2 ? 2.0 : 3.0;
  double jn = 0.0;
  for (int d = 0; d < s->dims; d++) jn += s->g.jnorm[d];
  const double per = wsum * jn + 1.0;
  const double total = std::max(qabs * per * 8.0, 1e-300);
  const double single = std::max(qmax * per * 2.0, 1e-300);
  int sh = std::min(62 - (int)std::ceil(std::log2(total)), 50 - (int)std::ceil(std::log2(single)));
  sh = std::max(0, std::min(sh, 80));
  s->g.jscale = std::ldexp(1.0, sh);
  s->g.jinv = std::ldexp(1.0, -sh);
}

// device descriptors of the species of s (v[0 .. nspecies)) and of its fields
void make_spec_descs(const osr_sim* s, SpecDesc* v) {
  for (size_t i = 0; i < s->sp.size(); i++) {
    const Species& p = *s->sp[i];
    SpecDesc& d = v[i];
    d = SpecDesc{};
    d.g = s->g;
    d.cell = p.cell.as<uint32_t>();
    d.x = p.x.as<float>();
    d.y = p.y.as<float>();
    d.z = p.z.as<float>();
    d.ux = p.ux.as<float>();
    d.uy = p.uy.as<float>();
    d.uz = p.uz.as<float>();
    d.q = p.q.as<float>();
    d.pid = p.pid.as<uint32_t>();
    d.mcell = p.mcell.as<uint32_t>();
    d.mx = p.mx.as<float>();
    d.my = p.my.as<float>();
    d.mz = p.mz.as<float>();
    d.mux = p.mux.as<float>();
    d.muy = p.muy.as<float>();
    d.muz = p.muz.as<float>();
    d.mq = p.mq.as<float>();
    d.mpid = p.mpid.as<uint32_t>();
    d.midx = p.midx.as<int>();
    d.off = p.off.as<int>();
    d.cnt = p.cnt.as<int>();
    d.cap = p.cap.as<int>();
    d.mv_cnt = p.mv_cnt.as<int>();
    d.mv_stride = p.mv_stride.as<int>();
    d.newcnt = p.newcnt.as<int>();
    d.inc = p.inc.as<int>();
    d.epart = p.epart.as<double>();
    d.tem = (float)(0.5 * s->dt / (double)p.rqm);
    d.waves = s->push_waves;
  }
}

void make_field_desc(const osr_sim* s, FieldDesc& f) {
  f = FieldDesc{};
  f.g = s->g;
  f.fe_part = s->fe_part.as<double>();
  f.nblk = s->nblk_e;
  f.nspec = (int)s->sp.size();
  double dvol = 1.0;
  for (int d = 0; d < s->dims; d++) dvol *= s->dx[d];
  f.fe_scale = 0.5 * dvol;
  for (size_t i = 0; i < s->sp.size(); i++) {
    f.ep[i] = s->sp[i]->epart.as<double>();
    f.ep_n[i] = s->sp[i]->epart_n;
    f.ep_scale[i] = (double)s->sp[i]->rqm * dvol;
  }
  f.hist = s->hist.as<double>();
  f.hist_cap = s->hist_cap;
  for (int d = 0; d < 3; d++) {
    f.bdt[d] = d < s->dims ? (float)((s->dt / 2.0) / s->dx[d]) : 0.f;
    f.edt[d] = d < s->dims ? (float)(s->dt / s->dx[d]) : 0.f;
  }
  f.dtif = (float)s->dt;
  f.Jf = s->plain ? s->J.as<float>() : nullptr;
  f.jf_zero = s->plain ? 1 : 0;
  f.sm_np = 0;
  f.jsrc_clear = 0;
  for (int d = 0; d < s->dims; d++) {
    if (s->sm_order[d] <= 0) continue;
    const int k = f.sm_np++;
    f.sm_dir[k] = d;
    f.sm_n[k] = s->sm_order[d];
    for (int c = 0; c < 31; c++) f.sm_coef[k][c] = s->sm_coef[d][c];
  }
  if (s->g.open_nx) {  // new_lindman, emf/boundary/os-emf-lindman.f90:235-241
    f.wall = s->wall.as<float>();
    f.vp_n = s->vp_n;  // setup_vpml, emf/boundary/os-emf-vpml.f03:468-470
    f.vp_dt2 = (float)(s->dt / s->dx[1]);
    f.vp_dt3 = s->dims > 2 ? (float)(s->dt / s->dx[2]) : 0.f;
    f.vp_diff[0] = s->desc.vpml_diffuse[0];
    f.vp_diff[1] = s->desc.vpml_diffuse[1];
    f.wall_dc = (float)((1.0 - s->dt / s->dx[0]) / (1.0 + s->dt / s->dx[0]));
    for (int d = 0; d < s->dims; d++) f.wall_dt[d] = (float)(s->dt / s->dx[d]) * 0.5f;
  }
  if (!s->zp.empty()) {
    f.las_k = zp_launches(s);
    f.las_rows = zp_rows(s, f.las_k);
    int k = 0;
    for (auto& w : s->zp)
      for (int r = 0; r < (w.pol_type != 0 ? 2 : 1); r++)
        f.las_col[k++] = w.backward ? s->g.open_nx - 1 : 0;  // x = nx or x = 1
    f.las_tab = s->zp_tab.as<float>();
    f.las_mask = s->zp_mask.as<int>();
    f.las_first = s->zp_first.as<long long>();
  }
  if (f.sm_np > 0) {
    f.sm_src = s->plain ? 1 : 0;
    f.sm_buf[0] = s->sm_buf[0].as<float>();
    f.sm_buf[1] = s->sm_buf[1].as<float>();
    f.Jf = f.sm_buf[(f.sm_np - 1) & 1];
    f.jf_zero = 0;
    f.jsrc_clear = s->plain ? 2 : 1;
  }
}

int upload_descs(osr_sim* s) {
  std::vector<SpecDesc> v(kMaxSpecies);
  make_spec_descs(s, v.data());
  TRY(copy_on(s->stream, s->d_spec.p, v.data(), sizeof(SpecDesc) * kMaxSpecies, hipMemcpyHostToDevice));
  make_field_desc(s, s->fd);
  TRY(copy_on(s->stream, s->d_fd.p, &s->fd, sizeof(FieldDesc), hipMemcpyHostToDevice));
  s->desc_version++;
  return OSR_OK;
}

// Cell coordinates (0-based) of a packed cell.
inline void cell_coords(const osr_sim* s, uint32_t c, int* ic) {
  if (s->dims == 2) {
    unpack2(c, ic[0], ic[1]);
    ic[2] = 0;
  } else
    unpack3(c, ic[0], ic[1], ic[2]);
}

// Allocate a species' arrays for the given per-tile counts and capacities.
int alloc_species(osr_sim* s, Species& p, const std::vector<int>& cnt) {
  const int nt = s->ntiles;
  std::vector<int> cap(nt), off(nt);
  int64_t total = 0;
  for (int t = 0; t < nt; t++) {
    const int c = cnt[t];
    int64_t cp = (int64_t)c + std::max(c / 4, 64) + 64 * kMaxPushWaves;
    cp = (cp + 63) / 64 * 64;
    if (total + cp > INT32_MAX) return fail(OSR_ERR_ARG, "species too large for int32 slots");
    cap[t] = (int)cp;
    off[t] = (int)total;
    total += cp;
  }
  p.cap_total = total;
  const size_t n = (size_t)total;
  TRY(p.cell.alloc(4 * n));
  TRY(p.x.alloc(4 * n));
  TRY(p.y.alloc(4 * n));
  TRY(p.z.alloc(s->dims == 3 ? 4 * n : 4));
  TRY(p.ux.alloc(4 * n));
  TRY(p.uy.alloc(4 * n));
  TRY(p.uz.alloc(4 * n));
  TRY(p.q.alloc(4 * n));
  TRY(p.pid.alloc(4 * n));
  TRY(p.mcell.alloc(4 * n));
  TRY(p.mx.alloc(4 * n));
  TRY(p.my.alloc(4 * n));
  TRY(p.mz.alloc(s->dims == 3 ? 4 * n : 4));
  TRY(p.mux.alloc(4 * n));
  TRY(p.muy.alloc(4 * n));
  TRY(p.muz.alloc(4 * n));
  TRY(p.mq.alloc(4 * n));
  TRY(p.mpid.alloc(4 * n));
  TRY(p.midx.alloc(4 * n));
  TRY(p.off.alloc(4 * (size_t)nt));
  TRY(p.cnt.alloc(4 * (size_t)nt));
  TRY(p.cap.alloc(4 * (size_t)nt));
  TRY(p.mv_cnt.alloc(4 * (size_t)nt * kMaxPushWaves));
  TRY(p.mv_stride.alloc(4 * (size_t)nt));
  TRY(p.newcnt.alloc(4 * (size_t)nt));
  TRY(p.inc.alloc(4 * (size_t)nt));
  p.epart_n = s->plain ? (int)((total + 255) / 256) : nt;
  TRY(p.epart.alloc(8 * (size_t)p.epart_n));
  if (s->plain) {
    TRY(p.pl_ep.alloc(12 * n));
    TRY(p.pl_bp.alloc(12 * n));
    TRY(p.pl_vp.alloc(7 * 3 * 4 * n));
    TRY(p.pl_nvp.alloc(4 * n));
  }
  TRY(copy_on(s->stream, p.off.p, off.data(), 4 * nt, hipMemcpyHostToDevice));
  TRY(copy_on(s->stream, p.cnt.p, cnt.data(), 4 * nt, hipMemcpyHostToDevice));
  TRY(copy_on(s->stream, p.cap.p, cap.data(), 4 * nt, hipMemcpyHostToDevice));
  HIPC(hipMemsetAsync(p.mv_cnt.p, 0, 4 * (size_t)nt * kMaxPushWaves, s->stream));
  HIPC(hipMemsetAsync(p.mv_stride.p, 0, 4 * (size_t)nt, s->stream));
  HIPC(hipMemsetAsync(p.epart.p, 0, 8 * (size_t)p.epart_n, s->stream));
  HIPC(hipMemsetAsync(p.inc.p, 0, 4 * (size_t)nt, s->stream));
  HIPC(hipStreamSynchronize(s->stream));
  p.h_off = off;
  p.h_cap = cap;
  int64_t np = 0;
  for (int c : cnt) np += c;
  p.np = np;
  return OSR_OK;
}

template <class T>
int upload(hipStream_t st, DBuf& b, const std::vector<T>& v) {
  return copy_on(st, b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
}

// Lay a host particle list out in tiles (order inside a tile: round robin over the
// tile's cells, so that neighbouring lanes work on different cells) and upload it.
int layout_upload(osr_sim* s, Species& p, const HostParts& hp) {
  const int nt = s->ntiles;
  const size_t n = hp.size();
  std::vector<int> tile(n), cnt(nt, 0);
  for (size_t i = 0; i < n; i++) {
    tile[i] = tile_of(s->g, hp.cell[i]);
    cnt[tile[i]]++;
  }
  TRY(alloc_species(s, p, cnt));
  // bucket by tile (stable), then by local cell, round robin
  std::vector<int64_t> tstart(nt + 1, 0);
  for (int t = 0; t < nt; t++) tstart[t + 1] = tstart[t] + cnt[t];
  std::vector<int64_t> byt(n);
  {
    std::vector<int64_t> fillp(tstart.begin(), tstart.end() - 1);
    for (size_t i = 0; i < n; i++) byt[fillp[tile[i]]++] = (int64_t)i;
  }
  const size_t C = (size_t)p.cap_total;
  std::vector<uint32_t> cell(C, 0), pid(C, 0);
  std::vector<float> x(C, 0.f), y(C, 0.f), z(s->dims == 3 ? C : 0, 0.f), ux(C, 0.f), uy(C, 0.f),
      uz(C, 0.f), q(C, 0.f);
  const int tcell = s->g.tx * s->g.ty * (s->dims == 3 ? s->g.tz : 1);
  std::vector<int> ccount(tcell), crank;
  std::vector<int64_t> order;
  for (int t = 0; t < nt; t++) {
    const int64_t b0 = tstart[t], b1 = tstart[t + 1];
    if (b1 == b0) continue;
    std::fill(ccount.begin(), ccount.end(), 0);
    crank.assign((size_t)(b1 - b0), 0);
    int maxc = 0;
    for (int64_t k = b0; k < b1; k++) {
      int ic[3];
      cell_coords(s, hp.cell[byt[k]], ic);
      const int lc = cell_key(ic[0] % s->g.tx, ic[1] % s->g.ty, s->dims == 3 ? ic[2] % s->g.tz : 0,
                              s->dims, s->g.tx, s->g.ty, s->dims == 3 ? s->g.tz : 1);
      crank[k - b0] = ccount[lc]++;
      maxc = std::max(maxc, ccount[lc]);
    }
    // order by (rank within cell, local cell, input order)
    order.resize((size_t)(b1 - b0));
    std::vector<int> bucket(maxc + 1, 0);
    for (int64_t k = b0; k < b1; k++) bucket[crank[k - b0] + 1]++;
    for (int r = 0; r < maxc; r++) bucket[r + 1] += bucket[r];
    // inside one rank, keep local-cell order: counting sort by cell first
    std::vector<int64_t> bycell((size_t)(b1 - b0));
    {
      std::vector<int> cstart(tcell + 1, 0);
      std::vector<int> lcv((size_t)(b1 - b0));
      for (int64_t k = b0; k < b1; k++) {
        int ic[3];
        cell_coords(s, hp.cell[byt[k]], ic);
        lcv[k - b0] = cell_key(ic[0] % s->g.tx, ic[1] % s->g.ty, s->dims == 3 ? ic[2] % s->g.tz : 0,
                               s->dims, s->g.tx, s->g.ty, s->dims == 3 ? s->g.tz : 1);
        cstart[lcv[k - b0] + 1]++;
      }
      for (int c = 0; c < tcell; c++) cstart[c + 1] += cstart[c];
      for (int64_t k = b0; k < b1; k++) bycell[cstart[lcv[k - b0]]++] = k;
    }
    for (int64_t kk : bycell) order[bucket[crank[kk - b0]]++] = byt[kk];
    const int64_t base = p.h_off[t];
    for (int64_t k = 0; k < b1 - b0; k++) {
      const int64_t i = order[k], o = base + k;
      cell[o] = hp.cell[i];
      pid[o] = hp.pid[i];
      x[o] = hp.x[i];
      y[o] = hp.y[i];
      if (s->dims == 3) z[o] = hp.z[i];
      ux[o] = hp.ux[i];
      uy[o] = hp.uy[i];
      uz[o] = hp.uz[i];
      q[o] = hp.q[i];
    }
  }
  TRY(upload(s->stream, p.cell, cell));
  TRY(upload(s->stream, p.pid, pid));
  TRY(upload(s->stream, p.x, x));
  TRY(upload(s->stream, p.y, y));
  if (s->dims == 3) TRY(upload(s->stream, p.z, z));
  TRY(upload(s->stream, p.ux, ux));
  TRY(upload(s->stream, p.uy, uy));
  TRY(upload(s->stream, p.uz, uz));
  TRY(upload(s->stream, p.q, q));
  double qa = 0.0, qm = 0.0;
  for (size_t i = 0; i < n; i++) {
    qa += std::fabs((double)hp.q[i]);
    qm = std::max(qm, std::fabs((double)hp.q[i]));
  }
  p.qabs = qa;
  p.qmax = qm;
  return OSR_OK;
}

template <class T>
int download(hipStream_t st, std::vector<T>& v, const DBuf& b, size_t n) {
  v.resize(n);
  return copy_on(st, v.data(), b.p, sizeof(T) * n, hipMemcpyDeviceToHost);
}

// Collect the valid particles of a species.  `pending`: the last push moved leavers
// into the mover buffer but the fill did not run (halted step).
int collect(osr_sim* s, Species& p, bool pending, HostParts& hp) {
  const int nt = s->ntiles;
  const int W = s->push_waves;  // wave count of the push that filled the mover buffers
  const size_t C = (size_t)p.cap_total;
  std::vector<int> cnt, mvc, mvs, midx;
  TRY(download(s->stream, cnt, p.cnt, nt));
  std::vector<uint32_t> cell, pid;
  std::vector<float> x, y, z, ux, uy, uz, q;
  TRY(download(s->stream, cell, p.cell, C));
  TRY(download(s->stream, pid, p.pid, C));
  TRY(download(s->stream, x, p.x, C));
  TRY(download(s->stream, y, p.y, C));
  if (s->dims == 3) TRY(download(s->stream, z, p.z, C));
  TRY(download(s->stream, ux, p.ux, C));
  TRY(download(s->stream, uy, p.uy, C));
  TRY(download(s->stream, uz, p.uz, C));
  TRY(download(s->stream, q, p.q, C));
  std::vector<uint32_t> mcell, mpid;
  std::vector<float> mx, my, mz, mux, muy, muz, mq;
  if (pending) {
    TRY(download(s->stream, mvc, p.mv_cnt, (size_t)nt * W));
    TRY(download(s->stream, mvs, p.mv_stride, nt));
    TRY(download(s->stream, midx, p.midx, C));
    TRY(download(s->stream, mcell, p.mcell, C));
    TRY(download(s->stream, mpid, p.mpid, C));
    TRY(download(s->stream, mx, p.mx, C));
    TRY(download(s->stream, my, p.my, C));
    if (s->dims == 3) TRY(download(s->stream, mz, p.mz, C));
    TRY(download(s->stream, mux, p.mux, C));
    TRY(download(s->stream, muy, p.muy, C));
    TRY(download(s->stream, muz, p.muz, C));
    TRY(download(s->stream, mq, p.mq, C));
  }
  hp = HostParts();
  hp.reserve((size_t)p.np);
  auto push = [&](const std::vector<uint32_t>& c_, const std::vector<uint32_t>& id_,
                  const std::vector<float>& x_, const std::vector<float>& y_,
                  const std::vector<float>& z_, const std::vector<float>& ux_,
                  const std::vector<float>& uy_, const std::vector<float>& uz_,
                  const std::vector<float>& q_, size_t i) {
    hp.cell.push_back(c_[i]);
    hp.pid.push_back(id_[i]);
    hp.x.push_back(x_[i]);
    hp.y.push_back(y_[i]);
    hp.z.push_back(s->dims == 3 ? z_[i] : 0.f);
    hp.ux.push_back(ux_[i]);
    hp.uy.push_back(uy_[i]);
    hp.uz.push_back(uz_[i]);
    hp.q.push_back(q_[i]);
  };
  // after a halted step, tiles whose fill ran have inc >= 0 and already hold their
  // incoming movers; the others still have their holes and their movers are pending
  std::vector<int> inc;
  if (pending) TRY(download(s->stream, inc, p.inc, nt));
  std::vector<char> hole;
  for (int t = 0; t < nt; t++) {
    const int base = p.h_off[t];
    hole.assign((size_t)cnt[t], 0);
    if (pending && inc[t] < 0)
      for (int w = 0; w < W; w++)
        for (int k = 0; k < mvc[t * W + w]; k++) {
          const int h = midx[base + w * mvs[t] + k];
          if (h >= 0 && h < cnt[t]) hole[h] = 1;
        }
    for (int k = 0; k < cnt[t]; k++)
      if (!hole[k]) push(cell, pid, x, y, z, ux, uy, uz, q, (size_t)base + k);
  }
  if (pending)
    for (int t = 0; t < nt; t++)
      for (int w = 0; w < W; w++)
        for (int k = 0; k < mvc[t * W + w]; k++) {
          const size_t m = (size_t)p.h_off[t] + w * mvs[t] + k;
          if (inc[tile_of(s->g, mcell[m])] < 0)
            push(mcell, mpid, mx, my, mz, mux, muy, muz, mq, m);
        }
  return OSR_OK;
}

// Stable order by pid, used to return particles in the order they were given.
void sort_by_pid(HostParts& hp) {
  const size_t n = hp.size();
  std::vector<size_t> idx(n);
  for (size_t i = 0; i < n; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return hp.pid[a] < hp.pid[b]; });
  HostParts o;
  o.reserve(n);
  for (size_t i : idx) {
    o.cell.push_back(hp.cell[i]);
    o.pid.push_back(hp.pid[i]);
    o.x.push_back(hp.x[i]);
    o.y.push_back(hp.y[i]);
    o.z.push_back(hp.z[i]);
    o.ux.push_back(hp.ux[i]);
    o.uy.push_back(hp.uy[i]);
    o.uz.push_back(hp.uz[i]);
    o.q.push_back(hp.q[i]);
  }
  hp = std::move(o);
}

// Fill kernel and its launch shape (4 waves per tile; LDS lists for up to 4096 incoming
// particles and 32768 holes left in a tile's tail).
typedef void (*fill_fn)(const SpecDesc*, const int2*);
struct FillKernel {
  fill_fn f;
  int nt;
};
FillKernel pick_fill() { return {tile_fill<256, 4096, 32768>, 256}; }

// Device re-layout of species `sid` after a halted step (a tile fill overflowed): every
// tile gets its count after the step plus half of it again as capacity (no tile
// shrinks), every tile's particles and pending mover lists move to the
// new arrays on the device (regrow_move), and the fill of the overflowed tiles runs again
// there.  *ok = false when that cannot finish the step (a tile's incoming or tail lists
// exceed the fill kernel's LDS tables, or slots would exceed int32): the species is then
// left in the halted state, possibly in the larger layout, for the host re-layout.
int regrow_species(osr_sim* s, int sid, bool* ok) {
  Species& p = *s->sp[sid];
  const int nt = s->ntiles;
  *ok = false;
  std::vector<int> cnt, inc, need;
  TRY(download(s->stream, cnt, p.cnt, nt));
  TRY(download(s->stream, inc, p.inc, nt));
  TRY(download(s->stream, need, p.newcnt, nt));
  std::vector<int2> over;
  for (int t = 0; t < nt; t++)
    if (inc[t] < 0) {
      if (need[t] < 0) return OSR_OK;
      over.push_back(make_int2(sid, t));
    }
  if (over.empty()) {
    *ok = true;
    return OSR_OK;
  }
  std::vector<int> cap(nt), off(nt);
  int64_t total = 0;
  for (int t = 0; t < nt; t++) {
    // every tile gets room for half its count again (a density front that overflows one
    // tile usually reaches its neighbours next); no tile shrinks
    const int64_t r = inc[t] < 0 ? need[t] : cnt[t];
    int64_t c = std::max<int64_t>(p.h_cap[t], r + std::max<int64_t>(r / 2, 64) + 64 * kMaxPushWaves);
    c = (c + 63) / 64 * 64;
    if (total + c > INT32_MAX) return OSR_OK;
    cap[t] = (int)c;
    off[t] = (int)total;
    total += c;
  }
  const size_t n = (size_t)total;
  const bool three = s->dims == 3;
  DBuf* arr[19] = {&p.cell, &p.x, &p.y, &p.z, &p.ux, &p.uy, &p.uz, &p.q, &p.pid,
                   &p.mcell, &p.mx, &p.my, &p.mz, &p.mux, &p.muy, &p.muz, &p.mq, &p.mpid,
                   &p.midx};
  DBuf nb[19];
  RegrowArrays src{}, dst{};
  for (int a = 0; a < 19; a++) {
    const bool dummy = !three && (a == 3 || a == 12);  // z and mz of a 2D species
    TRY(nb[a].alloc(dummy ? 4 : 4 * n));
    src.a[a] = arr[a]->as<uint32_t>();
    dst.a[a] = nb[a].as<uint32_t>();
  }
  DBuf d_off, d_work;
  TRY(d_off.alloc(4 * (size_t)nt));
  TRY(d_work.alloc(sizeof(int2) * over.size()));
  TRY(copy_on(s->stream, d_off.p, off.data(), 4 * (size_t)nt, hipMemcpyHostToDevice));
  TRY(copy_on(s->stream, d_work.p, over.data(), sizeof(int2) * over.size(), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(regrow_move, dim3((unsigned)nt), dim3(256), 0, s->stream,
                     s->d_spec.as<SpecDesc>(), sid, src, dst, d_off.as<int>());
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(s->stream));
  for (int a = 0; a < 19; a++) arr[a]->swap(nb[a]);  // the old arrays go with nb
  TRY(copy_on(s->stream, p.off.p, off.data(), 4 * (size_t)nt, hipMemcpyHostToDevice));
  TRY(copy_on(s->stream, p.cap.p, cap.data(), 4 * (size_t)nt, hipMemcpyHostToDevice));
  p.h_off = off;
  p.h_cap = cap;
  p.cap_total = total;
  TRY(upload_descs(s));
  // the fill of the overflowed tiles, on the new layout (the halt flag cleared for it)
  Ctrl c{};
  TRY(copy_on(s->stream, &c, s->ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost));
  Ctrl z = c;
  z.halt = 0;
  TRY(copy_on(s->stream, s->ctrl.p, &z, sizeof(Ctrl), hipMemcpyHostToDevice));
  const FillKernel fk = pick_fill();
  hipLaunchKernelGGL(fk.f, dim3((unsigned)over.size()), dim3(fk.nt), 0, s->stream,
                     s->d_spec.as<SpecDesc>(), d_work.as<int2>());
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(s->stream));
  Ctrl after{};
  TRY(copy_on(s->stream, &after, s->ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost));
  TRY(copy_on(s->stream, s->ctrl.p, &c, sizeof(Ctrl), hipMemcpyHostToDevice));
  if (after.halt != 0 || after.error != 0) return OSR_OK;
  TRY(download(s->stream, cnt, p.cnt, nt));
  int64_t np = 0;
  for (int v : cnt) np += v;
  p.np = np;
  *ok = true;
  return OSR_OK;
}

// Re-layout of every species.  After a halted step (`pending`: the fill overflowed) the
// device re-layout is tried first and the host one (collect + layout_upload) is the
// fallback; otherwise the particles are laid out again from the host.
int relayout(osr_sim* s, bool pending) {
  for (size_t i = 0; i < s->sp.size(); i++) {
    if (pending) {
      bool ok = false;
      TRY(regrow_species(s, (int)i, &ok));
      if (ok) continue;
    }
    HostParts hp;
    TRY(collect(s, *s->sp[i], pending, hp));
    TRY(layout_upload(s, *s->sp[i], hp));
  }
  TRY(upload_descs(s));
  s->relayouts++;
  return OSR_OK;
}

// Tile processed by the workgroup at position p of a species' launch.  Workgroups are
// handed to the chip's 8 XCDs round robin (p mod 8), and each XCD has its own L2.  In 3D
// the tiles of one XCD are a contiguous run of the tile order (x fastest): the field halo
// a tile stages and the current halo it adds into (4.5 and 5.9 times its interior for
// 8x8x4 cubic tiles) are then mostly lines its neighbours on the same L2 also touch,
// which cuts the push's HBM reads by 40% on a uniform plasma.  In 2D the halos are small
// and the plain order is kept: it spreads the tiles of a dense region (a laser-heated
// channel) over all XCDs.  Tiles are independent work items, so the order changes no
// result.
constexpr int kXcds = 8;
int xcd_tile(int dims, int p, int ntiles) {
  if (dims != 3) return p;
  const int q = ntiles / kXcds, r = ntiles % kXcds, k = p % kXcds;
  const int start = k * q + std::min(k, r);  // XCD k runs q tiles, one more if k < r
  return start + p / kXcds;
}

// ---- kernel dispatch -------------------------------------------------------------
typedef void (*push_fn)(const SpecDesc*, const int2*);

push_fn pick_push(int dims, int order, int pusher) {
#define OSR_P(D, O, P) \
  if (dims == D && order == O && pusher == P) return push_deposit_##D##d<O, P, kNT>;
  OSR_P(2, 1, 0) OSR_P(2, 2, 0) OSR_P(2, 3, 0) OSR_P(2, 1, 1) OSR_P(2, 2, 1) OSR_P(2, 3, 1)
  OSR_P(3, 1, 0) OSR_P(3, 2, 0) OSR_P(3, 3, 0) OSR_P(3, 1, 1) OSR_P(3, 2, 1) OSR_P(3, 3, 1)
#undef OSR_P
  return nullptr;
}

// 2D kernels with compile-time tile shape (static LDS) and `waves` waves per workgroup;
// nullptr if not instantiated
push_fn pick_push2d_t(int order, int pusher, int tx, int ty, int waves) {
#define OSR_T(O, P, X, Y, W)                                                 \
  if (order == O && pusher == P && tx == X && ty == Y && waves == W)         \
    return push_deposit_2d_t<O, P, X, Y, W>;
#define OSR_TT(X, Y, W)                                                      \
  OSR_T(1, 0, X, Y, W) OSR_T(2, 0, X, Y, W) OSR_T(3, 0, X, Y, W)             \
  OSR_T(1, 1, X, Y, W) OSR_T(2, 1, X, Y, W) OSR_T(3, 1, X, Y, W)
  OSR_TT(8, 8, 4) OSR_TT(16, 16, 4) OSR_TT(16, 16, 8) OSR_TT(32, 16, 4) OSR_TT(32, 16, 8)
#undef OSR_TT
#undef OSR_T
  return nullptr;
}

// 3D kernels with compile-time tile shape (static LDS); nullptr if not instantiated
push_fn pick_push3d_t(int order, int pusher, int tx, int ty, int tz) {
#define OSR_T3(O, P, X, Y, Z)                                               \
  if (order == O && pusher == P && tx == X && ty == Y && tz == Z)           \
    return push_deposit_3d_t<O, P, X, Y, Z>;
#define OSR_TT3(X, Y, Z)                                                     \
  OSR_T3(1, 0, X, Y, Z) OSR_T3(2, 0, X, Y, Z) OSR_T3(3, 0, X, Y, Z)          \
  OSR_T3(1, 1, X, Y, Z) OSR_T3(2, 1, X, Y, Z) OSR_T3(3, 1, X, Y, Z)
  OSR_TT3(8, 8, 4)
#undef OSR_TT3
#undef OSR_T3
  return nullptr;
}

// The fused push kernel of a simulation (or batch) kind: the compile-time tile variant
// with `waves` waves per workgroup when one is instantiated and the generic kernel was
// not asked for, else the generic kernel (4 waves) with `lds` bytes of dynamic LDS.  Sets
// the threads per workgroup and the dynamic LDS of the launch.
push_fn pick_push_kernel(int dims, int order, int pusher, int kernel, const int* tile,
                         int waves, size_t lds, size_t* lds_launch, int* nt) {
  push_fn f = nullptr;
  *lds_launch = lds;
  *nt = kNT;
  if (kernel != 1) {
    if (dims == 2) {
      f = pick_push2d_t(order, pusher, tile[0], tile[1], waves);
      if (f) *nt = 64 * waves;
    } else {
      f = pick_push3d_t(order, pusher, tile[0], tile[1], tile[2]);
    }
    if (f) *lds_launch = 0;
  }
  if (!f) f = pick_push(dims, order, pusher);
  return f;
}

// One push launch for species first .. first+count-1 (same pusher): the work list holds
// their (species, tile) items back to back, so the workgroups of the later species fill
// the tail of the earlier one's.  The species only share the read-only fields and the
// integer current, so the result does not depend on the grouping.
int launch_push(osr_sim* s, int first, int count = 1) {
  const int pusher = s->sp[first]->pusher;
  const int tile[3] = {s->g.tx, s->g.ty, s->g.tz};
  size_t lds = 0;
  int nt = kNT;
  push_fn f = pick_push_kernel(s->dims, s->order, pusher, s->desc.kernel, tile, s->push_waves,
                               s->lds_push, &lds, &nt);
  if (!f) return fail(OSR_ERR_ARG, "no push kernel for dims=%d order=%d", s->dims, s->order);
  if (nt != 64 * s->push_waves) return fail(OSR_ERR_STATE, "push wave count mismatch");
  hipLaunchKernelGGL(f, dim3((unsigned)(count * s->ntiles)), dim3(nt), lds, s->stream,
                     s->d_spec.as<SpecDesc>(), s->d_work.as<int2>() + (size_t)first * s->ntiles);
  HIPC(hipGetLastError());
  return OSR_OK;
}

// every species, one launch per run of consecutive species with the same pusher
int launch_pushes(osr_sim* s) {
  const int ns = (int)s->sp.size();
  for (int a = 0; a < ns;) {
    int b = a + 1;
    while (b < ns && s->sp[b]->pusher == s->sp[a]->pusher) b++;
    TRY(launch_push(s, a, b - a));
    a = b;
  }
  return OSR_OK;
}


// the in-tile particle order restored every s->reorder steps (the kernel itself checks the
// step and returns at once on the other steps)
int launch_reorder(osr_sim* s) {
  if (s->reorder <= 0 || s->plain) return OSR_OK;
  const int nw = (int)s->sp.size() * s->ntiles;
  hipLaunchKernelGGL((tile_reorder<256, 64, 512, 4096>), dim3(nw), dim3(256), 0, s->stream,
                     s->d_spec.as<SpecDesc>(), s->d_work.as<int2>(), s->reorder);
  HIPC(hipGetLastError());
  return OSR_OK;
}

int launch_fill(osr_sim* s) {
  const int nw = (int)s->sp.size() * s->ntiles;
  const FillKernel k = pick_fill();
  hipLaunchKernelGGL(k.f, dim3(nw), dim3(k.nt), 0, s->stream, s->d_spec.as<SpecDesc>(),
                     s->d_work.as<int2>());
  HIPC(hipGetLastError());
  return OSR_OK;
}

// E and B on the x walls from the Lindman buffers (the state every push starts from)
int launch_wall_restore(osr_sim* s) {
  if (s->vp_n) return OSR_OK;  // VPML walls put nothing back on the box
  const unsigned nb = (unsigned)ceil_div(s->g.ny * s->g.nz, 256);
  hipLaunchKernelGGL(wall_restore, dim3(nb), dim3(256), 0, s->stream, s->d_fd.as<FieldDesc>());
  HIPC(hipGetLastError());
  return OSR_OK;
}

int launch_yee(osr_sim* s) {
  const FieldDesc* fd = s->d_fd.as<FieldDesc>();
  const int2* w = s->d_ywork.as<int2>();
  const int2* we = s->d_ework.as<int2>();
  if (s->vp_n) {  // advance_yee with VPML x walls: update_boundary_bfld step 1 advances
                   // the wall E, update_boundary_efld the wall B (os-emf-boundary.f90:930-935,
                   // 1036-1041)
    const unsigned nb = (unsigned)((2ll * s->vp_n * s->g.ny * s->g.nz + 255) / 256);
    if (s->dims == 2) {
      hipLaunchKernelGGL((yee_dbdt<2, true>), dim3(s->nblk_e), dim3(256), 0, s->stream, fd, we);
      if (s->desc.vpml_diffuse[0] || s->desc.vpml_diffuse[1])
        hipLaunchKernelGGL(vpml_diffuse, dim3(nb), dim3(256), 0, s->stream, fd);
      hipLaunchKernelGGL((vpml_update<0, 2>), dim3(nb), dim3(256), 0, s->stream, fd);
      hipLaunchKernelGGL((yee_dedt<2>), dim3(s->nblk), dim3(256), 0, s->stream, fd, w);
      hipLaunchKernelGGL((vpml_update<1, 2>), dim3(nb), dim3(256), 0, s->stream, fd);
      hipLaunchKernelGGL((yee_dbdt<2, false>), dim3(s->nblk), dim3(256), 0, s->stream, fd, w);
    } else {
      hipLaunchKernelGGL((yee_dbdt<3, true>), dim3(s->nblk_e), dim3(256), 0, s->stream, fd, we);
      hipLaunchKernelGGL((vpml_update<0, 3>), dim3(nb), dim3(256), 0, s->stream, fd);
      hipLaunchKernelGGL((yee_dedt<3>), dim3(s->nblk), dim3(256), 0, s->stream, fd, w);
      hipLaunchKernelGGL((vpml_update<1, 3>), dim3(nb), dim3(256), 0, s->stream, fd);
      hipLaunchKernelGGL((yee_dbdt<3, false>), dim3(s->nblk), dim3(256), 0, s->stream, fd, w);
    }
    HIPC(hipGetLastError());
    return OSR_OK;
  }
  if (s->g.open_nx) {  // advance_yee with Lindman x walls (os-emf-solver-yee.f03:542-568)
    const unsigned nb = (unsigned)ceil_div(s->g.ny * s->g.nz, 256);
    const unsigned nbf = (unsigned)ceil_div(s->g.ny * s->g.nz, kFinNT);
    // the work list of the first B half step carries the wall's first B update
    if (s->dims == 2) {
      hipLaunchKernelGGL((yee_dbdt<2, true>), dim3(s->nwork_e), dim3(256), 0, s->stream, fd, we);
      hipLaunchKernelGGL((yee_dedt<2>), dim3(s->nblk), dim3(256), 0, s->stream, fd, w);
      hipLaunchKernelGGL(wall_e, dim3(nb), dim3(256), 0, s->stream, fd);
      hipLaunchKernelGGL((yee_dbdt<2, false>), dim3(s->nblk), dim3(256), 0, s->stream, fd, w);
      hipLaunchKernelGGL((wall_b2_end<2>), dim3(nbf), dim3(kFinNT), 0, s->stream, fd);
    } else {
      hipLaunchKernelGGL((yee_dbdt<3, true>), dim3(s->nwork_e), dim3(256), 0, s->stream, fd, we);
      hipLaunchKernelGGL((yee_dedt<3>), dim3(s->nblk), dim3(256), 0, s->stream, fd, w);
      hipLaunchKernelGGL(wall_e, dim3(nb), dim3(256), 0, s->stream, fd);
      hipLaunchKernelGGL((yee_dbdt<3, false>), dim3(s->nblk), dim3(256), 0, s->stream, fd, w);
      hipLaunchKernelGGL((wall_b2_end<3>), dim3(nbf), dim3(kFinNT), 0, s->stream, fd);
    }
    // the last wall kernel also restored the wall values and ended the step
    HIPC(hipGetLastError());
    return OSR_OK;
  }
  if (s->dims == 2) {
    hipLaunchKernelGGL((yee_dbdt<2, true>), dim3(s->nblk_e), dim3(256), 0, s->stream, fd, we);
    hipLaunchKernelGGL((yee_dedt<2>), dim3(s->nblk), dim3(256), 0, s->stream, fd, w);
    hipLaunchKernelGGL((yee_dbdt<2, false>), dim3(s->nblk), dim3(256), 0, s->stream, fd, w);
  } else {
    hipLaunchKernelGGL((yee_dbdt<3, true>), dim3(s->nblk_e), dim3(256), 0, s->stream, fd, we);
    hipLaunchKernelGGL((yee_dedt<3>), dim3(s->nblk), dim3(256), 0, s->stream, fd, w);
    hipLaunchKernelGGL((yee_dbdt<3, false>), dim3(s->nblk), dim3(256), 0, s->stream, fd, w);
  }
  HIPC(hipGetLastError());
  return OSR_OK;
}

// End of the step: history row, step counter, halt protocol.  With Lindman walls the last
// wall kernel of launch_yee has done it.
int launch_finalize(osr_sim* s) {
  if (s->g.open_nx && !s->vp_n) return OSR_OK;
  hipLaunchKernelGGL(finalize_step, dim3(1), dim3(kFinNT), 0, s->stream, s->d_fd.as<FieldDesc>(),
                     s->d_fwork.as<int>());
  HIPC(hipGetLastError());
  return OSR_OK;
}

int launch_plain(osr_sim* s, int sid) {
  Species& p = *s->sp[sid];
  if (p.np == 0) return OSR_OK;
  PlainBufs b;
  b.ep = p.pl_ep.as<float>();
  b.bp = p.pl_bp.as<float>();
  b.vp = p.pl_vp.as<float>();
  b.nvp = p.pl_nvp.as<int>();
  b.Jf = s->J.as<float>();
  b.cap = p.cap_total;
  const SpecDesc* d = s->d_spec.as<SpecDesc>();
  const unsigned nb = (unsigned)((p.np + 255) / 256), nb3 = (unsigned)((3 * p.np + 255) / 256);
  typedef void (*pk)(const SpecDesc*, int, PlainBufs);
  pk gth = s->order == 1 ? plain_gather<1> : (s->order == 2 ? plain_gather<2> : plain_gather<3>);
  pk psh = p.pusher == OSR_BORIS ? plain_push<0> : plain_push<1>;
  pk dep = s->order == 1 ? plain_deposit<1> : (s->order == 2 ? plain_deposit<2> : plain_deposit<3>);
  hipLaunchKernelGGL(gth, dim3(nb), dim3(256), 0, s->stream, d, sid, b);
  hipLaunchKernelGGL(psh, dim3(nb), dim3(256), 0, s->stream, d, sid, b);
  hipLaunchKernelGGL(plain_move, dim3(nb), dim3(256), 0, s->stream, d, sid, b);
  hipLaunchKernelGGL(dep, dim3(nb3), dim3(256), 0, s->stream, d, sid, b);
  HIPC(hipGetLastError());
  return OSR_OK;
}

// current smoothing passes between the deposit and the field advance
int launch_smooth(osr_sim* s) {
  int np = 0;
  for (int d = 0; d < s->dims; d++) np += s->sm_order[d] > 0;
  for (int p = 0; p < np; p++)
    hipLaunchKernelGGL(smooth_pass, dim3(s->nblk), dim3(256), 0, s->stream,
                       s->d_fd.as<FieldDesc>(), s->d_ywork.as<int2>(), p);
  HIPC(hipGetLastError());
  return OSR_OK;
}

// laser walls of the step, before the push (os-simulation.f03:528-531)
int launch_laser(osr_sim* s) {
  if (s->zp.empty()) return OSR_OK;
  hipLaunchKernelGGL(laser_wall, dim3(ceil_div(s->g.ny * s->g.nz, 256)), dim3(256), 0, s->stream,
                     s->d_fd.as<FieldDesc>());
  HIPC(hipGetLastError());
  return OSR_OK;
}

// Host evaluation of the laser columns of steps first .. first+n-1 (t = step dt) into the
// device table the laser kernel reads.
int zp_fill(osr_sim* s, int64_t first, int n) {
  HIPC(hipStreamSynchronize(s->stream));  // the copies below are not ordered with the stream
  const int K = zp_launches(s), rows = zp_rows(s, K);
  const size_t R = (size_t)s->g.ny * s->g.nz;
  n = std::min(n, rows);
  std::vector<float> tab((size_t)n * K * 2 * R);
  std::vector<int> mask((size_t)rows * K, 0);
  for (int r = 0; r < n; r++) {
    const double t = (double)(first + r) * s->dt;
    int k = 0;
    for (auto& w : s->zp)
      for (int sec = 0; sec < (w.pol_type != 0 ? 2 : 1); sec++, k++) {
        const size_t e = (size_t)r * K + k;
        if (!osr_zp::active(w, t)) continue;
        mask[e] = 1;
        osr_zp::column(w, s->dims, s->g.ny, s->g.nz, s->dx, s->g.dt_dx[0], s->order, t, sec == 1,
                       &tab[e * 2 * R], &tab[e * 2 * R + R]);
      }
  }
  const long long f = first;
  TRY(copy_on(s->stream, s->zp_tab.p, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice));
  TRY(copy_on(s->stream, s->zp_mask.p, mask.data(), sizeof(int) * mask.size(), hipMemcpyHostToDevice));
  TRY(copy_on(s->stream, s->zp_first.p, &f, sizeof(f), hipMemcpyHostToDevice));
  return OSR_OK;
}

int enqueue_step(osr_sim* s) {
  TRY(launch_laser(s));
  if (s->plain) {
    for (size_t i = 0; i < s->sp.size(); i++) TRY(launch_plain(s, (int)i));
    TRY(launch_smooth(s));
    TRY(launch_yee(s));
    TRY(launch_finalize(s));
    return OSR_OK;
  }
  TRY(launch_pushes(s));
  if (!s->sp.empty()) TRY(launch_fill(s));
  if (!s->sp.empty()) TRY(launch_reorder(s));
  TRY(launch_smooth(s));
  TRY(launch_yee(s));
  TRY(launch_finalize(s));
  return OSR_OK;
}

// Capture of `nsteps` consecutive steps into one executable graph.  A replay boundary
// costs a few microseconds of idle device time that the kernels of one graph do not pay
// between them, so steps are replayed kGraphSteps at a time.  The kernels read the step
// and the halt flag on the device, so a step that halts leaves the later steps of the
// same replay idle, as separate replays would be.
int capture_steps(osr_sim* s, int nsteps, hipGraphExec_t* out) {
  hipGraph_t graph = nullptr;
  HIPC(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  int r = OSR_OK;
  for (int k = 0; k < nsteps && r == OSR_OK; k++) r = enqueue_step(s);
  hipError_t e = hipStreamEndCapture(s->stream, &graph);
  if (r != OSR_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return r;
  }
  if (e != hipSuccess) return fail(OSR_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(OSR_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  return OSR_OK;
}

int build_graph(osr_sim* s) {
  for (hipGraphExec_t* g : {&s->gexec, &s->gexec_k})
    if (*g) {
      (void)hipGraphExecDestroy(*g);
      *g = nullptr;
    }
  s->graph_ok = false;
  TRY(capture_steps(s, 1, &s->gexec));
  TRY(capture_steps(s, kGraphSteps, &s->gexec_k));
  s->graph_ok = true;
  return OSR_OK;
}

int ensure_hist(osr_sim* s, int64_t rows) {
  if (rows <= s->hist_cap) return OSR_OK;
  int64_t cap = std::max<int64_t>(rows, std::max<int64_t>(2 * s->hist_cap, 1024));
  const int ncol = 6 + (int)s->sp.size();
  DBuf nb;
  TRY(nb.alloc(sizeof(double) * (size_t)cap * ncol));
  if (s->hist_cap > 0 && s->hist.p)
    TRY(copy_on(s->stream, nb.p, s->hist.p, sizeof(double) * (size_t)s->hist_cap * ncol,
                   hipMemcpyDeviceToDevice));
  s->hist.swap(nb);
  s->hist_cap = cap;
  return upload_descs(s);
}

int read_ctrl(osr_sim* s, Ctrl& c) {
  if (!s->h_ctrl) HIPC(hipHostMalloc((void**)&s->h_ctrl, sizeof(Ctrl), hipHostMallocDefault));
  HIPC(hipMemcpyAsync(s->h_ctrl, s->ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, s->stream));
  HIPC(hipStreamSynchronize(s->stream));
  c = *s->h_ctrl;
  return OSR_OK;
}

int check_sim(osr_sim* s) {
  if (!s) return fail(OSR_ERR_ARG, "null simulation handle");
  return OSR_OK;
}

}  // namespace

// ====================================================================================
// C ABI
// ====================================================================================
extern "C" {

const char* osr_last_error(void) { return g_err.c_str(); }
int osr_version(void) { return 100; }

int osr_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int osr_sim_create(const osr_sim_desc* desc, osr_sim** out) {
  if (!desc || !out) return fail(OSR_ERR_ARG, "null argument");
  *out = nullptr;
  const int dims = desc->dims;
  if (dims != 2 && dims != 3) return fail(OSR_ERR_ARG, "dims must be 2 or 3");
  if (desc->order < 1 || desc->order > 3) return fail(OSR_ERR_ARG, "order must be 1..3");
  for (int d = 0; d < dims; d++)
    if (desc->nx[d] < 1 || !(desc->dx[d] > 0)) return fail(OSR_ERR_ARG, "bad nx/dx");
  if (dims == 2 && (desc->nx[0] > 65535 || desc->nx[1] > 65535))
    return fail(OSR_ERR_ARG, "2D grids are limited to 65535 cells per dimension");
  if (dims == 3 && (desc->nx[0] > 2047 || desc->nx[1] > 2047 || desc->nx[2] > 1023))
    return fail(OSR_ERR_ARG, "3D grids are limited to 2047x2047x1023 cells");
  if (!(desc->dt > 0)) return fail(OSR_ERR_ARG, "dt must be positive");
  {  // test_stability_yee, os-emf-solver-yee.f03:430-456
    double cour = 0;
    for (int d = 0; d < dims; d++) cour += 1.0 / (desc->dx[d] * desc->dx[d]);
    cour = std::sqrt(1.0 / cour);
    if (desc->dt > cour)
      return fail(OSR_ERR_ARG, "Yee stability condition violated: dt=%g > %g", desc->dt, cour);
  }
  if (osr_device_count() < 1) return fail(OSR_ERR_HIP, "no GPU visible");
  if (desc->path != OSR_PATH_FUSED && desc->path != OSR_PATH_PLAIN)
    return fail(OSR_ERR_ARG, "unknown path");
  if (desc->path == OSR_PATH_PLAIN && dims != 2)
    return fail(OSR_ERR_ARG, "the plain path is 2D only");
  if (desc->kernel < 0 || desc->kernel > 2) return fail(OSR_ERR_ARG, "unknown kernel variant");
  if (desc->xbnd != OSR_BND_PERIODIC && desc->xbnd != OSR_BND_LINDMAN &&
      desc->xbnd != OSR_BND_VPML)
    return fail(OSR_ERR_ARG, "unknown x boundary type");
  const bool open_x = desc->xbnd != OSR_BND_PERIODIC;
  const int vp_n = desc->xbnd == OSR_BND_VPML ? (desc->vpml_size > 0 ? desc->vpml_size : 8) : 0;
  if ((desc->vpml_diffuse[0] || desc->vpml_diffuse[1]) && (!vp_n || dims != 2))
    return fail(OSR_ERR_ARG, "vpml_diffuse needs 2D VPML x walls");
  // the wall's lower guard rows are max(field gc, 1) (vdf/os-wall.f90:190): one with linear
  // interpolation (get_min_gc_part, particles/os-particles.f03:321-360), and then the
  // reference's mixing skips row 0, so that its wall rows 1 and ny+1 part and the box
  // reads the guard one
  if ((desc->vpml_diffuse[0] || desc->vpml_diffuse[1]) && desc->order == 1)
    return fail(OSR_ERR_ARG, "vpml_diffuse is built for interpolation orders 2 and 3");
  if (vp_n && (vp_n < 4 || vp_n > 64))
    return fail(OSR_ERR_ARG, "vpml_size must be in 4..64 (setup_att_coeff divides by n - 3)");
  if (open_x && desc->path != OSR_PATH_FUSED)
    return fail(OSR_ERR_ARG, "open x walls need the fused path");
  std::unique_ptr<osr_sim> s(new osr_sim());
  s->desc = *desc;
  s->plain = desc->path == OSR_PATH_PLAIN;
  s->dims = dims;
  s->order = desc->order;
  s->dt = desc->dt;
  int tile[3] = {1, 1, 1};
  for (int d = 0; d < dims; d++) {
    s->dx[d] = desc->dx[d];
    tile[d] = desc->tile[d] > 0 ? desc->tile[d] : (dims == 2 ? 16 : (d < 2 ? 8 : 4));
    tile[d] = std::min(tile[d], desc->nx[d]);
    if (desc->path == OSR_PATH_PLAIN) tile[d] = desc->nx[d];  // one "tile": no sorting
  }
  GridDesc& g = s->g;
  g.dims = dims;
  g.nx = desc->nx[0];
  if (open_x) {
    // guard band of whole tiles: the reference's guard cells of the order
    // (get_min_gc_part, particles/os-particles.f03:321-360) fit with room to spare
    const int gc = s->order == 1 ? 3 : (s->order == 2 ? 5 : 7);
    if (desc->nx[0] % tile[0] != 0)
      return fail(OSR_ERR_ARG, "open x walls need nx to be a multiple of the tile width");
    g.open_nx = desc->nx[0];
    g.nx = desc->nx[0] + ceil_div(gc, tile[0]) * tile[0];
    if (g.nx > (dims == 2 ? 65535 : 2047))
      return fail(OSR_ERR_ARG, "open x walls: nx plus the guard band exceeds the grid limit");
  }
  g.ny = desc->nx[1];
  g.nz = dims == 3 ? desc->nx[2] : 1;
  g.ncell = (long long)g.nx * g.ny * g.nz;  // < 2^32 within the extents accepted above
  set_grid_divisors(g);
  g.tx = tile[0];
  g.ty = tile[1];
  g.tz = tile[2];
  g.ntx = ceil_div(g.nx, g.tx);
  g.nty = ceil_div(g.ny, g.ty);
  g.ntz = ceil_div(g.nz, g.tz);
  g.ntiles = g.ntx * g.nty * g.ntz;
  s->ntiles = g.ntiles;
  for (int d = 0; d < 3; d++) {
    g.dt_dx[d] = d < dims ? (float)(desc->dt / desc->dx[d]) : 0.f;
    g.jnorm[d] = d < dims ? (float)(desc->dx[d] / desc->dt / (dims == 2 ? 2.0 : 3.0)) : 0.f;
  }
  s->lds_push = s->plain ? 0 : push_lds_bytes(dims, s->order, tile);
  if (s->lds_push > 160 * 1024) return fail(OSR_ERR_ARG, "tile too large for LDS");
  s->reorder = desc->reorder > 0 ? desc->reorder : (desc->reorder == 0 && dims == 3 ? 4 : 0);
  s->push_waves = kWaves;
  if (desc->push_waves != 0 && desc->push_waves != 4 && desc->push_waves != 8)
    return fail(OSR_ERR_ARG, "push_waves must be 0, 4 or 8");
  if (!s->plain && dims == 2 && desc->kernel != 1) {
    const int w = desc->push_waves > 0 ? desc->push_waves : push2d_default_waves(g.tx, g.ty);
    if (pick_push2d_t(s->order, 0, g.tx, g.ty, w)) s->push_waves = w;
  }
  if (desc->push_waves > 0 && desc->push_waves != s->push_waves)
    return fail(OSR_ERR_ARG, "no push kernel with %d waves per workgroup for this tile and kernel",
                desc->push_waves);
  HIPC(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  const size_t nc = (size_t)g.ncell;
  TRY(s->E.alloc(3 * nc * 4));
  TRY(s->B.alloc(3 * nc * 4));
  TRY(s->J.alloc(3 * nc * 8));
  TRY(s->ctrl.alloc(sizeof(Ctrl)));
  s->vp_n = vp_n;
  if (open_x && vp_n) {
    // both walls zero at creation (vdf/os-wall.f90:189-205), coefficients behind them
    const size_t nw = (size_t)4 * (dims == 3 ? 6 : 4) * vp_n * g.ny * g.nz;
    std::vector<float> coef(12 * (size_t)vp_n);
    vpml_coefs(vp_n, (float)(desc->dt / desc->dx[0]), coef.data());
    TRY(s->wall.alloc(sizeof(float) * (nw + coef.size())));
    HIPC(hipMemsetAsync(s->wall.p, 0, sizeof(float) * nw, s->stream));
    HIPC(hipMemcpyAsync(s->wall.as<float>() + nw, coef.data(), sizeof(float) * coef.size(),
                        hipMemcpyHostToDevice, s->stream));
  } else if (open_x) {
    const size_t wb = sizeof(float) * ((size_t)kWallEntries * g.ny * g.nz + g.nz + g.ny);
    TRY(s->wall.alloc(wb));
    HIPC(hipMemsetAsync(s->wall.p, 0, wb, s->stream));
  }
  HIPC(hipMemsetAsync(s->E.p, 0, 3 * nc * 4, s->stream));
  HIPC(hipMemsetAsync(s->B.p, 0, 3 * nc * 4, s->stream));
  HIPC(hipMemsetAsync(s->J.p, 0, 3 * nc * 8, s->stream));
  HIPC(hipMemsetAsync(s->ctrl.p, 0, sizeof(Ctrl), s->stream));
  HIPC(hipStreamSynchronize(s->stream));
  g.E = s->E.as<float>();
  g.B = s->B.as<float>();
  g.J = s->J.as<unsigned long long>();
  g.ctrl = s->ctrl.as<Ctrl>();
  s->nblk = (int)((nc + 255) / 256);
  s->nblk_e = (s->nblk + kEneBlk - 1) / kEneBlk;
  TRY(s->fe_part.alloc(sizeof(double) * 6 * (size_t)s->nblk_e));
  HIPC(hipMemsetAsync(s->fe_part.p, 0, sizeof(double) * 6 * (size_t)s->nblk_e, s->stream));
  HIPC(hipStreamSynchronize(s->stream));
  TRY(s->d_spec.alloc(sizeof(SpecDesc) * kMaxSpecies));
  TRY(s->d_fd.alloc(sizeof(FieldDesc)));
  TRY(s->d_ywork.alloc(sizeof(int2) * (size_t)s->nblk));
  const int nwall_e = (g.open_nx && !vp_n) ? ceil_div(g.ny * g.nz, 256) : 0;
  s->nwork_e = s->nblk_e + nwall_e;
  TRY(s->d_ework.alloc(sizeof(int2) * (size_t)s->nwork_e));
  TRY(s->d_fwork.alloc(sizeof(int)));
  {
    std::vector<int2> yw((size_t)s->nblk);
    for (int b = 0; b < s->nblk; b++) yw[b] = make_int2(0, b);
    TRY(copy_on(s->stream, s->d_ywork.p, yw.data(), sizeof(int2) * yw.size(), hipMemcpyHostToDevice));
    std::vector<int2> ew(yw.begin(), yw.begin() + s->nblk_e);
    for (int b = 0; b < nwall_e; b++) ew.push_back(make_int2(0, -1 - b));
    TRY(copy_on(s->stream, s->d_ework.p, ew.data(), sizeof(int2) * ew.size(), hipMemcpyHostToDevice));
    int z = 0;
    TRY(copy_on(s->stream, s->d_fwork.p, &z, sizeof(int), hipMemcpyHostToDevice));
  }
  update_jscale(s.get());
  TRY(ensure_hist(s.get(), 1024));
  if (s->lds_push > 64 * 1024) {
    for (int o = 1; o <= 3; o++)
      for (int p = 0; p <= 1; p++) {
        push_fn f = pick_push(dims, o, p);
        HIPC(hipFuncSetAttribute((const void*)f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)s->lds_push));
      }
  }
  *out = s.release();
  return OSR_OK;
}

int osr_sim_destroy(osr_sim* s) {
  delete s;  // ~osr_sim waits for the stream and releases everything
  return OSR_OK;
}

int osr_sim_add_species(osr_sim* s, float rqm, int pusher, int* sid) {
  TRY(check_sim(s));
  if ((int)s->sp.size() >= kMaxSpecies) return fail(OSR_ERR_ARG, "too many species");
  if (pusher != OSR_BORIS && pusher != OSR_VAY) return fail(OSR_ERR_ARG, "unknown pusher");
  if (!(rqm != 0.f)) return fail(OSR_ERR_ARG, "rqm must be non-zero");
  std::unique_ptr<Species> p(new Species());
  p->rqm = rqm;
  p->pusher = pusher;
  std::vector<int> cnt(s->ntiles, 0);
  TRY(alloc_species(s, *p, cnt));
  s->sp.push_back(std::move(p));
  // the work list covers (species, tile) pairs, each species' tiles in XCD order
  const int ns = (int)s->sp.size();
  TRY(s->d_work.alloc(sizeof(int2) * (size_t)ns * s->ntiles));
  std::vector<int2> w((size_t)ns * s->ntiles);
  for (int i = 0; i < ns; i++)
    for (int t = 0; t < s->ntiles; t++) w[(size_t)i * s->ntiles + t] = make_int2(i, xcd_tile(s->dims, t, s->ntiles));
  TRY(copy_on(s->stream, s->d_work.p, w.data(), sizeof(int2) * w.size(), hipMemcpyHostToDevice));
  // history rows gain a column: restart the history buffer
  s->hist.release();
  s->hist_cap = 0;
  TRY(ensure_hist(s, std::max<int64_t>(1024, s->steps + 1)));
  s->graph_ok = false;
  if (sid) *sid = ns - 1;
  return OSR_OK;
}

int osr_species_set(osr_sim* s, int sid, int64_t np, const int* ix, const float* x,
                    const float* u, const float* q) {
  TRY(check_sim(s));
  if (sid < 0 || sid >= (int)s->sp.size()) return fail(OSR_ERR_ARG, "bad species id");
  if (np < 0 || np >= (int64_t)UINT32_MAX) return fail(OSR_ERR_ARG, "bad particle count");
  if (np > 0 && (!ix || !x || !u || !q)) return fail(OSR_ERR_ARG, "null particle array");
  const int D = s->dims;
  const int nxs[3] = {interior_nx(s->g), s->g.ny, s->g.nz};
  HostParts hp;
  hp.reserve((size_t)np);
  for (int64_t i = 0; i < np; i++) {
    int c[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) {
      c[d] = ix[D * i + d] - 1;
      const float xv = x[D * i + d];
      if (c[d] < 0 || c[d] >= nxs[d] || !(xv >= -0.5f && xv < 0.5f))
        return fail(OSR_ERR_ARG, "particle %lld outside the grid or x not in [-0.5,0.5)",
                    (long long)i);
    }
    hp.cell.push_back(D == 2 ? pack2(c[0], c[1]) : pack3(c[0], c[1], c[2]));
    hp.pid.push_back((uint32_t)i);
    hp.x.push_back(x[D * i]);
    hp.y.push_back(x[D * i + 1]);
    hp.z.push_back(D == 3 ? x[D * i + 2] : 0.f);
    hp.ux.push_back(u[3 * i]);
    hp.uy.push_back(u[3 * i + 1]);
    hp.uz.push_back(u[3 * i + 2]);
    hp.q.push_back(q[i]);
  }
  HIPC(hipStreamSynchronize(s->stream));
  TRY(layout_upload(s, *s->sp[sid], hp));
  update_jscale(s);
  TRY(upload_descs(s));
  s->graph_ok = false;
  return OSR_OK;
}

int osr_species_count(osr_sim* s, int sid, int64_t* np) {
  TRY(check_sim(s));
  if (sid < 0 || sid >= (int)s->sp.size() || !np) return fail(OSR_ERR_ARG, "bad argument");
  *np = s->sp[sid]->np;
  return OSR_OK;
}

int osr_species_get(osr_sim* s, int sid, int* ix, float* x, float* u, float* q) {
  TRY(check_sim(s));
  if (sid < 0 || sid >= (int)s->sp.size()) return fail(OSR_ERR_ARG, "bad species id");
  HIPC(hipStreamSynchronize(s->stream));
  HostParts hp;
  TRY(collect(s, *s->sp[sid], false, hp));
  sort_by_pid(hp);
  const int D = s->dims;
  for (size_t i = 0; i < hp.size(); i++) {
    int c[3];
    cell_coords(s, hp.cell[i], c);
    if (ix)
      for (int d = 0; d < D; d++) ix[D * i + d] = c[d] + 1;
    if (x) {
      x[D * i] = hp.x[i];
      x[D * i + 1] = hp.y[i];
      if (D == 3) x[D * i + 2] = hp.z[i];
    }
    if (u) {
      u[3 * i] = hp.ux[i];
      u[3 * i + 1] = hp.uy[i];
      u[3 * i + 2] = hp.uz[i];
    }
    if (q) q[i] = hp.q[i];
  }
  return OSR_OK;
}

int osr_species_generate(osr_sim* s, int sid, const int* ppc, double density,
                         const double* ufl, const double* uth, uint64_t seed, int random_pos) {
  TRY(check_sim(s));
  if (sid < 0 || sid >= (int)s->sp.size() || !ppc || !ufl || !uth)
    return fail(OSR_ERR_ARG, "bad argument");
  int tot = 1;
  for (int d = 0; d < s->dims; d++) {
    if (ppc[d] < 1) return fail(OSR_ERR_ARG, "ppc must be >= 1");
    tot *= ppc[d];
  }
  if ((double)tot * (double)s->g.ncell >= 4294967295.0)
    return fail(OSR_ERR_ARG, "too many particles for 32-bit particle ids");
  Species& p = *s->sp[sid];
  std::vector<int> cnt(s->ntiles);
  for (int t = 0; t < s->ntiles; t++) {
    const int tx = t % s->g.ntx, ty = (t / s->g.ntx) % s->g.nty, tz = t / (s->g.ntx * s->g.nty);
    const int wx = std::max(0, std::min(s->g.tx, interior_nx(s->g) - tx * s->g.tx));
    const int wy = std::min(s->g.ty, s->g.ny - ty * s->g.ty);
    const int wz = s->dims == 3 ? std::min(s->g.tz, s->g.nz - tz * s->g.tz) : 1;
    cnt[t] = wx * wy * wz * tot;
  }
  HIPC(hipStreamSynchronize(s->stream));
  TRY(alloc_species(s, p, cnt));
  const float q = (float)((p.rqm < 0 ? -1.0 : 1.0) * density / tot);
  p.qabs = std::fabs((double)q) * (double)p.np;
  p.qmax = std::fabs((double)q);
  update_jscale(s);
  TRY(upload_descs(s));
  GenArgs a{};
  for (int d = 0; d < 3; d++) {
    a.ppc[d] = d < s->dims ? ppc[d] : 1;
    a.ufl[d] = ufl[d];
    a.uth[d] = uth[d];
  }
  a.ppc_tot = tot;
  a.q = q;
  a.seed = seed;
  a.random_pos = random_pos;
  int maxc = 0;
  for (int v : cnt) maxc = std::max(maxc, v);
  const unsigned ych = (unsigned)std::max(1, std::min(65535, maxc / 4096));
  hipLaunchKernelGGL(generate_particles, dim3(s->ntiles, ych), dim3(256), 0, s->stream,
                     s->d_spec.as<SpecDesc>(), sid, a);
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(s->stream));
  s->graph_ok = false;
  return OSR_OK;
}

// open x walls: the wall values of the buffers, as the reference's boundary update puts
// them before every push (also after the host changed the fields)
static int restore_walls(osr_sim* s) {
  if (!s->g.open_nx) return OSR_OK;
  TRY(launch_wall_restore(s));
  HIPC(hipStreamSynchronize(s->stream));
  return OSR_OK;
}

static int fields_xfer(osr_sim* s, float* e, float* b, const int* ld, const int* gcl, bool set) {
  const int D = s->dims;
  // nxs: device extents (with open x walls the x ring includes the guard band);
  // nin: interior cells
  const int nxs[3] = {s->g.nx, s->g.ny, s->g.nz};
  const int nin[3] = {interior_nx(s->g), s->g.ny, s->g.nz};
  for (int d = 0; d < D; d++)
    if (ld[d] < nin[d] + gcl[d] || gcl[d] < 0) return fail(OSR_ERR_ARG, "bad field extents");
  const size_t nc = (size_t)s->g.ncell;
  std::vector<float> pe(3 * nc), pb(3 * nc);
  if (set && s->g.open_nx) {  // the guard band keeps its values
    HIPC(hipStreamSynchronize(s->stream));
    TRY(copy_on(s->stream, pe.data(), s->E.p, 12 * nc, hipMemcpyDeviceToHost));
    TRY(copy_on(s->stream, pb.data(), s->B.p, 12 * nc, hipMemcpyDeviceToHost));
  }
  const long ld0 = ld[0], ld1 = ld[1];
  auto ref = [&](int i, int j, int k) -> size_t {
    size_t o = (size_t)(i + gcl[0]) + (size_t)ld0 * (size_t)(j + gcl[1]);
    if (D == 3) o += (size_t)ld0 * ld1 * (size_t)(k + gcl[2]);
    return 3 * o;
  };
  if (set) {
    for (int k = 0; k < nin[2]; k++)
      for (int j = 0; j < nin[1]; j++)
        for (int i = 0; i < nin[0]; i++) {
          const size_t c = (size_t)i + (size_t)nxs[0] * (j + (size_t)nxs[1] * k);
          const size_t r = ref(i, j, k);
          for (int m = 0; m < 3; m++) {
            pe[m * nc + c] = e ? e[r + m] : 0.f;
            pb[m * nc + c] = b ? b[r + m] : 0.f;
          }
        }
    HIPC(hipStreamSynchronize(s->stream));
    if (e) TRY(copy_on(s->stream, s->E.p, pe.data(), 12 * nc, hipMemcpyHostToDevice));
    if (b) TRY(copy_on(s->stream, s->B.p, pb.data(), 12 * nc, hipMemcpyHostToDevice));
    return restore_walls(s);
  }
  HIPC(hipStreamSynchronize(s->stream));
  TRY(copy_on(s->stream, pe.data(), s->E.p, 12 * nc, hipMemcpyDeviceToHost));
  TRY(copy_on(s->stream, pb.data(), s->B.p, 12 * nc, hipMemcpyDeviceToHost));
  // every array cell (guards included) gets its periodic image; with open x walls the x
  // guard cells come from the guard band of the ring
  const int ldk = D == 3 ? ld[2] : 1;
  for (int k = 0; k < ldk; k++)
    for (int j = 0; j < ld[1]; j++)
      for (int i = 0; i < ld[0]; i++) {
        const int ii = ((i - gcl[0]) % nxs[0] + nxs[0]) % nxs[0];
        const int jj = ((j - gcl[1]) % nxs[1] + nxs[1]) % nxs[1];
        const int kk = D == 3 ? ((k - gcl[2]) % nxs[2] + nxs[2]) % nxs[2] : 0;
        const size_t c = (size_t)ii + (size_t)nxs[0] * (jj + (size_t)nxs[1] * kk);
        const size_t r = 3 * ((size_t)i + (size_t)ld0 * (j + (size_t)ld1 * k));
        for (int m = 0; m < 3; m++) {
          if (e) e[r + m] = pe[m * nc + c];
          if (b) b[r + m] = pb[m * nc + c];
        }
      }
  return OSR_OK;
}

int osr_fields_set(osr_sim* s, const float* e, const float* b, const int* ld, const int* gcl) {
  TRY(check_sim(s));
  if (!ld || !gcl) return fail(OSR_ERR_ARG, "null extents");
  return fields_xfer(s, (float*)e, (float*)b, ld, gcl, true);
}

int osr_fields_get(osr_sim* s, float* e, float* b, const int* ld, const int* gcl) {
  TRY(check_sim(s));
  if (!ld || !gcl) return fail(OSR_ERR_ARG, "null extents");
  return fields_xfer(s, e, b, ld, gcl, false);
}

int osr_fields_fill(osr_sim* s, const double* e0, const double* b0) {
  TRY(check_sim(s));
  const size_t nc = (size_t)s->g.ncell;
  std::vector<float> v(3 * nc);
  HIPC(hipStreamSynchronize(s->stream));
  for (int w = 0; w < 2; w++) {
    const double* f = w == 0 ? e0 : b0;
    if (!f) continue;
    for (int m = 0; m < 3; m++) std::fill(v.begin() + m * nc, v.begin() + (m + 1) * nc, (float)f[m]);
    TRY(copy_on(s->stream, w == 0 ? s->E.p : s->B.p, v.data(), 12 * nc, hipMemcpyHostToDevice));
  }
  return restore_walls(s);
}

// particles left after absorbing walls removed some: the sum of the tile counts
static int refresh_counts(osr_sim* s) {
  for (auto& p : s->sp) {
    std::vector<int> cnt((size_t)s->ntiles);
    TRY(copy_on(s->stream, cnt.data(), p->cnt.p, sizeof(int) * cnt.size(), hipMemcpyDeviceToHost));
    int64_t n = 0;
    for (int c : cnt) n += c;
    p->np = n;
  }
  return OSR_OK;
}

int osr_sim_run(osr_sim* s, int nsteps) {
  TRY(check_sim(s));
  if (nsteps < 0) return fail(OSR_ERR_ARG, "negative step count");
  if (nsteps == 0) return OSR_OK;
  TRY(ensure_hist(s, s->steps + nsteps));
  if (!s->graph_ok) TRY(build_graph(s));
  int64_t left = nsteps;
  int batch = 8;
  while (left > 0) {
    int nb = (int)std::min<int64_t>(left, batch);
    if (!s->zp.empty()) {  // a batch never outruns the laser table
      nb = std::min(nb, zp_rows(s, zp_launches(s)));
      TRY(zp_fill(s, s->steps, nb));
    }
    int i = 0;
    for (; i + kGraphSteps <= nb; i += kGraphSteps) HIPC(hipGraphLaunch(s->gexec_k, s->stream));
    for (; i < nb; i++) HIPC(hipGraphLaunch(s->gexec, s->stream));
    Ctrl c;
    TRY(read_ctrl(s, c));
    if (c.error & 1) return fail(OSR_ERR_LAYOUT, "particle found outside its tile");
    if (c.error) return fail(OSR_ERR_STATE, "mover lists exceed the tile fill tables");
    const int64_t done = c.steps - s->steps;
    s->steps = c.steps;
    left -= done;
    if (c.halt == 2) {
      TRY(relayout(s, true));
      Ctrl z = c;
      z.halt = 0;
      TRY(copy_on(s->stream, s->ctrl.p, &z, sizeof(Ctrl), hipMemcpyHostToDevice));
      TRY(build_graph(s));
      batch = 8;
    } else if (done != nb) {
      return fail(OSR_ERR_STATE, "step counter advanced by %lld instead of %d",
                  (long long)done, nb);
    } else {
      batch = std::min(batch * 2, 256);
    }
  }
  if (s->g.open_nx) TRY(refresh_counts(s));
  return OSR_OK;
}

int osr_sim_steps(osr_sim* s, int64_t* steps) {
  TRY(check_sim(s));
  if (!steps) return fail(OSR_ERR_ARG, "null argument");
  *steps = s->steps;
  return OSR_OK;
}

int osr_sim_history(osr_sim* s, int64_t first, int64_t count, double* fld, double* par) {
  TRY(check_sim(s));
  if (first < 0 || count < 0 || first + count > s->steps)
    return fail(OSR_ERR_ARG, "history rows out of range");
  const int ns = (int)s->sp.size(), ncol = 6 + ns;
  std::vector<double> rows((size_t)count * ncol);
  HIPC(hipStreamSynchronize(s->stream));
  if (count)
    TRY(copy_on(s->stream, rows.data(), s->hist.as<double>() + first * ncol,
                   sizeof(double) * rows.size(), hipMemcpyDeviceToHost));
  for (int64_t r = 0; r < count; r++) {
    if (fld)
      for (int c = 0; c < 6; c++) fld[6 * r + c] = rows[r * ncol + c];
    if (par)
      for (int c = 0; c < ns; c++) par[ns * r + c] = rows[r * ncol + 6 + c];
  }
  return OSR_OK;
}

int osr_sim_write_hist(osr_sim* s, const char* dir, const char* const* names, int64_t first,
                       int64_t count) {
  TRY(check_sim(s));
  const int ns = (int)s->sp.size();
  std::vector<double> fld((size_t)(count > 0 ? count : 0) * 6), par((size_t)(count > 0 ? count : 0) * ns);
  TRY(osr_sim_history(s, first, count, fld.data(), par.data()));
  std::vector<int64_t> iter((size_t)count), npar((size_t)ns);
  std::vector<double> t((size_t)count);
  for (int64_t r = 0; r < count; r++) {
    iter[r] = first + r;
    t[r] = (double)(first + r) * s->dt;
  }
  for (int k = 0; k < ns; k++) npar[k] = s->sp[k]->np;
  return osr_hist_append(dir, ns, names, count, iter.data(), t.data(), fld.data(), par.data(),
                         npar.data());
}

int osr_sim_save_field(osr_sim* s, int quant, const char* ms_dir, const double* xmin) {
  TRY(check_sim(s));
  if (quant < 0 || quant > 5 || !ms_dir) return fail(OSR_ERR_ARG, "bad argument");
  const size_t nc = (size_t)s->g.ncell;
  std::vector<float> v(nc);
  HIPC(hipStreamSynchronize(s->stream));
  const DBuf& src = quant < 3 ? s->E : s->B;
  TRY(copy_on(s->stream, v.data(), src.as<float>() + (size_t)(quant % 3) * nc, 4 * nc,
                 hipMemcpyDeviceToHost));
  if (s->g.open_nx) {  // interior columns of the x ring only
    const int nxi = s->g.open_nx, nxr = s->g.nx;
    for (size_t r = 0; r < nc / nxr; r++)
      std::copy(v.begin() + r * nxr, v.begin() + r * nxr + nxi, v.begin() + r * nxi);
  }
  // os-emf-diag-define.f03:164-190: names, labels and units of the field reports
  static const char* qname[6] = {"e1", "e2", "e3", "b1", "b2", "b3"};
  static const char* qlabel[6] = {"E_1", "E_2", "E_3", "B_1", "B_2", "B_3"};
  static const char* xname[3] = {"x1", "x2", "x3"};
  static const char* xlabel[3] = {"x_1", "x_2", "x_3"};
  osr_zdf_grid_axis ax[3];
  osr_zdf_grid_info gi{};
  gi.name = qname[quant];
  gi.ndims = s->dims;
  gi.label = qlabel[quant];
  gi.units = "m_e c \\omega_p e^{-1}";
  const int nxs[3] = {interior_nx(s->g), s->g.ny, s->g.nz};
  for (int d = 0; d < s->dims; d++) {
    gi.count[d] = (uint64_t)nxs[d];
    const double x0 = xmin ? xmin[d] : 0.0;
    ax[d] = osr_zdf_grid_axis{xname[d], 0, x0, x0 + nxs[d] * s->dx[d], xlabel[d], "c / \\omega_p"};
  }
  gi.axis = ax;
  osr_zdf_iteration it{"ITERATION", (int32_t)s->steps, (double)s->steps * s->dt, "1 / \\omega_p"};
  std::string dir = std::string(ms_dir) + "/FLD/" + qname[quant];
  if (osr_out::make_dirs(dir)) return fail(OSR_ERR_ARG, "cannot create directory %s", dir.c_str());
  char fn[64];
  snprintf(fn, sizeof fn, "/%s-%06u.zdf", qname[quant], (unsigned)s->steps);
  return osr_zdf_write_grid_file((dir + fn).c_str(), v.data(), OSR_ZDF_FLOAT32, &gi, &it);
}

int osr_sim_gauss(osr_sim* s, double* res, double* rho_out, double* maxabs) {
  TRY(check_sim(s));
  if (s->g.open_nx) return fail(OSR_ERR_STATE, "the Gauss residual needs periodic x walls");
  const size_t nc = (size_t)s->g.ncell;
  DBuf rho, dres;
  TRY(rho.alloc(8 * nc));
  TRY(dres.alloc(8 * nc));
  HIPC(hipStreamSynchronize(s->stream));
  HIPC(hipMemsetAsync(rho.p, 0, 8 * nc, s->stream));
  double qabs = 0;
  for (auto& p : s->sp) qabs += p->qabs;
  const int sh = 62 - (int)std::ceil(std::log2(std::max(qabs * 1.01, 1e-300)));
  const double scale = std::ldexp(1.0, std::max(0, std::min(sh, 100)));
  for (size_t i = 0; i < s->sp.size(); i++) {
    typedef void (*rf)(const SpecDesc*, const int2*, unsigned long long*, double);
    rf f = s->order == 1 ? rho_deposit<1> : (s->order == 2 ? rho_deposit<2> : rho_deposit<3>);
    hipLaunchKernelGGL(f, dim3(s->ntiles), dim3(256), 0, s->stream, s->d_spec.as<SpecDesc>(),
                       s->d_work.as<int2>() + i * s->ntiles, rho.as<unsigned long long>(), scale);
    HIPC(hipGetLastError());
  }
  hipLaunchKernelGGL(gauss_residual, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s->stream,
                     s->g, rho.as<unsigned long long>(), 1.0 / scale, 1.0 / s->dx[0],
                     1.0 / s->dx[1], s->dims == 3 ? 1.0 / s->dx[2] : 0.0, dres.as<double>());
  HIPC(hipGetLastError());
  std::vector<double> h(nc);
  HIPC(hipStreamSynchronize(s->stream));
  TRY(copy_on(s->stream, h.data(), dres.p, 8 * nc, hipMemcpyDeviceToHost));
  double m = 0;
  for (double v : h) m = std::max(m, std::fabs(v));
  if (maxabs) *maxabs = m;
  if (res) std::memcpy(res, h.data(), 8 * nc);
  if (rho_out) {
    std::vector<long long> r(nc);
    TRY(copy_on(s->stream, r.data(), rho.p, 8 * nc, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nc; i++) rho_out[i] = (double)r[i] / scale;
  }
  return OSR_OK;
}

int osr_sim_init_fields(osr_sim* s, int sid, float ufl1) {
  TRY(check_sim(s));
  if (s->dims != 2 && ufl1 == 0.0f)
    return fail(OSR_ERR_ARG, "the electrostatic solver of a species at rest is 2D only (the "
                             "reference's es_solver has no 3D form)");
  if (sid < 0 || sid >= (int)s->sp.size()) return fail(OSR_ERR_ARG, "bad species id");
  if (s->plain) return fail(OSR_ERR_STATE, "init_fields needs the fused path");
  const size_t nc = (size_t)s->g.ncell;
  const long long nt = (long long)interior_nx(s->g) * s->g.ny * s->g.nz;
  DBuf rho, charge;
  TRY(rho.alloc(8 * nc));
  TRY(charge.alloc(8 * (size_t)nt));
  HIPC(hipStreamSynchronize(s->stream));
  HIPC(hipMemsetAsync(rho.p, 0, 8 * nc, s->stream));
  const Species& p = *s->sp[sid];
  const int sh = 62 - (int)std::ceil(std::log2(std::max(p.qabs * 1.01, 1e-300)));
  const double scale = std::ldexp(1.0, std::max(0, std::min(sh, 100)));
  typedef void (*rf)(const SpecDesc*, const int2*, unsigned long long*, double);
  rf f = s->order == 1 ? rho_deposit<1> : (s->order == 2 ? rho_deposit<2> : rho_deposit<3>);
  hipLaunchKernelGGL(f, dim3(s->ntiles), dim3(256), 0, s->stream, s->d_spec.as<SpecDesc>(),
                     s->d_work.as<int2>() + (size_t)sid * s->ntiles, rho.as<unsigned long long>(),
                     scale);
  const unsigned nb = (unsigned)((nt + 255) / 256);
  hipLaunchKernelGGL(es_charge, dim3(nb), dim3(256), 0, s->stream, s->g,
                     rho.as<unsigned long long>(), 1.0 / scale, charge.as<double>());
  if (ufl1 == 0.0f) {  // es_solver
    hipLaunchKernelGGL(es_direct_2d, dim3(nb), dim3(256), 0, s->stream, s->g, charge.as<double>(),
                       s->dx[0], s->dx[1], s->dx[0] * s->dx[1]);
  } else if (s->dims == 3) {  // es_solver_3d_beam
    const float u1sq = ufl1 * ufl1;
    const double gamma = std::sqrt((double)u1sq + 1.0);
    const double cn = s->dx[1] * s->dx[2] / gamma / 2 / 3.14159265358979323846;
    DBuf e23u;
    TRY(e23u.alloc(8 * (size_t)nt));
    const long long ns = (long long)s->g.ny * s->g.nz;
    hipLaunchKernelGGL(es_beam3d_e, dim3((unsigned)((ns + 255) / 256), (unsigned)interior_nx(s->g)),
                       dim3(256), 0, s->stream, s->g, charge.as<double>(), cn, s->dx[1], s->dx[2],
                       e23u.as<float>());
    hipLaunchKernelGGL(es_beam3d_apply, dim3(nb), dim3(256), 0, s->stream, s->g, e23u.as<float>(),
                       ufl1, gamma);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(s->stream));
  } else {  // es_solver_beam: gamma from u1^2 in the particle precision
    const float u1sq = ufl1 * ufl1;
    const double gamma = std::sqrt((double)u1sq + 1.0);
    const double cn = s->dx[1] / gamma / 2;
    DBuf e2u;
    TRY(e2u.alloc(4 * (size_t)nt));
    hipLaunchKernelGGL(es_beam_e2, dim3(nb), dim3(256), 0, s->stream, s->g, charge.as<double>(),
                       cn, e2u.as<float>());
    hipLaunchKernelGGL(es_beam_apply, dim3(nb), dim3(256), 0, s->stream, s->g, e2u.as<float>(),
                       ufl1, gamma);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(s->stream));
  }
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(s->stream));
  return OSR_OK;
}

int osr_sim_relayouts(osr_sim* s, int64_t* count) {
  TRY(check_sim(s));
  if (!count) return fail(OSR_ERR_ARG, "null argument");
  *count = s->relayouts;
  return OSR_OK;
}

int osr_sim_footprint(osr_sim* s, int64_t* bytes) {
  TRY(check_sim(s));
  if (!bytes) return fail(OSR_ERR_ARG, "null argument");
  size_t b = s->E.n + s->B.n + s->J.n + s->fe_part.n + s->hist.n;
  for (auto& p : s->sp) {
    const DBuf* all[] = {&p->cell, &p->x, &p->y, &p->z, &p->ux, &p->uy, &p->uz, &p->q, &p->pid,
                         &p->mcell, &p->mx, &p->my, &p->mz, &p->mux, &p->muy, &p->muz, &p->mq,
                         &p->mpid, &p->midx, &p->off, &p->cnt, &p->cap, &p->mv_cnt,
                         &p->mv_stride, &p->newcnt, &p->inc, &p->epart};
    for (const DBuf* d : all) b += d->n;
  }
  *bytes = (int64_t)b;
  return OSR_OK;
}

int osr_sim_launch_config(osr_sim* s, int* tile, int* waves, int* compiled_tile) {
  TRY(check_sim(s));
  if (!tile || !waves || !compiled_tile) return fail(OSR_ERR_ARG, "null argument");
  tile[0] = s->g.tx;
  tile[1] = s->g.ty;
  tile[2] = s->g.tz;
  *waves = s->push_waves;
  *compiled_tile = 0;
  if (!s->plain && s->desc.kernel != 1)
    *compiled_tile = (s->dims == 2 ? pick_push2d_t(s->order, 0, s->g.tx, s->g.ty, s->push_waves)
                                   : pick_push3d_t(s->order, 0, s->g.tx, s->g.ty, s->g.tz)) != nullptr;
  return OSR_OK;
}

int osr_sim_time_kernels(osr_sim* s, int nsteps, double* ms) {
  TRY(check_sim(s));
  if (nsteps < 1 || !ms) return fail(OSR_ERR_ARG, "bad argument");
  TRY(ensure_hist(s, s->steps + nsteps));
  const int nev = 6;
  hipEvent_t ev[nev];
  for (int i = 0; i < nev; i++) HIPC(hipEventCreate(&ev[i]));
  double acc[6] = {0, 0, 0, 0, 0, 0};
  int done = 0;
  for (int k = 0; k < nsteps; k++) {
    if (!s->zp.empty()) {  // the laser walls of the step, not timed
      TRY(zp_fill(s, s->steps, 1));
      TRY(launch_laser(s));
    }
    HIPC(hipEventRecord(ev[0], s->stream));
    if (s->plain)
      for (size_t i = 0; i < s->sp.size(); i++) TRY(launch_plain(s, (int)i));
    else
      TRY(launch_pushes(s));
    HIPC(hipEventRecord(ev[1], s->stream));
    if (!s->sp.empty() && !s->plain) {
      TRY(launch_fill(s));
      TRY(launch_reorder(s));
    }
    HIPC(hipEventRecord(ev[2], s->stream));
    TRY(launch_smooth(s));
    TRY(launch_yee(s));
    HIPC(hipEventRecord(ev[3], s->stream));
    TRY(launch_finalize(s));
    HIPC(hipEventRecord(ev[4], s->stream));
    HIPC(hipEventSynchronize(ev[4]));
    Ctrl c;
    TRY(read_ctrl(s, c));
    if (c.error & 1) return fail(OSR_ERR_LAYOUT, "particle found outside its tile");
    if (c.error) return fail(OSR_ERR_STATE, "mover lists exceed the tile fill tables");
    s->steps = c.steps;
    if (c.halt == 2) {
      TRY(relayout(s, true));
      Ctrl z = c;
      z.halt = 0;
      TRY(copy_on(s->stream, s->ctrl.p, &z, sizeof(Ctrl), hipMemcpyHostToDevice));
      continue;  // a re-laid-out step is not representative
    }
    float t;
    for (int i = 0; i < 4; i++) {
      HIPC(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
      acc[i] += t;
    }
    HIPC(hipEventElapsedTime(&t, ev[0], ev[4]));
    acc[4] += t;
    done++;
  }
  for (int i = 0; i < nev; i++) (void)hipEventDestroy(ev[i]);
  for (int i = 0; i < 6; i++) ms[i] = done ? acc[i] / done : 0.0;
  if (s->g.open_nx) TRY(refresh_counts(s));
  return OSR_OK;
}

__global__ __launch_bounds__(256) void copy_kernel(const float4* __restrict__ src,
                                                   float4* __restrict__ dst, long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256)
    dst[i] = src[i];
}

int osr_copy_probe(int64_t bytes, int reps, double* ms) {
  if (bytes < 16 || reps < 1 || !ms) return fail(OSR_ERR_ARG, "bad argument");
  const long long n = bytes / 16;
  DBuf a, b;
  TRY(a.alloc((size_t)n * 16));
  TRY(b.alloc((size_t)n * 16));
  HIPC(hipMemset(a.p, 0, (size_t)n * 16));
  HIPC(hipDeviceSynchronize());
  hipEvent_t e0, e1;
  HIPC(hipEventCreate(&e0));
  HIPC(hipEventCreate(&e1));
  const int grid = 256 * 64;
  hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, 0, a.as<float4>(), b.as<float4>(), n);
  HIPC(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, 0, a.as<float4>(), b.as<float4>(), n);
  HIPC(hipEventRecord(e1, 0));
  HIPC(hipEventSynchronize(e1));
  float t = 0;
  HIPC(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = t / reps;
  return OSR_OK;
}

// ------------------------------------------------------------------------------------
// stateless drop-ins
// ------------------------------------------------------------------------------------
namespace {
struct Pool {
  DBuf b[8];
  int get(int i, size_t bytes) {
    if (b[i].n >= bytes && b[i].p) return OSR_OK;
    return b[i].alloc(bytes);
  }
};
thread_local Pool g_pool;
}  // namespace

int osr_advance_deposit(int dims, int order, int pusher, int* ix, float* x, float* u,
                        const float* q, int i0, int i1, float rqm, const float* efield,
                        const float* bfield, const int* emf_size, const int* emf_offset,
                        float* j, const int* j_size, const int* j_offset, const double* dx,
                        double dt, double* ene) {
  if (dims != 2 && dims != 3) return fail(OSR_ERR_ARG, "dims must be 2 or 3");
  if (order < 1 || order > 3) return fail(OSR_ERR_ARG, "order must be 1..3");
  if (pusher != OSR_BORIS && pusher != OSR_VAY) return fail(OSR_ERR_ARG, "unknown pusher");
  if (!ix || !x || !u || !q || !efield || !bfield || !emf_size || !emf_offset || !j || !j_size ||
      !j_offset || !dx || !ene)
    return fail(OSR_ERR_ARG, "null argument");
  if (i0 < 1 || i1 < i0 - 1) return fail(OSR_ERR_ARG, "bad particle range");
  const int np = i1 - i0 + 1;
  const int lp = order == 1 ? 0 : -1;
  size_t ne = 3, nj = 3;
  for (int d = 0; d < dims; d++) {
    ne *= (size_t)emf_size[d];
    nj *= (size_t)j_size[d];
  }
  // range check of cell indices on the host: gather reaches ix-1+lp .. ix+lp+order,
  // the deposit ix-1+lp .. ix+1+lp+order
  const int* pix = ix + (size_t)dims * (i0 - 1);
  for (int i = 0; i < np; i++)
    for (int d = 0; d < dims; d++) {
      const int c = pix[(size_t)dims * i + d];
      const int glo = c - 1 + lp, ghi = c + lp + order, jlo = c - 1 + lp, jhi = c + 1 + lp + order;
      if (glo + emf_offset[d] < 1 || ghi + emf_offset[d] > emf_size[d] || jlo + j_offset[d] < 1 ||
          jhi + j_offset[d] > j_size[d])
        return fail(OSR_ERR_ARG, "particle %d stencil leaves the guarded arrays", i0 + i);
    }
  Pool& P = g_pool;
  TRY(P.get(0, 4 * (size_t)dims * np + 4));
  TRY(P.get(1, 4 * (size_t)dims * np + 4));
  TRY(P.get(2, 12 * (size_t)np + 4));
  TRY(P.get(3, 4 * (size_t)np + 4));
  TRY(P.get(4, 4 * ne));
  TRY(P.get(5, 4 * ne));
  TRY(P.get(6, 8 * nj));
  const int nblk = (np + 255) / 256;
  TRY(P.get(7, 8 * (size_t)(nblk > 0 ? nblk : 1) + 4 * nj));
  float* dj = (float*)((char*)P.b[7].p + 8 * (size_t)(nblk > 0 ? nblk : 1));
  if (np > 0) {
    HIPC(hipMemcpy(P.b[0].p, pix, 4 * (size_t)dims * np, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(P.b[1].p, x + (size_t)dims * (i0 - 1), 4 * (size_t)dims * np, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(P.b[2].p, u + (size_t)3 * (i0 - 1), 12 * (size_t)np, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(P.b[3].p, q + (i0 - 1), 4 * (size_t)np, hipMemcpyHostToDevice));
  }
  HIPC(hipMemcpy(P.b[4].p, efield, 4 * ne, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(P.b[5].p, bfield, 4 * ne, hipMemcpyHostToDevice));
  HIPC(hipMemset(P.b[6].p, 0, 8 * nj));
  HIPC(hipMemcpy(dj, j, 4 * nj, hipMemcpyHostToDevice));
  DropinArgs a{};
  a.ix = P.b[0].as<int>();
  a.x = P.b[1].as<float>();
  a.u = P.b[2].as<float>();
  a.q = P.b[3].as<float>();
  a.np = np;
  a.e = P.b[4].as<float>();
  a.b = P.b[5].as<float>();
  for (int d = 0; d < 3; d++) {
    a.emf_ld[d] = d < dims ? emf_size[d] : 1;
    a.emf_off[d] = d < dims ? emf_offset[d] : 0;
    a.j_ld[d] = d < dims ? j_size[d] : 1;
    a.j_off[d] = d < dims ? j_offset[d] : 0;
    a.dt_dx[d] = d < dims ? (float)(dt / dx[d]) : 0.f;
    a.jnorm[d] = d < dims ? (float)(dx[d] / dt / (dims == 2 ? 2.0 : 3.0)) : 0.f;
  }
  a.jfix = P.b[6].as<unsigned long long>();
  a.tem = (float)(0.5 * dt / (double)rqm);
  // fixed-point scale from this call's charges (same bound as the resident path)
  double qabs = 0, qmax = 0;
  for (int i = 0; i < np; i++) {
    qabs += std::fabs((double)q[i0 - 1 + i]);
    qmax = std::max(qmax, std::fabs((double)q[i0 - 1 + i]));
  }
  double jn = 0;
  for (int d = 0; d < dims; d++) jn += a.jnorm[d];
  const double per = (dims == 2 ? 2.0 : 3.0) * jn + 1.0;
  int sh = std::min(62 - (int)std::ceil(std::log2(std::max(qabs * per * 8.0, 1e-300))),
                    50 - (int)std::ceil(std::log2(std::max(qmax * per * 2.0, 1e-300))));
  sh = std::max(0, std::min(sh, 80));
  a.jscale = std::ldexp(1.0, sh);
  a.epart = P.b[7].as<double>();
  if (np > 0) {
    typedef void (*df)(DropinArgs);
    df f = nullptr;
#define OSR_D(D, O, PP) \
  if (dims == D && order == O && pusher == PP) f = dropin_push<O, PP, D>;
    OSR_D(2, 1, 0) OSR_D(2, 2, 0) OSR_D(2, 3, 0) OSR_D(2, 1, 1) OSR_D(2, 2, 1) OSR_D(2, 3, 1)
    OSR_D(3, 1, 0) OSR_D(3, 2, 0) OSR_D(3, 3, 0) OSR_D(3, 1, 1) OSR_D(3, 2, 1) OSR_D(3, 3, 1)
#undef OSR_D
    hipLaunchKernelGGL(f, dim3(nblk), dim3(256), 0, 0, a);
    HIPC(hipGetLastError());
  }
  hipLaunchKernelGGL(dropin_jadd, dim3((unsigned)((nj + 255) / 256)), dim3(256), 0, 0, dj,
                     a.jfix, (long long)nj, std::ldexp(1.0, -sh));
  HIPC(hipGetLastError());
  HIPC(hipDeviceSynchronize());
  if (np > 0) {
    HIPC(hipMemcpy(ix + (size_t)dims * (i0 - 1), P.b[0].p, 4 * (size_t)dims * np, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(x + (size_t)dims * (i0 - 1), P.b[1].p, 4 * (size_t)dims * np, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(u + (size_t)3 * (i0 - 1), P.b[2].p, 12 * (size_t)np, hipMemcpyDeviceToHost));
    std::vector<double> ep((size_t)nblk);
    HIPC(hipMemcpy(ep.data(), P.b[7].p, 8 * (size_t)nblk, hipMemcpyDeviceToHost));
    double e = 0;
    for (double v : ep) e += v;
    *ene += e;
  }
  HIPC(hipMemcpy(j, dj, 4 * nj, hipMemcpyDeviceToHost));
  return OSR_OK;
}

#define OSR_FORTRAN_DROPIN(D, O)                                                            \
  void hip_vadvance_deposit_##D##d_s##O##_(int* ix, float* x, float* u, float* q, int* i0, \
                                           int* i1, float* rqm, float* efield, float* bfield, \
                                           int* emf_size, int* emf_offset, float* j,          \
                                           int* j_size, int* j_offset, double* dx, double* dt, \
                                           double* ene) {                                     \
    int r = osr_advance_deposit(D, O, OSR_BORIS, ix, x, u, q, *i0, *i1, *rqm, efield, bfield, \
                                emf_size, emf_offset, j, j_size, j_offset, dx, *dt, ene);     \
    if (r != OSR_OK) {                                                                      \
      printf("(*error*) hip_vadvance_deposit_%dd_s%d: %s\n", D, O, osr_last_error());        \
      exit(1);                                                                              \
    }                                                                                       \
  }
OSR_FORTRAN_DROPIN(2, 1)
OSR_FORTRAN_DROPIN(2, 2)
OSR_FORTRAN_DROPIN(2, 3)
OSR_FORTRAN_DROPIN(3, 1)
OSR_FORTRAN_DROPIN(3, 2)
OSR_FORTRAN_DROPIN(3, 3)

int osr_yee_advance(int dims, const int* nx, const int* ld, const int* gcl, const double* dx,
                    double dt, float* e, float* b, const float* jay) {
  if ((dims != 2 && dims != 3) || !nx || !ld || !gcl || !dx || !e || !b || !jay)
    return fail(OSR_ERR_ARG, "bad argument");
  const int nxs[3] = {nx[0], nx[1], dims == 3 ? nx[2] : 1};
  for (int d = 0; d < dims; d++)
    if (nx[d] < 1 || gcl[d] < 1 || ld[d] < nx[d] + gcl[d] + 2)
      return fail(OSR_ERR_ARG, "Yee needs 1 lower and 2 upper guard cells");
  const size_t nc = (size_t)nxs[0] * nxs[1] * nxs[2];
  if (nc >= (1ull << 32)) return fail(OSR_ERR_ARG, "grid of 2^32 cells or more");
  std::vector<float> pe(3 * nc), pb(3 * nc), pj(3 * nc);
  const size_t ld0 = ld[0], ld1 = ld[1];
  auto ref = [&](int i, int j, int k) {
    return 3 * ((size_t)(i + gcl[0]) + ld0 * ((size_t)(j + gcl[1]) + (dims == 3 ? ld1 * (size_t)(k + gcl[2]) : 0)));
  };
  for (int k = 0; k < nxs[2]; k++)
    for (int j = 0; j < nxs[1]; j++)
      for (int i = 0; i < nxs[0]; i++) {
        const size_t c = (size_t)i + (size_t)nxs[0] * (j + (size_t)nxs[1] * k), r = ref(i, j, k);
        for (int m = 0; m < 3; m++) {
          pe[m * nc + c] = e[r + m];
          pb[m * nc + c] = b[r + m];
          pj[m * nc + c] = jay[r + m];
        }
      }
  DBuf dE, dB, dJ, dctrl, dfd, dw, dfe;
  TRY(dE.alloc(12 * nc));
  TRY(dB.alloc(12 * nc));
  TRY(dJ.alloc(12 * nc));
  TRY(dctrl.alloc(sizeof(Ctrl)));
  TRY(dfd.alloc(sizeof(FieldDesc)));
  const int nblk = (int)((nc + 255) / 256);
  TRY(dw.alloc(sizeof(int2) * (size_t)nblk));
  TRY(dfe.alloc(48 * (size_t)nblk));
  HIPC(hipMemcpy(dE.p, pe.data(), 12 * nc, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(dB.p, pb.data(), 12 * nc, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(dJ.p, pj.data(), 12 * nc, hipMemcpyHostToDevice));
  HIPC(hipMemset(dctrl.p, 0, sizeof(Ctrl)));
  FieldDesc f{};
  f.g.dims = dims;
  f.g.nx = nxs[0];
  f.g.ny = nxs[1];
  f.g.nz = nxs[2];
  f.g.ncell = (long long)nc;
  set_grid_divisors(f.g);
  f.g.E = dE.as<float>();
  f.g.B = dB.as<float>();
  f.g.ctrl = dctrl.as<Ctrl>();
  f.Jf = dJ.as<float>();
  f.fe_part = dfe.as<double>();
  f.nblk = nblk;
  for (int d = 0; d < 3; d++) {
    f.bdt[d] = d < dims ? (float)((dt / 2.0) / dx[d]) : 0.f;
    f.edt[d] = d < dims ? (float)(dt / dx[d]) : 0.f;
  }
  f.dtif = (float)dt;
  HIPC(hipMemcpy(dfd.p, &f, sizeof(FieldDesc), hipMemcpyHostToDevice));
  std::vector<int2> w((size_t)nblk);
  for (int i = 0; i < nblk; i++) w[i] = make_int2(0, i);
  HIPC(hipMemcpy(dw.p, w.data(), sizeof(int2) * w.size(), hipMemcpyHostToDevice));
  const FieldDesc* fd = dfd.as<FieldDesc>();
  const int2* wk = dw.as<int2>();
  if (dims == 2) {
    hipLaunchKernelGGL((yee_dbdt<2, false>), dim3(nblk), dim3(256), 0, 0, fd, wk);
    hipLaunchKernelGGL((yee_dedt<2>), dim3(nblk), dim3(256), 0, 0, fd, wk);
    hipLaunchKernelGGL((yee_dbdt<2, false>), dim3(nblk), dim3(256), 0, 0, fd, wk);
  } else {
    hipLaunchKernelGGL((yee_dbdt<3, false>), dim3(nblk), dim3(256), 0, 0, fd, wk);
    hipLaunchKernelGGL((yee_dedt<3>), dim3(nblk), dim3(256), 0, 0, fd, wk);
    hipLaunchKernelGGL((yee_dbdt<3, false>), dim3(nblk), dim3(256), 0, 0, fd, wk);
  }
  HIPC(hipGetLastError());
  HIPC(hipDeviceSynchronize());
  HIPC(hipMemcpy(pe.data(), dE.p, 12 * nc, hipMemcpyDeviceToHost));
  HIPC(hipMemcpy(pb.data(), dB.p, 12 * nc, hipMemcpyDeviceToHost));
  // write back interior and periodic images into every guard cell
  const int ldk = dims == 3 ? ld[2] : 1;
  for (int k = 0; k < ldk; k++)
    for (int j = 0; j < ld[1]; j++)
      for (int i = 0; i < ld[0]; i++) {
        const int ii = ((i - gcl[0]) % nxs[0] + nxs[0]) % nxs[0];
        const int jj = ((j - gcl[1]) % nxs[1] + nxs[1]) % nxs[1];
        const int kk = dims == 3 ? ((k - gcl[2]) % nxs[2] + nxs[2]) % nxs[2] : 0;
        const size_t c = (size_t)ii + (size_t)nxs[0] * (jj + (size_t)nxs[1] * kk);
        const size_t r = 3 * ((size_t)i + ld0 * (j + ld1 * (size_t)k));
        for (int m = 0; m < 3; m++) {
          e[r + m] = pe[m * nc + c];
          b[r + m] = pb[m * nc + c];
        }
      }
  return OSR_OK;
}

}  // extern "C"

// ---- current smoothing -----------------------------------------------------------------
namespace {

// besseli0_rdoub, os-math.f90:450-472 (Abramowitz & Stegun 9.8.1/9.8.2).  Literals without
// a kind suffix in the reference are default (single precision) reals: (double)<float>.
double besseli0(double x) {
  const double ax = std::fabs(x);
  if (ax < (double)3.75f) {
    const double y = (x / (double)3.75f) * (x / (double)3.75f);
    return (double)1.0f +
           y * ((double)3.5156229f +
                y * ((double)3.0899424f +
                     y * ((double)1.2067492f + y * ((double)0.2659732f + y * (0.360768e-1 + y * 0.45813e-2)))));
  }
  const double y = (double)3.75f / ax;
  return (std::exp(ax) / std::sqrt(ax)) *
         ((double)0.39894228f +
          y * (0.1328592e-1 +
               y * (0.225319e-2 +
                    y * (-0.157565e-2 +
                         y * (0.916281e-2 +
                              y * (-0.2057706e-1 +
                                   y * (0.2635537e-1 + y * (-0.1647633e-1 + y * 0.392377e-2))))))));
}

// combine_kernel3, os-vdf-smooth.f90:142-172: n 3-point kernels k3[j] = kern3(1:3, j+1) into
// one kernel of 2 nmax + 1 points (double); the first entry of a 3-point kernel weighs the
// upper neighbour, the third the lower one.
void combine_kernel3(const double (*k3)[3], int n, int nmax, double* out) {
  for (int i = 0; i < 2 * nmax + 1; i++) out[i] = 0.0;
  out[nmax] = 1.0;
  for (int k = 1; k <= n; k++) {
    const double a = k3[k - 1][0], b = k3[k - 1][1], c = k3[k - 1][2];
    double f1 = 0.0;
    for (int i = nmax - (k - 1); i <= nmax + k; i++) {
      out[i - 1] = c * out[i] + out[i - 1];
      const double f2 = a * out[i];
      out[i] = b * out[i] + f1;
      f1 = f2;
    }
  }
}

constexpr double kPi = 3.14159265358979323846264;  // os-math.f90:13

// digital_filter, os-vdf-smooth.f90:110-137: Kaiser-windowed sinc of 2n+1 points
void digital_filter(double fc, double A, int n, double* out) {
  double alpha;
  if (A <= 21.0)
    alpha = 0.0;
  else if (A >= 50.0)
    alpha = 0.1102 * (A - 8.7);
  else
    alpha = 0.5842 * std::pow(A - 21.0, 0.4) + 0.07886 * (A - 21.0);
  for (int i = 1; i <= n; i++) {
    const float r = (float)i / (float)n;
    out[n + i] = besseli0(alpha * std::sqrt(1.0 - (double)(r * r))) / besseli0(alpha) *
                 std::sin(i * kPi * fc) / (i * kPi);
  }
  out[n] = fc;
  for (int i = 1; i <= n; i++) out[n - i] = out[n + i];
}


int check_smooth_args(int dims, const int* order, const float* coef) {
  if ((dims != 2 && dims != 3) || !order || !coef) return fail(OSR_ERR_ARG, "bad argument");
  for (int d = 0; d < dims; d++)
    if (order[d] < 0 || order[d] > OSR_SMOOTH_ORDER_MAX)
      return fail(OSR_ERR_ARG, "smoothing order must be between 0 and %d", OSR_SMOOTH_ORDER_MAX);
  return OSR_OK;
}

}  // namespace

extern "C" {

// read_nml_smooth (os-vdf-smooth.f90:183-361) + setup_smooth (:366-470) in the reference's
// single-precision build: scoef is an FP32 array, the normalisation sums those values in
// double and rounds each quotient to FP32.
int osr_smooth_setup(int dims, const osr_smooth_params* p, const double* dx, double gamma,
                     int* order_out, float* coef_out) {
  if ((dims < 1 || dims > 3) || !p || !dx || !order_out || !coef_out)
    return fail(OSR_ERR_ARG, "bad argument");
  constexpr int OM = OSR_SMOOTH_ORDER_MAX;
  int type[3] = {OSR_SMOOTH_NONE, OSR_SMOOTH_NONE, OSR_SMOOTH_NONE};
  int order[3] = {0, 0, 0};
  const bool lkdx = p->type[0] == OSR_SMOOTH_LASERKDX;
  if (lkdx) {
    if (p->laserkdx_dir < 1 || p->laserkdx_dir > dims)
      return fail(OSR_ERR_ARG, "laserkdx_dir must be between 1 and %d", dims);
    for (int i = 0; i < dims; i++) {
      type[i] = OSR_SMOOTH_LASERKDX;
      order[i] = i + 1 == p->laserkdx_dir ? 2 : 0;
    }
    if (p->laserkdx_omega0 <= 0.0) return fail(OSR_ERR_ARG, "laserkdx_omega0 must be > 0");
  }
  for (int i = 0; i < dims; i++) {
    switch (p->type[i]) {
      case OSR_SMOOTH_NONE:
        type[i] = OSR_SMOOTH_NONE;
        order[i] = 0;
        break;
      case OSR_SMOOTH_CUSTOM:
      case OSR_SMOOTH_BINOMIAL:
        type[i] = p->type[i];
        if (p->order[i] < 1 || p->order[i] > OM)
          return fail(OSR_ERR_ARG, "smoothing order must be between 1 and %d", OM);
        order[i] = p->order[i];
        break;
      case OSR_SMOOTH_COMPENSATED:
        type[i] = p->type[i];
        order[i] = p->order[i] < 2 ? 2 : p->order[i];
        if (order[i] > OM) return fail(OSR_ERR_ARG, "smoothing order above %d", OM);
        break;
      case OSR_SMOOTH_5PASS:
        type[i] = p->type[i];
        order[i] = 5;
        break;
      case OSR_SMOOTH_TRI6:
        type[i] = p->type[i];
        order[i] = 3;
        break;
      case OSR_SMOOTH_TRI8:
        type[i] = p->type[i];
        order[i] = 4;
        break;
      case OSR_SMOOTH_LASERKDX:
        break;
      case OSR_SMOOTH_DIGITAL:
        type[i] = p->type[i];
        if (p->order[i] < 1 || p->order[i] > OM)
          return fail(OSR_ERR_ARG, "digital smoothing order must be between 1 and %d", OM);
        order[i] = p->order[i];
        break;
      default:
        return fail(OSR_ERR_ARG, "invalid smoothing type %d", p->type[i]);
    }
  }
  int mo = 0;
  for (int i = 0; i < dims; i++) mo = std::max(mo, order[i]);
  float sc[3][2 * OM + 1] = {};
  double k[2 * OM + 1];
  double k3[OM][3];
  if (mo > 0) {
    for (int i = 0; i < dims; i++)
      if (type[i] == OSR_SMOOTH_CUSTOM) {
        combine_kernel3(p->swfj[i], order[i], mo, k);
        for (int j = 0; j < 2 * mo + 1; j++) sc[i][j] = (float)k[j];
      }
    if (type[0] == OSR_SMOOTH_LASERKDX) {
      const int d = p->laserkdx_dir - 1;
      const double lk = p->laserkdx_omega0 * dx[d] / (gamma + std::sqrt(gamma * gamma - 1));
      const double kl[2][3] = {{1.0, 2.0, 1.0}, {-1.0, 4 + 2 * std::cos(lk), -1.0}};
      combine_kernel3(kl, 2, mo, k);
      for (int j = 0; j < 2 * mo + 1; j++) sc[d][j] = (float)k[j];
    } else {
      for (int i = 0; i < dims; i++) {
        const int n = order[i];
        const int c = mo;
        switch (type[i]) {
          case OSR_SMOOTH_BINOMIAL:
            for (int j = 0; j < n; j++) k3[j][0] = 1.0, k3[j][1] = 2.0, k3[j][2] = 1.0;
            combine_kernel3(k3, n, mo, k);
            for (int j = 0; j < 2 * mo + 1; j++) sc[i][j] = (float)k[j];
            break;
          case OSR_SMOOTH_COMPENSATED: {
            for (int j = 0; j < n - 1; j++) k3[j][0] = 1.0, k3[j][1] = 2.0, k3[j][2] = 1.0;
            // (4.0 + 2.0 * (order - 1)) / (order - 1) in default (single) precision
            const double comp = (double)((4.0f + 2.0f * (float)(n - 1)) / (float)(n - 1));
            k3[n - 1][0] = -1.0, k3[n - 1][1] = comp, k3[n - 1][2] = -1.0;
            combine_kernel3(k3, n, mo, k);
            for (int j = 0; j < 2 * mo + 1; j++) sc[i][j] = (float)k[j];
            break;
          }
          case OSR_SMOOTH_5PASS:
            for (int j = 0; j < 4; j++) k3[j][0] = 1.0, k3[j][1] = 2.0, k3[j][2] = 1.0;
            k3[4][0] = -5.0, k3[4][1] = 15.0, k3[4][2] = -5.0;
            combine_kernel3(k3, 5, mo, k);
            for (int j = 0; j < 2 * mo + 1; j++) sc[i][j] = (float)k[j];
            break;
          case OSR_SMOOTH_TRI6: {
            const float t[7] = {0.015625f, -0.09375f, 0.234375f, 0.6875f, 0.234375f, -0.09375f, 0.015625f};
            for (int j = 0; j < 7; j++) sc[i][c - 3 + j] = t[j];
            break;
          }
          case OSR_SMOOTH_TRI8: {
            const float t[9] = {-0.00390625f, 0.03125f, -0.109375f, 0.21875f, 0.7265625f,
                                0.21875f, -0.109375f, 0.03125f, -0.00390625f};
            for (int j = 0; j < 9; j++) sc[i][c - 4 + j] = t[j];
            break;
          }
          case OSR_SMOOTH_DIGITAL:
            digital_filter(p->digital_fc[i], p->digital_A[i], n, k);
            for (int j = 0; j < 2 * n + 1; j++) sc[i][c - n + j] = (float)k[j];
            break;
          default:
            break;
        }
      }
    }
    for (int i = 0; i < dims; i++) {
      const int n = order[i];
      if (n <= 0) continue;
      double total = 0.0;
      for (int j = mo - n; j <= mo + n; j++) total = total + (double)sc[i][j];
      for (int j = mo - n; j <= mo + n; j++) sc[i][j] = (float)((double)sc[i][j] / total);
    }
  }
  for (int i = 0; i < 3; i++) {
    order_out[i] = i < dims ? order[i] : 0;
    for (int j = 0; j < OSR_SMOOTH_NCOEF; j++) coef_out[i * OSR_SMOOTH_NCOEF + j] = 0.f;
    if (i < dims && order[i] > 0)
      for (int j = -order[i]; j <= order[i]; j++) coef_out[i * OSR_SMOOTH_NCOEF + OM + j] = sc[i][mo + j];
  }
  return OSR_OK;
}

int osr_smooth_vdf(int dims, const int* ld, int fdim, const int* order, const float* coef,
                   float* f) {
  TRY(check_smooth_args(dims, order, coef));
  if (!ld || !f || fdim < 1) return fail(OSR_ERR_ARG, "bad argument");
  const int l[3] = {ld[0], ld[1], dims == 3 ? ld[2] : 1};
  for (int d = 0; d < 3; d++)
    if (l[d] < 1) return fail(OSR_ERR_ARG, "bad extents");
  const size_t n = (size_t)fdim * l[0] * l[1] * l[2];
  DBuf a, b;
  TRY(a.alloc(4 * n));
  TRY(b.alloc(4 * n));
  HIPC(hipMemcpy(a.p, f, 4 * n, hipMemcpyHostToDevice));
  const unsigned nb = (unsigned)((n + 255) / 256);
  for (int d = 0; d < dims; d++) {
    if (order[d] <= 0) continue;
    SmoothCoef c;
    for (int j = 0; j < 31; j++) c.s[j] = coef[d * OSR_SMOOTH_NCOEF + j];
    hipLaunchKernelGGL(smooth_vdf_pass, dim3(nb), dim3(256), 0, 0, a.as<float>(), b.as<float>(),
                       fdim, l[0], l[1], l[2], d, order[d], c);
    HIPC(hipGetLastError());
    a.swap(b);
  }
  HIPC(hipDeviceSynchronize());
  HIPC(hipMemcpy(f, a.p, 4 * n, hipMemcpyDeviceToHost));
  return OSR_OK;
}

void osr_zpulse_wall_defaults(osr_zpulse_wall* zp) {
  if (!zp) return;
  *zp = osr_zpulse_wall{};
  zp->a0 = 1.0;
  zp->omega0 = 10.0;
  zp->phase = 0.0;
  zp->pol = 90.0;
  zp->pol_type = 0;
  zp->propagation = OSR_PROP_FORWARD;
  zp->tenv_type = OSR_TENV_POLYNOMIAL;
  zp->per_type = OSR_PER_PLANE;
  zp->tenv_fwhm = -1.0;
  zp->tenv_range = OSR_ZP_UNSET;
  zp->per_center[0] = zp->per_center[1] = OSR_ZP_UNSET;
  zp->per_focus = OSR_ZP_UNSET;
}

int osr_sim_add_zpulse_wall(osr_sim* s, const osr_zpulse_wall* zp) {
  TRY(check_sim(s));
  if (!zp) return fail(OSR_ERR_ARG, "null argument");
  if (!s->g.open_nx)
    return fail(OSR_ERR_STATE, "the launch wall must be open (xbnd = OSR_BND_LINDMAN or OSR_BND_VPML)");
  if ((int)s->zp.size() >= osr_zp::kMaxPulses) return fail(OSR_ERR_ARG, "at most 8 laser walls");
  osr_zp::Wall w;
  const int nx[3] = {s->g.open_nx, s->g.ny, s->g.nz};
  const char* err = osr_zp::resolve(*zp, s->dims, nx, s->dx, w);
  if (err) return fail(OSR_ERR_ARG, "zpulse_wall: %s", err);
  HIPC(hipStreamSynchronize(s->stream));
  s->zp.push_back(w);
  const int K = zp_launches(s), rows = zp_rows(s, K);
  TRY(s->zp_tab.alloc(sizeof(float) * (size_t)rows * K * 2 * s->g.ny * s->g.nz));
  TRY(s->zp_mask.alloc(sizeof(int) * (size_t)rows * K));
  TRY(s->zp_first.alloc(sizeof(long long)));
  HIPC(hipMemsetAsync(s->zp_mask.p, 0, sizeof(int) * (size_t)rows * K, s->stream));
  const long long none = -(1ll << 62);  // no table yet: every step falls outside it
  TRY(copy_on(s->stream, s->zp_first.p, &none, sizeof(none), hipMemcpyHostToDevice));
  TRY(upload_descs(s));
  s->graph_ok = false;
  return OSR_OK;
}

void osr_zpulse_defaults(osr_zpulse* zp) {
  if (!zp) return;
  *zp = osr_zpulse{};
  zp->a0 = 1.0;
  zp->omega0 = 10.0;
  zp->phase = 0.0;
  zp->pol = 90.0;
  zp->pol_type = 0;
  zp->propagation = OSR_PROP_FORWARD;
  zp->lon_type = OSR_TENV_POLYNOMIAL;
  zp->per_type = OSR_PER_PLANE;
  zp->lon_start = OSR_ZP_UNSET;
  zp->lon_fwhm = -1.0;
  zp->lon_x0 = zp->lon_range = OSR_ZP_UNSET;
  zp->per_center[0] = zp->per_center[1] = OSR_ZP_UNSET;
  zp->per_focus = OSR_ZP_UNSET;
}

// launch_zpulse (os-zpulse-std.f03:2350-2469) of an in-box pulse: built on the host in the
// reference's guarded layout (osr_zp::box_build) and added to E and B, cell by cell in
// FP32 (add(emf%b, b_pulse), add(emf%e, e_pulse)); a circular pulse adds its second half
// after the first.  Interior cells always; with open x walls also the x guard cells, which
// the ring's guard band holds.
int osr_sim_add_zpulse(osr_sim* s, const osr_zpulse* zp) {
  TRY(check_sim(s));
  if (!zp) return fail(OSR_ERR_ARG, "null argument");
  const int nxi = interior_nx(s->g);
  const int nx[3] = {nxi, s->g.ny, s->g.nz};
  osr_zp::Box w;
  const char* err = osr_zp::resolve_box(*zp, s->dims, nx, s->dx, w);
  if (err) return fail(OSR_ERR_ARG, "zpulse: %s", err);
  // guard cells of the reference's fields (get_min_gc_part, particles/os-particles.f03:
  // 321-360): linear (1, 2), quadratic (2, 3), cubic (3, 4)
  const int gl = s->order, gu = s->order + 1;
  osr_zp::GGrid g{s->dims, {nx[0], nx[1], s->dims == 3 ? nx[2] : 1}, {gl, gl, s->dims == 3 ? gl : 0},
                  {gu, gu, s->dims == 3 ? gu : 0}};
  const size_t nc = (size_t)s->g.ncell, gn = g.cells();
  std::vector<float> pe(3 * nc), pb(3 * nc), eb(3 * gn), bb(3 * gn);
  HIPC(hipStreamSynchronize(s->stream));
  TRY(copy_on(s->stream, pe.data(), s->E.p, 12 * nc, hipMemcpyDeviceToHost));
  TRY(copy_on(s->stream, pb.data(), s->B.p, 12 * nc, hipMemcpyDeviceToHost));
  const int ring = s->g.nx;
  double phase0 = w.phase0, pol = w.pol;
  for (int k = 0; k < (w.pol_type != 0 ? 2 : 1); k++) {
    if (k == 1) {  // :2431-2437
      phase0 = phase0 + (w.pol_type > 0 ? osr_zp::kPi2 : -osr_zp::kPi2);
      pol = pol + osr_zp::kPi2;
    }
    std::fill(eb.begin(), eb.end(), 0.0f);
    std::fill(bb.begin(), bb.end(), 0.0f);
    osr_zp::box_build(w, g, s->order, s->dx, phase0, pol, eb.data(), bb.data());
    for (int cz = 0; cz < s->g.nz; cz++)
      for (int cy = 0; cy < s->g.ny; cy++)
        for (int cx = 0; cx < ring; cx++) {
          int i;  // Fortran x index of the ring column
          if (cx < nxi)
            i = cx + 1;
          else if (!s->g.open_nx)
            continue;
          else if (cx < nxi + gu)
            i = cx + 1;  // x = nx + 1 .. nx + gu
          else if (cx >= ring - gl)
            i = cx - ring + 1;  // x = 1 - gl .. 0
          else
            continue;
          const size_t c = (size_t)cx + (size_t)ring * (cy + (size_t)s->g.ny * cz);
          for (int m = 0; m < 3; m++) {
            const size_t r = g.at(m, i, cy + 1, cz + 1);
            pb[m * nc + c] = pb[m * nc + c] + bb[r];
            pe[m * nc + c] = pe[m * nc + c] + eb[r];
          }
        }
  }
  TRY(copy_on(s->stream, s->E.p, pe.data(), 12 * nc, hipMemcpyHostToDevice));
  TRY(copy_on(s->stream, s->B.p, pb.data(), 12 * nc, hipMemcpyHostToDevice));
  return restore_walls(s);
}

int osr_sim_zpulse_column(osr_sim* s, int k, int second, double t, float* b12, float* b3,
                          int* active) {
  TRY(check_sim(s));
  if (k < 0 || k >= (int)s->zp.size() || !b12 || !b3 || !active)
    return fail(OSR_ERR_ARG, "bad argument");
  const osr_zp::Wall& w = s->zp[k];
  if (second && w.pol_type == 0) return fail(OSR_ERR_ARG, "a linear pulse launches once");
  *active = osr_zp::active(w, t) ? 1 : 0;
  osr_zp::column(w, s->dims, s->g.ny, s->g.nz, s->dx, s->g.dt_dx[0], s->order, t, second != 0,
                 b12, b3);
  return OSR_OK;
}

int osr_sim_set_smooth(osr_sim* s, const int* order, const float* coef) {
  TRY(check_sim(s));
  TRY(check_smooth_args(s->dims, order, coef));
  if (s->g.open_nx) {
    bool any = false;
    for (int d = 0; d < s->dims; d++) any = any || order[d] > 0;
    if (any) return fail(OSR_ERR_STATE, "current smoothing needs periodic x walls");
  }
  HIPC(hipStreamSynchronize(s->stream));
  bool any = false;
  for (int d = 0; d < 3; d++) {
    s->sm_order[d] = d < s->dims ? order[d] : 0;
    any = any || s->sm_order[d] > 0;
    for (int j = 0; j < 31; j++) s->sm_coef[d][j] = d < s->dims ? coef[d * OSR_SMOOTH_NCOEF + j] : 0.f;
  }
  const size_t bytes = 12 * (size_t)s->g.ncell;
  if (any && !s->sm_buf[0].p) {
    TRY(s->sm_buf[0].alloc(bytes));
    TRY(s->sm_buf[1].alloc(bytes));
  }
  TRY(upload_descs(s));
  s->graph_ok = false;
  return OSR_OK;
}

}  // extern "C"

#include "ensemble.hip"
#include "noalloc.hip"
