// noalloc.hip -- osr_advance_deposit_noalloc: the stateless push + deposit of
// osr_advance_deposit on arrays that are already in device memory, without any device
// allocation.  Included at the end of capi.hip.
//
// Its working memory (the int64 fixed-point current, the per-block energy partials, the
// range-check flag and the charge sums that set the fixed-point scale) lives in the
// file-scope device array `scratch` of kScratchBytes.  Each GPU holds its own copy for the
// life of the process; its address on a device comes from hipGetSymbolAddress, taken once
// per device at the first call made with that device current, and only that address is
// ever handed to kernels and copies (the array's name is a host-side stand-in).  Every
// call on a GPU shares that GPU's array and clears what must start at zero, so calls that
// overlap on one GPU (several host threads, or several streams) race: serialise them.
// Kernels run on the null stream, as the host-pointer entry point's do, and the call
// returns after they finished.
#include <climits>
#include <mutex>

namespace {
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kScratchHdr = 8192;  // header: flag, charge partials, energy
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

struct NoallocHdr {
  int bad;        // first particle whose stencil leaves the arrays (INT_MAX: none)
  int pad;
  double qabs;    // sum |q| (fixed-order tree of 256 strided partial sums)
  double qmax;    // max |q|
  double ene;     // sum of the per-block energy partials, in block order
};

constexpr int kMaxScratchDev = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDev] = {};

// address of the current device's scratch array (looked up once per device)
int scratch_address(void** out) {
  int dev = 0;
  HIPC(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxScratchDev) return fail(OSR_ERR_HIP, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(OSR_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return OSR_OK;
}

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

// zero the fixed-point current, reset the flag
__global__ __launch_bounds__(256) void na_prepare(NoallocHdr* h, unsigned long long* jfix,
                                                  long long nj) {
  const long long stride = (long long)gridDim.x * 256;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nj; i += stride) jfix[i] = 0ull;
  if (blockIdx.x == 0 && threadIdx.x == 0) h->bad = INT_MAX;
}

// the host entry point's range check: the gather reaches ix-1+lp .. ix+lp+order, the
// deposit ix-1+lp .. ix+1+lp+order (1-based Fortran cells, offsets = lower guards)
struct NaCheckArgs {
  const int* ix;
  int np, dims, order;
  int emf_ld[3], emf_off[3], j_ld[3], j_off[3];
};
__global__ __launch_bounds__(256) void na_check(NoallocHdr* h, NaCheckArgs a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.np) return;
  const int lp = a.order == 1 ? 0 : -1;
  bool bad = false;
  for (int d = 0; d < a.dims; d++) {
    const int c = a.ix[(long long)a.dims * i + d];
    const int glo = c - 1 + lp, ghi = c + lp + a.order, jlo = c - 1 + lp, jhi = c + 1 + lp + a.order;
    if (glo + a.emf_off[d] < 1 || ghi + a.emf_off[d] > a.emf_ld[d] || jlo + a.j_off[d] < 1 ||
        jhi + a.j_off[d] > a.j_ld[d])
      bad = true;
  }
  if (bad) atomicMin(&h->bad, i);
}

// sum and maximum of |q| in one workgroup: 256 strided partial sums, then the partials in
// thread order (deterministic)
__global__ __launch_bounds__(256) void na_charge(NoallocHdr* h, const float* q, int np) {
  __shared__ double s_sum[256], s_max[256];
  double s = 0.0, m = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) {
    const double v = fabs((double)q[i]);
    s += v;
    m = fmax(m, v);
  }
  s_sum[threadIdx.x] = s;
  s_max[threadIdx.x] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double ts = 0.0, tm = 0.0;
    for (int t = 0; t < 256; t++) {
      ts += s_sum[t];
      tm = fmax(tm, s_max[t]);
    }
    h->qabs = ts;
    h->qmax = tm;
  }
}

// energy partials summed in block order, as the host entry point sums them
__global__ void na_energy(NoallocHdr* h, const double* ep, int nblk) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double e = 0.0;
  for (int b = 0; b < nblk; b++) e += ep[b];
  h->ene = e;
}

}  // namespace

extern "C" {

int osr_advance_deposit_noalloc(int dims, int order, int pusher, int* ix, float* x, float* u,
                                const float* q, int i0, int i1, float rqm, const float* efield,
                                const float* bfield, const int* emf_size, const int* emf_offset,
                                float* j, const int* j_size, const int* j_offset,
                                const double* dx, double dt, double* ene) {
  if (dims != 2 && dims != 3) return fail(OSR_ERR_ARG, "dims must be 2 or 3");
  if (order < 1 || order > 3) return fail(OSR_ERR_ARG, "order must be 1..3");
  if (pusher != OSR_BORIS && pusher != OSR_VAY) return fail(OSR_ERR_ARG, "unknown pusher");
  if (!ix || !x || !u || !q || !efield || !bfield || !emf_size || !emf_offset || !j || !j_size ||
      !j_offset || !dx || !ene)
    return fail(OSR_ERR_ARG, "null argument");
  if (i0 < 1 || i1 < i0 - 1) return fail(OSR_ERR_ARG, "bad particle range");
  for (int d = 0; d < dims; d++)
    if (emf_size[d] < 1 || j_size[d] < 1 || !(dx[d] > 0))
      return fail(OSR_ERR_ARG, "bad array extents or cell size");
  const int np = i1 - i0 + 1;
  size_t nj = 3;
  for (int d = 0; d < dims; d++) nj *= (size_t)j_size[d];
  const int nblk = (np + 255) / 256;
  const size_t off_ep = kScratchHdr;
  const size_t off_j = align256(off_ep + 8 * (size_t)std::max(nblk, 1));
  const size_t need = off_j + 8 * nj;
  if (need > kScratchBytes)
    return fail(OSR_ERR_NOMEM,
                "osr_advance_deposit_noalloc needs %zu bytes of working memory (%d particles, "
                "%zu current values), more than its fixed %zu-byte device array",
                need, np, nj, kScratchBytes);
  void* base = nullptr;
  TRY(scratch_address(&base));
  NoallocHdr* hdr = (NoallocHdr*)base;
  double* epart = (double*)((char*)base + off_ep);
  unsigned long long* jfix = (unsigned long long*)((char*)base + off_j);

  const unsigned zb = (unsigned)std::min<size_t>((nj + 255) / 256, 4096);
  hipLaunchKernelGGL(na_prepare, dim3(std::max(zb, 1u)), dim3(256), 0, 0, hdr, jfix, (long long)nj);
  HIPC(hipGetLastError());
  int* pix = ix + (size_t)dims * (i0 - 1);
  const float* pq = q + (i0 - 1);
  NaCheckArgs ca{};
  ca.ix = pix;
  ca.np = np;
  ca.dims = dims;
  ca.order = order;
  for (int d = 0; d < 3; d++) {
    ca.emf_ld[d] = d < dims ? emf_size[d] : 1;
    ca.emf_off[d] = d < dims ? emf_offset[d] : 0;
    ca.j_ld[d] = d < dims ? j_size[d] : 1;
    ca.j_off[d] = d < dims ? j_offset[d] : 0;
  }
  if (np > 0) {
    hipLaunchKernelGGL(na_check, dim3(nblk), dim3(256), 0, 0, hdr, ca);
    HIPC(hipGetLastError());
    hipLaunchKernelGGL(na_charge, dim3(1), dim3(256), 0, 0, hdr, pq, np);
    HIPC(hipGetLastError());
  }
  NoallocHdr h{};
  HIPC(hipStreamSynchronize(nullptr));
  HIPC(hipMemcpy(&h, hdr, sizeof(h), hipMemcpyDeviceToHost));
  if (np > 0 && h.bad != INT_MAX)
    return fail(OSR_ERR_ARG, "particle %d stencil leaves the guarded arrays", i0 + h.bad);

  DropinArgs a{};
  a.ix = pix;
  a.x = x + (size_t)dims * (i0 - 1);
  a.u = u + (size_t)3 * (i0 - 1);
  a.q = pq;
  a.np = np;
  a.e = efield;
  a.b = bfield;
  for (int d = 0; d < 3; d++) {
    a.emf_ld[d] = ca.emf_ld[d];
    a.emf_off[d] = ca.emf_off[d];
    a.j_ld[d] = ca.j_ld[d];
    a.j_off[d] = ca.j_off[d];
    a.dt_dx[d] = d < dims ? (float)(dt / dx[d]) : 0.f;
    a.jnorm[d] = d < dims ? (float)(dx[d] / dt / (dims == 2 ? 2.0 : 3.0)) : 0.f;
  }
  a.jfix = jfix;
  a.tem = (float)(0.5 * dt / (double)rqm);
  // fixed-point scale from this call's charges: the host entry point's bound, with sum |q|
  // from the device (it can differ from the host's sequential sum in the last bit, which
  // moves the scale only when the bound lies within round-off of a power of two)
  double jn = 0;
  for (int d = 0; d < dims; d++) jn += a.jnorm[d];
  const double per = (dims == 2 ? 2.0 : 3.0) * jn + 1.0;
  const double qabs = np > 0 ? h.qabs : 0.0, qmax = np > 0 ? h.qmax : 0.0;
  int sh = std::min(62 - (int)std::ceil(std::log2(std::max(qabs * per * 8.0, 1e-300))),
                    50 - (int)std::ceil(std::log2(std::max(qmax * per * 2.0, 1e-300))));
  sh = std::max(0, std::min(sh, 80));
  a.jscale = std::ldexp(1.0, sh);
  a.epart = epart;
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
    hipLaunchKernelGGL(na_energy, dim3(1), dim3(1), 0, 0, hdr, (const double*)epart, nblk);
    HIPC(hipGetLastError());
  }
  hipLaunchKernelGGL(dropin_jadd, dim3((unsigned)((nj + 255) / 256)), dim3(256), 0, 0, j,
                     (const unsigned long long*)jfix, (long long)nj, std::ldexp(1.0, -sh));
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(nullptr));
  if (np > 0) {
    double e = 0.0;
    HIPC(hipMemcpy(&e, &hdr->ene, sizeof(double), hipMemcpyDeviceToHost));
    *ene += e;
  }
  return OSR_OK;
}

}  // extern "C"
