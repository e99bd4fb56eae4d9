// bgw_ctx.cpp -- context lifecycle, errors, pools, timing and the host-pointer
// (synchronous, reference-like) forms of the C ABI.  The host forms validate
// their arguments on the host, stage arrays into context-owned device buffers,
// run the `_device` path and copy results back, so a caller sees the
// reference's blocking semantics (GPP/gpp.cpp:283-285, FF/ff.cpp:700-728).
#include <omp.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "bgw_internal.h"

namespace bgw {

int fail(bgw_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int fail(bgw_ctx* ctx, int code, const char* msg) {
  if (ctx) ctx->err.assign(msg);
  return code;
}

int hip_fail(bgw_ctx* ctx, hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return fail(ctx, e == hipErrorOutOfMemory ? BGW_ENOMEM : BGW_EHIP, m);
}

int pool_get(bgw_ctx* ctx, const char* name, size_t bytes, void** out) {
  if (bytes == 0) bytes = 16;
  if (ctx->arena) {
    const size_t off = (ctx->arena_used + 255) & ~size_t(255);
    if (off > ctx->arena_bytes || bytes > ctx->arena_bytes - off)
      return fail(ctx, BGW_EINVAL, "workspace exceeds the scratch array of the allocation-free call");
    *out = ctx->arena + off;
    ctx->arena_used = off + bytes;
    return BGW_OK;
  }
  DevBuf& b = ctx->pool[name];
  if (b.bytes < bytes) {
    if (b.ptr) {
      hipError_t e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
      (void)hipFree(b.ptr);
      b.ptr = nullptr;
      b.bytes = 0;
    }
    size_t want = (bytes + 255) & ~size_t(255);
    hipError_t e = hipMalloc(&b.ptr, want);
    if (e != hipSuccess) {
      b.ptr = nullptr;
      return hip_fail(ctx, e, (std::string("hipMalloc ") + name).c_str());
    }
    b.bytes = want;
  }
  *out = b.ptr;
  return BGW_OK;
}

// no timing inside allocation-free calls (it creates events)
KernelTimer::KernelTimer(bgw_ctx* c, const char* name) : ctx(c), on(c->timing && !c->arena) {
  if (!on) return;
  tl.name = name;
  for (hipEvent_t* ev : {&tl.start, &tl.stop}) {
    if (!ctx->free_events.empty()) {
      *ev = ctx->free_events.back();
      ctx->free_events.pop_back();
    } else if (hipEventCreate(ev) != hipSuccess) {
      on = false;
      return;
    }
  }
  (void)hipEventRecord(tl.start, ctx->stream);
}

KernelTimer::~KernelTimer() {
  if (!on) return;
  (void)hipEventRecord(tl.stop, ctx->stream);
  ctx->pending.push_back(tl);
}

int check_launch(bgw_ctx* ctx, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(ctx, e, what);
  return BGW_OK;
}

bool mul_ok(size_t a, size_t b, size_t* out) {
  if (a != 0 && b > SIZE_MAX / a) return false;
  *out = a * b;
  return true;
}

// ------------------------------------------------------------ pinned staging ring
namespace {
constexpr size_t kSlotBytes = size_t(32) << 20;  // 32 MiB per pinned buffer
constexpr int kSlots = 4;
constexpr size_t kDirectBytes = size_t(4) << 20;  // smaller copies go straight from pageable memory

// memcpy split over the host threads (pageable <-> pinned runs at host memory bandwidth)
void parallel_copy(void* dst, const void* src, size_t n) {
  const int nt = std::max(1, std::min(16, omp_get_max_threads()));
  const size_t piece = (n + nt - 1) / nt;
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int t = 0; t < nt; ++t) {
    const size_t b = std::min(n, (size_t)t * piece), e = std::min(n, b + piece);
    if (e > b) std::memcpy(static_cast<char*>(dst) + b, static_cast<const char*>(src) + b, e - b);
  }
}

// dst = a - b elementwise over n doubles, split over the host threads
void parallel_sub(double* dst, const double* a, const double* b, size_t n) {
  const int nt = std::max(1, std::min(16, omp_get_max_threads()));
  const size_t piece = (n + nt - 1) / nt;
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int t = 0; t < nt; ++t) {
    const size_t lo = std::min(n, (size_t)t * piece), hi = std::min(n, lo + piece);
    for (size_t i = lo; i < hi; ++i) dst[i] = a[i] - b[i];
  }
}

// Makes the ring usable; false if pinned memory is unavailable (copies then go direct).
bool ring_ready(bgw_ctx* ctx) {
  if (!ctx->ring.empty()) return true;
  std::vector<PinnedSlot> r(kSlots);
  for (auto& s : r) {
    if (hipHostMalloc(&s.host, kSlotBytes, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) {
      for (auto& t : r) {
        if (t.host) (void)hipHostFree(t.host);
        if (t.done) (void)hipEventDestroy(t.done);
      }
      return false;
    }
  }
  ctx->ring = std::move(r);
  return true;
}

// Next ring slot, after the DMA that last used it has finished.
int take_slot(bgw_ctx* ctx, PinnedSlot** out) {
  PinnedSlot& s = ctx->ring[ctx->ring_next++ % ctx->ring.size()];
  if (s.busy) {
    hipError_t e = hipEventSynchronize(s.done);
    if (e != hipSuccess) return hip_fail(ctx, e, "staging event");
    s.busy = false;
  }
  *out = &s;
  return BGW_OK;
}
}  // namespace

int upload_host(bgw_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return upload_host(ctx, dst, src, bytes, ctx->stream);
}

int upload_host(bgw_ctx* ctx, void* dst, const void* src, size_t bytes, hipStream_t stream) {
  if (bytes < kDirectBytes || !ring_ready(ctx)) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream);
    return e == hipSuccess ? BGW_OK : hip_fail(ctx, e, "upload");
  }
  for (size_t off = 0; off < bytes; off += kSlotBytes) {
    const size_t n = std::min(kSlotBytes, bytes - off);
    PinnedSlot* s;
    int rc = take_slot(ctx, &s);
    if (rc) return rc;
    parallel_copy(s->host, static_cast<const char*>(src) + off, n);
    hipError_t e = hipMemcpyAsync(static_cast<char*>(dst) + off, s->host, n, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipEventRecord(s->done, stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "staged upload");
    s->busy = true;
  }
  return BGW_OK;
}

int upload_host_diff(bgw_ctx* ctx, double* dst, const double* a, const double* b, size_t n) {
  if (!ring_ready(ctx)) {
    // no pinned memory: form the difference in pageable memory, copy it, and wait
    std::vector<double> tmp(n);
    parallel_sub(tmp.data(), a, b, n);
    hipError_t e = hipMemcpyAsync(dst, tmp.data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? BGW_OK : hip_fail(ctx, e, "difference upload");
  }
  const size_t per = kSlotBytes / sizeof(double);
  for (size_t off = 0; off < n; off += per) {
    const size_t m = std::min(per, n - off);
    PinnedSlot* s;
    int rc = take_slot(ctx, &s);
    if (rc) return rc;
    parallel_sub(static_cast<double*>(s->host), a + off, b + off, m);
    hipError_t e = hipMemcpyAsync(dst + off, s->host, m * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(s->done, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "staged difference upload");
    s->busy = true;
  }
  return BGW_OK;
}

int download_host(bgw_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes < kDirectBytes || !ring_ready(ctx)) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? BGW_OK : hip_fail(ctx, e, "download");
  }
  // DMA of chunk c runs while the host copies chunk c - (slots - 1) out of its buffer
  struct Pending {
    PinnedSlot* s;
    size_t off, n;
  };
  std::vector<Pending> q;
  auto drain_one = [&]() -> int {
    Pending p = q.front();
    q.erase(q.begin());
    hipError_t e = hipEventSynchronize(p.s->done);
    if (e != hipSuccess) return hip_fail(ctx, e, "staging event");
    p.s->busy = false;
    parallel_copy(static_cast<char*>(dst) + p.off, p.s->host, p.n);
    return BGW_OK;
  };
  for (size_t off = 0; off < bytes; off += kSlotBytes) {
    if (q.size() + 1 >= ctx->ring.size()) {
      int rc = drain_one();
      if (rc) return rc;
    }
    const size_t n = std::min(kSlotBytes, bytes - off);
    PinnedSlot* s;
    int rc = take_slot(ctx, &s);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(s->host, static_cast<const char*>(src) + off, n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(s->done, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "staged download");
    s->busy = true;
    q.push_back({s, off, n});
  }
  while (!q.empty()) {
    int rc = drain_one();
    if (rc) return rc;
  }
  return BGW_OK;
}

int copy_stream(bgw_ctx* ctx, hipStream_t* out) {
  if (!ctx->copy_stream) {
    hipError_t e = hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking);
    if (e == hipSuccess && !ctx->xfer_event) e = hipEventCreateWithFlags(&ctx->xfer_event, hipEventDisableTiming);
    if (e != hipSuccess) return hip_fail(ctx, e, "copy stream");
  }
  *out = ctx->copy_stream;
  return BGW_OK;
}

int cross_stream_wait(bgw_ctx* ctx, hipStream_t from, hipStream_t to) {
  hipError_t e = hipEventRecord(ctx->xfer_event, from);
  if (e == hipSuccess) e = hipStreamWaitEvent(to, ctx->xfer_event, 0);
  return e == hipSuccess ? BGW_OK : hip_fail(ctx, e, "stream hand-over");
}

void gather_igp_columns(const double2* am, const double2* an, const int* inv, const int* indinv, int nb,
                        int ngpown, int ncouls, double2* gm, double2* gn) {
  std::vector<int> igp(ngpown);
  for (int p = 0; p < ngpown; ++p) igp[p] = indinv[inv[p]];
  const int nt = std::max(1, std::min(16, omp_get_max_threads()));
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int n1 = 0; n1 < nb; ++n1) {
    const double2 *rm = am + (size_t)n1 * ncouls, *rn = an + (size_t)n1 * ncouls;
    double2 *om = gm + (size_t)n1 * ngpown, *on = gn + (size_t)n1 * ngpown;
    for (int p = 0; p < ngpown; ++p) {
      om[p] = rm[igp[p]];
      on[p] = rn[igp[p]];
    }
  }
}

void ring_release(bgw_ctx* ctx) {
  for (auto& s : ctx->ring) {
    if (s.host) (void)hipHostFree(s.host);
    if (s.done) (void)hipEventDestroy(s.done);
  }
  ctx->ring.clear();
}

}  // namespace bgw

using namespace bgw;

extern "C" {

int bgw_create(bgw_ctx** out, int device) {
  if (!out) return BGW_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return BGW_ENODEV;
  if (device < 0 || device >= n) return BGW_EINVAL;
  bgw_ctx* ctx = new bgw_ctx();
  ctx->device = device;
  ctx->err.reserve(512);
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ctx->cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e == hipSuccess) e = hipMalloc(&ctx->d_flag, 256);
  ctx->stream = ctx->own_stream;
  if (e == hipSuccess) {
    reset_flag(ctx->d_flag, ctx->stream);
    e = hipStreamSynchronize(ctx->stream);
  }
  if (e != hipSuccess) {
    const int code = e == hipErrorOutOfMemory ? BGW_ENOMEM : BGW_ENODEV;
    bgw_destroy(ctx);  // releases whatever was created
    return code;
  }
  // the allocation-free entry point's scratch array; a failure here only disables that
  // entry point (it then returns the error), not the context
  (void)gpp_scratch_attach(ctx);
  ctx->err.clear();
  *out = ctx;
  return BGW_OK;
}

void bgw_destroy(bgw_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  // both streams drain before any buffer they may still read or fill is freed
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
  for (auto& kv : ctx->pool)
    if (kv.second.ptr) (void)hipFree(kv.second.ptr);
  for (auto& tl : ctx->pending) {
    (void)hipEventDestroy(tl.start);
    (void)hipEventDestroy(tl.stop);
  }
  for (hipEvent_t ev : ctx->free_events) (void)hipEventDestroy(ev);
  if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
  if (ctx->xfer_event) (void)hipEventDestroy(ctx->xfer_event);
  ring_release(ctx);
  if (ctx->d_flag) (void)hipFree(ctx->d_flag);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
}

const char* bgw_last_error(const bgw_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int bgw_set_stream(bgw_ctx* ctx, void* s) {
  if (!ctx) return BGW_EINVAL;
  hipStream_t next = s ? static_cast<hipStream_t>(s) : ctx->own_stream;
  hipError_t e = hipSuccess;
  if (next != ctx->stream) {
    // work queued on the previous stream may still use the context's buffers, which
    // the next call on the new stream reuses: let it finish first.  The switch happens
    // even when that fails (a caller's stream already destroyed), and the error is reported
    e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  }
  ctx->stream = next;
  return e == hipSuccess ? BGW_OK : hip_fail(ctx, e, "hipStreamSynchronize (stream switch)");
}

void* bgw_get_stream(bgw_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int bgw_set_precision(bgw_ctx* ctx, int precision) {
  if (!ctx) return BGW_EINVAL;
  if (precision != BGW_FP64 && precision != BGW_FP32_BULK && precision != BGW_FP64_MOMENTS)
    return fail(ctx, BGW_EINVAL, "unknown precision");
  ctx->precision = precision;
  return BGW_OK;
}

int bgw_synchronize(bgw_ctx* ctx) {
  if (!ctx) return BGW_EINVAL;
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
  int flag = 0;
  e = hipMemcpy(&flag, ctx->d_flag, sizeof(int), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(ctx, e, "flag readback");
  if (flag) {
    reset_flag(ctx->d_flag, ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    return fail(ctx, BGW_EINVAL, "index out of range in inv_igp_index/indinv (clamped)");
  }
  return BGW_OK;
}

int bgw_set_timing(bgw_ctx* ctx, int on) {
  if (!ctx) return BGW_EINVAL;
  ctx->timing = on != 0;
  return BGW_OK;
}

static int resolve_timing(bgw_ctx* ctx) {
  if (ctx->pending.empty()) return BGW_OK;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
  for (auto& tl : ctx->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, tl.start, tl.stop) == hipSuccess) {
      auto& slot = ctx->ktime[tl.name];
      slot.first += ms;
      slot.second += 1;
    }
    ctx->free_events.push_back(tl.start);
    ctx->free_events.push_back(tl.stop);
  }
  ctx->pending.clear();
  return BGW_OK;
}

int bgw_kernel_time(bgw_ctx* ctx, const char* kernel, double* total_ms, long* launches) {
  if (!ctx || !kernel) return BGW_EINVAL;
  int rc = resolve_timing(ctx);
  if (rc) return rc;
  auto it = ctx->ktime.find(kernel);
  if (total_ms) *total_ms = it == ctx->ktime.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == ctx->ktime.end() ? 0 : it->second.second;
  return BGW_OK;
}

int bgw_reset_timing(bgw_ctx* ctx) {
  if (!ctx) return BGW_EINVAL;
  int rc = resolve_timing(ctx);
  ctx->ktime.clear();
  return rc;
}

int bgw_device_cus(bgw_ctx* ctx) { return ctx ? ctx->cus : 0; }

}  // extern "C"

// ------------------------------------------------------------------ host forms
namespace {

struct Stage {
  bgw_ctx* ctx;
  int rc = BGW_OK;
  explicit Stage(bgw_ctx* c) : ctx(c) {}
  // uploads `bytes` from host `src` into pool slot `name`
  template <typename T>
  T* up(const char* name, const T* src, size_t count) {
    if (rc) return nullptr;
    size_t bytes;
    if (!mul_ok(count, sizeof(T), &bytes)) {
      rc = fail(ctx, BGW_EINVAL, "size overflow");
      return nullptr;
    }
    void* d = nullptr;
    rc = pool_get(ctx, name, bytes, &d);
    if (rc) return nullptr;
    rc = upload_host(ctx, d, src, bytes);
    return rc ? nullptr : static_cast<T*>(d);
  }
  void* buf(const char* name, size_t bytes) {
    if (rc) return nullptr;
    void* d = nullptr;
    rc = pool_get(ctx, name, bytes, &d);
    return rc ? nullptr : d;
  }
  double* out(const char* name, size_t count) {
    if (rc) return nullptr;
    void* d = nullptr;
    rc = pool_get(ctx, name, count * sizeof(double), &d);
    return rc ? nullptr : static_cast<double*>(d);
  }
  void down(double* host, const double* dev, size_t count) {
    if (rc || !host) return;
    rc = download_host(ctx, host, dev, count * sizeof(double));
  }
};

int check_ptrs(bgw_ctx* ctx, std::initializer_list<const void*> ps) {
  for (const void* p : ps)
    if (!p) return fail(ctx, BGW_EINVAL, "null array argument");
  return BGW_OK;
}

// GPP index domain: inv_igp_index in [0,ncouls] (the reference's own init reaches
// ncouls, GPP/gpp.cpp:263-264), indinv[ncouls+1] in [0,ncouls).
int check_gpp_indices(bgw_ctx* ctx, size_t ngpown, size_t ncouls, const int* inv, const int* indinv) {
  for (size_t p = 0; p < ngpown; ++p)
    if (inv[p] < 0 || (size_t)inv[p] > ncouls)
      return fail(ctx, BGW_EINVAL, "inv_igp_index[" + std::to_string(p) + "] out of [0,ncouls]");
  for (size_t g = 0; g <= ncouls; ++g)
    if (indinv[g] < 0 || (size_t)indinv[g] >= ncouls)
      return fail(ctx, BGW_EINVAL, "indinv[" + std::to_string(g) + "] out of [0,ncouls)");
  return BGW_OK;
}

int gpp_sizes(bgw_ctx* ctx, size_t nb, size_t ngpown, size_t ncouls, size_t* n_bg, size_t* n_pg) {
  if (nb == 0 || ngpown == 0 || ncouls == 0) return fail(ctx, BGW_EINVAL, "zero size");
  if (ncouls >= (size_t)INT32_MAX || ngpown >= (size_t)INT32_MAX || nb >= (size_t)INT32_MAX)
    return fail(ctx, BGW_EINVAL, "size exceeds int range");
  if (!mul_ok(nb, ncouls, n_bg) || !mul_ok(ngpown, ncouls, n_pg) || *n_pg > SIZE_MAX / 16 ||
      *n_bg > SIZE_MAX / 16)
    return fail(ctx, BGW_EINVAL, "size overflow");
  return BGW_OK;
}

int finish(bgw_ctx* ctx, int rc) {
  if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
  if (rc) {
    (void)hipStreamSynchronize(ctx->stream);
    return rc;
  }
  return bgw_synchronize(ctx);
}

int gpp_host(bgw_ctx* ctx, size_t nb, size_t nvband, size_t ngpown, size_t ncouls, const int* inv,
             const int* indinv, const double* wx, int wx_per_band, const double* wtilde,
             const double* aqsm, const double* aqsn, const double* I_eps, const double* vcoul,
             double occ, double* achtemp, double* acht_n1_loc, double* achstemp, double* asxtemp) {
  if (!ctx) return BGW_EINVAL;
  size_t n_bg, n_pg;
  int rc = check_ptrs(ctx, {inv, indinv, wx, wtilde, aqsm, aqsn, I_eps, vcoul, achtemp});
  if (!rc) rc = gpp_sizes(ctx, nb, ngpown, ncouls, &n_bg, &n_pg);
  if (!rc && nvband > nb) rc = fail(ctx, BGW_EINVAL, "nvband > number_bands");
  if (!rc) rc = check_gpp_indices(ctx, ngpown, ncouls, inv, indinv);
  if (rc) return rc;
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, BGW_ENODEV, "hipSetDevice");
  // device buffers (grow-only pools), filled by the overlapped staging of gpp_forward_staged
  Stage st(ctx);
  GppDevBufs d{};
  d.inv = static_cast<int*>(st.buf("h_inv", (size_t)ngpown * sizeof(int)));
  d.indinv = static_cast<int*>(st.buf("h_indinv", (ncouls + 1) * sizeof(int)));
  d.wx = static_cast<double*>(st.buf("h_wx", (wx_per_band ? nb * 3 : 3) * sizeof(double)));
  d.wtilde = static_cast<double*>(st.buf("h_wtilde", 16 * n_pg));
  d.aqsm = static_cast<double*>(st.buf("h_aqsm", 16 * n_bg));
  d.aqsn = static_cast<double*>(st.buf("h_aqsn", 16 * n_bg));
  d.I_eps = static_cast<double*>(st.buf("h_ieps", 16 * n_pg));
  d.vcoul = static_cast<double*>(st.buf("h_vcoul", ncouls * sizeof(double)));
  double* d_ach = st.out("h_achtemp", 6);
  double* d_loc = acht_n1_loc ? st.out("h_acht_n1_loc", 6 * nb) : nullptr;
  double* d_achs = achstemp ? st.out("h_achstemp", 2) : nullptr;
  double* d_asx = asxtemp ? st.out("h_asxtemp", 6) : nullptr;
  if (st.rc) return finish(ctx, st.rc);
  const GppHostArgs h{inv, indinv, wx, wx_per_band, wtilde, aqsm, aqsn, I_eps, vcoul};
  rc = gpp_forward_staged(ctx, (int)nb, (int)nvband, (int)ngpown, (int)ncouls, h, d, occ, d_ach, d_loc, d_achs,
                          d_asx);
  if (rc) return finish(ctx, rc);
  st.down(achtemp, d_ach, 6);
  st.down(acht_n1_loc, d_loc, 6 * nb);
  st.down(achstemp, d_achs, 2);
  st.down(asxtemp, d_asx, 6);
  return finish(ctx, st.rc);
}

}  // namespace

extern "C" {

int bgw_noflagOCC_solver(bgw_ctx* ctx, size_t number_bands, size_t ngpown, size_t ncouls,
                         const int* inv_igp_index, const int* indinv, const double* wx_array,
                         const double* wtilde_array, const double* aqsmtemp,
                         const double* aqsntemp, const double* I_eps_array, const double* vcoul,
                         double* achtemp) {
  return gpp_host(ctx, number_bands, 0, ngpown, ncouls, inv_igp_index, indinv, wx_array, 0,
                  wtilde_array, aqsmtemp, aqsntemp, I_eps_array, vcoul, 0.0, achtemp, nullptr,
                  nullptr, nullptr);
}

int bgw_gpp_solver(bgw_ctx* ctx, size_t number_bands, size_t nvband, size_t ngpown, size_t ncouls,
                   const int* inv_igp_index, const int* indinv, const double* wx_array,
                   const double* wtilde_array, const double* aqsmtemp, const double* aqsntemp,
                   const double* I_eps_array, const double* vcoul, double occ, double* achtemp,
                   double* acht_n1_loc, double* achstemp, double* asxtemp) {
  return gpp_host(ctx, number_bands, nvband, ngpown, ncouls, inv_igp_index, indinv, wx_array, 1,
                  wtilde_array, aqsmtemp, aqsntemp, I_eps_array, vcoul, occ, achtemp, acht_n1_loc,
                  achstemp, asxtemp);
}

int bgw_gpp_solver_device(bgw_ctx* ctx, size_t number_bands, size_t nvband, size_t ngpown,
                          size_t ncouls, const int* inv_igp_index, const int* indinv,
                          const double* wx_array, const double* wtilde_array,
                          const double* aqsmtemp, const double* aqsntemp,
                          const double* I_eps_array, const double* vcoul, double occ,
                          double* achtemp, double* acht_n1_loc, double* achstemp,
                          double* asxtemp) {
  if (!ctx) return BGW_EINVAL;
  size_t n_bg, n_pg;
  int rc = check_ptrs(ctx, {inv_igp_index, indinv, wx_array, wtilde_array, aqsmtemp, aqsntemp,
                            I_eps_array, vcoul, achtemp});
  if (!rc) rc = gpp_sizes(ctx, number_bands, ngpown, ncouls, &n_bg, &n_pg);
  if (!rc && nvband > number_bands) rc = fail(ctx, BGW_EINVAL, "nvband > number_bands");
  if (rc) return rc;
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, BGW_ENODEV, "hipSetDevice");
  return gpp_forward_device(ctx, number_bands, nvband, ngpown, ncouls, inv_igp_index, indinv,
                            wx_array, 1, wtilde_array, aqsmtemp, aqsntemp, I_eps_array, vcoul, occ,
                            achtemp, acht_n1_loc, achstemp, asxtemp);
}

int bgw_gpp_backward(bgw_ctx* ctx, size_t nb, size_t nvband, size_t ngpown, size_t ncouls,
                     const int* inv, const int* indinv, const double* wx, const double* wtilde,
                     const double* aqsm, const double* aqsn, const double* I_eps,
                     const double* vcoul, double occ, double* dsigma, double* dwt, double* dI) {
  if (!ctx) return BGW_EINVAL;
  size_t n_bg, n_pg;
  int rc = check_ptrs(ctx, {inv, indinv, wx, wtilde, aqsm, aqsn, I_eps, vcoul});
  if (!rc) rc = gpp_sizes(ctx, nb, ngpown, ncouls, &n_bg, &n_pg);
  if (!rc && nvband > nb) rc = fail(ctx, BGW_EINVAL, "nvband > number_bands");
  if (!rc) rc = check_gpp_indices(ctx, ngpown, ncouls, inv, indinv);
  if (rc) return rc;
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, BGW_ENODEV, "hipSetDevice");
  Stage st(ctx);
  const int* d_inv = st.up("h_inv", inv, ngpown);
  const int* d_indinv = st.up("h_indinv", indinv, ncouls + 1);
  const double* d_wx = st.up("h_wx", wx, nb * 3);
  const double* d_wt = st.up("h_wtilde", wtilde, 2 * n_pg);
  const double* d_am = st.up("h_aqsm", aqsm, 2 * n_bg);
  const double* d_an = st.up("h_aqsn", aqsn, 2 * n_bg);
  const double* d_I = st.up("h_ieps", I_eps, 2 * n_pg);
  const double* d_vc = st.up("h_vcoul", vcoul, ncouls);
  double* d_ds = dsigma ? st.out("h_dsigma", 6) : nullptr;
  double* d_dw = dwt ? st.out("h_dwt", 2 * n_pg) : nullptr;
  double* d_dI = dI ? st.out("h_dI", 2 * n_pg) : nullptr;
  if (st.rc) return finish(ctx, st.rc);
  rc = gpp_backward_device(ctx, nb, nvband, ngpown, ncouls, d_inv, d_indinv, d_wx, d_wt, d_am, d_an,
                           d_I, d_vc, occ, d_ds, d_dw, d_dI);
  if (rc) return finish(ctx, rc);
  st.down(dsigma, d_ds, 6);
  st.down(dwt, d_dw, 2 * n_pg);
  st.down(dI, d_dI, 2 * n_pg);
  return finish(ctx, st.rc);
}

int bgw_gpp_backward_device(bgw_ctx* ctx, size_t nb, size_t nvband, size_t ngpown, size_t ncouls,
                            const int* inv, const int* indinv, const double* wx,
                            const double* wtilde, const double* aqsm, const double* aqsn,
                            const double* I_eps, const double* vcoul, double occ, double* dsigma,
                            double* dwt, double* dI) {
  if (!ctx) return BGW_EINVAL;
  size_t n_bg, n_pg;
  int rc = check_ptrs(ctx, {inv, indinv, wx, wtilde, aqsm, aqsn, I_eps, vcoul});
  if (!rc) rc = gpp_sizes(ctx, nb, ngpown, ncouls, &n_bg, &n_pg);
  if (!rc && nvband > nb) rc = fail(ctx, BGW_EINVAL, "nvband > number_bands");
  if (rc) return rc;
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, BGW_ENODEV, "hipSetDevice");
  return gpp_backward_device(ctx, nb, nvband, ngpown, ncouls, inv, indinv, wx, wtilde, aqsm, aqsn,
                             I_eps, vcoul, occ, dsigma, dwt, dI);
}

}  // extern "C"
