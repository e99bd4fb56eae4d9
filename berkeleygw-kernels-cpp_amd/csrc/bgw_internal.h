// bgw_internal.h -- context, error handling and device-buffer plumbing shared by
// the host-side translation units of libbgw (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "bgw.h"

namespace bgw {

// Grow-only device allocation owned by a context.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

struct TimedLaunch {
  std::string name;
  hipEvent_t start, stop;
};

// One buffer of the pinned staging ring of the host-pointer entry points.
struct PinnedSlot {
  void* host = nullptr;
  hipEvent_t done = nullptr;  // recorded after the DMA that reads or fills `host`
  bool busy = false;
};

}  // namespace bgw

struct bgw_ctx {
  int device = 0;
  int cus = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  int precision = BGW_FP64;
  std::string err;
  int* d_flag = nullptr;  // device words: [0] bad-index report of *_device calls,
                          // [1] the GPP moment form's "some chunk takes the direct sum"
  // named grow-only pools: workspace and host-API staging of inputs/outputs
  std::map<std::string, bgw::DevBuf> pool;
  // timing
  bool timing = false;
  std::vector<bgw::TimedLaunch> pending;
  std::vector<hipEvent_t> free_events;
  std::map<std::string, std::pair<double, long>> ktime;
  // pinned staging ring for large host<->device copies (allocated on first use)
  std::vector<bgw::PinnedSlot> ring;
  size_t ring_next = 0;
  // copy stream of the overlapped host-pointer GPP path and its hand-over event
  hipStream_t copy_stream = nullptr;
  hipEvent_t xfer_event = nullptr;
  // host buffers of the gathered aqsm / aqsn columns: a copy from them may still be
  // queued when the staging function returns, so they live until the next call
  std::vector<double2> gather_m, gather_n;
  // device address of this device's static scratch array (allocation-free entry point),
  // looked up once per device when the first context on it is created; null if unavailable
  void* scratch = nullptr;
  // while an allocation-free call runs, pool_get carves workspace from this arena (the
  // scratch array) instead of the grow-only pools
  char* arena = nullptr;
  size_t arena_bytes = 0, arena_used = 0;
};

namespace bgw {

int fail(bgw_ctx* ctx, int code, const std::string& msg);
// Same, copying a C string into the context's message buffer (reserved at creation, so
// short messages cost no host allocation).
int fail(bgw_ctx* ctx, int code, const char* msg);
int hip_fail(bgw_ctx* ctx, hipError_t e, const char* what);

// Returns a device buffer of at least `bytes` from pool slot `name`, or, during an
// allocation-free call, the next 256-byte-aligned piece of the context's arena.
int pool_get(bgw_ctx* ctx, const char* name, size_t bytes, void** out);
// Looks up (once per device, with ctx->device current) the device address of the
// scratch array of the allocation-free entry point and records it in ctx->scratch.
int gpp_scratch_attach(bgw_ctx* ctx);

// Brackets a launch with HIP events when timing is on.
struct KernelTimer {
  bgw_ctx* ctx;
  TimedLaunch tl;
  bool on;
  KernelTimer(bgw_ctx* c, const char* name);
  ~KernelTimer();
};

int check_launch(bgw_ctx* ctx, const char* what);

// Overflow-checked product of sizes; returns false on overflow.
bool mul_ok(size_t a, size_t b, size_t* out);

// Copies between caller (pageable) host memory and device memory on ctx->stream.
// Large copies go through the context's pinned ring: a multi-threaded host copy
// fills one pinned buffer while the DMA engine drains the previous one.  Uploads
// return once the source has been consumed (the caller may reuse it); downloads
// return with the data in `dst`.
int upload_host(bgw_ctx* ctx, void* dst, const void* src, size_t bytes);
int upload_host(bgw_ctx* ctx, void* dst, const void* src, size_t bytes, hipStream_t s);
int download_host(bgw_ctx* ctx, void* dst, const void* src, size_t bytes);
// Uploads a - b (n doubles, elementwise, IEEE double subtraction) on ctx->stream: the
// difference is formed by the host copy into the pinned ring, so only n doubles cross
// the bus.  Returns once a and b have been consumed.
int upload_host_diff(bgw_ctx* ctx, double* dst, const double* a, const double* b, size_t n);
// Releases the pinned ring (the stream must be idle).
void ring_release(bgw_ctx* ctx);
// The context's copy stream (created on first use, default priority).
int copy_stream(bgw_ctx* ctx, hipStream_t* out);
// Makes `to` wait for the work enqueued so far on `from`.
int cross_stream_wait(bgw_ctx* ctx, hipStream_t from, hipStream_t to);
// Host gather of the columns igp(p) = indinv[inv[p]] of two [nb][ncouls] complex
// arrays into [nb][ngpown] (indices already checked), on the host threads.
void gather_igp_columns(const double2* am, const double2* an, const int* inv, const int* indinv, int nb,
                        int ngpown, int ncouls, double2* gm, double2* gn);

// Zero-fill kernels usable inside captured graphs (no memset nodes).
void fill_zero(void* p, size_t bytes, hipStream_t s);
void reset_flag(int* d_flag, hipStream_t s);

// Device-side index validation: out-of-range indices set *flag (the kernels
// clamp them so no access leaves its array).
void launch_check_indices(const int* inv_igp_index, size_t ngpown, const int* indinv,
                          size_t n_indinv, int inv_max, int ncouls, int* flag, hipStream_t s);

}  // namespace bgw

// Internal entry points implemented in the kernel translation units.
namespace bgw {
// Host arrays of a GPP call and the device buffers they are staged into.
struct GppHostArgs {
  const int *inv, *indinv;
  const double* wx;
  int wx_per_band;
  const double *wtilde, *aqsm, *aqsn, *I_eps, *vcoul;
};
struct GppDevBufs {
  int *inv, *indinv;
  double *wx, *wtilde, *aqsm, *aqsn, *I_eps, *vcoul;
};
int gpp_forward_staged(bgw_ctx* ctx, int nb, int nvband, int ngpown, int ncouls, const GppHostArgs& h,
                       const GppDevBufs& d, double occ, double* d_ach, double* d_loc, double* d_achs,
                       double* d_asx);
int gpp_forward_device(bgw_ctx* ctx, size_t nb, size_t nvband, size_t ngpown, size_t ncouls,
                       const int* inv, const int* indinv, const double* wx, int wx_per_band,
                       const double* wtilde, const double* aqsm, const double* aqsn,
                       const double* I_eps, const double* vcoul, double occ, double* achtemp,
                       double* acht_n1_loc, double* achstemp, double* asxtemp);
int gpp_backward_device(bgw_ctx* ctx, size_t nb, size_t nvband, size_t ngpown, size_t ncouls,
                        const int* inv, const int* indinv, const double* wx,
                        const double* wtilde, const double* aqsm, const double* aqsn,
                        const double* I_eps, const double* vcoul, double occ,
                        double* dsigma, double* dwt, double* dI);

// FF: op 0 = achsDtemp, 1 = asxDtemp, 2 = achDtemp_cor, 3 = achDtemp (frequency sum,
// reads pref[nFreq]; I_epsA null means I_epsR holds R - A).  slice_map (device, ops 1
// and 2, nullable): the slice of I_epsR / I_epsA each plan job reads (asxDtemp: jobs
// [0, nFreq) on I_epsR, [nFreq, 2 nFreq) on I_epsA), for arrays that hold only the
// touched slices, packed; null: job j reads slice j % nFreq.
int ff_device(bgw_ctx* ctx, int op, int nb, int nvband, int nfreqeval, int ncouls, int ngpown,
              int nFreq, double freqevalmin, double freqevalstep, double occ, const double* ekq,
              const double* dFreqGrid, const int* inv, const int* indinv, const double* aqsm,
              const double* aqsn, const double* vcoul, const double* I_epsR,
              const double* I_epsA, double* out, const double* pref = nullptr,
              const int* slice_map = nullptr);
}  // namespace bgw
