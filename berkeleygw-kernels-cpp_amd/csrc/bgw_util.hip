// bgw_util.hip -- small device utilities: graph-safe zero fill, error flag, index check.
#include "bgw_dd.h"
#include "bgw_internal.h"
#include "contract.h"

namespace {

__global__ void k_fill_zero(uint4* p, size_t n16) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_uint4(0, 0, 0, 0);
}

__global__ void k_fill_zero_bytes(unsigned char* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

__global__ void k_reset_flag(int* f) {
  if (threadIdx.x == 0) f[0] = 0;
}

// inv_igp_index[p] in [0, inv_max]; indinv[j] in [0, n_indinv_max) where the
// caller passes n_indinv entries and the valid targets are [0, ncouls).
__global__ void k_check_indices(const int* __restrict__ inv, int ngpown, const int* __restrict__ indinv,
                                int n_indinv, int inv_max, int ncouls, int* flag) {
  int bad = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ngpown + n_indinv;
       i += gridDim.x * blockDim.x) {
    if (i < ngpown) {
      int v = inv[i];
      bad |= (v < 0 || v > inv_max);
    } else {
      int v = indinv[i - ngpown];
      bad |= (v < 0 || v >= ncouls);
    }
  }
  if (bad) atomicOr(flag, 1);
}


__global__ void k_reduce_cdd(const double4* __restrict__ partial, int W, int nout, double2* __restrict__ out) {
  __shared__ bgw::cdd sh[16];
  for (int k = 0; k < nout; ++k) {
    bgw::cdd a = bgw::cdd_zero();
    for (int w = threadIdx.x; w < W; w += blockDim.x) a = bgw::cdd_add(a, bgw::cdd_load(&partial[(size_t)w * nout + k]));
    bgw::cdd r = bgw::cdd_block_reduce(a, sh);
    if (threadIdx.x == 0) out[k] = make_double2(r.re.hi + r.re.lo, r.im.hi + r.im.lo);
    __syncthreads();
  }
}

__global__ void k_identity_rows(int* rows, double2* wts, int* cnt, int* slice, int n, double2 w, int slice_idx) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    rows[i] = i;
    wts[i] = w;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    cnt[0] = n;
    slice[0] = slice_idx;
  }
}

__global__ void k_replicate(double2* out, int n) {
  for (int i = 1 + threadIdx.x; i < n; i += blockDim.x) out[i] = out[0];
}

}  // namespace

namespace bgw {

void fill_zero(void* p, size_t bytes, hipStream_t s) {
  if (!p || !bytes) return;
  if ((reinterpret_cast<uintptr_t>(p) & 15) == 0 && (bytes & 15) == 0) {
    size_t n16 = bytes / 16;
    int blocks = (int)((n16 + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, s, static_cast<uint4*>(p), n16);
  } else {
    int blocks = (int)((bytes + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_zero_bytes, dim3(blocks), dim3(256), 0, s,
                       static_cast<unsigned char*>(p), bytes);
  }
}

void reset_flag(int* d_flag, hipStream_t s) {
  hipLaunchKernelGGL(k_reset_flag, dim3(1), dim3(64), 0, s, d_flag);
}

void launch_check_indices(const int* inv, size_t ngpown, const int* indinv, size_t n_indinv,
                          int inv_max, int ncouls, int* flag, hipStream_t s) {
  int total = (int)(ngpown + n_indinv);
  int blocks = (total + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_check_indices, dim3(blocks), dim3(256), 0, s, inv, (int)ngpown, indinv,
                     (int)n_indinv, inv_max, ncouls, flag);
}

void launch_reduce_cdd(const double4* partial, int W, int nout, double2* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_cdd, dim3(1), dim3(256), 0, s, partial, W, nout, out);
}

void launch_identity_rows(int* rows, double2* wts, int* cnt, int* slice, int n, double2 w, int slice_idx,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_identity_rows, dim3((n + 255) / 256), dim3(256), 0, s, rows, wts, cnt, slice, n, w,
                     slice_idx);
}

void launch_replicate(double2* out, int n, hipStream_t s) {
  if (n > 1) hipLaunchKernelGGL(k_replicate, dim3(1), dim3(256), 0, s, out, n);
}

}  // namespace bgw
