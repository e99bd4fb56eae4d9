// contract.h -- arguments of the MFMA trace contraction and the helpers around it.
#pragma once

#include <hip/hip_runtime.h>

#include "bgw_internal.h"

namespace bgw {

constexpr int kModePlain = 0;  // B = X
constexpr int kModeDiff = 1;   // B = X - Y   (I_epsR - I_epsA, FF/ff.cpp:399-404)
constexpr int kModeMask = 2;   // B = X where |X|^2 > thr2, else 0

struct ContractArgs {
  const double2* A;        // [rows of A][K]   aqsntemp
  int K;                   // ncouls
  const double2* Xa;       // slices of X for jobs [0, jobs_a)
  const double2* Xb;       // slices of X for jobs [jobs_a, njobs) (nullable)
  const double2* Y;        // subtracted slices (nullable)
  long slice_stride;       // elements between frequency slices (ngpown*ncouls)
  int jobs_a;
  int N;                   // ngpown
  const int* slice_of_job; // [njobs]
  const int* rows;         // [njobs][max_rows] band indices
  const double2* wts;      // [njobs][max_rows] complex row weights
  const int* cnt;          // [njobs] active rows
  int max_rows;
  const int* inv;          // inv_igp_index
  const int* indinv;
  int gpp_index;           // 1: indinv has ncouls+1 entries (GPP); 0: FF
  const double2* aqsm;
  const double* vcoul;
  int ncouls;
  double thr2;             // mask threshold (kModeMask)
  int ksplit;              // set by contract_launch
  int tiles;               // set by contract_launch: row-tile slots per (G' tile, K split)
  int njobs;               // set by contract_launch
  const int* tile_start;   // set by contract_launch: [njobs+1] prefix of ceil(cnt/row tile)
  int items;               // set by contract_launch: workgroup tasks (= partials)
  int chunk;               // set by contract_launch: consecutive items kept on one XCD
  int xcd_ks;              // set by contract_launch: sparse plan dealt one K split per XCD
  int job_tiles;           // set by contract_launch: row tiles per job of a dense plan dealt
                           //   one K split per XCD (0: chunked order)
  double4* partial;        // set by contract_launch
  double2* out;            // c128[1] result (device)
};

// bm: the row tile, a value of contract_row_tile; max_tiles: a bound on
// sum_j ceil(cnt[j] / bm) known on the host, exact_tiles: whether the bound is met
// with equality (identity row lists), est_tiles: a tighter estimate of the live
// tiles, used to size the K split.
int contract_launch(bgw_ctx* ctx, ContractArgs a, int njobs, int bm, int max_tiles, bool exact_tiles,
                    int est_tiles, const char* tname);
// Rows (bands) per contraction tile.  Dense row lists (sparse = false) take the
// square 64-row tile.  A sparse plan (frequency-slice band lists) of jobs with at
// most max_job_rows rows takes the tall 80-row tile in FP64, and in the FP32 form
// (f32) when a job can hold more rows than the square tile.  In the tall tile a
// slice's band list fits one tile, so the slice streams once, and every wave has
// MFMA work even for short lists; the FP32 tall tile runs its k-steps two at a time
// (registers), which makes short FP32 lists faster in the square tile.
int contract_row_tile(bool f32, bool sparse, int max_job_rows);
// Sparse plans with at most kStreamRows rows per job (and no mask): the streaming
// form, FP64 in both precision forms (it is bound by reading the slices).  live_bound:
// a host bound on the jobs with rows; the grid covers that many jobs, not njobs.
constexpr int kStreamRows = 4;
int contract_stream(bgw_ctx* ctx, ContractArgs a, int njobs, int max_job_rows, int live_bound, const char* tname);

// rows[i] = i, wts[i] = w, cnt[0] = n, slice[0] = slice.
void launch_identity_rows(int* rows, double2* wts, int* cnt, int* slice, int n, double2 w, int slice_idx,
                          hipStream_t s);
// out[k] = sum_w partial[w*nout + k] in fixed order (double-double), k < nout.
void launch_reduce_cdd(const double4* partial, int W, int nout, double2* out, hipStream_t s);
// out[1..n-1] = out[0]
void launch_replicate(double2* out, int n, hipStream_t s);

}  // namespace bgw
