// ff.hip -- FF full-frequency kernels (FF/ff.cpp:103-451) on the MFMA contraction.
//
// Each reference kernel is a sum over bands of an interpolated frequency slice
// contracted with aqsntemp over ig and weighted by vcoul*conj(aqsm) over my_igp.
// A one-block plan kernel evaluates compute_fact per band on the device
// (FF/ff.cpp:245-252, :337-342) and lists, for every frequency slice, the bands
// that use it with their interpolation weight; the MFMA contraction then reads
// each touched slice once for all bands of its list.  Because wx has no iw
// dependence in the reference (FF/ff.cpp:245,337) each output is computed once
// and written to all nfreqeval entries.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "bgw_fact.h"
#include "bgw_internal.h"
#include "contract.h"
#include "gpp_common.h"

namespace bgw {
namespace {

// op 1 (asxDtemp): bands n1 < nvband, loop 1, jobs [0,nFreq) on I_epsR, [nFreq,2nFreq) on I_epsA.
// op 2 (achDtemp_cor): all bands, loop 2, jobs [0,nFreq) on I_epsR - I_epsA;
//   wx>0 & unoccupied -> schDttt_corKernel1 (weights fact1,fact2),
//   wx<=0 & occupied  -> schDttt_corKernel2 (weights -1/2 fact1, -1/2 fact2).
__global__ __launch_bounds__(256) void k_ff_plan(int op, int nb, int nvband, int nFreq, double freqevalmin,
                                                 double freqevalstep, double occ, const double* __restrict__ ekq,
                                                 const double* __restrict__ grid, int* __restrict__ rows,
                                                 double2* __restrict__ wts, int* __restrict__ cnt,
                                                 int* __restrict__ slice_of_job, int* __restrict__ band_job,
                                                 double2* __restrict__ band_w, const int* __restrict__ slice_map) {
  const int njobs = op == 1 ? 2 * nFreq : nFreq;
  const int nbands = op == 1 ? nvband : nb;
  // 1) compute_fact for every band in parallel: band_job = (first job, or -1), band_w = (w1, w2)
  for (int n1 = threadIdx.x; n1 < nbands; n1 += blockDim.x) {
    const double wx = freqevalmin - ekq[n1] + freqevalstep;
    double f1 = 0.0, f2 = 0.0;
    int ifreq = 0;
    int job = -1;
    double scale = 0.0;
    if (op == 1) {
      compute_fact(wx, nFreq, grid, f1, f2, ifreq, 1, false);
      scale = occ;
      job = (wx > 0 ? 0 : nFreq) + ifreq;
    } else {
      const bool occd = n1 < nvband;
      compute_fact(wx, nFreq, grid, f1, f2, ifreq, 2, occd);
      if (wx > 0 && !occd) {
        scale = 1.0;
        job = ifreq;
      } else if (!(wx > 0) && occd) {
        scale = -0.5;
        job = ifreq;
      }
    }
    band_job[n1] = job;
    band_w[n1] = make_double2(scale * f1, scale * f2);
  }
  __syncthreads();
  // 2) one wave per job.  compute_fact keeps ifreq in [0, nFreq-2], so a band with
  //    first job ja is a row of ja (weight fact1) and of ja+1 (weight fact2), both in
  //    the same half of the asxDtemp job list.  Job j's rows are therefore the bands
  //    with band_job == j or band_job == j-1, appended in ascending band order (a
  //    fixed order, so the contraction's summation order is reproducible).
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  for (int j = wave; j < njobs; j += nwaves) {
    int c = 0;
    for (int base = 0; base < nbands; base += 64) {
      const int n1 = base + lane;
      const int ja = n1 < nbands ? band_job[n1] : -1;
      const bool first = ja == j, second = ja >= 0 && ja + 1 == j;
      const unsigned long long m = __ballot(first || second);
      if (first || second) {
        const int pos = c + __popcll(m & ((1ull << lane) - 1ull));
        const double2 w = band_w[n1];
        rows[(size_t)j * nb + pos] = n1;
        wts[(size_t)j * nb + pos] = make_double2(first ? w.x : w.y, 0.0);
      }
      c += __popcll(m);
    }
    if (lane == 0) {
      cnt[j] = c;
      slice_of_job[j] = slice_map ? slice_map[j] : j % nFreq;
    }
  }
}

// op 3 (achDtemp, the real-axis frequency sum): job f = slice f of R - A, rows = every
// band in ascending order, weight w(n1,f) = -i/2 pref(f) / (wx(n1) - s(n1) dFreqGrid(f)),
// s = -1 for occupied bands (n1 < nvband), +1 otherwise; zero where the denominator is
// exactly zero.  One block per slice.
__global__ __launch_bounds__(256) void k_ff_freq_plan(int nb, int nvband, double freqevalmin, double freqevalstep,
                                                      const double* __restrict__ ekq,
                                                      const double* __restrict__ grid,
                                                      const double* __restrict__ pref, int* __restrict__ rows,
                                                      double2* __restrict__ wts, int* __restrict__ cnt,
                                                      int* __restrict__ slice_of_job) {
  const int f = blockIdx.x;
  const double wf = grid[f], pf = pref[f];
  for (int n1 = threadIdx.x; n1 < nb; n1 += blockDim.x) {
    const double wx = freqevalmin - ekq[n1] + freqevalstep;
    const double den = n1 < nvband ? wx + wf : wx - wf;
    rows[(size_t)f * nb + n1] = n1;
    wts[(size_t)f * nb + n1] = make_double2(0.0, den == 0.0 ? 0.0 : -0.5 * (pf / den));
  }
  if (threadIdx.x == 0) {
    cnt[f] = nb;
    slice_of_job[f] = f;
  }
}

// One band's term of achDtemp_cor (schDttt_corKernel1/2, FF/ff.cpp:381-451) as two
// one-row jobs: band n1 on slice ifreq with weight w1 and on slice ifreq + 1 with w2.
__global__ void k_cor_band_plan(int* __restrict__ rows, double2* __restrict__ wts, int* __restrict__ cnt,
                                int* __restrict__ slice_of_job, int n1, int ifreq, double w1, double w2) {
  if (threadIdx.x < 2) {
    const int j = threadIdx.x;
    rows[j] = n1;
    wts[j] = make_double2(j == 0 ? w1 : w2, 0.0);
    cnt[j] = 1;
    slice_of_job[j] = ifreq + j;
  }
}

}  // namespace

int ff_device(bgw_ctx* ctx, int op, int nb, int nvband, int nfreqeval, int ncouls, int ngpown, int nFreq,
              double freqevalmin, double freqevalstep, double occ, const double* ekq, const double* dFreqGrid,
              const int* inv, const int* indinv, const double* aqsm, const double* aqsn, const double* vcoul,
              const double* I_epsR, const double* I_epsA, double* out, const double* pref,
              const int* slice_map) {
  hipStream_t s = ctx->stream;
  launch_check_indices(inv, (size_t)ngpown, indinv, (size_t)ncouls, ncouls - 1, ncouls, ctx->d_flag, s);
  ContractArgs a{};
  a.A = reinterpret_cast<const double2*>(aqsn);
  a.K = ncouls;
  a.N = ngpown;
  a.slice_stride = (long)ngpown * ncouls;
  a.inv = inv;
  a.indinv = indinv;
  a.gpp_index = 0;
  a.aqsm = reinterpret_cast<const double2*>(aqsm);
  a.vcoul = vcoul;
  a.ncouls = ncouls;
  a.thr2 = 0.0;
  a.out = reinterpret_cast<double2*>(out);
  const bool f32 = ctx->precision == BGW_FP32_BULK;
  void *prow, *pw, *pcnt, *pslice;
  const int njobs = op == 0 ? 1 : (op == 1 ? 2 * nFreq : nFreq);  // op 2 and 3: one job per slice
  const int maxrows = op == 0 ? nb : (op == 1 ? nvband : nb);
  int rc = pool_get(ctx, "ws_ff_rows", (size_t)njobs * nb * sizeof(int), &prow);
  if (!rc) rc = pool_get(ctx, "ws_ff_wts", (size_t)njobs * nb * 16, &pw);
  if (!rc) rc = pool_get(ctx, "ws_ff_cnt", (size_t)njobs * sizeof(int), &pcnt);
  if (!rc) rc = pool_get(ctx, "ws_ff_slice", (size_t)njobs * sizeof(int), &pslice);
  if (rc) return rc;
  a.rows = static_cast<const int*>(prow);
  a.wts = static_cast<const double2*>(pw);
  a.cnt = static_cast<const int*>(pcnt);
  a.slice_of_job = static_cast<const int*>(pslice);
  a.max_rows = nb;
  if (op == 0) {
    // achsDtemp: slice 1 of I_epsR (FF/ff.cpp:200), every band, weight -1/2
    launch_identity_rows(static_cast<int*>(prow), static_cast<double2*>(pw), static_cast<int*>(pcnt),
                         static_cast<int*>(pslice), nb, make_double2(-0.5, 0.0), 1, s);
    a.Xa = reinterpret_cast<const double2*>(I_epsR);
    a.jobs_a = 1;
    const int BMd = contract_row_tile(f32, false, nb);
    const int tiles = (nb + BMd - 1) / BMd;
    return contract_launch(ctx, a, 1, BMd, tiles, true, tiles, "contract_ff_achs");
  }
  if (op == 3) {
    // the frequency sum: every band on every slice of R - A (dense row lists).  With
    // I_epsA null, I_epsR already holds R - A (the host path forms it while staging).
    hipLaunchKernelGGL(k_ff_freq_plan, dim3(nFreq), dim3(256), 0, s, nb, nvband, freqevalmin, freqevalstep, ekq,
                       dFreqGrid, pref, static_cast<int*>(prow), static_cast<double2*>(pw), static_cast<int*>(pcnt),
                       static_cast<int*>(pslice));
    if ((rc = check_launch(ctx, "k_ff_freq_plan"))) return rc;
    a.Xa = reinterpret_cast<const double2*>(I_epsR);
    a.Y = reinterpret_cast<const double2*>(I_epsA);
    a.jobs_a = nFreq;
    const int BMd = contract_row_tile(f32, false, nb);
    const long long tiles = (long long)nFreq * ((nb + BMd - 1) / BMd);
    if (tiles > INT32_MAX) return fail(ctx, BGW_EINVAL, "frequency sum too large");
    rc = contract_launch(ctx, a, nFreq, BMd, (int)tiles, true, (int)tiles, "contract_ff_achD");
    if (rc) return rc;
    launch_replicate(a.out, nfreqeval, s);
    return check_launch(ctx, "replicate");
  }
  void *pbj, *pbw;
  if ((rc = pool_get(ctx, "ws_ff_band_job", (size_t)nb * sizeof(int), &pbj))) return rc;
  if ((rc = pool_get(ctx, "ws_ff_band_w", (size_t)nb * 16, &pbw))) return rc;
  hipLaunchKernelGGL(k_ff_plan, dim3(1), dim3(256), 0, s, op, nb, nvband, nFreq, freqevalmin, freqevalstep, occ,
                     ekq, dFreqGrid, static_cast<int*>(prow), static_cast<double2*>(pw), static_cast<int*>(pcnt),
                     static_cast<int*>(pslice), static_cast<int*>(pbj), static_cast<double2*>(pbw), slice_map);
  if ((rc = check_launch(ctx, "k_ff_plan"))) return rc;
  // every listed band (at most maxrows of them) is a row of exactly two jobs, so at most
  // min(njobs, 2*maxrows) jobs have rows and sum_j ceil(cnt_j/BM) is at most that plus
  // ceil(2*maxrows/BM), and never above njobs * ceil(maxrows/BM)
  const int BMr = contract_row_tile(f32, true, maxrows);
  const long long live_jobs = std::min<long long>(njobs, 2LL * maxrows);
  const long long tiles_dense = (long long)njobs * ((maxrows + BMr - 1) / BMr);
  const long long tiles_rows = live_jobs + (2LL * maxrows + BMr - 1) / BMr;
  const int max_tiles = (int)(tiles_dense < tiles_rows ? tiles_dense : tiles_rows);
  // live tiles: at most one per listed row when rows are few (2*maxrows rows in all)
  const int est_tiles = (int)std::min<long long>(max_tiles, 2LL * maxrows);
  if (op == 1) {
    a.Xa = reinterpret_cast<const double2*>(I_epsR);
    a.Xb = reinterpret_cast<const double2*>(I_epsA);
    a.jobs_a = nFreq;
    // a job's rows are valence bands: few of them make the slices' reading the bound
    if (nvband <= kStreamRows)
      rc = contract_stream(ctx, a, njobs, nvband, std::min(njobs, 2 * nvband), "contract_ff_asx");
    else
      rc = contract_launch(ctx, a, njobs, BMr, max_tiles, false, est_tiles, "contract_ff_asx");
  } else {
    a.Xa = reinterpret_cast<const double2*>(I_epsR);
    a.Y = reinterpret_cast<const double2*>(I_epsA);
    a.jobs_a = nFreq;
    rc = contract_launch(ctx, a, njobs, BMr, max_tiles, false, est_tiles, "contract_ff_cor");
  }
  if (rc) return rc;
  launch_replicate(a.out, nfreqeval, s);
  return check_launch(ctx, "replicate");
}

// schDttt_corKernel1 (scale 1) / schDttt_corKernel2 (scale -1/2) on device arrays:
// sum_{p,g} aqsn(n1,g) conj(aqsm(n1,igp)) vcoul(igp) scale (D(ifreq) fact1 + D(ifreq+1) fact2),
// D = I_epsR - I_epsA (or D = I_epsR when I_epsA is null), on the MFMA contraction.
int cor_band_device(bgw_ctx* ctx, double scale, int ncouls, int ngpown, int ifreq, int n1, double fact1,
                    double fact2, const int* inv, const int* indinv, const double* I_epsR, const double* I_epsA,
                    const double* aqsm, const double* aqsn, const double* vcoul, double* out) {
  hipStream_t s = ctx->stream;
  launch_check_indices(inv, (size_t)ngpown, indinv, (size_t)ncouls, ncouls - 1, ncouls, ctx->d_flag, s);
  void *prow, *pw, *pcnt, *pslice;
  int rc = pool_get(ctx, "ws_cb_rows", 2 * sizeof(int), &prow);
  if (!rc) rc = pool_get(ctx, "ws_cb_wts", 2 * 16, &pw);
  if (!rc) rc = pool_get(ctx, "ws_cb_cnt", 2 * sizeof(int), &pcnt);
  if (!rc) rc = pool_get(ctx, "ws_cb_slice", 2 * sizeof(int), &pslice);
  if (rc) return rc;
  hipLaunchKernelGGL(k_cor_band_plan, dim3(1), dim3(64), 0, s, static_cast<int*>(prow), static_cast<double2*>(pw),
                     static_cast<int*>(pcnt), static_cast<int*>(pslice), n1, ifreq, scale * fact1, scale * fact2);
  if ((rc = check_launch(ctx, "k_cor_band_plan"))) return rc;
  ContractArgs a{};
  a.A = reinterpret_cast<const double2*>(aqsn);
  a.K = ncouls;
  a.N = ngpown;
  a.slice_stride = (long)ngpown * ncouls;
  a.Xa = reinterpret_cast<const double2*>(I_epsR);
  a.Y = reinterpret_cast<const double2*>(I_epsA);
  a.jobs_a = 2;
  a.slice_of_job = static_cast<const int*>(pslice);
  a.rows = static_cast<const int*>(prow);
  a.wts = static_cast<const double2*>(pw);
  a.cnt = static_cast<const int*>(pcnt);
  a.max_rows = 1;
  a.inv = inv;
  a.indinv = indinv;
  a.gpp_index = 0;
  a.aqsm = reinterpret_cast<const double2*>(aqsm);
  a.vcoul = vcoul;
  a.ncouls = ncouls;
  a.thr2 = 0.0;
  a.out = reinterpret_cast<double2*>(out);
  return contract_stream(ctx, a, 2, 1, 2, "contract_ff_corband");
}

}  // namespace bgw

// ===================================================================== C ABI
using namespace bgw;

namespace {

int ff_validate(bgw_ctx* ctx, int op, int nb, int nvband, int nfreqeval, int ncouls, int ngpown, int nFreq,
                std::initializer_list<const void*> ptrs) {
  if (!ctx) return BGW_EINVAL;
  for (const void* p : ptrs)
    if (!p) return fail(ctx, BGW_EINVAL, "null array argument");
  if (nb <= 0 || ncouls <= 0 || ngpown <= 0) return fail(ctx, BGW_EINVAL, "zero or negative size");
  if (ngpown > ncouls) return fail(ctx, BGW_EINVAL, "ngpown cannot be greater than ncouls");  // FF/ff.cpp:492-499
  if (op != 0 && (nFreq < 2 || nfreqeval <= 0)) return fail(ctx, BGW_EINVAL, "nFreq < 2 or nfreqeval <= 0");
  if (op != 0 && (nvband < 0 || nvband > nb)) return fail(ctx, BGW_EINVAL, "nvband out of [0, number_bands]");
  size_t t;
  if (!mul_ok((size_t)ngpown, (size_t)ncouls, &t) || !mul_ok(t, (size_t)(nFreq > 2 ? nFreq : 2), &t) ||
      t > SIZE_MAX / 16)
    return fail(ctx, BGW_EINVAL, "size overflow");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, BGW_ENODEV, "hipSetDevice");
  return BGW_OK;
}

int ff_host_indices(bgw_ctx* ctx, int ngpown, int ncouls, const int* inv, const int* indinv) {
  for (int p = 0; p < ngpown; ++p)
    if (inv[p] < 0 || inv[p] >= ncouls) return fail(ctx, BGW_EINVAL, "inv_igp_index out of [0,ncouls)");
  for (int g = 0; g < ncouls; ++g)
    if (indinv[g] < 0 || indinv[g] >= ncouls) return fail(ctx, BGW_EINVAL, "indinv out of [0,ncouls)");
  return BGW_OK;
}

template <typename T>
int upload(bgw_ctx* ctx, const char* slot, const T* src, size_t count, const T** dst) {
  void* d;
  int rc = pool_get(ctx, slot, count * sizeof(T), &d);
  if (rc) return rc;
  if ((rc = upload_host(ctx, d, src, count * sizeof(T)))) return rc;
  *dst = static_cast<const T*>(d);
  return BGW_OK;
}

int ff_host(bgw_ctx* ctx, int op, int nb, int nvband, int nfreqeval, int ncouls, int ngpown, int nFreq,
            double fmin, double fstep, double occ, const double* ekq, const double* grid, const int* inv,
            const int* indinv, const double* aqsm, const double* aqsn, const double* vcoul,
            const double* R, const double* A, double* out, const double* pref = nullptr) {
  int rc = ff_host_indices(ctx, ngpown, ncouls, inv, indinv);
  if (rc) return rc;
  const size_t npg = (size_t)ngpown * ncouls;
  const int *d_inv, *d_indinv;
  const double *d_ekq = nullptr, *d_grid = nullptr, *d_pref = nullptr, *d_am, *d_an, *d_vc, *d_R = nullptr,
               *d_A = nullptr;
  const int* d_map = nullptr;
  std::vector<int> map;  // host copy of the packed-slice map
  if ((rc = upload(ctx, "h_inv", inv, (size_t)ngpown, &d_inv))) return rc;
  if ((rc = upload(ctx, "h_indinv", indinv, (size_t)ncouls, &d_indinv))) return rc;
  // asxDtemp reads only the valence rows n1 < nvband of aqsm / aqsn (its row lists
  // hold those bands only); the other kernels read every band
  const size_t arows = op == 1 ? (size_t)nvband : (size_t)nb;
  if ((rc = upload(ctx, "h_aqsm", aqsm, 2 * arows * ncouls, &d_am))) return rc;
  if ((rc = upload(ctx, "h_aqsn", aqsn, 2 * arows * ncouls, &d_an))) return rc;
  if ((rc = upload(ctx, "h_vcoul", vcoul, (size_t)ncouls, &d_vc))) return rc;
  if (op == 0) {
    // only slice 1 of I_epsR is read (FF/ff.cpp:200): stage it as slice 1 of a 2-slice array
    void* d;
    if ((rc = pool_get(ctx, "h_epsR", 2 * npg * 16, &d))) return rc;
    if ((rc = upload_host(ctx, static_cast<char*>(d) + npg * 16, R + 2 * npg, npg * 16))) return rc;
    d_R = static_cast<const double*>(d);
  } else {
    if ((rc = upload(ctx, "h_ekq", ekq, (size_t)(op == 1 ? nvband : nb), &d_ekq))) return rc;
    if ((rc = upload(ctx, "h_grid", grid, (size_t)nFreq, &d_grid))) return rc;
    if (op == 3 && (rc = upload(ctx, "h_pref", pref, (size_t)nFreq, &d_pref))) return rc;
    // Only the slices the device plan (k_ff_plan) will read are transferred.  The
    // slice of a band is ifreq (and ifreq + 1) of compute_fact, the same function
    // on host and device, and wx is formed with the same two exact-order operations,
    // so the sets agree.  The frequency sum (op 3) reads every slice.
    std::vector<char> needR(nFreq, op == 3), needA(nFreq, op == 3);
    const int nbands = op == 1 ? nvband : (op == 2 ? nb : 0);
    for (int n1 = 0; n1 < nbands; ++n1) {
      const double wx = fmin - ekq[n1] + fstep;
      double f1 = 0.0, f2 = 0.0;
      int ifreq = 0;
      if (op == 1) {
        compute_fact(wx, nFreq, grid, f1, f2, ifreq, 1, false);
        std::vector<char>& need = wx > 0 ? needR : needA;
        need[ifreq] = need[ifreq + 1] = 1;
      } else {
        const bool occd = n1 < nvband;
        compute_fact(wx, nFreq, grid, f1, f2, ifreq, 2, occd);
        if ((wx > 0 && !occd) || (!(wx > 0) && occd)) needR[ifreq] = needR[ifreq + 1] = needA[ifreq] = needA[ifreq + 1] = 1;
      }
    }
    void *pR, *pA = nullptr;
    const size_t slice = npg * 16;
    if (op == 3) {
      // the frequency sum reads every slice, as I_epsR - I_epsA (FF/ff.cpp:436-441 form):
      // the host forms the difference while staging, so half the bytes cross the bus, and
      // the contraction runs on it in its plain mode (the same IEEE subtraction as on the
      // device, so the result equals the device entry point's bit for bit)
      if ((rc = pool_get(ctx, "h_epsR", slice * nFreq, &pR))) return rc;
      for (int f = 0; f < nFreq; ++f)
        if ((rc = upload_host_diff(ctx, static_cast<double*>(pR) + 2 * npg * f, R + 2 * npg * f, A + 2 * npg * f,
                                   2 * npg)))
          return rc;
    } else {
      // The touched slices are staged packed, in ascending order, so the device holds
      // only them; slice_map gives every plan job its packed slice (a job whose slice is
      // untouched has no rows and reads nothing).  asxDtemp stages R and A slices apart;
      // achDtemp_cor reads its slices only as I_epsR - I_epsA (FF/ff.cpp:399-404), which
      // the host forms while staging, as for the frequency sum.
      map.assign(op == 1 ? 2 * (size_t)nFreq : (size_t)nFreq, 0);
      int nR = 0, nA = 0;
      for (int f = 0; f < nFreq; ++f)
        if (needR[f]) map[f] = nR++;
      if (op == 1)
        for (int f = 0; f < nFreq; ++f)
          if (needA[f]) map[nFreq + f] = nA++;
      if ((rc = pool_get(ctx, "h_epsR", slice * std::max(nR, 1), &pR))) return rc;
      if (op == 1 && (rc = pool_get(ctx, "h_epsA", slice * std::max(nA, 1), &pA))) return rc;
      for (int f = 0; f < nFreq; ++f) {
        if (!needR[f]) continue;
        char* dst = static_cast<char*>(pR) + map[f] * slice;
        rc = op == 2 ? upload_host_diff(ctx, reinterpret_cast<double*>(dst), R + 2 * npg * f, A + 2 * npg * f, 2 * npg)
                     : upload_host(ctx, dst, R + 2 * npg * f, slice);
        if (rc) return rc;
      }
      if (op == 1)
        for (int f = 0; f < nFreq; ++f)
          if (needA[f] &&
              (rc = upload_host(ctx, static_cast<char*>(pA) + map[nFreq + f] * slice, A + 2 * npg * f, slice)))
            return rc;
      if ((rc = upload(ctx, "h_smap", map.data(), map.size(), &d_map))) return rc;
      // the map (4 B per plan job) comes from a host vector: wait for its copy here, so no later
      // early return can leave the copy reading freed memory
      const hipError_t e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) return hip_fail(ctx, e, "slice map upload");
    }
    d_R = static_cast<const double*>(pR);
    d_A = static_cast<const double*>(pA);  // null for achDtemp_cor and achDtemp: d_R holds R - A
  }
  void* d_out;
  if ((rc = pool_get(ctx, "h_ffout", (size_t)(nfreqeval > 1 ? nfreqeval : 1) * 16, &d_out))) return rc;
  rc = ff_device(ctx, op, nb, nvband, nfreqeval, ncouls, ngpown, nFreq, fmin, fstep, occ, d_ekq, d_grid,
                 d_inv, d_indinv, d_am, d_an, d_vc, d_R, d_A, static_cast<double*>(d_out), d_pref, d_map);
  if (rc) {
    (void)hipStreamSynchronize(ctx->stream);
    return rc;
  }
  const size_t nout = op == 0 ? 1 : (size_t)nfreqeval;
  hipError_t e = hipMemcpyAsync(out, d_out, nout * 16, hipMemcpyDeviceToHost, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "download");
  return bgw_synchronize(ctx);
}

// schDttt_corKernel1/2 arguments: sizes as FF (ngpown <= ncouls, FF/ff.cpp:492-499),
// ifreq >= 0 and n1 >= 0 (the arrays must hold slices ifreq, ifreq + 1 and row n1, as
// the reference reads them), and the element count of the slices read within size_t.
int cor_band_validate(bgw_ctx* ctx, int ncouls, int ngpown, int ifreq, int n1,
                      std::initializer_list<const void*> ptrs) {
  if (!ctx) return BGW_EINVAL;
  for (const void* p : ptrs)
    if (!p) return fail(ctx, BGW_EINVAL, "null array argument");
  if (ncouls <= 0 || ngpown <= 0) return fail(ctx, BGW_EINVAL, "zero or negative size");
  if (ngpown > ncouls) return fail(ctx, BGW_EINVAL, "ngpown cannot be greater than ncouls");
  if (ifreq < 0 || n1 < 0) return fail(ctx, BGW_EINVAL, "negative ifreq or n1");
  size_t t;
  if (!mul_ok((size_t)ngpown, (size_t)ncouls, &t) || !mul_ok(t, (size_t)ifreq + 2, &t) || t > SIZE_MAX / 16 ||
      !mul_ok((size_t)n1 + 1, (size_t)ncouls, &t) || t > SIZE_MAX / 16)
    return fail(ctx, BGW_EINVAL, "size overflow");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, BGW_ENODEV, "hipSetDevice");
  return BGW_OK;
}

// Host arrays: transfers slices ifreq and ifreq + 1 of I_epsR and I_epsA, row n1 of
// aqsmtemp / aqsntemp and the index arrays, then runs cor_band_device on them.
int cor_band_host(bgw_ctx* ctx, double scale, int ncouls, int ngpown, int ifreq, int n1, double fact1, double fact2,
                  const int* inv, const int* indinv, const double* R, const double* A, const double* aqsm,
                  const double* aqsn, const double* vcoul, double* out) {
  int rc = ff_host_indices(ctx, ngpown, ncouls, inv, indinv);
  if (rc) return rc;
  const size_t npg = (size_t)ngpown * ncouls;
  const int *d_inv, *d_indinv;
  const double *d_am, *d_an, *d_vc, *d_R, *d_A;
  if ((rc = upload(ctx, "h_inv", inv, (size_t)ngpown, &d_inv))) return rc;
  if ((rc = upload(ctx, "h_indinv", indinv, (size_t)ncouls, &d_indinv))) return rc;
  if ((rc = upload(ctx, "h_aqsm", aqsm + 2 * (size_t)n1 * ncouls, 2 * (size_t)ncouls, &d_am))) return rc;
  if ((rc = upload(ctx, "h_aqsn", aqsn + 2 * (size_t)n1 * ncouls, 2 * (size_t)ncouls, &d_an))) return rc;
  if ((rc = upload(ctx, "h_vcoul", vcoul, (size_t)ncouls, &d_vc))) return rc;
  // slices ifreq, ifreq + 1 of I_epsR - I_epsA, formed by the host while staging (the
  // kernels read the arrays only through that difference, FF/ff.cpp:399-404, :436-441)
  void* pD;
  if ((rc = pool_get(ctx, "h_epsR", 4 * npg * sizeof(double), &pD))) return rc;
  if ((rc = upload_host_diff(ctx, static_cast<double*>(pD), R + 2 * npg * (size_t)ifreq, A + 2 * npg * (size_t)ifreq,
                             4 * npg)))
    return rc;
  d_R = static_cast<const double*>(pD);
  d_A = nullptr;
  void* d_out;
  if ((rc = pool_get(ctx, "h_ffout", 16, &d_out))) return rc;
  // the staged arrays hold row n1 as row 0 and slices ifreq, ifreq + 1 of R - A as slices 0, 1
  rc = cor_band_device(ctx, scale, ncouls, ngpown, 0, 0, fact1, fact2, d_inv, d_indinv, d_R, d_A, d_am, d_an, d_vc,
                       static_cast<double*>(d_out));
  if (rc) {
    (void)hipStreamSynchronize(ctx->stream);
    return rc;
  }
  hipError_t e = hipMemcpyAsync(out, d_out, 16, hipMemcpyDeviceToHost, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "download");
  return bgw_synchronize(ctx);
}

}  // namespace

extern "C" {

int bgw_compute_fact(double wx, int nFreq, const double* dFreqGrid, double* fact1, double* fact2, int* ifreq,
                     int loop, int flag_occ) {
  if (!dFreqGrid || !fact1 || !fact2 || !ifreq || nFreq < 2) return BGW_EINVAL;
  return compute_fact(wx, nFreq, dFreqGrid, *fact1, *fact2, *ifreq, loop, flag_occ != 0) ? BGW_OK : BGW_EINVAL;
}

int bgw_achsDtemp_Kernel(bgw_ctx* ctx, int number_bands, int ngpown, int ncouls, const int* inv_igp_index,
                         const int* indinv, const double* aqsntemp, const double* aqsmtemp,
                         const double* I_epsR_array, const double* vcoul, double* achsDtemp) {
  int rc = ff_validate(ctx, 0, number_bands, 0, 1, ncouls, ngpown, 2,
                       {inv_igp_index, indinv, aqsntemp, aqsmtemp, I_epsR_array, vcoul, achsDtemp});
  if (rc) return rc;
  return ff_host(ctx, 0, number_bands, 0, 1, ncouls, ngpown, 2, 0.0, 0.0, 0.0, nullptr, nullptr, inv_igp_index,
                 indinv, aqsmtemp, aqsntemp, vcoul, I_epsR_array, nullptr, achsDtemp);
}

int bgw_asxDtemp_Kernel(bgw_ctx* ctx, int nvband, int nfreqeval, int ncouls, int ngpown, int nFreq,
                        double freqevalmin, double freqevalstep, double occ, const double* ekq,
                        const double* dFreqGrid, const int* inv_igp_index, const int* indinv,
                        const double* aqsmtemp, const double* aqsntemp, const double* vcoul,
                        const double* I_epsR_array, const double* I_epsA_array, double* asxDtemp) {
  // asxDtemp_Kernel reads bands n1 < nvband only (FF/ff.cpp:242): the band arrays
  // passed here need nvband rows.
  int rc = ff_validate(ctx, 1, nvband > 0 ? nvband : 1, nvband, nfreqeval, ncouls, ngpown, nFreq,
                       {ekq, dFreqGrid, inv_igp_index, indinv, aqsmtemp, aqsntemp, vcoul, I_epsR_array,
                        I_epsA_array, asxDtemp});
  if (rc) return rc;
  if (nvband == 0) {
    std::memset(asxDtemp, 0, sizeof(double) * 2 * nfreqeval);
    return BGW_OK;
  }
  return ff_host(ctx, 1, nvband, nvband, nfreqeval, ncouls, ngpown, nFreq, freqevalmin, freqevalstep, occ, ekq,
                 dFreqGrid, inv_igp_index, indinv, aqsmtemp, aqsntemp, vcoul, I_epsR_array, I_epsA_array,
                 asxDtemp);
}

int bgw_achDtemp_cor_Kernel(bgw_ctx* ctx, int number_bands, int nvband, int nfreqeval, int ncouls, int ngpown,
                            int nFreq, double freqevalmin, double freqevalstep, const double* ekq,
                            const double* dFreqGrid, const int* inv_igp_index, const int* indinv,
                            const double* aqsmtemp, const double* aqsntemp, const double* vcoul,
                            const double* I_epsR_array, const double* I_epsA_array, double* achDtemp_cor) {
  int rc = ff_validate(ctx, 2, number_bands, nvband, nfreqeval, ncouls, ngpown, nFreq,
                       {ekq, dFreqGrid, inv_igp_index, indinv, aqsmtemp, aqsntemp, vcoul, I_epsR_array,
                        I_epsA_array, achDtemp_cor});
  if (rc) return rc;
  return ff_host(ctx, 2, number_bands, nvband, nfreqeval, ncouls, ngpown, nFreq, freqevalmin, freqevalstep, 0.0,
                 ekq, dFreqGrid, inv_igp_index, indinv, aqsmtemp, aqsntemp, vcoul, I_epsR_array, I_epsA_array,
                 achDtemp_cor);
}

int bgw_achDtemp_Kernel(bgw_ctx* ctx, int number_bands, int nvband, int nfreqeval, int ncouls, int ngpown,
                        int nFreq, double freqevalmin, double freqevalstep, const double* ekq,
                        const double* dFreqGrid, const double* pref, const int* inv_igp_index, const int* indinv,
                        const double* aqsmtemp, const double* aqsntemp, const double* vcoul,
                        const double* I_epsR_array, const double* I_epsA_array, double* achDtemp) {
  int rc = ff_validate(ctx, 3, number_bands, nvband, nfreqeval, ncouls, ngpown, nFreq,
                       {ekq, dFreqGrid, pref, inv_igp_index, indinv, aqsmtemp, aqsntemp, vcoul, I_epsR_array,
                        I_epsA_array, achDtemp});
  if (rc) return rc;
  return ff_host(ctx, 3, number_bands, nvband, nfreqeval, ncouls, ngpown, nFreq, freqevalmin, freqevalstep, 0.0,
                 ekq, dFreqGrid, inv_igp_index, indinv, aqsmtemp, aqsntemp, vcoul, I_epsR_array, I_epsA_array,
                 achDtemp, pref);
}

int bgw_achDtemp_Kernel_device(bgw_ctx* ctx, int number_bands, int nvband, int nfreqeval, int ncouls, int ngpown,
                               int nFreq, double freqevalmin, double freqevalstep, const double* ekq,
                               const double* dFreqGrid, const double* pref, const int* inv_igp_index,
                               const int* indinv, const double* aqsmtemp, const double* aqsntemp,
                               const double* vcoul, const double* I_epsR_array, const double* I_epsA_array,
                               double* achDtemp) {
  int rc = ff_validate(ctx, 3, number_bands, nvband, nfreqeval, ncouls, ngpown, nFreq,
                       {ekq, dFreqGrid, pref, inv_igp_index, indinv, aqsmtemp, aqsntemp, vcoul, I_epsR_array,
                        I_epsA_array, achDtemp});
  if (rc) return rc;
  return ff_device(ctx, 3, number_bands, nvband, nfreqeval, ncouls, ngpown, nFreq, freqevalmin, freqevalstep, 0.0,
                   ekq, dFreqGrid, inv_igp_index, indinv, aqsmtemp, aqsntemp, vcoul, I_epsR_array, I_epsA_array,
                   achDtemp, pref);
}

int bgw_achsDtemp_Kernel_device(bgw_ctx* ctx, int number_bands, int ngpown, int ncouls, const int* inv_igp_index,
                                const int* indinv, const double* aqsntemp, const double* aqsmtemp,
                                const double* I_epsR_array, const double* vcoul, double* achsDtemp) {
  int rc = ff_validate(ctx, 0, number_bands, 0, 1, ncouls, ngpown, 2,
                       {inv_igp_index, indinv, aqsntemp, aqsmtemp, I_epsR_array, vcoul, achsDtemp});
  if (rc) return rc;
  return ff_device(ctx, 0, number_bands, 0, 1, ncouls, ngpown, 2, 0.0, 0.0, 0.0, nullptr, nullptr, inv_igp_index,
                   indinv, aqsmtemp, aqsntemp, vcoul, I_epsR_array, nullptr, achsDtemp);
}

int bgw_asxDtemp_Kernel_device(bgw_ctx* ctx, int nvband, int nfreqeval, int ncouls, int ngpown, int nFreq,
                               double freqevalmin, double freqevalstep, double occ, const double* ekq,
                               const double* dFreqGrid, const int* inv_igp_index, const int* indinv,
                               const double* aqsmtemp, const double* aqsntemp, const double* vcoul,
                               const double* I_epsR_array, const double* I_epsA_array, double* asxDtemp) {
  int rc = ff_validate(ctx, 1, nvband > 0 ? nvband : 1, nvband, nfreqeval, ncouls, ngpown, nFreq,
                       {ekq, dFreqGrid, inv_igp_index, indinv, aqsmtemp, aqsntemp, vcoul, I_epsR_array,
                        I_epsA_array, asxDtemp});
  if (rc) return rc;
  if (nvband == 0) {
    fill_zero(asxDtemp, sizeof(double) * 2 * nfreqeval, ctx->stream);
    return check_launch(ctx, "fill");
  }
  return ff_device(ctx, 1, nvband, nvband, nfreqeval, ncouls, ngpown, nFreq, freqevalmin, freqevalstep, occ, ekq,
                   dFreqGrid, inv_igp_index, indinv, aqsmtemp, aqsntemp, vcoul, I_epsR_array, I_epsA_array,
                   asxDtemp);
}

int bgw_achDtemp_cor_Kernel_device(bgw_ctx* ctx, int number_bands, int nvband, int nfreqeval, int ncouls,
                                   int ngpown, int nFreq, double freqevalmin, double freqevalstep,
                                   const double* ekq, const double* dFreqGrid, const int* inv_igp_index,
                                   const int* indinv, const double* aqsmtemp, const double* aqsntemp,
                                   const double* vcoul, const double* I_epsR_array, const double* I_epsA_array,
                                   double* achDtemp_cor) {
  int rc = ff_validate(ctx, 2, number_bands, nvband, nfreqeval, ncouls, ngpown, nFreq,
                       {ekq, dFreqGrid, inv_igp_index, indinv, aqsmtemp, aqsntemp, vcoul, I_epsR_array,
                        I_epsA_array, achDtemp_cor});
  if (rc) return rc;
  return ff_device(ctx, 2, number_bands, nvband, nfreqeval, ncouls, ngpown, nFreq, freqevalmin, freqevalstep, 0.0,
                   ekq, dFreqGrid, inv_igp_index, indinv, aqsmtemp, aqsntemp, vcoul, I_epsR_array, I_epsA_array,
                   achDtemp_cor);
}

int bgw_schDttt_corKernel1(bgw_ctx* ctx, double* schDttt_cor, const int* inv_igp_index, const int* indinv,
                           const double* I_epsR_array, const double* I_epsA_array, const double* aqsmtemp,
                           const double* aqsntemp, const double* vcoul, int ncouls, int ifreq, int ngpown, int n1,
                           double fact1, double fact2) {
  int rc = cor_band_validate(ctx, ncouls, ngpown, ifreq, n1,
                             {schDttt_cor, inv_igp_index, indinv, I_epsR_array, I_epsA_array, aqsmtemp, aqsntemp,
                              vcoul});
  if (rc) return rc;
  return cor_band_host(ctx, 1.0, ncouls, ngpown, ifreq, n1, fact1, fact2, inv_igp_index, indinv, I_epsR_array,
                       I_epsA_array, aqsmtemp, aqsntemp, vcoul, schDttt_cor);
}

int bgw_schDttt_corKernel2(bgw_ctx* ctx, double* schDttt_cor, const int* inv_igp_index, const int* indinv,
                           const double* I_epsR_array, const double* I_epsA_array, const double* aqsmtemp,
                           const double* aqsntemp, const double* vcoul, int ncouls, int ifreq, int ngpown, int n1,
                           double fact1, double fact2) {
  int rc = cor_band_validate(ctx, ncouls, ngpown, ifreq, n1,
                             {schDttt_cor, inv_igp_index, indinv, I_epsR_array, I_epsA_array, aqsmtemp, aqsntemp,
                              vcoul});
  if (rc) return rc;
  return cor_band_host(ctx, -0.5, ncouls, ngpown, ifreq, n1, fact1, fact2, inv_igp_index, indinv, I_epsR_array,
                       I_epsA_array, aqsmtemp, aqsntemp, vcoul, schDttt_cor);
}

int bgw_schDttt_corKernel1_device(bgw_ctx* ctx, double* schDttt_cor, const int* inv_igp_index, const int* indinv,
                                  const double* I_epsR_array, const double* I_epsA_array, const double* aqsmtemp,
                                  const double* aqsntemp, const double* vcoul, int ncouls, int ifreq, int ngpown,
                                  int n1, double fact1, double fact2) {
  int rc = cor_band_validate(ctx, ncouls, ngpown, ifreq, n1,
                             {schDttt_cor, inv_igp_index, indinv, I_epsR_array, I_epsA_array, aqsmtemp, aqsntemp,
                              vcoul});
  if (rc) return rc;
  return cor_band_device(ctx, 1.0, ncouls, ngpown, ifreq, n1, fact1, fact2, inv_igp_index, indinv, I_epsR_array,
                         I_epsA_array, aqsmtemp, aqsntemp, vcoul, schDttt_cor);
}

int bgw_schDttt_corKernel2_device(bgw_ctx* ctx, double* schDttt_cor, const int* inv_igp_index, const int* indinv,
                                  const double* I_epsR_array, const double* I_epsA_array, const double* aqsmtemp,
                                  const double* aqsntemp, const double* vcoul, int ncouls, int ifreq, int ngpown,
                                  int n1, double fact1, double fact2) {
  int rc = cor_band_validate(ctx, ncouls, ngpown, ifreq, n1,
                             {schDttt_cor, inv_igp_index, indinv, I_epsR_array, I_epsA_array, aqsmtemp, aqsntemp,
                              vcoul});
  if (rc) return rc;
  return cor_band_device(ctx, -0.5, ncouls, ngpown, ifreq, n1, fact1, fact2, inv_igp_index, indinv, I_epsR_array,
                         I_epsA_array, aqsmtemp, aqsntemp, vcoul, schDttt_cor);
}

}  // extern "C"
