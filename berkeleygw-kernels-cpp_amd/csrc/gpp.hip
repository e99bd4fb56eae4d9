// gpp.hip -- GPP self-energy forward on gfx950 (CDNA4).
//
// Reference: noflagOCC_solver, GPP/gpp.cpp:307-402.  The quadruple loop
// n1 x my_igp x ig x iw (:339-391) is mapped as
//   * lanes   = bands n1 (one or two bands per lane), so the band factor
//               sch_store1(n1,p) (:369-371) leaves the G-vector sum and is applied
//               once per (n1, p, G-chunk) instead of once per term;
//   * workgroup = one (p, G-chunk) work item at a time; the (p, g) data of a
//               G-tile is read once from HBM, coalesced along ig, turned into the
//               per-g constants of the pole term and staged in LDS, from where all
//               band-waves read it as a broadcast;
//   * iw (3 frequencies) unrolled in registers, the three reciprocals of a
//               (p,g,n1) triple formed from one v_rcp_f64 (batched inversion).
// Partial sums stay in registers per lane and per work item, then go to
// per-workgroup partials [W][nb][3]; a fixed-order double-double pass produces
// acht_n1_loc and achtemp (run-to-run bitwise reproducible).
//
// Moment form (BGW_FP64_MOMENTS): the G-sum of a work item is a sum of poles
// I/(wx - wtilde), so k_gpp_moments reduces each chunk's far poles to 16 moments of a
// Moebius expansion, k_gpp_moments_row re-expands a row's far chunks about one centre,
// k_gpp_far_eval evaluates them at every band frequency on the main kernel's grid, and
// k_gpp_main<..., FAR> sums the rest directly (whole chunks, or a chunk's listed near poles)
// into the same per-workgroup partials.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <mutex>
#include <vector>

#include "bgw_dd.h"
#include "bgw_internal.h"
#include "contract.h"
#include "gpp_common.h"

namespace bgw {
namespace {

// ------------------------------------------------------------------ prep kernels
// S[p][n1] = conj(aqsm(n1,igp))*aqsn(n1,igp)*0.5*vcoul(igp)*wtilde(p,igp)  (GPP/gpp.cpp:369-371)
// wxb[n1][iw] = wx(n1,iw) (per-band form) or wx[iw] (reference form, GPP/gpp.cpp:374-375).
__global__ void k_gpp_prep(const int* __restrict__ inv, const int* __restrict__ indinv,
                           const double2* __restrict__ wt, const double2* __restrict__ aqsm,
                           const double2* __restrict__ aqsn, const double* __restrict__ vcoul,
                           const double* __restrict__ wx, int wx_per_band, int nb, int ncouls,
                           int p0, int gathered, int ngpown, double2* __restrict__ S,
                           double2* __restrict__ C, double* __restrict__ wxb) {
  const int n1 = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = p0 + blockIdx.y;  // rows [p0, p0 + gridDim.y)
  if (n1 >= nb) return;
  const int igp = gpp_igp(inv, indinv, p, ncouls);
  // gathered: aqsm/aqsn hold only the columns igp(p), as [nb][ngpown]
  const size_t at = gathered ? (size_t)n1 * ngpown + p : (size_t)n1 * ncouls + igp;
  const double2 am = aqsm[at];
  const double2 an = aqsn[at];
  const double vc = vcoul[igp];
  const double2 w = wt[(size_t)p * ncouls + igp];
  // conj(am) * an, then * 0.5, * vcoul, * wtilde(p,igp), in the reference's order
  double2 t = cmul(make_double2(am.x, -am.y), an);
  t.x *= 0.5;
  t.y *= 0.5;
  t.x *= vc;
  t.y *= vc;
  if (C) C[(size_t)p * nb + n1] = t;
  S[(size_t)p * nb + n1] = cmul(t, w);
  if (p == 0) {
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) wxb[n1 * 3 + iw] = wx_per_band ? wx[n1 * 3 + iw] : wx[iw];
  }
}

// [min, max] of all wx(n1,iw): bounds every pole denominator of a G-tile at once.
__global__ void k_wx_range(const double* __restrict__ wxb, int n, double* __restrict__ range) {
  __shared__ double smin[256], smax[256];
  double lo = INFINITY, hi = -INFINITY;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    lo = fmin(lo, wxb[i]);
    hi = fmax(hi, wxb[i]);
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // NaN anywhere in wx makes the range NaN, so every tile fails the range test
    // and takes the reference-arithmetic path, propagating the NaN like the reference.
    range[0] = smin[0];
    range[1] = smax[0];
  }
  __syncthreads();
  bool bad = false;
  for (int i = threadIdx.x; i < n; i += blockDim.x) bad |= !(wxb[i] == wxb[i]);
  if (__syncthreads_or(bad) && threadIdx.x == 0) {
    range[0] = NAN;
    range[1] = NAN;
  }
}

// ------------------------------------------------------------------ main kernel
// One LDS entry per G-vector of a tile: the per-(p,g) constants of the term
//   delw*I = (u + i wi)(Ir + i Ii)/d,  u = wx - wr,  d = u^2 + wi^2
//          = (fma(Ir,u,a) + i fma(Ii,u,b)) / d,  a = -wi*Ii,  b = wi*Ir.
struct __align__(16) TileEntry {
  double wr, wi2, Ir, Ii, a, b;
};

// Bounds of the pole denominators over all bands: every d in [2^-150, 2^150]
// keeps the batched products of up to six denominators and their reciprocal normal.
constexpr double kDMin = 0x1p-150;
constexpr double kDMax = 0x1p+150;

// Reciprocals of the N denominators of one lane from a single hardware
// reciprocal (batched inversion): prefix products, one v_rcp of the total,
// then back-substitution.  FP64 adds one Newton step: v_rcp_f64 alone is good to
// ~2^-23 relative, after the step the reciprocal is within ~11 ulp (mean -0.6
// ulp), measured over 2^20 arguments spanning 2^-40..2^40.  The products stay
// normal because every denominator is inside the tile guard's range.
__device__ inline double hw_rcp(double m) {
  const double r = __builtin_amdgcn_rcp(m);
  return fma(r, fma(-m, r, 1.0), r);
}
__device__ inline float hw_rcp(float m) { return __builtin_amdgcn_rcpf(m); }

template <typename T, int N>
__device__ inline void batch_inverse(const T (&d)[N], T (&r)[N]) {
  T p[N];
  p[0] = d[0];
#pragma unroll
  for (int i = 1; i < N; ++i) p[i] = p[i - 1] * d[i];
  T inv = hw_rcp(p[N - 1]);
#pragma unroll
  for (int i = N - 1; i > 0; --i) {
    r[i] = inv * p[i - 1];
    inv = inv * d[i];
  }
  r[0] = inv;
}

__device__ inline TileEntry make_entry(double2 w, double2 I, double wxmin, double wxmax, bool& ok) {
  TileEntry e;
  e.wr = w.x;
  e.wi2 = w.y * w.y;
  e.Ir = I.x;
  e.Ii = I.y;
  e.a = -w.y * I.y;
  e.b = w.y * I.x;
  const double lo = w.x < wxmin ? wxmin - w.x : (w.x > wxmax ? w.x - wxmax : 0.0);
  const double hi = fmax(fabs(wxmin - w.x), fabs(wxmax - w.x));
  const double dlo = fma(lo, lo, e.wi2), dhi = fma(hi, hi, e.wi2);
  ok = (dlo >= kDMin) && (dhi <= kDMax);  // false for NaN
  return e;
}

// Fast path, all FP64: the 3*BPL reciprocals of a lane's (p,g) terms come from
// one batched inversion.
template <int BPL>
__device__ inline void tile_fast(const TileEntry* __restrict__ tile, int n, const double (&x)[BPL][3],
                                 double (&acc)[BPL][3][2]) {
#pragma unroll 2
  for (int j = 0; j < n; ++j) {
    const TileEntry e = tile[j];
    double u[BPL * 3], d[BPL * 3], r[BPL * 3];
#pragma unroll
    for (int k = 0; k < BPL * 3; ++k) {
      u[k] = x[k / 3][k % 3] - e.wr;
      d[k] = fma(u[k], u[k], e.wi2);
    }
    batch_inverse(d, r);
#pragma unroll
    for (int k = 0; k < BPL * 3; ++k) {
      acc[k / 3][k % 3][0] = fma(fma(e.Ir, u[k], e.a), r[k], acc[k / 3][k % 3][0]);
      acc[k / 3][k % 3][1] = fma(fma(e.Ii, u[k], e.b), r[k], acc[k / 3][k % 3][1]);
    }
  }
}

// Reference-arithmetic path (GPP/gpp.cpp:374-383): IEEE division per term.  Used
// for a tile whose denominators are not all safely inside [2^-150, 2^150]
// (poles, zeros, NaN/Inf), so such inputs give what the reference gives.
template <int BPL>
__device__ inline void tile_exact(const TileEntry* __restrict__ tile, int n, const double (&x)[BPL][3],
                                  double (&acc)[BPL][3][2]) {
  for (int j = 0; j < n; ++j) {
    const TileEntry e = tile[j];
#pragma unroll
    for (int bb = 0; bb < BPL; ++bb) {
#pragma unroll
      for (int iw = 0; iw < 3; ++iw) {
        const double u = x[bb][iw] - e.wr;
        const double d = u * u + e.wi2;
        const double rden = 1.0 / d;
        // delw*I with delw = (u, wi_signed)*rden; a = -wi*Ii, b = wi*Ir carry the sign
        const double pr = e.Ir * u + e.a;
        const double pi = e.Ii * u + e.b;
        acc[bb][iw][0] += pr * rden;
        acc[bb][iw][1] += pi * rden;
      }
    }
  }
}

// FP32-bulk form: the per-(p,g) constants and the complex products in FP32, sums
// over runs of G-vectors in FP32 folded into the FP64 totals.  Used for a tile only
// when every denominator of it is far from the pole: |wx - wr| >= max|wx|/4 for all
// bands, so u = wx - wr loses at most 2 bits in FP32, and every d in [2^-20, 2^20], so
// products of six stay inside FP32 range; other tiles take the FP64 paths.
// The numerators are formed from wx itself: P = Ir u + a = Ir wx + c1 with
// c1 = a - Ir wr (c2 likewise), computed in FP64 at staging, so u is only needed for d.
struct TileEntry32 {
  float wr, wi2, Ir, Ii, c1, c2;
};

__device__ inline TileEntry32 entry32(const TileEntry& e) {
  TileEntry32 f;
  f.wr = (float)e.wr;
  f.wi2 = (float)e.wi2;
  f.Ir = (float)e.Ir;
  f.Ii = (float)e.Ii;
  f.c1 = (float)(e.a - e.Ir * e.wr);
  f.c2 = (float)(e.b - e.Ii * e.wr);
  return f;
}

__device__ inline bool far_from_pole(double2 w, double wxmin, double wxmax) {
  const double lo = w.x < wxmin ? wxmin - w.x : (w.x > wxmax ? w.x - wxmax : 0.0);
  const double hi = fmax(fabs(wxmin - w.x), fabs(wxmax - w.x));
  const double wi2 = w.y * w.y;
  const double scale = fmax(fabs(wxmin), fabs(wxmax));
  return lo >= 0.25 * scale && fma(lo, lo, wi2) >= 0x1p-20 && fma(hi, hi, wi2) <= 0x1p+20;
}

// One band per lane.  A step takes two G-vectors j, j+1: the six denominators of the
// lane's (G-vector, frequency) terms share one batched inversion.  The terms of j go to
// sr and those of j+1 to sr2, so each FP32 sum holds at most 32 terms of a run of 64
// (the same rounding as single sums over runs of 32, at half the FP64 folds); the run's
// two sums are added and folded into the per-lane FP64 totals through sch_store1.
// A tile of odd length reads one padding entry past its end: every entry of the tile
// array is written, and a padding entry contributes exactly zero.
constexpr int kRun32 = 64;  // G-vectors per FP32 run

__device__ inline void tile_f32(const TileEntry32* __restrict__ tile, int n, const float (&x)[3], double2 s,
                                double (&tot)[3][2]) {
  for (int j0 = 0; j0 < n; j0 += kRun32) {
    const int j1 = min(n, j0 + kRun32);
    float sr[3][2], sr2[3][2];
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) sr[iw][0] = sr[iw][1] = sr2[iw][0] = sr2[iw][1] = 0.f;
#pragma unroll 2
    for (int j = j0; j < j1; j += 2) {
      const TileEntry32 e[2] = {tile[j], tile[j + 1]};
      float d[6], p[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const float u = x[k % 3] - e[k / 3].wr;
        d[k] = fmaf(u, u, e[k / 3].wi2);
      }
      // batched inversion: prefix products, one reciprocal, back-substitution
      p[0] = d[0];
#pragma unroll
      for (int i = 1; i < 6; ++i) p[i] = p[i - 1] * d[i];
      float inv = hw_rcp(p[5]);
#pragma unroll
      for (int i = 5; i >= 0; --i) {
        const float r = i > 0 ? inv * p[i > 0 ? i - 1 : 0] : inv;
        if (i > 0) inv = inv * d[i];
        const TileEntry32& ei = e[i / 3];
        float (&acc)[3][2] = i >= 3 ? sr2 : sr;
        acc[i % 3][0] = fmaf(fmaf(ei.Ir, x[i % 3], ei.c1), r, acc[i % 3][0]);
        acc[i % 3][1] = fmaf(fmaf(ei.Ii, x[i % 3], ei.c2), r, acc[i % 3][1]);
      }
    }
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) {
      const double re = sr[iw][0] + sr2[iw][0], im = sr[iw][1] + sr2[iw][1];
      tot[iw][0] = fma(s.x, re, fma(-s.y, im, tot[iw][0]));
      tot[iw][1] = fma(s.x, im, fma(s.y, re, tot[iw][1]));
    }
  }
}

// ------------------------------------------------------------ moment (far-field) form
// The G-sum of one work item (p, G-chunk) is a sum of simple poles: conj(x - w)/|x - w|^2
// = 1/(x - w) (GPP/gpp.cpp:376-380), so with w = wtilde(p,g) and I = I_eps(p,g)
//   T(x) = sum_g I_g / (x - w_g),
// to be evaluated at the band frequencies x, all in [xmin, xmax].  With e = (xmin + xmax)/2
// and a centre c of the chunk's poles, the Moebius variable phi(z) = (z - c)/(z - e) gives
//   phi(x) - phi(w) = (c - e)(x - w) / ((x - e)(w - e)),
// so 1/(x - w) = (c - e) / ((x - c)(w - e)) sum_{m>=0} (phi(w)/phi(x))^m, and the tail of a
// pole beyond m = M - 1 is exactly (phi(w)/phi(x))^M times its term.  Hence
//   T(x) = (c - e)/(x - c) sum_{m<M} nu_m psi^m + E,   psi = 1/phi(x) = (x - e)/(x - c),
//   nu_m = sum_g I_g phi(w_g)^m / (w_g - e),   |E| <= rho^M sum_g |I_g / (x - w_g)|,
// with rho = P h / D: P bounds |phi(w)| over the chunk, h = max |x - e| = (xmax - xmin)/2
// and D = min |x - c| (so |phi(x)| >= D/h).  A chunk is "far" when rho <= kFarRho, which
// holds E below 2^-54 times the sum of the magnitudes of its terms -- under the rounding
// error bound of the direct sum.  (Poles 5 away from a frequency range of width 4, as in
// the benchmark workload, give rho ~ 0.06; the Taylor expansion about c, e at infinity,
// would give R/D ~ 0.2 and need 28 moments.)  The far chunks of a G' row are re-expanded
// about one centre and added, so the row is evaluated once per band frequency; every
// other chunk (poles near a band frequency, a non-finite pole, weight or moment) takes the
// direct sum.  The moments cost O(M) per (p, g) instead of O(3 nb).
constexpr int kMom = 16;
constexpr double kFarRho = 0.095;  // 0.095^16 = 4.4e-17 < 2^-54
constexpr int kMomSample = 256;    // the first poles of a chunk, whose bounding box gives its centre

constexpr int kChunkDirect = 0;  // the direct sum
constexpr int kChunkFar = 1;     // re-expanded into its row's moments, evaluated with the row
constexpr int kChunkEval = 2;    // far, and evaluates its row's moments
constexpr int kChunkSelf = 3;    // far, but evaluated from its own moments (outside the row's bound)
constexpr int kModeMask = 7;
constexpr int kNearFlag = 8;     // mode bit: the chunk's near poles go through the direct pass
constexpr int kNearCap = 128;    // near poles listed per chunk; more: the whole chunk takes the direct sum

// Moments are summed over a chunk's far poles, |phi(w)| <= kFarRho D/h about the chunk's
// centre; the others ("near": close to a band frequency, or non-finite) are listed, up to
// kNearCap of them, and summed directly.
struct __align__(16) MomChunk {
  double cx, cy;  // centre c
  double P;       // bound of |phi(w)| over the chunk's far poles; +inf: not usable
  int mode;       // set by the row pass
  int nnear;      // listed near poles
  double2 nu[kMom];
};
struct __align__(16) MomRow {
  double cx, cy;
  double2 nu[kMom];  // sum over the row's kChunkFar/kChunkEval chunks, about centre (cx, cy)
};

// e, h and the far test of a centre c with bound P, from the range [xmin, xmax] of the band
// frequencies: rho = P h / D <= kFarRho with D the distance from c to the range, and every
// |x - c|^2 inside the batched inversion's [2^-150, 2^150] (false for NaN).
struct FarRange {
  double e, h, xmin, xmax;
};
__device__ inline FarRange far_range(const double* __restrict__ wxrange) {
  FarRange f;
  f.xmin = wxrange[0];
  f.xmax = wxrange[1];
  f.e = 0.5 * (f.xmin + f.xmax);
  f.h = fmax(f.xmax - f.e, f.e - f.xmin) * (1.0 + 0x1p-40);
  return f;
}
__device__ inline bool far_ok(double cx, double cy, double P, const FarRange& f) {
  const double dx = cx < f.xmin ? f.xmin - cx : (cx > f.xmax ? cx - f.xmax : 0.0);
  const double D = sqrt(fma(dx, dx, cy * cy)) * (1.0 - 0x1p-40);
  const double e0 = f.xmin - cx, e1 = f.xmax - cx;
  const double dmax2 = fmax(e0 * e0, e1 * e1) + cy * cy;
  return P * f.h <= kFarRho * D && D * D >= kDMin && dmax2 <= kDMax;
}

// Reciprocal for the moments: v_rcp_f64 and one Newton step (within ~11 ulp).
__device__ inline double rcp_nr(double q) {
  const double r = __builtin_amdgcn_rcp(q);
  return fma(r, fma(-q, r, 1.0), r);
}

// One wave per work item (p, G-chunk [gb, ge)), four items per workgroup.  The chunk's
// centre is the middle of the bounding box of sample poles (its first kMomSample) that lie
// more than 2h from e, on the side of e that holds more of them (all sample poles when none
// does): a few poles near the band frequencies then neither move the centre nor, being
// listed as near, spoil the moments of the rest.  A non-finite weight of a far pole makes a
// moment non-finite (NaN and Inf survive every sum and product), which marks the chunk as
// not usable.
constexpr int kMomWaves = 4;

__device__ inline void box_add(double (&b)[4], double2 w) {
  b[0] = fmin(b[0], w.x);
  b[1] = fmax(b[1], w.x);
  b[2] = fmin(b[2], w.y);
  b[3] = fmax(b[3], w.y);
}
__device__ inline void box_wave(double (&b)[4]) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
    b[0] = fmin(b[0], __shfl_xor(b[0], m, 64));
    b[1] = fmax(b[1], __shfl_xor(b[1], m, 64));
    b[2] = fmin(b[2], __shfl_xor(b[2], m, 64));
    b[3] = fmax(b[3], __shfl_xor(b[3], m, 64));
  }
}

__global__ __launch_bounds__(64 * kMomWaves) void k_gpp_moments(const double2* __restrict__ wt,
                                                                const double2* __restrict__ Ieps,
                                                                const double* __restrict__ wxrange, int ncouls,
                                                                int chunk, int nchunks, long item0, long item1,
                                                                MomChunk* __restrict__ tab,
                                                                unsigned short* __restrict__ near_idx,
                                                                int* __restrict__ any_direct, int reset) {
  const int lane = threadIdx.x & 63;
  if (reset && blockIdx.x == 0 && threadIdx.x == 0) *any_direct = 0;  // before the row passes of this call
  const long item = item0 + (long)blockIdx.x * kMomWaves + (threadIdx.x >> 6);
  if (item >= item1) return;  // the whole wave: the kernel has no workgroup barrier
  const int p = (int)(item / nchunks);
  const int gb = (int)(item - (long)p * nchunks) * chunk;
  const int ge = min(ncouls, gb + chunk);
  const double2* __restrict__ wrow = wt + (size_t)p * ncouls;
  const double2* __restrict__ irow = Ieps + (size_t)p * ncouls;
  const FarRange fr = far_range(wxrange);
  const double e = fr.e;

  double ball[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
  double bl[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY}, br[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
  int nl = 0, nr = 0;
  const double clear2 = 4.0 * fr.h * fr.h;
#pragma unroll
  for (int k = 0; k < kMomSample / 64; ++k) {
    const int g = gb + lane + 64 * k;
    if (g < ge) {
      const double2 w = wrow[g];
      box_add(ball, w);
      const double dr = w.x - e;
      if (fma(dr, dr, w.y * w.y) > clear2) {
        if (w.x < e) {
          box_add(bl, w);
          ++nl;
        } else {
          box_add(br, w);
          ++nr;
        }
      }
    }
  }
  box_wave(ball);
  box_wave(bl);
  box_wave(br);
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
    nl += __shfl_xor(nl, m, 64);
    nr += __shfl_xor(nr, m, 64);
  }
  const double* b = nl + nr == 0 ? ball : (nl >= nr ? bl : br);
  double cx = 0.5 * (b[0] + b[1]), cy = 0.5 * (b[2] + b[3]);
  if (!(fabs(cx) <= 0x1p+1000 && fabs(cy) <= 0x1p+1000)) cx = cy = 0.0;  // a sample of NaN/Inf poles
  // far poles: |phi| <= kFarRho D/h, a 2^-30 margin inside the row pass's test of the same
  // centre (far_ok), so the bound P formed below passes it; 0 when no pole can be far
  double t2 = 0.0;
  {
    const double dx = cx < fr.xmin ? fr.xmin - cx : (cx > fr.xmax ? cx - fr.xmax : 0.0);
    const double D = sqrt(fma(dx, dx, cy * cy)) * (1.0 - 0x1p-40);
    const double t = kFarRho * D / fr.h * (1.0 - 0x1p-30);
    if (t > 0.0) t2 = t * t;  // false for NaN; +inf for a single band frequency (h = 0)
  }

  double v[2 * kMom];
#pragma unroll
  for (int j = 0; j < 2 * kMom; ++j) v[j] = 0.0;
  double f2 = 0.0;  // max |phi|^2 over the far poles
  int nn = 0;       // near poles so far (every lane holds the same count)
  unsigned short* __restrict__ list = near_idx + (size_t)item * kNearCap;
  // rounds of 64 poles, the same number in every lane (lanes past the end carry a zero
  // weight); one round ahead: the next round's loads are in flight while this one is summed
  // (indices past the end re-read the last pole, in bounds)
  const int rounds = (ge - gb + 63) / 64;
  double2 wn = wrow[min(gb + lane, ge - 1)], In = irow[min(gb + lane, ge - 1)];
  for (int k = 0; k < rounds; ++k) {
    const int g = gb + lane + 64 * k;
    const bool valid = g < ge;
    const double2 w = wn, I = In;
    const int gn = min(g + 64, ge - 1);
    wn = wrow[gn];
    In = irow[gn];
    // keeps the loads issued here, ahead of this pole's arithmetic (no wait: the
    // barrier has no register operands), instead of sunk to their use
    asm volatile("" ::: "memory");
    const double dr = w.x - e, di = w.y;                 // w - e
    const double r = rcp_nr(fma(dr, dr, di * di));
    const double gr = dr * r, gi = -di * r;              // 1/(w - e)
    const double zr = w.x - cx, zi = w.y - cy;           // w - c
    const double fr_ = fma(zr, gr, -zi * gi), fi = fma(zr, gi, zi * gr);  // phi
    const double q = fma(fr_, fr_, fi * fi);
    const bool far = q <= t2;  // false for NaN
    const bool near = valid && !far;
    f2 = (valid && far) ? fmax(f2, q) : f2;
    const double wgt = (valid && far) ? 1.0 : 0.0;
    double pr = wgt * fma(I.x, gr, -I.y * gi), pi = wgt * fma(I.x, gi, I.y * gr);  // I/(w - e)
    const unsigned long long bal = __ballot(near);
    if (near) {
      const int pos = nn + __builtin_popcountll(bal & ((1ull << lane) - 1));
      if (pos < kNearCap) list[pos] = (unsigned short)(g - gb);
    }
    nn += __builtin_popcountll(bal);
#pragma unroll
    for (int m = 0; m < kMom; ++m) {
      v[2 * m] += pr;
      v[2 * m + 1] += pi;
      if (m + 1 < kMom) {
        const double t = fma(pr, fr_, -pi * fi);
        pi = fma(pr, fi, pi * fr_);
        pr = t;
      }
    }
  }
  int base = 0;
  bool writer = true;
  wave_sum_scatter<2 * kMom, 32>(v, lane, base, writer);
  constexpr int NF = scatter_len(2 * kMom);
  bool fin = true;
#pragma unroll
  for (int j = 0; j < NF; ++j) fin &= fabs(v[j]) <= DBL_MAX;
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) f2 = fmax(f2, __shfl_xor(f2, m, 64));
  const bool ok = __all(fin) && nn <= kNearCap;
  MomChunk* __restrict__ me = tab + item;
  double* __restrict__ out = reinterpret_cast<double*>(me->nu);
  if (writer) {
#pragma unroll
    for (int j = 0; j < NF; ++j) out[base + j] = v[j];
  }
  if (lane == 0) {
    me->cx = cx;
    me->cy = cy;
    // sqrt(f2) is within a few tens of ulp of max |phi| over the far poles: the factor
    // makes P a bound
    me->P = ok ? sqrt(f2) * (1.0 + 0x1p-40) : INFINITY;
    me->nnear = ok ? nn : 0;
  }
}

// Row pass, one workgroup per G' row p0 + blockIdx.x.  The row's centre is that of its
// first far chunk, taken from chunk floor(p nchunks / W) mod nchunks on, which becomes
// kChunkEval: with W the main kernel's grid width (item i runs on workgroup i mod W), this
// spreads the rows' evaluations over all of its workgroups.  Each other far chunk is
// re-expanded about the row's centre c' when its bound there stays far (kChunkFar), and
// evaluates its own moments otherwise (kChunkSelf); the rest take the direct sum.  About
// c', phi' = b + a phi with b = (c - c')/(c - e), a = 1 - b, so
//   nu'_m = sum_{j<=m} C(m,j) b^(m-j) a^j nu_j,   |phi'| <= |b| + |a| P,
// formed in lane m from a^m nu_m by a Taylor shift in b (kMom - 1 rounds of
// mu_m += b mu_(m-1), no binomial coefficients).  Wave w takes the chunks c = w mod
// kRowWaves in chunk order, and the waves' sums are added in wave order.
constexpr int kRowWaves = 4;

__global__ __launch_bounds__(64 * kRowWaves) void k_gpp_moments_row(int nchunks, int p0, int W,
                                                                   const double* __restrict__ wxrange,
                                                                   MomChunk* __restrict__ tab,
                                                                   MomRow* __restrict__ rows,
                                                                   int* __restrict__ any_direct) {
  __shared__ double2 part[kRowWaves][kMom];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int p = p0 + blockIdx.x;
  const FarRange fr = far_range(wxrange);
  MomChunk* __restrict__ ch = tab + (size_t)p * nchunks;
  // the first far chunk from chunk `first` on, 64 chunks tested at a time (every wave alike)
  const int first = (int)(((long)p * nchunks / W) % nchunks);
  int d = -1;
  for (int k0 = 0; k0 < nchunks && d < 0; k0 += 64) {
    const int k = k0 + lane;
    const int c = first + k < nchunks ? first + k : first + k - nchunks;
    const bool far = k < nchunks && far_ok(ch[c].cx, ch[c].cy, ch[c].P, fr);
    const unsigned long long b = __ballot(far);
    if (b) {
      const int kk = k0 + __builtin_ctzll(b);
      d = first + kk < nchunks ? first + kk : first + kk - nchunks;
    }
  }
  if (d < 0) {  // no far chunk: every chunk takes the direct sum
    for (int c = threadIdx.x; c < nchunks; c += 64 * kRowWaves) ch[c].mode = kChunkDirect;
    if (threadIdx.x == 0) *any_direct = 1;
    return;
  }
  const double pcx = ch[d].cx, pcy = ch[d].cy;
  const int m = lane < kMom ? lane : kMom - 1;
  double sr = 0.0, si = 0.0;
  for (int c = wave; c < nchunks; c += kRowWaves) {
    const double cx = ch[c].cx, cy = ch[c].cy, P = ch[c].P;
    int mode = kChunkDirect;
    if (c == d) {
      mode = kChunkEval;
      sr += ch[c].nu[m].x;
      si += ch[c].nu[m].y;
    } else if (far_ok(cx, cy, P, fr)) {
      mode = kChunkSelf;
      // b = (c - c')/(c - e), a = 1 - b
      const double ur = cx - fr.e, ui = cy;
      const double q = 1.0 / fma(ur, ur, ui * ui);
      const double vr = cx - pcx, vi = cy - pcy;
      const double br = (vr * ur + vi * ui) * q, bi = (vi * ur - vr * ui) * q;
      const double ar = 1.0 - br, ai = -bi;
      const double Pw = (sqrt(fma(br, br, bi * bi)) + sqrt(fma(ar, ar, ai * ai)) * P) * (1.0 + 0x1p-40);
      if (far_ok(pcx, pcy, Pw, fr)) {
        // lane m: a^m by binary powering, mu_m = a^m nu_m, then the Taylor shift in b
        double tr = 1.0, ti = 0.0, kr = ar, ki = ai;
#pragma unroll
        for (int bit = 1; bit < kMom; bit <<= 1) {
          if (m & bit) {
            const double nr = fma(tr, kr, -ti * ki);
            ti = fma(tr, ki, ti * kr);
            tr = nr;
          }
          const double nr = fma(kr, kr, -ki * ki);
          ki = 2.0 * kr * ki;
          kr = nr;
        }
        const double2 nj = ch[c].nu[m];
        double mr = fma(tr, nj.x, -ti * nj.y), mi = fma(tr, nj.y, ti * nj.x);
        for (int k = 1; k < kMom; ++k) {
          const double qr = __shfl_up(mr, 1, 64), qi = __shfl_up(mi, 1, 64);
          if (lane >= k && lane < kMom) {
            const double nr = fma(br, qr, fma(-bi, qi, mr));
            mi = fma(br, qi, fma(bi, qr, mi));
            mr = nr;
          }
        }
        if (__all(fabs(mr) <= DBL_MAX && fabs(mi) <= DBL_MAX)) {
          sr += mr;
          si += mi;
          mode = kChunkFar;
        }
      }
    }
    if (lane == 0) {
      // a far chunk's near poles, if any, go through the direct pass
      if (mode != kChunkDirect && ch[c].nnear > 0) mode |= kNearFlag;
      ch[c].mode = mode;
      if (mode == kChunkDirect || (mode & kNearFlag)) *any_direct = 1;
    }
  }
  if (lane < kMom) part[wave][lane] = make_double2(sr, si);
  __syncthreads();
  if (threadIdx.x < kMom) {
    double2 t = part[0][threadIdx.x];
#pragma unroll
    for (int w = 1; w < kRowWaves; ++w) {
      t.x += part[w][threadIdx.x].x;
      t.y += part[w][threadIdx.x].y;
    }
    rows[p].nu[threadIdx.x] = t;
  }
  if (threadIdx.x == 0) {
    rows[p].cx = pcx;
    rows[p].cy = pcy;
  }
}

// Far-field value of a set of moments about centre c (a row's summed far chunks, or one
// chunk) for a lane's 3*BPL frequencies: acc[b][iw] = T(x[b][iw]).  nu: the moments (LDS).
// The three reciprocals 1/(x - c) of a band come from one batched inversion, as in the
// direct form; the bands' Horner chains are unrolled together (33 against 40 us for the
// far pass at the benchmark workload, with a few spill slots).
template <int BPL>
__device__ inline void row_far(const double2* __restrict__ nu, double cx, double cy, double e,
                               const double (&x)[BPL][3], double (&acc)[BPL][3][2]) {
  const double cer = cx - e, cei = cy;
#pragma unroll
  for (int bb = 0; bb < BPL; ++bb) {
    double ur[3], ui[3], sr[3], si[3];  // (c - e) t and psi = (x - e) t, t = 1/(x - c)
    {
      double a[3], d[3], r[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        a[k] = x[bb][k] - cx;
        d[k] = fma(a[k], a[k], cy * cy);
      }
      batch_inverse(d, r);
      // 1/(x - c) = 1/((x - cx) - i cy) = ((x - cx) + i cy) / |x - c|^2
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double tr = a[k] * r[k], ti = cy * r[k];
        const double xe = x[bb][k] - e;
        sr[k] = xe * tr;
        si[k] = xe * ti;
        ur[k] = fma(cer, tr, -cei * ti);
        ui[k] = fma(cer, ti, cei * tr);
      }
    }
    double hr[3], hi[3];
    const double2 top = nu[kMom - 1];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      hr[k] = top.x;
      hi[k] = top.y;
    }
#pragma unroll 3
    for (int m = kMom - 2; m >= 0; --m) {
      const double2 c = nu[m];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double nr = fma(hr[k], sr[k], fma(-hi[k], si[k], c.x));
        hi[k] = fma(hr[k], si[k], fma(hi[k], sr[k], c.y));
        hr[k] = nr;
      }
    }
    // T = (c - e) t h
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      acc[bb][k][0] = fma(ur[k], hr[k], -ui[k] * hi[k]);
      acc[bb][k][1] = fma(ur[k], hi[k], ui[k] * hr[k]);
    }
  }
}

// Waves per SIMD: FP64 four (128 VGPRs); FP32 eight (64 VGPRs), since FP32 VALU
// instructions of two waves of a SIMD issue together and the pair rate grows with the
// number of waves.  FAR: the direct pass of the moment form (FP64), launched after
// k_gpp_far_eval over the same items and grid and continuing its totals: it sums the
// items of kChunkDirect chunks exactly as without FAR and skips the others (all of them
// when the row passes found no such chunk).
template <int NT, int BPL, bool F32, bool FAR>
__global__ __launch_bounds__(NT, F32 ? 8 : 4) void k_gpp_main(const double2* __restrict__ wt,
                                                 const double2* __restrict__ Ieps,
                                                 const double2* __restrict__ S,
                                                 const double* __restrict__ wxb,
                                                 const double* __restrict__ wxrange, int nb,
                                                 int ngpown, int ncouls, int chunk, int nchunks,
                                                 long item0, long item1, int carry,
                                                 const MomChunk* __restrict__ mom,
                                                 const unsigned short* __restrict__ near_idx,
                                                 const int* __restrict__ any_direct,
                                                 double2* __restrict__ partial) {
  static_assert(!F32 || BPL == 1, "the FP32-bulk form runs one band per lane");
  static_assert(!(F32 && FAR), "the moment form is FP64");
  // no chunk of the call takes the direct sum: the far pass's totals stand as they are
  if constexpr (FAR) {
    if (*any_direct == 0) return;
  }
  __shared__ TileEntry tile[NT];
  __shared__ TileEntry32 tile32[F32 ? NT : 1];
  const int tid = threadIdx.x;
  const int band0 = blockIdx.y * NT * BPL;
  int n1[BPL];
  double x[BPL][3];
  double tot[BPL][3][2];
#pragma unroll
  for (int bb = 0; bb < BPL; ++bb) {
    n1[bb] = band0 + tid + bb * NT;
    const int nc = n1[bb] < nb ? n1[bb] : nb - 1;  // idle lanes mirror the last band
    // carry: a launch continuing an earlier one (same grid) resumes from the totals
    // that launch left in this workgroup's partial slot
    const double2* src = partial + (size_t)blockIdx.x * nb * 3 + (size_t)nc * 3;
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) {
      x[bb][iw] = wxb[nc * 3 + iw];
      const double2 t0 = carry && n1[bb] < nb ? src[iw] : make_double2(0.0, 0.0);
      tot[bb][iw][0] = t0.x;
      tot[bb][iw][1] = t0.y;
    }
  }
  const double wxmin = wxrange[0], wxmax = wxrange[1];
  // items [item0, item1): workgroup b takes item0 + b, item0 + b + gridDim.x, ... so
  // that launches over consecutive ranges starting at multiples of gridDim.x give
  // every workgroup the items, in the order, of one launch over all of them
  const long nitems = item1;
  float x32[BPL][3];
#pragma unroll
  for (int bb = 0; bb < BPL; ++bb)
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) x32[bb][iw] = (float)x[bb][iw];

  // FAR: the chunk modes of the workgroup's next 64 items are loaded at once (lane k:
  // item + k * gridDim.x) and read lane by lane
  int modes = kChunkDirect, k = 0;
  for (long item = item0 + blockIdx.x; item < nitems; item += gridDim.x, k = (k + 1) & 63) {
    const int p = (int)(item / nchunks);
    const int c = (int)(item - (long)p * nchunks);
    const int gb = c * chunk;
    const int ge = min(ncouls, gb + chunk);
    const size_t rowoff = (size_t)p * ncouls;
    int mode = kChunkDirect;
    if constexpr (FAR) {
      if (k == 0) {
        const long it = item + (long)(tid & 63) * gridDim.x;
        modes = it < nitems ? mom[it].mode : kChunkDirect;
      }
      mode = __builtin_amdgcn_readlane(modes, k);  // uniform: the item's
      // far poles are summed by k_gpp_far_eval; a far chunk's near poles, listed, here
      if ((mode & kModeMask) != kChunkDirect && !(mode & kNearFlag)) continue;
    }
    // sch_store1(n1,p) of this lane's bands (zero for idle lanes)
    double2 sv[BPL];
#pragma unroll
    for (int bb = 0; bb < BPL; ++bb) sv[bb] = n1[bb] < nb ? S[(size_t)p * nb + n1[bb]] : make_double2(0.0, 0.0);
    // FP64 form: G-chunk sums in acc, folded with sch_store1 at the end of the item.
    // FP32 form: every FP32 run is folded into tot at once (no per-item state).
    double acc[BPL][3][2];
#pragma unroll
    for (int bb = 0; bb < BPL; ++bb)
#pragma unroll
      for (int iw = 0; iw < 3; ++iw) acc[bb][iw][0] = acc[bb][iw][1] = 0.0;

    // the chunk's G-vectors gb + j, j < cnt, or (FAR, a far chunk) its listed near poles
    int cnt = ge - gb;
    const unsigned short* __restrict__ list = nullptr;
    if constexpr (FAR) {
      if ((mode & kModeMask) != kChunkDirect) {
        list = near_idx + (size_t)item * kNearCap;
        cnt = mom[item].nnear;
      }
    }
    for (int t0 = 0; t0 < cnt; t0 += NT) {
      const int j = t0 + tid;
      int g = j < cnt ? gb + j : -1;
      if constexpr (FAR) {
        if (list && j < cnt) g = gb + list[j];
      }
      bool ok;
      TileEntry e;
      const double2 wv = g >= 0 ? wt[rowoff + g] : make_double2(0.0, 1.0);
      const double2 Iv = g >= 0 ? Ieps[rowoff + g] : make_double2(0.0, 0.0);
      // padding entries (g < 0) contribute exactly zero, denominators wx^2+1
      e = make_entry(wv, Iv, wxmin, wxmax, ok);
      __syncthreads();  // previous tile fully consumed
      tile[tid] = e;
      bool ok32 = false;
      if (F32) {
        ok32 = ok && far_from_pole(wv, wxmin, wxmax);
        tile32[tid] = entry32(e);
      }
      const int unsafe = __syncthreads_or(!ok);
      const int n = min(NT, cnt - t0);
      if constexpr (F32) {
        // FP32-bulk form: FP32 tiles far from every pole, FP64 (IEEE division) otherwise
        if (!unsafe && !__syncthreads_or(!ok32)) {
          if constexpr (BPL == 1) tile_f32(tile32, n, x32[0], sv[0], tot[0]);
        } else {
          double xe[BPL][3], at[BPL][3][2];
#pragma unroll
          for (int bb = 0; bb < BPL; ++bb) {
            const int nc = n1[bb] < nb ? n1[bb] : nb - 1;
#pragma unroll
            for (int iw = 0; iw < 3; ++iw) {
              xe[bb][iw] = wxb[nc * 3 + iw];
              at[bb][iw][0] = at[bb][iw][1] = 0.0;
            }
          }
          tile_exact<BPL>(tile, n, xe, at);
#pragma unroll
          for (int bb = 0; bb < BPL; ++bb)
#pragma unroll
            for (int iw = 0; iw < 3; ++iw) {
              tot[bb][iw][0] = fma(sv[bb].x, at[bb][iw][0], fma(-sv[bb].y, at[bb][iw][1], tot[bb][iw][0]));
              tot[bb][iw][1] = fma(sv[bb].x, at[bb][iw][1], fma(sv[bb].y, at[bb][iw][0], tot[bb][iw][1]));
            }
        }
      } else {
        if (unsafe)
          tile_exact<BPL>(tile, n, x, acc);
        else
          tile_fast<BPL>(tile, n, x, acc);
      }
    }
    if constexpr (!F32) {
      // fold the G-chunk sums with sch_store1(n1,p)
#pragma unroll
      for (int bb = 0; bb < BPL; ++bb)
#pragma unroll
        for (int iw = 0; iw < 3; ++iw) {
          tot[bb][iw][0] = fma(sv[bb].x, acc[bb][iw][0], fma(-sv[bb].y, acc[bb][iw][1], tot[bb][iw][0]));
          tot[bb][iw][1] = fma(sv[bb].x, acc[bb][iw][1], fma(sv[bb].y, acc[bb][iw][0], tot[bb][iw][1]));
        }
    }
  }
#pragma unroll
  for (int bb = 0; bb < BPL; ++bb) {
    if (n1[bb] < nb) {
      double2* dst = partial + (size_t)blockIdx.x * nb * 3 + (size_t)n1[bb] * 3;
#pragma unroll
      for (int iw = 0; iw < 3; ++iw) dst[iw] = make_double2(tot[bb][iw][0], tot[bb][iw][1]);
    }
  }
}

// Far pass of the moment form, on the main kernel's grid and item order (lanes = bands):
// workgroup b sums, over its items b, b + W, ..., the far-field values of the kChunkEval
// items (the row's summed moments, `mrows`) and kChunkSelf items (the chunk's own, `mom`),
// each folded with sch_store1(n1,p), into the per-workgroup totals; carry resumes them
// from an earlier launch over the items before item0.  The direct pass (k_gpp_main with
// FAR) continues these totals.
template <int NT, int BPL>
__global__ __launch_bounds__(NT, 4) void k_gpp_far_eval(const double2* __restrict__ S, const double* __restrict__ wxb,
                                                    const double* __restrict__ wxrange, int nb, int nchunks,
                                                    long item0, long item1, int carry,
                                                    const MomChunk* __restrict__ mom,
                                                    const MomRow* __restrict__ mrows, double2* __restrict__ partial) {
  __shared__ double2 nu_s[kMom];
  const int tid = threadIdx.x;
  const int band0 = blockIdx.y * NT * BPL;
  int n1[BPL];
  double x[BPL][3];
  double tot[BPL][3][2];
#pragma unroll
  for (int bb = 0; bb < BPL; ++bb) {
    n1[bb] = band0 + tid + bb * NT;
    const int nc = n1[bb] < nb ? n1[bb] : nb - 1;  // idle lanes mirror the last band
    const double2* src = partial + (size_t)blockIdx.x * nb * 3 + (size_t)nc * 3;
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) {
      x[bb][iw] = wxb[nc * 3 + iw];
      const double2 t0 = carry && n1[bb] < nb ? src[iw] : make_double2(0.0, 0.0);
      tot[bb][iw][0] = t0.x;
      tot[bb][iw][1] = t0.y;
    }
  }
  const double e = 0.5 * (wxrange[0] + wxrange[1]);
  // the chunk modes of the workgroup's next 64 items are loaded at once (lane k: item +
  // k * gridDim.x) and read lane by lane
  int modes = kChunkDirect, k = 0;
  for (long item = item0 + blockIdx.x; item < item1; item += gridDim.x, k = (k + 1) & 63) {
    if (k == 0) {
      const long it = item + (long)(tid & 63) * gridDim.x;
      modes = it < item1 ? mom[it].mode : kChunkDirect;
    }
    const int mode = __builtin_amdgcn_readlane(modes, k) & kModeMask;  // uniform: the item's
    if (mode != kChunkEval && mode != kChunkSelf) continue;
    const int p = (int)(item / nchunks);
    const bool row = mode == kChunkEval;
    const MomRow* __restrict__ rw = mrows + p;
    const MomChunk* __restrict__ mc = mom + item;
    const double2* __restrict__ src = row ? rw->nu : mc->nu;
    const double cx = row ? rw->cx : mc->cx, cy = row ? rw->cy : mc->cy;
    double2 sv[BPL];
#pragma unroll
    for (int bb = 0; bb < BPL; ++bb) sv[bb] = n1[bb] < nb ? S[(size_t)p * nb + n1[bb]] : make_double2(0.0, 0.0);
    __syncthreads();  // the previous item's readers of nu_s are done
    if (tid < kMom) nu_s[tid] = src[tid];
    __syncthreads();
    double acc[BPL][3][2];
    row_far<BPL>(nu_s, cx, cy, e, x, acc);
#pragma unroll
    for (int bb = 0; bb < BPL; ++bb)
#pragma unroll
      for (int iw = 0; iw < 3; ++iw) {
        tot[bb][iw][0] = fma(sv[bb].x, acc[bb][iw][0], fma(-sv[bb].y, acc[bb][iw][1], tot[bb][iw][0]));
        tot[bb][iw][1] = fma(sv[bb].x, acc[bb][iw][1], fma(sv[bb].y, acc[bb][iw][0], tot[bb][iw][1]));
      }
  }
#pragma unroll
  for (int bb = 0; bb < BPL; ++bb) {
    if (n1[bb] < nb) {
      double2* dst = partial + (size_t)blockIdx.x * nb * 3 + (size_t)n1[bb] * 3;
#pragma unroll
      for (int iw = 0; iw < 3; ++iw) dst[iw] = make_double2(tot[bb][iw][0], tot[bb][iw][1]);
    }
  }
}

// ------------------------------------------------------------------ reductions
// acht_n1_loc[n1][iw] = sum_w partial[w][n1][iw] (double-double).  A block owns 4
// consecutive (n1,iw) entries; its 64 row groups each sum a strided set of w, then a
// fixed tree over the groups adds them, so the result is bitwise stable.
constexpr int kRedEntries = 4, kRedGroups = 64;

__global__ __launch_bounds__(kRedEntries * kRedGroups) void k_reduce_bands(const double2* __restrict__ partial, int W,
                                                                         int nb, double2* __restrict__ acht_n1_loc,
                                                                         double4* __restrict__ band_dd) {
  __shared__ cdd sh[kRedGroups][kRedEntries];
  const int e = threadIdx.x % kRedEntries, grp = threadIdx.x / kRedEntries;
  const int k = blockIdx.x * kRedEntries + e;
  const int n = nb * 3;
  cdd a = cdd_zero();
  if (k < n)
    for (int w = grp; w < W; w += kRedGroups) {
      const double2 v = partial[(size_t)w * n + k];
      a = cdd_add(a, v.x, v.y);
    }
  sh[grp][e] = a;
  for (int h = kRedGroups / 2; h > 0; h >>= 1) {
    __syncthreads();
    if (grp < h) sh[grp][e] = cdd_add(sh[grp][e], sh[grp + h][e]);
  }
  if (grp == 0 && k < n) {
    const cdd t = sh[0][e];
    if (acht_n1_loc) acht_n1_loc[k] = make_double2(t.re.hi + t.re.lo, t.im.hi + t.im.lo);
    cdd_store(&band_dd[k], t);
  }
}

// achtemp[iw] = sum_n1 band_dd[n1][iw], fixed order.
__global__ void k_reduce_achtemp(const double4* __restrict__ band_dd, int nb,
                                 double2* __restrict__ achtemp) {
  __shared__ cdd sh[16];
  for (int iw = 0; iw < 3; ++iw) {
    cdd a = cdd_zero();
    for (int n1 = threadIdx.x; n1 < nb; n1 += blockDim.x) a = cdd_add(a, cdd_load(&band_dd[n1 * 3 + iw]));
    cdd r = cdd_block_reduce(a, sh);
    if (threadIdx.x == 0) achtemp[iw] = make_double2(r.re.hi + r.re.lo, r.im.hi + r.im.lo);
    __syncthreads();
  }
}

// ------------------------------------------------------------- screened exchange
// asxtemp[iw] = sum_{n1<nvband,p} occ vcoul(igp) sum_g aqsn(n1,g) conj(aqsm(n1,igp)) ssx(n1,p,g,iw)
template <int NT>
__global__ __launch_bounds__(NT) void k_gpp_asx(const int* __restrict__ inv, const int* __restrict__ indinv,
                                                const double2* __restrict__ wt,
                                                const double2* __restrict__ Ieps,
                                                const double2* __restrict__ aqsm,
                                                const double2* __restrict__ aqsn,
                                                const double* __restrict__ vcoul,
                                                const double* __restrict__ wxb, int nvband, int ngpown,
                                                int ncouls, int chunk, int nchunks, double occ,
                                                double4* __restrict__ partial /*[W][3]*/) {
  __shared__ cdd sh[NT / 64];
  cdd tot[3] = {cdd_zero(), cdd_zero(), cdd_zero()};
  const long nitems = (long)ngpown * nchunks;
  for (long item = blockIdx.x; item < nitems; item += gridDim.x) {
    const int p = (int)(item / nchunks);
    const int c = (int)(item - (long)p * nchunks);
    const int gb = c * chunk, ge = min(ncouls, gb + chunk);
    const int igp = gpp_igp(inv, indinv, p, ncouls);
    const double vc = vcoul[igp];
    // each (wtilde, I_eps) is read once and its wx-independent parts formed once for
    // every occupied band and frequency; the band weight m(n1) = conj(aqsm) occ vcoul
    // multiplies each term
    double s[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    for (int g = gb + threadIdx.x; g < ge; g += NT) {
      const SsxG G = gpp_ssx_g(wt[(size_t)p * ncouls + g], Ieps[(size_t)p * ncouls + g]);
      for (int n1 = 0; n1 < nvband; ++n1) {
        const double2 am = aqsm[(size_t)n1 * ncouls + igp];
        const double2 m = make_double2(am.x * occ * vc, -am.y * occ * vc);
        const double2 mn = cmul(m, aqsn[(size_t)n1 * ncouls + g]);
        double2 v[3];
        gpp_ssx3(wxb + n1 * 3, G, v);
#pragma unroll
        for (int iw = 0; iw < 3; ++iw) {
          const double2 t = cmul(mn, v[iw]);
          s[iw][0] += t.x;
          s[iw][1] += t.y;
        }
      }
    }
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) tot[iw] = cdd_add(tot[iw], s[iw][0], s[iw][1]);
  }
#pragma unroll
  for (int iw = 0; iw < 3; ++iw) {
    cdd r = cdd_block_reduce(tot[iw], sh);
    if (threadIdx.x == 0) cdd_store(&partial[(size_t)blockIdx.x * 3 + iw], r);
    __syncthreads();
  }
}

template <int NT, int BPL, bool F32, bool FAR = false>
int launch_main(bgw_ctx* ctx, const double2* wt, const double2* I, const double2* S,
                const double* wxb, const double* range, int nb, int ngpown, int ncouls,
                double2* partial_ws, int* W_out, int chunk, long item0, long item1) {
  const int nchunks = (ncouls + chunk - 1) / chunk;
  const int btiles = (nb + NT * BPL - 1) / (NT * BPL);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gpp_main<NT, BPL, F32, FAR>, NT, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  long items = (long)ngpown * nchunks;
  long W = (long)ctx->cus * per_cu / btiles;
  if (W < 1) W = 1;
  if (W > items) W = items;
  *W_out = (int)W;
  if (!partial_ws) return BGW_OK;  // sizing query
  if (item1 < 0) item1 = items;
  MomChunk* mom = nullptr;
  MomRow* mrows = nullptr;
  unsigned short* near = nullptr;
  if constexpr (FAR) {
    // Moments of every chunk of the rows that items [item0, item1) reach and that an
    // earlier launch of this call has not covered (whole rows: a row's pass needs all of
    // its chunks), the row pass over them, then the evaluation of the items.
    void *pm, *pr, *pn;
    int rc = pool_get(ctx, "ws_moments", (size_t)items * sizeof(MomChunk), &pm);
    if (!rc) rc = pool_get(ctx, "ws_moment_rows", (size_t)ngpown * sizeof(MomRow), &pr);
    if (!rc) rc = pool_get(ctx, "ws_near_idx", (size_t)items * kNearCap * sizeof(unsigned short), &pn);
    if (rc) return rc;
    mom = static_cast<MomChunk*>(pm);
    mrows = static_cast<MomRow*>(pr);
    near = static_cast<unsigned short*>(pn);
    const int r0 = item0 > 0 ? (int)((item0 - 1) / nchunks) + 1 : 0;
    const int r1 = (int)((item1 - 1) / nchunks) + 1;
    if (r1 > r0) {
      {
        KernelTimer kt(ctx, "gpp_moments");
        const long n = (long)(r1 - r0) * nchunks;
        hipLaunchKernelGGL(k_gpp_moments, dim3((unsigned)((n + kMomWaves - 1) / kMomWaves)), dim3(64 * kMomWaves), 0,
                           ctx->stream, wt, I, range, ncouls, chunk, nchunks, (long)r0 * nchunks, (long)r1 * nchunks,
                           mom, near, ctx->d_flag + 1, r0 == 0 ? 1 : 0);
      }
      hipLaunchKernelGGL(k_gpp_moments_row, dim3((unsigned)(r1 - r0)), dim3(64 * kRowWaves), 0, ctx->stream, nchunks,
                         r0, (int)W, range, mom, mrows, ctx->d_flag + 1);
      if ((rc = check_launch(ctx, "k_gpp_moments"))) return rc;
    }
  }
  if constexpr (FAR) {
    // the far pass first, then the direct pass continues its totals
    {
      KernelTimer kt(ctx, "gpp_far");
      hipLaunchKernelGGL((k_gpp_far_eval<NT, BPL>), dim3((unsigned)W, btiles), dim3(NT), 0, ctx->stream, S, wxb, range,
                         nb, nchunks, item0, item1, item0 > 0 ? 1 : 0, mom, mrows, partial_ws);
    }
    int rc = check_launch(ctx, "k_gpp_far_eval");
    if (rc) return rc;
  }
  KernelTimer kt(ctx, FAR ? "gpp_far_direct" : "gpp_main");
  hipLaunchKernelGGL((k_gpp_main<NT, BPL, F32, FAR>), dim3((unsigned)W, btiles), dim3(NT), 0, ctx->stream, wt, I, S,
                     wxb, range, nb, ngpown, ncouls, chunk, nchunks, item0, item1, (FAR || item0 > 0) ? 1 : 0, mom,
                     near, ctx->d_flag + 1, partial_ws);
  return check_launch(ctx, "k_gpp_main");
}

struct MainCfg {
  int nt, bpl;
};

MainCfg pick_main(int nb) {
  if (nb > 256) return {256, 2};
  if (nb > 128) return {128, 2};
  if (nb > 64) return {128, 1};
  return {64, 1};
}

constexpr int kGChunk = 2048;  // G-vectors per work item

template <bool F32>
int run_main_t(bgw_ctx* ctx, MainCfg cfg, const double2* wt, const double2* I, const double2* S, const double* wxb,
               const double* range, int nb, int ngpown, int ncouls, double2* ws, int* W, long item0, long item1) {
  const int chunk = kGChunk;
  if constexpr (F32) {
    // one band per lane, a workgroup covering up to 512 bands
    if (cfg.nt == 256 && cfg.bpl == 2)
      return launch_main<512, 1, true>(ctx, wt, I, S, wxb, range, nb, ngpown, ncouls, ws, W, chunk, item0, item1);
    if (cfg.nt == 128 && cfg.bpl == 2)
      return launch_main<256, 1, true>(ctx, wt, I, S, wxb, range, nb, ngpown, ncouls, ws, W, chunk, item0, item1);
    if (cfg.nt == 128)
      return launch_main<128, 1, true>(ctx, wt, I, S, wxb, range, nb, ngpown, ncouls, ws, W, chunk, item0, item1);
    return launch_main<64, 1, true>(ctx, wt, I, S, wxb, range, nb, ngpown, ncouls, ws, W, chunk, item0, item1);
  } else {
    if (cfg.nt == 256 && cfg.bpl == 2)
      return launch_main<256, 2, false>(ctx, wt, I, S, wxb, range, nb, ngpown, ncouls, ws, W, chunk, item0, item1);
    if (cfg.nt == 128 && cfg.bpl == 2)
      return launch_main<128, 2, false>(ctx, wt, I, S, wxb, range, nb, ngpown, ncouls, ws, W, chunk, item0, item1);
    if (cfg.nt == 128)
      return launch_main<128, 1, false>(ctx, wt, I, S, wxb, range, nb, ngpown, ncouls, ws, W, chunk, item0, item1);
    return launch_main<64, 1, false>(ctx, wt, I, S, wxb, range, nb, ngpown, ncouls, ws, W, chunk, item0, item1);
  }
}

int run_far(bgw_ctx* ctx, MainCfg cfg, const double2* wt, const double2* I, const double2* S, const double* wxb,
            const double* range, int nb, int ngpown, int ncouls, double2* ws, int* W, long item0, long item1) {
  const int chunk = kGChunk;
  if (cfg.nt == 256 && cfg.bpl == 2)
    return launch_main<256, 2, false, true>(ctx, wt, I, S, wxb, range, nb, ngpown, ncouls, ws, W, chunk, item0, item1);
  if (cfg.nt == 128 && cfg.bpl == 2)
    return launch_main<128, 2, false, true>(ctx, wt, I, S, wxb, range, nb, ngpown, ncouls, ws, W, chunk, item0, item1);
  if (cfg.nt == 128)
    return launch_main<128, 1, false, true>(ctx, wt, I, S, wxb, range, nb, ngpown, ncouls, ws, W, chunk, item0, item1);
  return launch_main<64, 1, false, true>(ctx, wt, I, S, wxb, range, nb, ngpown, ncouls, ws, W, chunk, item0, item1);
}

// Launches the main kernel over work items [item0, item1) (item1 < 0: all).  A
// launch with item0 > 0 continues the per-workgroup totals of the launches before
// it; item0 must then be a multiple of the grid width *W.
int run_main(bgw_ctx* ctx, MainCfg cfg, const double2* wt, const double2* I, const double2* S,
             const double* wxb, const double* range, int nb, int ngpown, int ncouls, double2* ws,
             int* W, long item0 = 0, long item1 = -1) {
  if (ctx->precision == BGW_FP32_BULK)
    return run_main_t<true>(ctx, cfg, wt, I, S, wxb, range, nb, ngpown, ncouls, ws, W, item0, item1);
  if (ctx->precision == BGW_FP64_MOMENTS)
    return run_far(ctx, cfg, wt, I, S, wxb, range, nb, ngpown, ncouls, ws, W, item0, item1);
  return run_main_t<false>(ctx, cfg, wt, I, S, wxb, range, nb, ngpown, ncouls, ws, W, item0, item1);
}

// Bytes of the moments table of a forward call (the moment form only).
size_t moments_bytes(bgw_ctx* ctx, int ngpown, int ncouls) {
  if (ctx->precision != BGW_FP64_MOMENTS) return 0;
  const size_t items = (size_t)ngpown * ((ncouls + kGChunk - 1) / kGChunk);
  return ((items * sizeof(MomChunk) + 255) & ~size_t(255)) + (((size_t)ngpown * sizeof(MomRow) + 255) & ~size_t(255)) +
         items * kNearCap * sizeof(unsigned short);
}

}  // namespace

// S[p][n1], C[p][n1] (sch_store1 without its wtilde(p,igp) factor; the backward
// reads it, the forward does not ask for it), the band frequencies wxb[n1][iw] and
// their [min,max]; buffers owned by the context.
namespace {

struct PrepTables {
  double2 *S, *C;
  double *wxb, *range;
};

int prep_alloc(bgw_ctx* ctx, int nb, int ngpown, bool want_C, PrepTables* t) {
  void *pS, *pC = nullptr, *pwx, *prange;
  int rc = pool_get(ctx, "ws_S", (size_t)nb * ngpown * 16, &pS);
  if (!rc && want_C) rc = pool_get(ctx, "ws_C", (size_t)nb * ngpown * 16, &pC);
  if (!rc) rc = pool_get(ctx, "ws_wxb", (size_t)nb * 3 * 8, &pwx);
  if (!rc) rc = pool_get(ctx, "ws_range", 64, &prange);
  if (rc) return rc;
  t->S = static_cast<double2*>(pS);
  t->C = static_cast<double2*>(pC);
  t->wxb = static_cast<double*>(pwx);
  t->range = static_cast<double*>(prange);
  return BGW_OK;
}

// S and C of rows [p0, p1); the rows containing p = 0 also write wxb.
// gathered != 0: am/an are the [nb][ngpown] gathers of the columns igp(p).
void prep_rows(bgw_ctx* ctx, const PrepTables& t, int nb, int p0, int p1, int ncouls, const int* inv,
               const int* indinv, const double* wx, int wx_per_band, const double2* wt, const double2* am,
               const double2* an, const double* vcoul, int gathered = 0, int ngpown = 0) {
  // one grid row per G' row, and a grid has at most 65535 rows: launch in chunks of that many
  for (int q0 = p0; q0 < p1; q0 += 65535) {
    const int q1 = std::min(p1, q0 + 65535);
    hipLaunchKernelGGL(k_gpp_prep, dim3((nb + 255) / 256, q1 - q0), dim3(256), 0, ctx->stream, inv, indinv, wt, am,
                       an, vcoul, wx, wx_per_band, nb, ncouls, q0, gathered, ngpown, t.S, t.C, t.wxb);
  }
}

// Work-item bookkeeping of one forward call: launch shape and the partial buffers.
struct Fwd {
  MainCfg cfg;
  int W = 0;
  double2* part = nullptr;
  double4* band = nullptr;
};

int fwd_begin(bgw_ctx* ctx, int nb, int ngpown, int ncouls, Fwd* f) {
  f->cfg = pick_main(nb);
  run_main(ctx, f->cfg, nullptr, nullptr, nullptr, nullptr, nullptr, nb, ngpown, ncouls, nullptr, &f->W);
  void *ppart, *pband;
  int rc = pool_get(ctx, "ws_partial", (size_t)f->W * nb * 3 * 16, &ppart);
  if (!rc) rc = pool_get(ctx, "ws_band_dd", (size_t)nb * 3 * 32, &pband);
  if (rc) return rc;
  f->part = static_cast<double2*>(ppart);
  f->band = static_cast<double4*>(pband);
  return BGW_OK;
}

// Band and frequency reductions, then the optional asxtemp and achstemp terms.
int fwd_end(bgw_ctx* ctx, const Fwd& f, int nb, int nvband, int ngpown, int ncouls, const int* inv,
            const int* indinv, const double2* wt, const double2* I, const double2* am, const double2* an,
            const double* vcoul, const double* wxb, double occ, double* achtemp, double* acht_n1_loc,
            double* achstemp, double* asxtemp) {
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(k_reduce_bands, dim3((nb * 3 + kRedEntries - 1) / kRedEntries), dim3(kRedEntries * kRedGroups), 0,
                     s, f.part, f.W, nb, reinterpret_cast<double2*>(acht_n1_loc), f.band);
  hipLaunchKernelGGL(k_reduce_achtemp, dim3(1), dim3(256), 0, s, f.band, nb, reinterpret_cast<double2*>(achtemp));
  int rc;
  if ((rc = check_launch(ctx, "k_reduce"))) return rc;
  if (asxtemp) {
    if (nvband == 0) {
      fill_zero(asxtemp, 6 * sizeof(double), s);
    } else {
      const int chunk = 4096, nchunks = (ncouls + chunk - 1) / chunk;
      long items = (long)ngpown * nchunks;
      long Wa = (long)ctx->cus * 4;
      if (Wa > items) Wa = items;
      void* pax;
      if ((rc = pool_get(ctx, "ws_asx", (size_t)Wa * 3 * 32, &pax))) return rc;
      {
        KernelTimer kt(ctx, "gpp_asx");
        hipLaunchKernelGGL(k_gpp_asx<256>, dim3((unsigned)Wa), dim3(256), 0, s, inv, indinv, wt, I, am, an,
                           vcoul, wxb, nvband, ngpown, ncouls, chunk, nchunks, occ,
                           static_cast<double4*>(pax));
      }
      launch_reduce_cdd(static_cast<const double4*>(pax), (int)Wa, 3, reinterpret_cast<double2*>(asxtemp), s);
      if ((rc = check_launch(ctx, "k_gpp_asx"))) return rc;
    }
  }
  if (achstemp) {
    rc = contract_achstemp(ctx, nb, ngpown, ncouls, inv, indinv, an, am, I, vcoul,
                           reinterpret_cast<double2*>(achstemp));
    if (rc) return rc;
  }
  return BGW_OK;
}

}  // namespace

int gpp_prep_tables(bgw_ctx* ctx, int nb, int ngpown, int ncouls, const int* inv, const int* indinv,
                    const double* wx, int wx_per_band, const double2* wt, const double2* am, const double2* an,
                    const double* vcoul, double2** S, double2** C, double** wxb, double** range) {
  PrepTables t;
  int rc = prep_alloc(ctx, nb, ngpown, C != nullptr, &t);
  if (rc) return rc;
  *S = t.S;
  if (C) *C = t.C;
  *wxb = t.wxb;
  *range = t.range;
  prep_rows(ctx, t, nb, 0, ngpown, ncouls, inv, indinv, wx, wx_per_band, wt, am, an, vcoul);
  hipLaunchKernelGGL(k_wx_range, dim3(1), dim3(256), 0, ctx->stream, t.wxb, nb * 3, t.range);
  return check_launch(ctx, "k_gpp_prep");
}

// Forward driver on device pointers; enqueues on ctx->stream.
int gpp_forward_device(bgw_ctx* ctx, size_t nb_, size_t nvband_, size_t ngpown_, size_t ncouls_,
                       const int* inv, const int* indinv, const double* wx, int wx_per_band,
                       const double* wtilde, const double* aqsm, const double* aqsn,
                       const double* I_eps, const double* vcoul, double occ, double* achtemp,
                       double* acht_n1_loc, double* achstemp, double* asxtemp) {
  const int nb = (int)nb_, nvband = (int)nvband_, ngpown = (int)ngpown_, ncouls = (int)ncouls_;
  const double2* wt = reinterpret_cast<const double2*>(wtilde);
  const double2* I = reinterpret_cast<const double2*>(I_eps);
  const double2* am = reinterpret_cast<const double2*>(aqsm);
  const double2* an = reinterpret_cast<const double2*>(aqsn);
  Fwd f;
  int rc = fwd_begin(ctx, nb, ngpown, ncouls, &f);
  if (rc) return rc;
  launch_check_indices(inv, ngpown_, indinv, ncouls_ + 1, ncouls, ncouls, ctx->d_flag, ctx->stream);
  double2* S;
  double *wxb, *range;
  rc = gpp_prep_tables(ctx, nb, ngpown, ncouls, inv, indinv, wx, wx_per_band, wt, am, an, vcoul, &S, nullptr, &wxb,
                       &range);
  if (rc) return rc;
  rc = run_main(ctx, f.cfg, wt, I, S, wxb, range, nb, ngpown, ncouls, f.part, &f.W);
  if (rc) return rc;
  return fwd_end(ctx, f, nb, nvband, ngpown, ncouls, inv, indinv, wt, I, am, an, vcoul, wxb, occ, achtemp,
                 acht_n1_loc, achstemp, asxtemp);
}

// Forward driver on HOST pointers with the transfer of wtilde and I_eps overlapped:
// their rows go up in ranges on the context's copy stream while the main kernel
// runs over the work items of the rows already resident (one launch per range,
// each continuing the per-workgroup totals of the last, so the results equal
// gpp_forward_device's bit for bit).  `d` holds the device buffers the arrays are
// staged into; the other inputs are uploaded first.
int gpp_forward_staged(bgw_ctx* ctx, int nb, int nvband, int ngpown, int ncouls, const GppHostArgs& h,
                       const GppDevBufs& d, double occ, double* d_ach, double* d_loc, double* d_achs,
                       double* d_asx) {
  int rc;
  hipStream_t cs;
  if ((rc = copy_stream(ctx, &cs))) return rc;
  const size_t row = (size_t)ncouls * 16;
  // inputs other than wtilde / I_eps, then the compute stream waits for them
  if ((rc = upload_host(ctx, d.inv, h.inv, (size_t)ngpown * sizeof(int), cs))) return rc;
  if ((rc = upload_host(ctx, d.indinv, h.indinv, (size_t)(ncouls + 1) * sizeof(int), cs))) return rc;
  if ((rc = upload_host(ctx, d.wx, h.wx, (size_t)(h.wx_per_band ? nb * 3 : 3) * sizeof(double), cs))) return rc;
  if ((rc = upload_host(ctx, d.vcoul, h.vcoul, (size_t)ncouls * sizeof(double), cs))) return rc;
  // aqsm/aqsn: whole, when asxtemp or achstemp read them; otherwise only the columns
  // igp(p) that sch_store1 reads, gathered on the host as [nb][ngpown]
  const int gathered = !d_achs && !d_asx;
  if (gathered) {
    std::vector<double2>& gm = ctx->gather_m;
    std::vector<double2>& gn = ctx->gather_n;
    gm.resize((size_t)nb * ngpown);
    gn.resize((size_t)nb * ngpown);
    const double2* hm = reinterpret_cast<const double2*>(h.aqsm);
    const double2* hn = reinterpret_cast<const double2*>(h.aqsn);
    gather_igp_columns(hm, hn, h.inv, h.indinv, nb, ngpown, ncouls, gm.data(), gn.data());
    if ((rc = upload_host(ctx, d.aqsm, gm.data(), gm.size() * 16, cs))) return rc;
    if ((rc = upload_host(ctx, d.aqsn, gn.data(), gn.size() * 16, cs))) return rc;
  } else {
    if ((rc = upload_host(ctx, d.aqsm, h.aqsm, (size_t)nb * row, cs))) return rc;
    if ((rc = upload_host(ctx, d.aqsn, h.aqsn, (size_t)nb * row, cs))) return rc;
  }
  if ((rc = cross_stream_wait(ctx, cs, ctx->stream))) return rc;
  Fwd f;
  if ((rc = fwd_begin(ctx, nb, ngpown, ncouls, &f))) return rc;
  launch_check_indices(d.inv, ngpown, d.indinv, (size_t)ncouls + 1, ncouls, ncouls, ctx->d_flag, ctx->stream);
  PrepTables t;
  if ((rc = prep_alloc(ctx, nb, ngpown, false, &t))) return rc;
  const double2* wt = reinterpret_cast<const double2*>(d.wtilde);
  const double2* I = reinterpret_cast<const double2*>(d.I_eps);
  const double2* am = reinterpret_cast<const double2*>(d.aqsm);
  const double2* an = reinterpret_cast<const double2*>(d.aqsn);
  // ranges of work items: multiples of the grid width W, at most 32 ranges.  Finer
  // ranges shorten the transfer before the first launch and the last launch after the
  // final transfer (benchmark workload, 26 ranges of one grid width: 34 ms per call
  // against 37 ms with 7 ranges)
  const int nchunks = (ncouls + kGChunk - 1) / kGChunk;
  const long items = (long)ngpown * nchunks;
  const long per = std::max<long>(1, (items + 32L * f.W - 1) / (32L * f.W)) * f.W;
  int rows_up = 0;
  for (long i0 = 0; i0 < items; i0 += per) {
    const long i1 = std::min(items, i0 + per);
    const int rows_need = (int)((i1 + nchunks - 1) / nchunks);  // rows holding items < i1
    if (rows_need > rows_up) {
      const size_t off = (size_t)rows_up * row, n = (size_t)(rows_need - rows_up) * row;
      if ((rc = upload_host(ctx, static_cast<char*>(static_cast<void*>(d.wtilde)) + off,
                            static_cast<const char*>(static_cast<const void*>(h.wtilde)) + off, n, cs)))
        return rc;
      if ((rc = upload_host(ctx, static_cast<char*>(static_cast<void*>(d.I_eps)) + off,
                            static_cast<const char*>(static_cast<const void*>(h.I_eps)) + off, n, cs)))
        return rc;
      if ((rc = cross_stream_wait(ctx, cs, ctx->stream))) return rc;
      prep_rows(ctx, t, nb, rows_up, rows_need, ncouls, d.inv, d.indinv, d.wx, h.wx_per_band, wt, am, an, d.vcoul,
                gathered, ngpown);
      if (rows_up == 0)
        hipLaunchKernelGGL(k_wx_range, dim3(1), dim3(256), 0, ctx->stream, t.wxb, nb * 3, t.range);
      if ((rc = check_launch(ctx, "k_gpp_prep"))) return rc;
      rows_up = rows_need;
    }
    if ((rc = run_main(ctx, f.cfg, wt, I, t.S, t.wxb, t.range, nb, ngpown, ncouls, f.part, &f.W, i0, i1)))
      return rc;
  }
  return fwd_end(ctx, f, nb, nvband, ngpown, ncouls, d.inv, d.indinv, wt, I, am, an, d.vcoul, t.wxb, occ, d_ach,
                 d_loc, d_achs, d_asx);
}

// ------------------------------------------------------- allocation-free entry point
// Workspace of bgw_gpp_solver_device_noalloc: a file-scope device array, one copy per
// GPU for the life of the process.  Host code uses only the device address that
// hipGetSymbolAddress returns for the current device (ctx->scratch), never the name.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

size_t al256(size_t n) { return (n + 255) & ~size_t(255); }
}  // namespace

int gpp_scratch_attach(bgw_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return fail(ctx, BGW_ENODEV, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[ctx->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // ctx->device is current
    if (e != hipSuccess || !p)
      return hip_fail(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress");
    g_scratch_ptr[ctx->device] = p;
  }
  ctx->scratch = g_scratch_ptr[ctx->device];
  return BGW_OK;
}

}  // namespace bgw

// The forward's device workspace in the order gpp_forward_device requests it when only
// achtemp and acht_n1_loc are computed: per-workgroup partials, band sums, S table, band
// frequencies and their range, each 256-byte aligned.
extern "C" int bgw_gpp_solver_device_noalloc(bgw_ctx* ctx, size_t number_bands, size_t ngpown, size_t ncouls,
                                             const int* inv_igp_index, const int* indinv, const double* wx_array,
                                             const double* wtilde_array, const double* aqsmtemp,
                                             const double* aqsntemp, const double* I_eps_array, const double* vcoul,
                                             double* achtemp, double* acht_n1_loc) {
  using namespace bgw;
  if (!ctx) return BGW_EINVAL;
  if (!inv_igp_index || !indinv || !wx_array || !wtilde_array || !aqsmtemp || !aqsntemp || !I_eps_array || !vcoul ||
      !achtemp)
    return fail(ctx, BGW_EINVAL, "null array argument");
  if (number_bands == 0 || ngpown == 0 || ncouls == 0) return fail(ctx, BGW_EINVAL, "zero size");
  if (number_bands >= (size_t)INT32_MAX || ngpown >= (size_t)INT32_MAX || ncouls >= (size_t)INT32_MAX ||
      ngpown > SIZE_MAX / 16 / ncouls || number_bands > SIZE_MAX / 16 / ncouls)
    return fail(ctx, BGW_EINVAL, "size exceeds int range");
  if (!ctx->scratch) return fail(ctx, BGW_EHIP, "the scratch array is not available on this device");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, BGW_ENODEV, "hipSetDevice");
  const int nb = (int)number_bands;
  int W = 0;
  run_main(ctx, pick_main(nb), nullptr, nullptr, nullptr, nullptr, nullptr, nb, (int)ngpown, (int)ncouls, nullptr, &W);
  const double need = (double)al256((size_t)W * nb * 48) + (double)al256((size_t)nb * 96) +
                      (double)al256(16 * number_bands * ngpown) + (double)al256((size_t)nb * 24) + 256.0 +
                      (double)al256(moments_bytes(ctx, (int)ngpown, (int)ncouls));
  if (need > (double)kScratchBytes)
    return fail(ctx, BGW_EINVAL, "number_bands x ngpown too large for the 64 MiB scratch of the allocation-free call");
  ctx->arena = static_cast<char*>(ctx->scratch);
  ctx->arena_bytes = kScratchBytes;
  ctx->arena_used = 0;
  const int rc = gpp_forward_device(ctx, number_bands, 0, ngpown, ncouls, inv_igp_index, indinv, wx_array, 1,
                                    wtilde_array, aqsmtemp, aqsntemp, I_eps_array, vcoul, 0.0, achtemp, acht_n1_loc,
                                    nullptr, nullptr);
  ctx->arena = nullptr;
  return rc;
}
