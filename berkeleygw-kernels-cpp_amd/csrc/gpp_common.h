// gpp_common.h -- device helpers shared by the GPP/FF kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "bgw_internal.h"

namespace bgw {

// GPP constants: to1 and limittwo as declared in GPP/gpp.cpp:167-168, with the
// companion thresholds of BerkeleyGW's occupied-band branch.
constexpr double kTo1 = 1e-6;
constexpr double kLimitTwo = 0.25;
constexpr double kLimitOne = 1.0 / (kTo1 * 4.0);
constexpr double kSexcut = 4.0;

__host__ __device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double cnorm2(double2 a) { return a.x * a.x + a.y * a.y; }

// Sums of the N doubles v[] over the 64 lanes of a wave, reduce-scatter style: while N
// is even each exchange halves the values a lane carries (a lane keeps the lower or the
// upper half by its lane bit and adds its partner's copy of that half), then the rest is
// summed across the remaining lane bits.  On return v[0..scatter_len(N)) of a lane holds
// the wave totals of components [base, base + scatter_len(N)); `writer` is set in one lane
// of each component block.  The order of every addition is fixed, so the sums are bitwise
// reproducible.
constexpr int scatter_len(int n, int mask = 32) {
  return mask == 0 ? n : (n % 2 == 0 ? scatter_len(n / 2, mask / 2) : n);
}

template <int N, int MASK>
__device__ __forceinline__ void wave_sum_scatter(double* v, int lane, int& base, bool& writer) {
  if constexpr (MASK > 0) {
    if constexpr (N % 2 == 0) {
      const bool up = (lane & MASK) != 0;
#pragma unroll
      for (int j = 0; j < N / 2; ++j) {
        const double send = up ? v[j] : v[j + N / 2];
        const double keep = up ? v[j + N / 2] : v[j];
        v[j] = keep + __shfl_xor(send, MASK, 64);
      }
      if (up) base += N / 2;
      wave_sum_scatter<N / 2, MASK / 2>(v, lane, base, writer);
    } else {
      // a + b in one lane and b + a in its partner: the same value (addition commutes)
#pragma unroll
      for (int j = 0; j < N; ++j) v[j] += __shfl_xor(v[j], MASK, 64);
      if (lane & MASK) writer = false;
      wave_sum_scatter<N, MASK / 2>(v, lane, base, writer);
    }
  }
}

// igp = indinv[inv_igp_index[p]] (GPP/gpp.cpp:362-363), clamped into range so a bad
// index can never address outside the arrays (bad indices are reported separately).
__device__ inline int gpp_igp(const int* __restrict__ inv, const int* __restrict__ indinv, int p, int ncouls) {
  int a = inv[p];
  a = a < 0 ? 0 : (a > ncouls ? ncouls : a);
  int g = indinv[a];
  return g < 0 ? 0 : (g >= ncouls ? ncouls - 1 : g);
}
// FF form: indinv has ncouls entries (FF/ff.cpp:552), inv_igp_index < ncouls.
__device__ inline int ff_igp(const int* __restrict__ inv, const int* __restrict__ indinv, int p, int ncouls) {
  int a = inv[p];
  a = a < 0 ? 0 : (a >= ncouls ? ncouls - 1 : a);
  int g = indinv[a];
  return g < 0 ? 0 : (g >= ncouls ? ncouls - 1 : g);
}

// ssx of BerkeleyGW's occupied-band GPP branch for one (wx, wtilde, I_eps):
// regular region ssx = Omega2 conj(cden)/|cden|^2 with cden = wx^2 - wtilde^2,
// near-pole region ssx = -Omega2 conj(cden)/|cden|^2 delw with cden = 4 wtilde^2 (delw+1/2),
// zero otherwise (gpp_ssx_raw); gpp_ssx also applies the cut-off to zero when
// |ssx| > sexcut |I_eps| for wx < 0.
// The parts of the ssx term that do not depend on wx, for one (wtilde, I_eps): computed
// once and shared by every band and frequency of a G-vector (the same operations as
// inline, so the values are bitwise those of the per-term form).
struct SsxG {
  double2 wt, wt2, Omega2;
  double icut;  // sexcut |I_eps|
};

__device__ inline SsxG gpp_ssx_g(double2 wt, double2 I) {
  SsxG s;
  s.wt = wt;
  s.wt2 = cmul(wt, wt);
  s.Omega2 = cmul(s.wt2, I);
  s.icut = kSexcut * sqrt(cnorm2(I));
  return s;
}

// The branch of the ssx term: regular when wdiffr > limittwo and |delw|^2 < limitone, with
// delw = wt conj(wdiff) / wdiffr and wdiffr = |wdiff|^2.  Most terms are decided without the
// division: |wt conj(wdiff)|^2 < limitone wdiffr^2 (1 - 2^-40) implies that |delw|^2 formed
// with the division (a few roundings, each within 2^-53) is below limitone too.  Otherwise
// (near the threshold, an overflow, NaN) delw is formed and the test made exactly, so every
// term takes the branch the exact test gives; delw is valid whenever regular is false.
// gpp_ssx_surely_regular is the division-free test: true only where the exact test gives
// the regular branch.
__device__ inline bool gpp_ssx_surely_regular(double2 wt, double2 wdiff, double wdiffr) {
  const double2 a = cmul(wt, cconj(wdiff));
  return wdiffr > kLimitTwo && cnorm2(a) < kLimitOne * (wdiffr * wdiffr) * (1.0 - 0x1p-40);
}

struct SsxBranch {
  bool regular;
  double2 delw;
  double delwr;
};
__device__ inline SsxBranch gpp_ssx_branch(double2 wt, double2 wdiff, double wdiffr) {
  SsxBranch b;
  b.regular = gpp_ssx_surely_regular(wt, wdiff, wdiffr);
  b.delw = make_double2(0.0, 0.0);
  b.delwr = 0.0;
  if (!b.regular) {
    const double rden = 1.0 / wdiffr;
    b.delw = cscale(cmul(wt, cconj(wdiff)), rden);
    b.delwr = cnorm2(b.delw);
    b.regular = wdiffr > kLimitTwo && b.delwr < kLimitOne;
  }
  return b;
}

__device__ inline double2 gpp_ssx_raw(double wx, const SsxG& g) {
  const double2 wt = g.wt, wt2 = g.wt2, Omega2 = g.Omega2;
  const double2 wdiff = make_double2(wx - wt.x, -wt.y);
  const double wdiffr = cnorm2(wdiff);
  const SsxBranch br = gpp_ssx_branch(wt, wdiff, wdiffr);
  const double2 delw = br.delw;
  const double delwr = br.delwr;
  double2 ssx;
  if (br.regular) {
    const double2 cden = make_double2(wx * wx - wt2.x, -wt2.y);
    const double rd = 1.0 / cnorm2(cden);
    ssx = cscale(cmul(Omega2, cconj(cden)), rd);
  } else if (delwr > kTo1) {
    const double2 cden = cscale(cmul(wt2, make_double2(delw.x + 0.5, delw.y)), 4.0);
    const double rd = 1.0 / cnorm2(cden);
    const double2 t = cscale(cmul(make_double2(-Omega2.x, -Omega2.y), cconj(cden)), rd);
    ssx = cmul(t, delw);
  } else {
    ssx = make_double2(0.0, 0.0);
  }
  return ssx;
}

__device__ inline double2 gpp_ssx_raw(double wx, double2 wt, double2 I) { return gpp_ssx_raw(wx, gpp_ssx_g(wt, I)); }

// the |ssx| > sexcut |I| cut-off for wx < 0, decided on the value before the cut (wx is
// tested first: it is the same for every lane of a band and frequency)
__device__ inline bool gpp_ssx_cut(double wx, double2 ssx, double icut) { return wx < 0.0 && sqrt(cnorm2(ssx)) > icut; }

__device__ inline bool gpp_ssx_cut(double wx, double2 ssx, double2 I) {
  return gpp_ssx_cut(wx, ssx, kSexcut * sqrt(cnorm2(I)));
}

__device__ inline double2 gpp_ssx(double wx, const SsxG& g) {
  const double2 ssx = gpp_ssx_raw(wx, g);
  return gpp_ssx_cut(wx, ssx, g.icut) ? make_double2(0.0, 0.0) : ssx;
}

__device__ inline double2 gpp_ssx(double wx, double2 wt, double2 I) { return gpp_ssx(wx, gpp_ssx_g(wt, I)); }

// ssx of one G-vector at a band's three frequencies.  When every frequency is >= 0 (no
// cut-off) and surely regular, the three regular values are formed in straight-line code,
// so their divisions overlap instead of each waiting behind its own branch; otherwise each
// frequency takes gpp_ssx.  Either way each value is the one of its branch.
__device__ inline void gpp_ssx3(const double* __restrict__ wx, const SsxG& g, double2 (&out)[3]) {
  bool reg = true;
#pragma unroll
  for (int iw = 0; iw < 3; ++iw) {
    const double2 wdiff = make_double2(wx[iw] - g.wt.x, -g.wt.y);
    reg = reg && wx[iw] >= 0.0 && gpp_ssx_surely_regular(g.wt, wdiff, cnorm2(wdiff));
  }
  if (reg) {
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) {
      const double2 cden = make_double2(wx[iw] * wx[iw] - g.wt2.x, -g.wt2.y);
      const double rd = 1.0 / cnorm2(cden);
      out[iw] = cscale(cmul(g.Omega2, cconj(cden)), rd);
    }
  } else {
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) out[iw] = gpp_ssx(wx[iw], g);
  }
}

// MFMA trace contraction (contract.hip):
//   out = sum_{r} sum_p W(r,p) sum_g A(rows[r], g) * B(p, g),
//   W(r,p) = wts[r] * vcoul(igp(p)) * conj(aqsm(rows[r], igp(p))).
int contract_achstemp(bgw_ctx* ctx, int nb, int ngpown, int ncouls, const int* inv,
                      const int* indinv, const double2* aqsn, const double2* aqsm,
                      const double2* Ieps, const double* vcoul, double2* achstemp);

}  // namespace bgw
