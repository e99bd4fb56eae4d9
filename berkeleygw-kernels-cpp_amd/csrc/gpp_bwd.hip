// gpp_bwd.hip -- adjoint of the GPP forward for L = Re sum_iw (asxtemp[iw] + achtemp[iw]).
//
// With c(n1,p) = conj(aqsm(n1,igp)) aqsn(n1,igp) 0.5 vcoul(igp) (so that
// sch_store1 = c * wtilde(p,igp), GPP/gpp.cpp:369-371) and inv = 1/(wx(n1,iw) - wtilde(p,g)):
//   df/dI(p,g)      = wtilde(p,igp) * sum_{n1,iw} c inv               (+ SX part)
//   df/dwtilde(p,g) = wtilde(p,igp) * I(p,g) * sum_{n1,iw} c inv^2    (+ SX part)
//                     + [g == igp(p)] * T(p),  T(p) = sum_g I(p,g) sum_{n1,iw} c inv
//   dSigma/domega[iw] = -sum wtilde(p,igp) I sum_n1 c inv^2           (+ SX part)
// and dL/dz = conj(df/dz) (dL/dRe z + i dL/dIm z for L = Re f, f holomorphic).
// The SX part (bands n1 < nvband) differentiates the branch of BerkeleyGW's ssx
// that the forward takes for each term.
//
// Mapping: one lane per (p, g) (coalesced along ig), the band loop runs over LDS
// broadcasts of c(n1,p) and wx(n1,:); the pole terms are recomputed, never stored.
// Each (p,g) output is owned by one lane, so the G-vector outputs are written
// directly; dSigma/domega and T(p) go through fixed-order partial reductions.
//
// Moment form (BGW_FP64_MOMENTS): the band sums of a pole w are sums over the band
// frequencies x = wx(n1,iw),
//   F_iw(w) = sum_n1 c/(x - w),  G = sum_iw F_iw,  H_iw = dF_iw/dw = sum_n1 c/(x - w)^2,
// the transpose of the forward's pole sum.  With e the middle of the frequency range, c_p
// a centre of row p's poles and phi(z) = (z - c_p)/(z - e),
//   1/(x - w) = (c_p - e)/((x - c_p)(w - e)) sum_m (phi(w)/phi(x))^m
// and, squared, 1/(x - w)^2 = ((c_p - e)/((x - c_p)(w - e)))^2 sum_m (m + 1) (phi(w)/phi(x))^m.
// So with g = 1/(w - e) and psi = 1/phi(x) = (x - e)/(x - c_p),
//   G(w) = (c_p - e) g P_A(phi(w)),       A_m = sum_iw sum_n1 c(n1,p) psi^m / (x - c_p),
//   H_iw(w) = (c_p - e)^2 g^2 P_Q_iw(phi), Q_iw,m = (m + 1) sum_n1 c(n1,p) psi^m / (x - c_p)^2,
// with P_A, P_Q_iw the polynomials of the row's target moments.  The tail of a term beyond
// m = M - 1 is exactly q^M (resp. q^M ((M + 1) - M q)) times the term, q = phi(w)/phi(x),
// so the truncation error of a pole's G terms is at most rho^M and that of its H terms at
// most rho^M (M + 1 + M rho) times the sum of their magnitudes (rho = |phi(w)| h/D,
// h = max |x - e|, D = min |x - c_p|); a pole is "far" when the larger bound is below 2^-54,
// and a wave whose poles are all far takes the expansion (four Horner chains of M terms per
// pole) instead of the band loop (O(3 nb)).
#include <cfloat>
#include <cmath>

#include "bgw_dd.h"
#include "bgw_internal.h"
#include "contract.h"
#include "gpp_common.h"

namespace bgw {
int gpp_prep_tables(bgw_ctx* ctx, int nb, int ngpown, int ncouls, const int* inv, const int* indinv,
                    const double* wx, int wx_per_band, const double2* wt, const double2* am, const double2* an,
                    const double* vcoul, double2** S, double2** C, double** wxb, double** range);

namespace {

constexpr int BT = 256;  // threads per block
constexpr int GPT = 2;   // G-vectors per thread
constexpr int kMaxBandsLds = 1024;

constexpr int kBwdMom = 16;  // target moments per frequency
constexpr double kBwdBound = 0x1p-54;

// Row p's centre (from the bounding box of sample poles among its first 256: any centre is
// valid, each pole's bound is checked), its target moments and whether they are usable
// (finite, and the centre away from the frequency range):
//   A[m]     = sum_iw sum_n1 c psi^m / (x - c)      (the frequencies' F summed)
//   Q[iw][m] = (m + 1) sum_n1 c psi^m / (x - c)^2   (H_iw).
struct __align__(16) BwdRow {
  double cx, cy;
  int ok, pad_[3];
  double2 A[kBwdMom];
  double2 Q[3][kBwdMom];
};

// One workgroup per G' row: lanes take bands n1 = tid, tid + BT, ...; the moments are
// summed per lane, then over each wave by reduce-scatter and over the waves in wave order
// (fixed order: bitwise reproducible).
__global__ __launch_bounds__(BT) void k_gpp_bwd_rowmom(const double2* __restrict__ wt, const double2* __restrict__ C,
                                                      const double* __restrict__ wxb,
                                                      const double* __restrict__ wxrange, int nb, int ncouls,
                                                      BwdRow* __restrict__ rows) {
  __shared__ double sbox[BT / 64][3][5];
  __shared__ double part[BT / 64][2 * kBwdMom];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int p = blockIdx.x;
  // the bounding box of the sample poles more than 2h from e, on the side of e that holds
  // more of them (all sample poles when none is): a few poles near the band frequencies do
  // not move the centre
  const double xmin = wxrange[0], xmax = wxrange[1], e = 0.5 * (xmin + xmax);
  const double h = fmax(xmax - e, e - xmin);
  // boxes: [0] all, [1] left of e, [2] right of e; (min re, max re, min im, max im, count)
  double bx[3][4];
  double cnt[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 3; ++q) bx[q][0] = bx[q][2] = INFINITY, bx[q][1] = bx[q][3] = -INFINITY;
  if (tid < ncouls) {
    const double2 w = wt[(size_t)p * ncouls + tid];
    const double dr = w.x - e;
    const int side = fma(dr, dr, w.y * w.y) > 4.0 * h * h ? (w.x < e ? 1 : 2) : 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (q == 0 || q == side) {
        bx[q][0] = bx[q][1] = w.x;
        bx[q][2] = bx[q][3] = w.y;
        cnt[q] = 1.0;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
      bx[q][0] = fmin(bx[q][0], __shfl_xor(bx[q][0], m, 64));
      bx[q][1] = fmax(bx[q][1], __shfl_xor(bx[q][1], m, 64));
      bx[q][2] = fmin(bx[q][2], __shfl_xor(bx[q][2], m, 64));
      bx[q][3] = fmax(bx[q][3], __shfl_xor(bx[q][3], m, 64));
      cnt[q] += __shfl_xor(cnt[q], m, 64);
    }
    if (lane == 0) {
      sbox[wave][q][0] = bx[q][0];
      sbox[wave][q][1] = bx[q][1];
      sbox[wave][q][2] = bx[q][2];
      sbox[wave][q][3] = bx[q][3];
      sbox[wave][q][4] = cnt[q];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    bx[q][0] = sbox[0][q][0], bx[q][1] = sbox[0][q][1], bx[q][2] = sbox[0][q][2], bx[q][3] = sbox[0][q][3];
    cnt[q] = sbox[0][q][4];
#pragma unroll
    for (int w = 1; w < BT / 64; ++w) {
      bx[q][0] = fmin(bx[q][0], sbox[w][q][0]);
      bx[q][1] = fmax(bx[q][1], sbox[w][q][1]);
      bx[q][2] = fmin(bx[q][2], sbox[w][q][2]);
      bx[q][3] = fmax(bx[q][3], sbox[w][q][3]);
      cnt[q] += sbox[w][q][4];
    }
  }
  const int pick = cnt[1] + cnt[2] == 0.0 ? 0 : (cnt[1] >= cnt[2] ? 1 : 2);
  double cx = 0.5 * (bx[pick][0] + bx[pick][1]), cy = 0.5 * (bx[pick][2] + bx[pick][3]);
  if (!(fabs(cx) <= 0x1p+1000 && fabs(cy) <= 0x1p+1000)) cx = cy = 0.0;  // a sample of NaN/Inf poles
  const double dx = cx < xmin ? xmin - cx : (cx > xmax ? cx - xmax : 0.0);
  const double e0 = xmin - cx, e1 = xmax - cx;
  // every |x - c|^2 in [2^-150, 2^150] (false for NaN)
  bool ok = fma(dx, dx, cy * cy) >= 0x1p-150 && fmax(e0 * e0, e1 * e1) + cy * cy <= 0x1p+150;
  // one reduce-scatter of 2 kBwdMom doubles over the block, in fixed order; the total of
  // component j lands in tot[j] for threads j < 2 kBwdMom
  auto block_sum = [&](double (&v)[2 * kBwdMom], double* tot) {
    int base = 0;
    bool writer = true;
    wave_sum_scatter<2 * kBwdMom, 32>(v, lane, base, writer);
    constexpr int NF = scatter_len(2 * kBwdMom);
    if (writer) {
#pragma unroll
      for (int j = 0; j < NF; ++j) part[wave][base + j] = v[j];
    }
    __syncthreads();
    if (tid < 2 * kBwdMom) {
      double t = part[0][tid];
#pragma unroll
      for (int w = 1; w < BT / 64; ++w) t += part[w][tid];
      ok = ok && fabs(t) <= DBL_MAX;
      tot[tid] = t;
    }
    __syncthreads();
  };
  double va[2 * kBwdMom];  // A: every frequency's first-order terms
#pragma unroll
  for (int j = 0; j < 2 * kBwdMom; ++j) va[j] = 0.0;
  for (int iw = 0; iw < 3; ++iw) {
    double v[2 * kBwdMom];
#pragma unroll
    for (int j = 0; j < 2 * kBwdMom; ++j) v[j] = 0.0;
    for (int n1 = tid; n1 < nb; n1 += BT) {
      const double x = wxb[3 * n1 + iw];
      const double2 cn = C[(size_t)p * nb + n1];
      const double a = x - cx, r = 1.0 / fma(a, a, cy * cy);
      const double tr = a * r, ti = cy * r;  // 1/(x - c)
      const double sr = (x - e) * tr, si = (x - e) * ti;  // psi
      double qr = fma(cn.x, tr, -cn.y * ti), qi = fma(cn.x, ti, cn.y * tr);  // c/(x - c)
      double ur = fma(qr, tr, -qi * ti), ui = fma(qr, ti, qi * tr);          // c/(x - c)^2
#pragma unroll
      for (int m = 0; m < kBwdMom; ++m) {
        va[2 * m] += qr;
        va[2 * m + 1] += qi;
        v[2 * m] += ur;
        v[2 * m + 1] += ui;
        if (m + 1 < kBwdMom) {
          const double t = fma(qr, sr, -qi * si);
          qi = fma(qr, si, qi * sr);
          qr = t;
          const double t2 = fma(ur, sr, -ui * si);
          ui = fma(ur, si, ui * sr);
          ur = t2;
        }
      }
    }
    double* Q = reinterpret_cast<double*>(rows[p].Q[iw]);
    block_sum(v, Q);
    if (tid < 2 * kBwdMom) {
      const double q = Q[tid] * (double)(tid / 2 + 1);  // the factor m + 1
      Q[tid] = q;
      ok = ok && fabs(q) <= DBL_MAX;
    }
  }
  block_sum(va, reinterpret_cast<double*>(rows[p].A));
  const int all_ok = __syncthreads_and(ok ? 1 : 0);
  if (tid == 0) {
    rows[p].cx = cx;
    rows[p].cy = cy;
    rows[p].ok = all_ok;
  }
}

// Holomorphic derivatives of gpp_ssx with respect to I, wtilde and wx (branch-wise) for one
// frequency wx of a band.  The branch is decided by the forward's own operations.  Regular
// branch (ic = 1/(wx^2 - wt^2)): dI = wt^2 ic, dW = 2 wx^2 wt I ic^2, dX = -2 wx wt^2 I ic^2;
// its dI and dW are returned as the frequency-dependent factors ic and 2 wx^2 ic^2, summed
// over the band's frequencies into s1 and s2 (the caller applies wt^2 and wt I once per band);
// the near-pole branch's dI and dW are summed into oI and oW as they are.  wt2 = wt^2 and
// w2I = wt^2 I.
__device__ inline void ssx_grad_regular(double wx, double2 wt2, double2 w2I, double2& s1, double2& s2, double2& dX) {
  const double2 cden = make_double2(wx * wx - wt2.x, -wt2.y);
  const double2 ic = cscale(cconj(cden), 1.0 / cnorm2(cden));  // 1/cden
  const double2 ic2 = cmul(ic, ic);
  s1 = cadd(s1, ic);
  s2 = cadd(s2, cscale(ic2, 2.0 * wx * wx));
  dX = cscale(cmul(w2I, ic2), -2.0 * wx);
}

__device__ inline void ssx_grad(double wx, double2 wt, double2 wt2, double2 w2I, double2 I, double2& s1, double2& s2,
                                double2& oI, double2& oW, double2& dX) {
  dX = make_double2(0.0, 0.0);
  // cut terms are constant zero; the cut applies only for wx < 0 (tested first, so the
  // value it needs is formed only then)
  if (wx < 0.0 && gpp_ssx_cut(wx, gpp_ssx_raw(wx, wt, I), I)) return;
  const double2 wdiff = make_double2(wx - wt.x, -wt.y);
  const double wdiffr = cnorm2(wdiff);
  const SsxBranch br = gpp_ssx_branch(wt, wdiff, wdiffr);
  const double2 delw = br.delw;  // = wt / wdiff, formed when the branch is not regular
  const double delwr = br.delwr;
  // 1/z = conj(z)/|z|^2
  auto inv = [](double2 z) { return cscale(cconj(z), 1.0 / cnorm2(z)); };
  if (br.regular) {
    ssx_grad_regular(wx, wt2, w2I, s1, s2, dX);
  } else if (delwr > kTo1) {
    const double2 dp = make_double2(delw.x + 0.5, delw.y);
    const double2 idp = inv(dp);
    const double2 q = cmul(delw, idp);
    const double2 dq = cscale(cmul(idp, idp), 0.5);
    const double2 iwd = inv(wdiff);
    const double2 iwd2 = cmul(iwd, iwd);
    oI = cadd(oI, cscale(q, -0.25));
    oW = cadd(oW, cscale(cmul(cmul(I, dq), iwd2), -0.25 * wx));
    dX = cscale(cmul(cmul(cmul(I, dq), iwd2), wt), 0.25);
  }
}

// three workgroups per CU (168 VGPRs, three waves per SIMD: 2-3% faster than two at 194
// VGPRs; the few spill slots are in the screened-exchange loop over the valence bands)
template <bool FAR>
__global__ __launch_bounds__(BT, 3) void k_gpp_bwd(const double2* __restrict__ wt, const double2* __restrict__ Ieps,
                                                const double2* __restrict__ C, const double* __restrict__ wxb,
                                                const double* __restrict__ wxrange, const int* __restrict__ inv,
                                                const int* __restrict__ indinv, const double2* __restrict__ aqsm,
                                                const double2* __restrict__ aqsn, const double* __restrict__ vcoul,
                                                int nb, int nvband, int ngpown, int ncouls, double occ,
                                                double2* __restrict__ dL_dwt, double2* __restrict__ dL_dI,
                                                double4* __restrict__ ds_part /*[blocks][3]*/,
                                                double4* __restrict__ T_part /*[blocks]*/,
                                                const BwdRow* __restrict__ rows) {
  __shared__ double2 sc[kMaxBandsLds];
  __shared__ double sx[kMaxBandsLds * 3];
  __shared__ dd red[BT / 64][8];
  __shared__ double2 sA[FAR ? 4 * kBwdMom : 1];  // the row's A, then Q[0..2]
  const int nchunk = (ncouls + BT * GPT - 1) / (BT * GPT);
  const int p = blockIdx.x / nchunk;
  const int ch = blockIdx.x - p * nchunk;
  const int igp = gpp_igp(inv, indinv, p, ncouls);
  const double2 wt_igp = wt[(size_t)p * ncouls + igp];
  const double wxmin = wxrange[0], wxmax = wxrange[1];

  double2 w[GPT], I[GPT];
  bool valid[GPT], fast[GPT];
  double G[GPT][2], H[GPT][3][2], dIs[GPT][2], dWs[GPT][2], dsx[3][2];
#pragma unroll
  for (int iw = 0; iw < 3; ++iw) dsx[iw][0] = dsx[iw][1] = 0.0;
#pragma unroll
  for (int k = 0; k < GPT; ++k) {
    const int g = ch * BT * GPT + k * BT + threadIdx.x;
    valid[k] = g < ncouls;
    w[k] = valid[k] ? wt[(size_t)p * ncouls + g] : make_double2(0.0, 1.0);
    I[k] = valid[k] ? Ieps[(size_t)p * ncouls + g] : make_double2(0.0, 0.0);
    const double lo = w[k].x < wxmin ? wxmin - w[k].x : (w[k].x > wxmax ? w[k].x - wxmax : 0.0);
    const double hi = fmax(fabs(wxmin - w[k].x), fabs(wxmax - w[k].x));
    const double wi2 = w[k].y * w[k].y;
    fast[k] = (fma(lo, lo, wi2) >= 0x1p-300) && (fma(hi, hi, wi2) <= 0x1p+300);
    G[k][0] = G[k][1] = dIs[k][0] = dIs[k][1] = dWs[k][0] = dWs[k][1] = 0.0;
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) H[k][iw][0] = H[k][iw][1] = 0.0;
  }

  // FAR: a wave whose poles are all far takes the expansion; the band loop runs (its LDS
  // staging for the whole block) when any wave of the block needs it
  bool wave_far = false, need_loop = true;
  if constexpr (FAR) {
    // the wave's valid poles far from the band frequencies about the row's centre
    const BwdRow* __restrict__ rw = rows + p;
    const double cx = rw->cx, cy = rw->cy;
    if (threadIdx.x < 4 * kBwdMom)
      sA[threadIdx.x] = threadIdx.x < kBwdMom ? rw->A[threadIdx.x] : rw->Q[threadIdx.x / kBwdMom - 1][threadIdx.x % kBwdMom];
    const double xmin = wxrange[0], xmax = wxrange[1], e = 0.5 * (xmin + xmax);
    const double h = fmax(xmax - e, e - xmin) * (1.0 + 0x1p-40);
    const double dx = cx < xmin ? xmin - cx : (cx > xmax ? cx - xmax : 0.0);
    const double hD = h / (sqrt(fma(dx, dx, cy * cy)) * (1.0 - 0x1p-40));  // h/D
    double gr[GPT], gi[GPT], fr[GPT], fi[GPT];
    bool far = rw->ok != 0;
#pragma unroll
    for (int k = 0; k < GPT; ++k) {
      const double dr = w[k].x - e, di = w[k].y;
      const double q = 1.0 / fma(dr, dr, di * di);
      gr[k] = dr * q;
      gi[k] = -di * q;  // 1/(w - e)
      const double zr = w[k].x - cx, zi = w[k].y - cy;
      fr[k] = fma(zr, gr[k], -zi * gi[k]);
      fi[k] = fma(zr, gi[k], zi * gr[k]);  // phi(w)
      const double rho = sqrt(fma(fr[k], fr[k], fi[k] * fi[k])) * (1.0 + 0x1p-40) * hD;
      const double r2 = rho * rho, r4 = r2 * r2, r8 = r4 * r4, r16 = r8 * r8;
      // the second-order tail; it bounds the first-order one (rho^16) too
      const double bound = r16 * fma(kBwdMom, rho, kBwdMom + 1.0);
      far = far && (!valid[k] || bound <= kBwdBound);  // false for NaN
    }
    wave_far = __all(far) != 0;
    need_loop = __syncthreads_or(wave_far ? 0 : 1) != 0;  // also orders the sA writes before their reads
    if (wave_far) {
      // G = (c - e) g P_A(phi), H_iw = (c - e)^2 g^2 P_Q_iw(phi)
      const double cer = cx - e, cei = cy;
      // the 4 GPT Horner chains (P_A and the three P_Q_iw of each pole) advance together, so
      // consecutive FMAs are independent
      double pr[GPT][4], pi[GPT][4];
#pragma unroll
      for (int k = 0; k < GPT; ++k) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          pr[k][j] = sA[j * kBwdMom + kBwdMom - 1].x;
          pi[k][j] = sA[j * kBwdMom + kBwdMom - 1].y;
        }
      }
#pragma unroll 3
      for (int m = kBwdMom - 2; m >= 0; --m) {
        double2 a[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) a[j] = sA[j * kBwdMom + m];
#pragma unroll
        for (int k = 0; k < GPT; ++k) {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const double npr = fma(pr[k][j], fr[k], fma(-pi[k][j], fi[k], a[j].x));
            pi[k][j] = fma(pr[k][j], fi[k], fma(pi[k][j], fr[k], a[j].y));
            pr[k][j] = npr;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < GPT; ++k) {
        const double cgr = fma(cer, gr[k], -cei * gi[k]), cgi = fma(cer, gi[k], cei * gr[k]);  // (c - e) g
        const double c2r = fma(cgr, cgr, -cgi * cgi), c2i = 2.0 * cgr * cgi;                  // its square
        G[k][0] = fma(cgr, pr[k][0], -cgi * pi[k][0]);
        G[k][1] = fma(cgr, pi[k][0], cgi * pr[k][0]);
#pragma unroll
        for (int iw = 0; iw < 3; ++iw) {
          H[k][iw][0] = fma(c2r, pr[k][iw + 1], -c2i * pi[k][iw + 1]);
          H[k][iw][1] = fma(c2r, pi[k][iw + 1], c2i * pr[k][iw + 1]);
        }
      }
    }
  }

  for (int b0 = 0; b0 < (need_loop ? nb : 0); b0 += kMaxBandsLds) {
    const int nbt = min(kMaxBandsLds, nb - b0);
    __syncthreads();
    for (int i = threadIdx.x; i < nbt; i += BT) {
      sc[i] = C[(size_t)p * nb + b0 + i];
      sx[3 * i] = wxb[3 * (b0 + i)];
      sx[3 * i + 1] = wxb[3 * (b0 + i) + 1];
      sx[3 * i + 2] = wxb[3 * (b0 + i) + 2];
    }
    __syncthreads();
    for (int j = 0; j < (wave_far ? 0 : nbt); ++j) {
      const double2 c = sc[j];
      const double x0 = sx[3 * j], x1 = sx[3 * j + 1], x2 = sx[3 * j + 2];
#pragma unroll
      for (int k = 0; k < GPT; ++k) {
        const double wr = w[k].x, wi = w[k].y, wi2 = wi * wi;
        const double u0 = x0 - wr, u1 = x1 - wr, u2 = x2 - wr;
        const double d0 = fma(u0, u0, wi2), d1 = fma(u1, u1, wi2), d2 = fma(u2, u2, wi2);
        double r0, r1, r2, q0r, q0i, q1r, q1i, q2r, q2i;
        if (fast[k]) {
          const double m01 = d0 * d1, m = m01 * d2;
          double r = __builtin_amdgcn_rcp(m);
          r = fma(r, fma(-m, r, 1.0), r);
          r2 = r * m01;
          const double q = r * d2;
          r0 = q * d1;
          r1 = q * d0;
          // inv^2 = r^2 (u^2 - wi^2, 2 u wi) = (r - 2 wi^2 r^2, 2 wi u r^2), as u^2 + wi^2 = d = 1/r;
          // every d here lies in [2^-300, 2^300], so r^2 stays a normal double
          const double m2wi2 = -2.0 * wi2, twi = 2.0 * wi;
          const double t0 = r0 * r0, t1 = r1 * r1, t2 = r2 * r2;
          q0r = fma(m2wi2, t0, r0), q0i = twi * (u0 * t0);
          q1r = fma(m2wi2, t1, r1), q1i = twi * (u1 * t1);
          q2r = fma(m2wi2, t2, r2), q2i = twi * (u2 * t2);
        } else {
          r0 = 1.0 / d0;
          r1 = 1.0 / d1;
          r2 = 1.0 / d2;
          // inv^2 = (a^2 - b^2, 2ab) with (a, b) = (u, wi) r: no r^2, which can leave the
          // double range for the denominators this path takes
          const double a0 = u0 * r0, b0_ = wi * r0, a1 = u1 * r1, b1 = wi * r1, a2 = u2 * r2, b2 = wi * r2;
          q0r = fma(a0, a0, -b0_ * b0_), q0i = 2.0 * a0 * b0_;
          q1r = fma(a1, a1, -b1 * b1), q1i = 2.0 * a1 * b1;
          q2r = fma(a2, a2, -b2 * b2), q2i = 2.0 * a2 * b2;
        }
        // inv_iw = 1/(wx - wtilde) = (u, wi) * r; the band's sum over iw
        const double sr = fma(u0, r0, fma(u1, r1, u2 * r2)), si = wi * (r0 + r1 + r2);
        G[k][0] = fma(c.x, sr, fma(-c.y, si, G[k][0]));
        G[k][1] = fma(c.x, si, fma(c.y, sr, G[k][1]));
        H[k][0][0] = fma(c.x, q0r, fma(-c.y, q0i, H[k][0][0]));
        H[k][0][1] = fma(c.x, q0i, fma(c.y, q0r, H[k][0][1]));
        H[k][1][0] = fma(c.x, q1r, fma(-c.y, q1i, H[k][1][0]));
        H[k][1][1] = fma(c.x, q1i, fma(c.y, q1r, H[k][1][1]));
        H[k][2][0] = fma(c.x, q2r, fma(-c.y, q2i, H[k][2][0]));
        H[k][2][1] = fma(c.x, q2i, fma(c.y, q2r, H[k][2][1]));
      }
    }
  }

  // screened-exchange part, bands n1 < nvband
  for (int n1 = 0; n1 < nvband; ++n1) {
    const double2 am = aqsm[(size_t)n1 * ncouls + igp];
    const double vc = vcoul[igp];
    const double2 m = make_double2(am.x * occ * vc, -am.y * occ * vc);
#pragma unroll
    for (int k = 0; k < GPT; ++k) {
      if (!valid[k]) continue;
      const int g = ch * BT * GPT + k * BT + threadIdx.x;
      const double2 ma = cmul(m, aqsn[(size_t)n1 * ncouls + g]);
      const double2 wt2 = cmul(w[k], w[k]), w2I = cmul(wt2, I[k]);
      double2 s1 = make_double2(0.0, 0.0), s2 = s1, oI = s1, oW = s1, gX[3];
      // the band's three frequencies in straight-line code when all are >= 0 (no cut-off)
      // and surely regular, so their divisions overlap; otherwise branch by branch
      bool reg = true;
#pragma unroll
      for (int iw = 0; iw < 3; ++iw) {
        const double wx = wxb[3 * n1 + iw];
        const double2 wdiff = make_double2(wx - w[k].x, -w[k].y);
        reg = reg && wx >= 0.0 && gpp_ssx_surely_regular(w[k], wdiff, cnorm2(wdiff));
      }
      if (reg) {
#pragma unroll
        for (int iw = 0; iw < 3; ++iw) ssx_grad_regular(wxb[3 * n1 + iw], wt2, w2I, s1, s2, gX[iw]);
      } else {
#pragma unroll
        for (int iw = 0; iw < 3; ++iw) ssx_grad(wxb[3 * n1 + iw], w[k], wt2, w2I, I[k], s1, s2, oI, oW, gX[iw]);
      }
#pragma unroll
      for (int iw = 0; iw < 3; ++iw) {
        const double2 tX = cmul(ma, gX[iw]);
        dsx[iw][0] += tX.x;
        dsx[iw][1] += tX.y;
      }
      const double2 tI = cmul(ma, cadd(cmul(wt2, s1), oI));
      const double2 tW = cmul(ma, cadd(cmul(cmul(w[k], I[k]), s2), oW));
      dIs[k][0] += tI.x;
      dIs[k][1] += tI.y;
      dWs[k][0] += tW.x;
      dWs[k][1] += tW.y;
    }
  }

  cdd ds[3] = {cdd_zero(), cdd_zero(), cdd_zero()};
  cdd Tp = cdd_zero();
#pragma unroll
  for (int k = 0; k < GPT; ++k) {
    if (!valid[k]) continue;
    const int g = ch * BT * GPT + k * BT + threadIdx.x;
    const double2 Gc = make_double2(G[k][0], G[k][1]);
    const double2 dfdI = cadd(cmul(wt_igp, Gc), make_double2(dIs[k][0], dIs[k][1]));
    const double2 Hs = make_double2(H[k][0][0] + H[k][1][0] + H[k][2][0], H[k][0][1] + H[k][1][1] + H[k][2][1]);
    const double2 sI = cmul(wt_igp, I[k]);
    const double2 dfdW = cadd(cmul(sI, Hs), make_double2(dWs[k][0], dWs[k][1]));
    if (dL_dI) dL_dI[(size_t)p * ncouls + g] = cconj(dfdI);
    if (dL_dwt) dL_dwt[(size_t)p * ncouls + g] = cconj(dfdW);
#pragma unroll
    for (int iw = 0; iw < 3; ++iw) {
      const double2 t = cmul(sI, make_double2(H[k][iw][0], H[k][iw][1]));
      ds[iw] = cdd_add(ds[iw], -t.x, -t.y);
    }
    const double2 t = cmul(I[k], Gc);
    Tp = cdd_add(Tp, t.x, t.y);
  }
  // the block's eight double-double totals (re and im of dSigma/domega[0..2] and of T):
  // one reduce-scatter over each wave, then the waves in wave order
  dd v[8];
#pragma unroll
  for (int iw = 0; iw < 3; ++iw) {
    ds[iw] = cdd_add(ds[iw], dsx[iw][0], dsx[iw][1]);
    v[2 * iw] = ds[iw].re;
    v[2 * iw + 1] = ds[iw].im;
  }
  v[6] = Tp.re;
  v[7] = Tp.im;
  int comp = 0;
  bool writer = true;
  dd_wave_sum_scatter<8>(v, threadIdx.x & 63, comp, writer);
  if (writer) red[threadIdx.x >> 6][comp] = v[0];
  __syncthreads();
  if (threadIdx.x < 8) {
    dd t = red[0][threadIdx.x];
#pragma unroll
    for (int w = 1; w < BT / 64; ++w) t = dd_add(t, red[w][threadIdx.x]);
    // stored as cdd partials (re.hi, re.lo, im.hi, im.lo): component 2k is the real part
    double4* dst = threadIdx.x < 6 ? &ds_part[(size_t)blockIdx.x * 3 + threadIdx.x / 2] : &T_part[blockIdx.x];
    reinterpret_cast<double2*>(dst)[threadIdx.x & 1] = make_double2(t.hi, t.lo);
  }
}

// dL/dwtilde(p, igp(p)) += conj(T(p)), T(p) = sum over the G-chunks of p (fixed order).
__global__ void k_gpp_bwd_T(const double4* __restrict__ T_part, int nchunk, const int* __restrict__ inv,
                            const int* __restrict__ indinv, int ngpown, int ncouls, double2* __restrict__ dL_dwt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ngpown) return;
  cdd t = cdd_zero();
  for (int c = 0; c < nchunk; ++c) t = cdd_add(t, cdd_load(&T_part[(size_t)p * nchunk + c]));
  const int igp = gpp_igp(inv, indinv, p, ncouls);
  double2 v = dL_dwt[(size_t)p * ncouls + igp];
  v.x += t.re.hi + t.re.lo;
  v.y -= t.im.hi + t.im.lo;
  dL_dwt[(size_t)p * ncouls + igp] = v;
}

}  // namespace

int gpp_backward_device(bgw_ctx* ctx, size_t nb_, size_t nvband_, size_t ngpown_, size_t ncouls_, const int* inv,
                        const int* indinv, const double* wx, const double* wtilde, const double* aqsm,
                        const double* aqsn, const double* I_eps, const double* vcoul, double occ, double* dsigma,
                        double* dwt, double* dI) {
  const int nb = (int)nb_, nvband = (int)nvband_, ngpown = (int)ngpown_, ncouls = (int)ncouls_;
  hipStream_t s = ctx->stream;
  const double2* wt = reinterpret_cast<const double2*>(wtilde);
  const double2* I = reinterpret_cast<const double2*>(I_eps);
  const double2* am = reinterpret_cast<const double2*>(aqsm);
  const double2* an = reinterpret_cast<const double2*>(aqsn);
  double2 *S, *C;
  double *wxb, *range;
  launch_check_indices(inv, ngpown_, indinv, ncouls_ + 1, ncouls, ncouls, ctx->d_flag, s);
  int rc = gpp_prep_tables(ctx, nb, ngpown, ncouls, inv, indinv, wx, 1, wt, am, an, vcoul, &S, &C, &wxb, &range);
  if (rc) return rc;
  const int nchunk = (ncouls + BT * GPT - 1) / (BT * GPT);
  const size_t blocks = (size_t)ngpown * nchunk;
  void *pds, *pT;
  if ((rc = pool_get(ctx, "ws_bwd_ds", blocks * 3 * 32, &pds))) return rc;
  if ((rc = pool_get(ctx, "ws_bwd_T", blocks * 32, &pT))) return rc;
  if (ctx->precision == BGW_FP64_MOMENTS) {
    void* prow;
    if ((rc = pool_get(ctx, "ws_bwd_rows", (size_t)ngpown * sizeof(BwdRow), &prow))) return rc;
    BwdRow* rows = static_cast<BwdRow*>(prow);
    {
      KernelTimer kt(ctx, "gpp_bwd_moments");
      hipLaunchKernelGGL(k_gpp_bwd_rowmom, dim3((unsigned)ngpown), dim3(BT), 0, s, wt, C, wxb, range, nb, ncouls,
                         rows);
    }
    KernelTimer kt(ctx, "gpp_bwd");
    hipLaunchKernelGGL(k_gpp_bwd<true>, dim3((unsigned)blocks), dim3(BT), 0, s, wt, I, C, wxb, range, inv, indinv, am,
                       an, vcoul, nb, nvband, ngpown, ncouls, occ, reinterpret_cast<double2*>(dwt),
                       reinterpret_cast<double2*>(dI), static_cast<double4*>(pds), static_cast<double4*>(pT), rows);
  } else {
    KernelTimer kt(ctx, "gpp_bwd");
    hipLaunchKernelGGL(k_gpp_bwd<false>, dim3((unsigned)blocks), dim3(BT), 0, s, wt, I, C, wxb, range, inv, indinv,
                       am, an, vcoul, nb, nvband, ngpown, ncouls, occ, reinterpret_cast<double2*>(dwt),
                       reinterpret_cast<double2*>(dI), static_cast<double4*>(pds), static_cast<double4*>(pT),
                       nullptr);
  }
  if ((rc = check_launch(ctx, "k_gpp_bwd"))) return rc;
  if (dwt) {
    hipLaunchKernelGGL(k_gpp_bwd_T, dim3((ngpown + 255) / 256), dim3(256), 0, s, static_cast<const double4*>(pT),
                       nchunk, inv, indinv, ngpown, ncouls, reinterpret_cast<double2*>(dwt));
  }
  if (dsigma) launch_reduce_cdd(static_cast<const double4*>(pds), (int)blocks, 3, reinterpret_cast<double2*>(dsigma), s);
  return check_launch(ctx, "k_gpp_bwd_T");
}

}  // namespace bgw
