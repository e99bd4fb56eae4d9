/*
 * bgw_oracle.c -- CPU restatement of the reference GPP/FF kernels (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
 * library, as the checker.  The product library never calls it.
 *
 * Per-term arithmetic follows the reference expressions (cited per function);
 * sums are compensated: TwoSum lanes over the G-vector axis and double-double
 * accumulation over bands and G'-columns, combined in a fixed order.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   oracle_gpp_forward, achtemp     the reference's own noflagOCC_solver built from
 *                                   its sources (oracle/Makefile -> _ref/), whole and
 *                                   per band; its printed goldens (GPP/gpp.cpp:119-125);
 *                                   closed forms of the constant inputs; exact
 *                                   rational fixtures; committed reference vectors
 *     acht_n1_loc                   per band against the reference build; band-sum
 *     achstemp, asxtemp             no reference counterpart (parity unpinned against
 *                                   the reference): exact rational fixtures reaching
 *                                   every SX branch, closed forms, a numpy
 *                                   restatement of the regular SX branch
 *   oracle_gpp_grad_entries,        no reference counterpart: central finite
 *   oracle_gpp_dsigma_domega        differences of oracle_gpp_forward, including the
 *                                   near-pole and cut-off SX branches
 *   oracle_compute_fact             every branch and quirk against exact rationals
 *   oracle_achsDtemp, _asxDtemp,    the reference's FF kernels built from its sources;
 *   _achDtemp_cor                   its printed goldens (FF/ff.cpp:72-93); closed forms;
 *                                   exact rational fixtures; committed reference vectors
 */
#include "bgw_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* GPP constants, GPP/gpp.cpp:165-169 (to1, limittwo) and the BerkeleyGW
 * flagOCC branch constants that accompany them (limitone = 1/(4 to1), sexcut = 4). */
#define TO1 1e-6
#define LIMITTWO 0.25
#define LIMITONE (1.0 / (TO1 * 4.0))
#define SEXCUT 4.0
#define LANES 8

/* ---------------------------------------------------------------- dd arithmetic */
typedef struct { double hi, lo; } dd_t;

static inline void two_sum(double a, double b, double* s, double* e) {
  double x = a + b;
  double bb = x - a;
  *e = (a - (x - bb)) + (b - bb);
  *s = x;
}
static inline void dd_add(dd_t* a, double b) {
  double s, e;
  two_sum(a->hi, b, &s, &e);
  e += a->lo;
  a->hi = s + e;
  a->lo = e - (a->hi - s);
}
static inline void dd_add_dd(dd_t* a, dd_t b) {
  double s, e;
  two_sum(a->hi, b.hi, &s, &e);
  e += a->lo + b.lo;
  a->hi = s + e;
  a->lo = e - (a->hi - s);
}
/* a += x*y exactly up to dd precision */
static inline void dd_add_prod(dd_t* a, double x, double y) {
  double p = x * y;
  double e = fma(x, y, -p);
  dd_add(a, p);
  dd_add(a, e);
}
static inline double dd_val(dd_t a) { return a.hi + a.lo; }

/* complex double-double accumulator */
typedef struct { dd_t re, im; } cdd_t;

static inline void cdd_add_cprod(cdd_t* a, double xr, double xi, double yr, double yi) {
  dd_add_prod(&a->re, xr, yr);
  dd_add_prod(&a->re, -xi, yi);
  dd_add_prod(&a->im, xr, yi);
  dd_add_prod(&a->im, xi, yr);
}
/* a += x * (y_hi + y_lo) with y given as complex dd */
static inline void cdd_add_cprod_dd(cdd_t* a, double xr, double xi, cdd_t y) {
  cdd_add_cprod(a, xr, xi, y.re.hi, y.im.hi);
  cdd_add_cprod(a, xr, xi, y.re.lo, y.im.lo);
}
static inline void cdd_add(cdd_t* a, cdd_t b) {
  dd_add_dd(&a->re, b.re);
  dd_add_dd(&a->im, b.im);
}

/* TwoSum lanes over an axis: sum of LANES (hi,lo) pairs -> cdd */
typedef struct { double h[LANES], l[LANES]; } lanes_t;

static inline void lanes_zero(lanes_t* a) { memset(a, 0, sizeof(*a)); }
static inline dd_t lanes_fold(const lanes_t* a) {
  dd_t r = {0.0, 0.0};
  for (int l = 0; l < LANES; ++l) {
    dd_add(&r, a->h[l]);
    dd_add(&r, a->l[l]);
  }
  return r;
}

static int nthreads_of(int nthreads) {
#ifdef _OPENMP
  return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

static inline int igp_of(const int* inv_igp_index, const int* indinv, size_t p) {
  return indinv[inv_igp_index[p]]; /* GPP/gpp.cpp:362-363, FF/ff.cpp:194-195 */
}

/* ========================================================================== GPP */

/* sch_store1, GPP/gpp.cpp:369-371: conj(aqsm)*aqsn*0.5*vcoul*wtilde(p,igp),
 * evaluated left to right as the reference writes it. */
static inline double complex sch_store1(double complex aqsm, double complex aqsn, double vc,
                                        double complex wt_igp) {
  double complex t = conj(aqsm) * aqsn;
  t = t * 0.5;
  t = t * vc;
  return t * wt_igp;
}

/* ssx of the occupied-band (flagOCC) branch of BerkeleyGW's GPP kernel, before
 * the |ssx| > sexcut |I| cut-off for wx < 0. */
static inline double complex ssx_raw(double wx, double complex wt, double complex Ie) {
  double complex wt2 = wt * wt;
  double complex Omega2 = wt2 * Ie;
  double complex wdiff = wx - wt;
  double wdiffr = creal(wdiff * conj(wdiff));
  double rden = 1.0 / wdiffr;
  double complex delw = wt * conj(wdiff) * rden;
  double delwr = creal(delw * conj(delw));
  double complex ssx;
  if (wdiffr > LIMITTWO && delwr < LIMITONE) {
    double complex cden = wx * wx - wt2;
    double rd = 1.0 / creal(cden * conj(cden));
    ssx = Omega2 * conj(cden) * rd;
  } else if (delwr > TO1) {
    double complex cden = 4.0 * wt2 * (delw + 0.5);
    double rd = 1.0 / creal(cden * conj(cden));
    ssx = -Omega2 * conj(cden) * rd * delw;
  } else {
    ssx = 0.0;
  }
  return ssx;
}

static inline int ssx_cut(double wx, double complex ssx, double complex Ie) {
  return cabs(ssx) > SEXCUT * cabs(Ie) && wx < 0.0;
}

static inline double complex ssx_term(double wx, double complex wt, double complex Ie) {
  const double complex ssx = ssx_raw(wx, wt, Ie);
  return ssx_cut(wx, ssx, Ie) ? 0.0 : ssx;
}

/* Holomorphic derivatives of ssx_term with respect to Ie, wtilde and wx. */
static inline void ssx_grad(double wx, double complex wt, double complex Ie, double complex* dI,
                            double complex* dW, double complex* dX) {
  double complex wt2 = wt * wt;
  double complex wdiff = wx - wt;
  double wdiffr = creal(wdiff * conj(wdiff));
  double complex delw = wt / wdiff;
  double delwr = creal(delw * conj(delw));
  *dI = *dW = *dX = 0.0;
  if (ssx_cut(wx, ssx_raw(wx, wt, Ie), Ie)) return; /* cut terms are constant zero */
  if (wdiffr > LIMITTWO && delwr < LIMITONE) {
    double complex cden = wx * wx - wt2;
    *dI = wt2 / cden;
    *dW = 2.0 * wt * Ie * wx * wx / (cden * cden);
    *dX = -2.0 * wx * wt2 * Ie / (cden * cden);
  } else if (delwr > TO1) {
    double complex q = delw / (delw + 0.5);
    double complex dq = 0.5 / ((delw + 0.5) * (delw + 0.5));
    *dI = -q / 4.0;
    *dW = -Ie / 4.0 * dq * (wx / (wdiff * wdiff));
    *dX = -Ie / 4.0 * dq * (-wt / (wdiff * wdiff));
  }
}

void oracle_gpp_forward(size_t nb, size_t nvband, size_t ngpown, size_t ncouls,
                        const int* inv_igp_index, const int* indinv, const double* wx,
                        const double* wtilde, const double* aqsm, const double* aqsn,
                        const double* I_eps, const double* vcoul, double occ,
                        double* achtemp, double* acht_n1_loc, double* achstemp,
                        double* asxtemp, int nthreads) {
  const int nt = nthreads_of(nthreads);
  const size_t nacc = nb * 3;
  /* per-thread partials: [nt][nb*3] for acht_n1_loc, [nt] achstemp, [nt][3] asxtemp */
  cdd_t* part_ach = (cdd_t*)calloc((size_t)nt * nacc, sizeof(cdd_t));
  cdd_t* part_achs = (cdd_t*)calloc((size_t)nt, sizeof(cdd_t));
  cdd_t* part_asx = (cdd_t*)calloc((size_t)nt * 3, sizeof(cdd_t));
  const int want_achs = achstemp != NULL;
  const int want_asx = asxtemp != NULL && nvband > 0;

#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
    double* wr = (double*)malloc(ncouls * sizeof(double));
    double* wi = (double*)malloc(ncouls * sizeof(double));
    double* ir = (double*)malloc(ncouls * sizeof(double));
    double* ii = (double*)malloc(ncouls * sizeof(double));
    double* irm = (double*)malloc(ncouls * sizeof(double)); /* masked I for achstemp */
    double* iim = (double*)malloc(ncouls * sizeof(double));
    cdd_t* my_ach = part_ach + (size_t)tid * nacc;

#pragma omp for schedule(static)
    for (long lp = 0; lp < (long)ngpown; ++lp) {
      const size_t p = (size_t)lp;
      const int igp = igp_of(inv_igp_index, indinv, p);
      const double* wrow = wtilde + 2 * p * ncouls;
      const double* irow = I_eps + 2 * p * ncouls;
      for (size_t g = 0; g < ncouls; ++g) {
        wr[g] = wrow[2 * g];
        wi[g] = wrow[2 * g + 1];
        ir[g] = irow[2 * g];
        ii[g] = irow[2 * g + 1];
        /* upstream static-COH cut: |schs| > to1 with schs = -I_eps */
        int keep = (ir[g] * ir[g] + ii[g] * ii[g]) > TO1 * TO1;
        irm[g] = keep ? ir[g] : 0.0;
        iim[g] = keep ? ii[g] : 0.0;
      }
      const double complex wt_igp = wrow[2 * igp] + I * wrow[2 * igp + 1];
      const double vc = vcoul[igp];

      for (size_t n1 = 0; n1 < nb; ++n1) {
        const double complex am = aqsm[2 * (n1 * ncouls + igp)] + I * aqsm[2 * (n1 * ncouls + igp) + 1];
        const double complex an = aqsn[2 * (n1 * ncouls + igp)] + I * aqsn[2 * (n1 * ncouls + igp) + 1];
        const double complex s = sch_store1(am, an, vc, wt_igp);
        const double sr = creal(s), si = cimag(s);
        /* ---- noflagOCC term, GPP/gpp.cpp:373-384: sum_g delw*I_eps for each iw */
        for (int iw = 0; iw < 3; ++iw) {
          const double x = wx[n1 * 3 + iw];
          lanes_t are, aim;
          lanes_zero(&are);
          lanes_zero(&aim);
          size_t g0 = 0;
          for (; g0 + LANES <= ncouls; g0 += LANES) {
            for (int l = 0; l < LANES; ++l) {
              const size_t g = g0 + l;
              const double u = x - wr[g];               /* wdiff = (u, -wi) */
              const double d = u * u + wi[g] * wi[g];   /* (wdiff*conj(wdiff)).real() */
              const double rden = 1.0 / d;
              const double dr = u * rden, di = wi[g] * rden; /* delw = conj(wdiff)*rden */
              const double xr = dr * ir[g] - di * ii[g];
              const double xi = dr * ii[g] + di * ir[g];
              double s1, e1, s2, e2;
              two_sum(are.h[l], xr, &s1, &e1);
              are.h[l] = s1;
              are.l[l] += e1;
              two_sum(aim.h[l], xi, &s2, &e2);
              aim.h[l] = s2;
              aim.l[l] += e2;
            }
          }
          dd_t Ar = lanes_fold(&are), Ai = lanes_fold(&aim);
          for (; g0 < ncouls; ++g0) {
            const double u = x - wr[g0];
            const double d = u * u + wi[g0] * wi[g0];
            const double rden = 1.0 / d;
            const double dr = u * rden, di = wi[g0] * rden;
            dd_add(&Ar, dr * ir[g0] - di * ii[g0]);
            dd_add(&Ai, dr * ii[g0] + di * ir[g0]);
          }
          cdd_t A = {Ar, Ai};
          cdd_add_cprod_dd(&my_ach[n1 * 3 + iw], sr, si, A);
        }
        /* ---- static COH: -1/2 vcoul conj(aqsm(n1,igp)) sum_g aqsn(n1,g) I_eps(p,g) */
        if (want_achs) {
          const double* nrow = aqsn + 2 * n1 * ncouls;
          lanes_t mre, mim;
          lanes_zero(&mre);
          lanes_zero(&mim);
          size_t g0 = 0;
          for (; g0 + LANES <= ncouls; g0 += LANES) {
            for (int l = 0; l < LANES; ++l) {
              const size_t g = g0 + l;
              const double nr = nrow[2 * g], ni = nrow[2 * g + 1];
              const double xr = nr * irm[g] - ni * iim[g];
              const double xi = nr * iim[g] + ni * irm[g];
              double s1, e1, s2, e2;
              two_sum(mre.h[l], xr, &s1, &e1);
              mre.h[l] = s1;
              mre.l[l] += e1;
              two_sum(mim.h[l], xi, &s2, &e2);
              mim.h[l] = s2;
              mim.l[l] += e2;
            }
          }
          dd_t Mr = lanes_fold(&mre), Mi = lanes_fold(&mim);
          for (; g0 < ncouls; ++g0) {
            const double nr = nrow[2 * g0], ni = nrow[2 * g0 + 1];
            dd_add(&Mr, nr * irm[g0] - ni * iim[g0]);
            dd_add(&Mi, nr * iim[g0] + ni * irm[g0]);
          }
          const double complex c = -0.5 * vc * conj(am);
          cdd_t M = {Mr, Mi};
          cdd_add_cprod_dd(&part_achs[tid], creal(c), cimag(c), M);
        }
        /* ---- screened exchange over occupied bands */
        if (want_asx && n1 < nvband) {
          const double* nrow = aqsn + 2 * n1 * ncouls;
          const double complex m = conj(am) * (occ * vc);
          for (int iw = 0; iw < 3; ++iw) {
            const double x = wx[n1 * 3 + iw];
            cdd_t acc;
            memset(&acc, 0, sizeof(acc));
            for (size_t g = 0; g < ncouls; ++g) {
              const double complex ss = ssx_term(x, wr[g] + I * wi[g], ir[g] + I * ii[g]);
              const double complex a = nrow[2 * g] + I * nrow[2 * g + 1];
              const double complex t = a * ss;
              dd_add(&acc.re, creal(t));
              dd_add(&acc.im, cimag(t));
            }
            cdd_add_cprod_dd(&part_asx[tid * 3 + iw], creal(m), cimag(m), acc);
          }
        }
      }
    }
    free(wr);
    free(wi);
    free(ir);
    free(ii);
    free(irm);
    free(iim);
  }

  /* fixed-order combine over threads, then over bands */
  cdd_t tot[3];
  memset(tot, 0, sizeof(tot));
  for (size_t k = 0; k < nacc; ++k) {
    cdd_t a;
    memset(&a, 0, sizeof(a));
    for (int t = 0; t < nt; ++t) cdd_add(&a, part_ach[(size_t)t * nacc + k]);
    if (acht_n1_loc) {
      acht_n1_loc[2 * k] = dd_val(a.re);
      acht_n1_loc[2 * k + 1] = dd_val(a.im);
    }
    cdd_add(&tot[k % 3], a);
  }
  for (int iw = 0; iw < 3; ++iw) {
    achtemp[2 * iw] = dd_val(tot[iw].re);
    achtemp[2 * iw + 1] = dd_val(tot[iw].im);
  }
  if (want_achs) {
    cdd_t a;
    memset(&a, 0, sizeof(a));
    for (int t = 0; t < nt; ++t) cdd_add(&a, part_achs[t]);
    achstemp[0] = dd_val(a.re);
    achstemp[1] = dd_val(a.im);
  }
  if (asxtemp) {
    for (int iw = 0; iw < 3; ++iw) {
      cdd_t a;
      memset(&a, 0, sizeof(a));
      if (want_asx)
        for (int t = 0; t < nt; ++t) cdd_add(&a, part_asx[t * 3 + iw]);
      asxtemp[2 * iw] = dd_val(a.re);
      asxtemp[2 * iw + 1] = dd_val(a.im);
    }
  }
  free(part_ach);
  free(part_achs);
  free(part_asx);
}

/* T(p) = sum_{n1,iw,g} c(n1,p) I(p,g)/(wx(n1,iw)-wtilde(p,g)), c = s / wtilde(p,igp). */
static double complex gpp_T_of_p(size_t nb, size_t ncouls, size_t p, int igp, const double* wx,
                                 const double* wtilde, const double* aqsm, const double* aqsn,
                                 const double* I_eps, const double* vcoul) {
  cdd_t acc;
  memset(&acc, 0, sizeof(acc));
  const double* wrow = wtilde + 2 * p * ncouls;
  const double* irow = I_eps + 2 * p * ncouls;
  for (size_t n1 = 0; n1 < nb; ++n1) {
    const double complex am = aqsm[2 * (n1 * ncouls + igp)] + I * aqsm[2 * (n1 * ncouls + igp) + 1];
    const double complex an = aqsn[2 * (n1 * ncouls + igp)] + I * aqsn[2 * (n1 * ncouls + igp) + 1];
    double complex c = conj(am) * an;
    c = c * 0.5;
    c = c * vcoul[igp];
    cdd_t A;
    memset(&A, 0, sizeof(A));
    for (int iw = 0; iw < 3; ++iw) {
      const double x = wx[n1 * 3 + iw];
      for (size_t g = 0; g < ncouls; ++g) {
        const double complex t = (irow[2 * g] + I * irow[2 * g + 1]) / (x - (wrow[2 * g] + I * wrow[2 * g + 1]));
        dd_add(&A.re, creal(t));
        dd_add(&A.im, cimag(t));
      }
    }
    cdd_add_cprod_dd(&acc, creal(c), cimag(c), A);
  }
  return dd_val(acc.re) + I * dd_val(acc.im);
}

void oracle_gpp_grad_entries(size_t nb, size_t nvband, size_t ngpown, size_t ncouls,
                             const int* inv_igp_index, const int* indinv, const double* wx,
                             const double* wtilde, const double* aqsm, const double* aqsn,
                             const double* I_eps, const double* vcoul, double occ,
                             const long* entries, size_t n_entries,
                             double* dL_dwtilde, double* dL_dI_eps, int nthreads) {
  (void)ngpown;
  const int nt = nthreads_of(nthreads);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (long k = 0; k < (long)n_entries; ++k) {
    const long e = entries[k];
    const size_t p = (size_t)e / ncouls, g = (size_t)e % ncouls;
    const int igp = igp_of(inv_igp_index, indinv, p);
    const double complex w = wtilde[2 * e] + I * wtilde[2 * e + 1];
    const double complex Iv = I_eps[2 * e] + I * I_eps[2 * e + 1];
    const double complex wt_igp = wtilde[2 * (p * ncouls + igp)] + I * wtilde[2 * (p * ncouls + igp) + 1];
    cdd_t dI, dW;
    memset(&dI, 0, sizeof(dI));
    memset(&dW, 0, sizeof(dW));
    for (size_t n1 = 0; n1 < nb; ++n1) {
      const double complex am = aqsm[2 * (n1 * ncouls + igp)] + I * aqsm[2 * (n1 * ncouls + igp) + 1];
      const double complex an = aqsn[2 * (n1 * ncouls + igp)] + I * aqsn[2 * (n1 * ncouls + igp) + 1];
      const double complex s = sch_store1(am, an, vcoul[igp], wt_igp);
      for (int iw = 0; iw < 3; ++iw) {
        const double complex inv = 1.0 / (wx[n1 * 3 + iw] - w);
        const double complex a = s * inv;
        const double complex b = s * Iv * inv * inv;
        dd_add(&dI.re, creal(a));
        dd_add(&dI.im, cimag(a));
        dd_add(&dW.re, creal(b));
        dd_add(&dW.im, cimag(b));
      }
      if (n1 < nvband) {
        const double complex m = conj(am) * (occ * vcoul[igp]);
        const double complex an_g = aqsn[2 * (n1 * ncouls + g)] + I * aqsn[2 * (n1 * ncouls + g) + 1];
        for (int iw = 0; iw < 3; ++iw) {
          double complex gI, gW, gX;
          ssx_grad(wx[n1 * 3 + iw], w, Iv, &gI, &gW, &gX);
          const double complex a = m * an_g * gI;
          const double complex b = m * an_g * gW;
          dd_add(&dI.re, creal(a));
          dd_add(&dI.im, cimag(a));
          dd_add(&dW.re, creal(b));
          dd_add(&dW.im, cimag(b));
        }
      }
    }
    if ((int)g == igp) {
      const double complex T = gpp_T_of_p(nb, ncouls, p, igp, wx, wtilde, aqsm, aqsn, I_eps, vcoul);
      dd_add(&dW.re, creal(T));
      dd_add(&dW.im, cimag(T));
    }
    /* dL/dz = dL/dRe z + i dL/dIm z = conj(df/dz) for L = Re f, f holomorphic */
    dL_dI_eps[2 * k] = dd_val(dI.re);
    dL_dI_eps[2 * k + 1] = -dd_val(dI.im);
    dL_dwtilde[2 * k] = dd_val(dW.re);
    dL_dwtilde[2 * k + 1] = -dd_val(dW.im);
  }
}

void oracle_gpp_dsigma_domega(size_t nb, size_t nvband, size_t ngpown, size_t ncouls,
                              const int* inv_igp_index, const int* indinv, const double* wx,
                              const double* wtilde, const double* aqsm, const double* aqsn,
                              const double* I_eps, const double* vcoul, double occ,
                              double* dsigma_domega, int nthreads) {
  const int nt = nthreads_of(nthreads);
  cdd_t* part = (cdd_t*)calloc((size_t)nt * 3, sizeof(cdd_t));
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
#pragma omp for schedule(static)
    for (long lp = 0; lp < (long)ngpown; ++lp) {
      const size_t p = (size_t)lp;
      const int igp = igp_of(inv_igp_index, indinv, p);
      const double* wrow = wtilde + 2 * p * ncouls;
      const double* irow = I_eps + 2 * p * ncouls;
      const double complex wt_igp = wrow[2 * igp] + I * wrow[2 * igp + 1];
      for (size_t n1 = 0; n1 < nb; ++n1) {
        const double complex am = aqsm[2 * (n1 * ncouls + igp)] + I * aqsm[2 * (n1 * ncouls + igp) + 1];
        const double complex an = aqsn[2 * (n1 * ncouls + igp)] + I * aqsn[2 * (n1 * ncouls + igp) + 1];
        const double complex s = sch_store1(am, an, vcoul[igp], wt_igp);
        const double complex m = conj(am) * (occ * vcoul[igp]);
        for (int iw = 0; iw < 3; ++iw) {
          const double x = wx[n1 * 3 + iw];
          cdd_t A;
          memset(&A, 0, sizeof(A));
          for (size_t g = 0; g < ncouls; ++g) {
            const double complex w = wrow[2 * g] + I * wrow[2 * g + 1];
            const double complex Iv = irow[2 * g] + I * irow[2 * g + 1];
            const double complex inv = 1.0 / (x - w);
            const double complex t = -Iv * inv * inv;
            dd_add(&A.re, creal(t));
            dd_add(&A.im, cimag(t));
          }
          cdd_add_cprod_dd(&part[tid * 3 + iw], creal(s), cimag(s), A);
          if (n1 < nvband) {
            cdd_t B;
            memset(&B, 0, sizeof(B));
            for (size_t g = 0; g < ncouls; ++g) {
              const double complex w = wrow[2 * g] + I * wrow[2 * g + 1];
              const double complex Iv = irow[2 * g] + I * irow[2 * g + 1];
              double complex gI, gW, gX;
              ssx_grad(x, w, Iv, &gI, &gW, &gX);
              const double complex t = (aqsn[2 * (n1 * ncouls + g)] + I * aqsn[2 * (n1 * ncouls + g) + 1]) * gX;
              dd_add(&B.re, creal(t));
              dd_add(&B.im, cimag(t));
            }
            cdd_add_cprod_dd(&part[tid * 3 + iw], creal(m), cimag(m), B);
          }
        }
      }
    }
  }
  for (int iw = 0; iw < 3; ++iw) {
    cdd_t a;
    memset(&a, 0, sizeof(a));
    for (int t = 0; t < nt; ++t) cdd_add(&a, part[t * 3 + iw]);
    dsigma_domega[2 * iw] = dd_val(a.re);
    dsigma_domega[2 * iw + 1] = dd_val(a.im);
  }
  free(part);
}

/* =========================================================================== FF */

int oracle_compute_fact(double wx, int nFreq, const double* g, double* fact1, double* fact2,
                        int* ifreq, int loop, int flag_occ) {
  /* FF/ff.cpp:105-127: loop 1 brackets wx (or -wx) with strict '<' on both sides,
   * then remaps ifreq==0 to nFreq-2 (so wx in the first interval, or on a grid
   * point, uses the top interval).  A stale incoming ifreq outside [0, nFreq-2]
   * that no bracket overwrites would make the reference read the grid out of
   * bounds: reported as -1, nothing read or written. */
  if (loop == 1 && wx > 0.0) {
    for (int k = 0; k < nFreq - 1; ++k)
      if (wx > g[k] && wx < g[k + 1]) *ifreq = k;
    if (*ifreq == 0) *ifreq = nFreq - 2;
    if (*ifreq < 0 || *ifreq > nFreq - 2) return -1;
    *fact1 = (g[*ifreq + 1] - wx) / (g[*ifreq + 1] - g[*ifreq]);
    *fact2 = (wx - g[*ifreq]) / (g[*ifreq + 1] - g[*ifreq]);
  } else if (loop == 1) {
    for (int k = 0; k < nFreq - 1; ++k)
      if (-wx > g[k] && -wx < g[k + 1]) *ifreq = k;
    if (*ifreq == 0) *ifreq = nFreq - 2;
    if (*ifreq < 0 || *ifreq > nFreq - 2) return -1;
    *fact1 = (g[*ifreq + 1] + wx) / (g[*ifreq + 1] - g[*ifreq]);
    *fact2 = (-g[*ifreq] - wx) / (g[*ifreq + 1] - g[*ifreq]);
  }
  /* FF/ff.cpp:128-151: loop 2 with wx>0 weights by -1/2 and never remaps (its
   * test is ifreq==-1); loop 2 occupied mirrors loop 1 on -wx; loop 2 with an
   * unoccupied band and wx<=0 sets nothing. */
  if (loop == 2 && wx > 0.0) {
    for (int k = 0; k < nFreq - 1; ++k)
      if (wx > g[k] && wx < g[k + 1]) *ifreq = k;
    if (*ifreq == -1) *ifreq = nFreq - 2;
    if (*ifreq < 0 || *ifreq > nFreq - 2) return -1;
    *fact1 = -0.5 * (g[*ifreq + 1] - wx) / (g[*ifreq + 1] - g[*ifreq]);
    *fact2 = -0.5 * (wx - g[*ifreq]) / (g[*ifreq + 1] - g[*ifreq]);
  } else if (loop == 2 && flag_occ) {
    wx = -wx;
    *ifreq = 0;
    for (int k = 0; k < nFreq - 1; ++k)
      if (wx > g[k] && wx < g[k + 1]) *ifreq = k;
    if (*ifreq == 0) *ifreq = nFreq - 2;
    *fact1 = (g[*ifreq + 1] - wx) / (g[*ifreq + 1] - g[*ifreq]);
    *fact2 = (wx - g[*ifreq]) / (g[*ifreq + 1] - g[*ifreq]);
  }
  return 0;
}

/* sum_g aqsn(n1,g) * (f1*X(f,p,g) + f2*X(f+1,p,g) [- same of Y]) compensated;
 * X,Y are [nFreq][ngpown][ncouls] complex; Y may be NULL. */
static cdd_t ff_row_dot(const double* an_row, const double* X, const double* Y, int f, double f1,
                        double f2, size_t p, size_t ngpown, size_t ncouls) {
  const double* x0 = X + 2 * (((size_t)f * ngpown + p) * ncouls);
  const double* x1 = X + 2 * (((size_t)(f + 1) * ngpown + p) * ncouls);
  const double* y0 = Y ? Y + 2 * (((size_t)f * ngpown + p) * ncouls) : NULL;
  const double* y1 = Y ? Y + 2 * (((size_t)(f + 1) * ngpown + p) * ncouls) : NULL;
  lanes_t ar, ai;
  lanes_zero(&ar);
  lanes_zero(&ai);
  size_t g0 = 0;
  for (; g0 + LANES <= ncouls; g0 += LANES) {
    for (int l = 0; l < LANES; ++l) {
      const size_t g = g0 + l;
      double er, ei;
      if (Y) {
        er = (x0[2 * g] - y0[2 * g]) * f1 + (x1[2 * g] - y1[2 * g]) * f2;
        ei = (x0[2 * g + 1] - y0[2 * g + 1]) * f1 + (x1[2 * g + 1] - y1[2 * g + 1]) * f2;
      } else {
        er = x0[2 * g] * f1 + x1[2 * g] * f2;
        ei = x0[2 * g + 1] * f1 + x1[2 * g + 1] * f2;
      }
      const double nr = an_row[2 * g], ni = an_row[2 * g + 1];
      const double tr = nr * er - ni * ei, ti = nr * ei + ni * er;
      double s1, e1, s2, e2;
      two_sum(ar.h[l], tr, &s1, &e1);
      ar.h[l] = s1;
      ar.l[l] += e1;
      two_sum(ai.h[l], ti, &s2, &e2);
      ai.h[l] = s2;
      ai.l[l] += e2;
    }
  }
  cdd_t r = {lanes_fold(&ar), lanes_fold(&ai)};
  for (; g0 < ncouls; ++g0) {
    double er, ei;
    if (Y) {
      er = (x0[2 * g0] - y0[2 * g0]) * f1 + (x1[2 * g0] - y1[2 * g0]) * f2;
      ei = (x0[2 * g0 + 1] - y0[2 * g0 + 1]) * f1 + (x1[2 * g0 + 1] - y1[2 * g0 + 1]) * f2;
    } else {
      er = x0[2 * g0] * f1 + x1[2 * g0] * f2;
      ei = x0[2 * g0 + 1] * f1 + x1[2 * g0 + 1] * f2;
    }
    const double nr = an_row[2 * g0], ni = an_row[2 * g0 + 1];
    dd_add(&r.re, nr * er - ni * ei);
    dd_add(&r.im, nr * ei + ni * er);
  }
  return r;
}

void oracle_achsDtemp(int nb, int ngpown, int ncouls, const int* inv_igp_index, const int* indinv,
                      const double* aqsn, const double* aqsm, const double* I_epsR,
                      const double* vcoul, double* achsDtemp, int nthreads) {
  /* FF/ff.cpp:191-208: achsDtemp = sum_{n1,p} 0.5 vcoul(igp) * (-sum_g aqsn(n1,g)
   * conj(aqsm(n1,igp)) I_epsR(1,p,g)); computed as sum_{n1,p} c(n1,p) M(n1,p) with
   * M = sum_g aqsn(n1,g) I_epsR(1,p,g) and c = -0.5 vcoul conj(aqsm). */
  const int nt = nthreads_of(nthreads);
  cdd_t* part = (cdd_t*)calloc((size_t)nt, sizeof(cdd_t));
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
#pragma omp for schedule(static)
    for (long lp = 0; lp < (long)ngpown; ++lp) {
      const int igp = igp_of(inv_igp_index, indinv, (size_t)lp);
      for (int n1 = 0; n1 < nb; ++n1) {
        /* slice 1 only: f=1 weight 1, f+1 weight 0 -> pass f=0 with (0,1) */
        cdd_t M = ff_row_dot(aqsn + 2 * (size_t)n1 * ncouls, I_epsR, NULL, 0, 0.0, 1.0,
                             (size_t)lp, (size_t)ngpown, (size_t)ncouls);
        const double complex am = aqsm[2 * ((size_t)n1 * ncouls + igp)] + I * aqsm[2 * ((size_t)n1 * ncouls + igp) + 1];
        const double complex c = -0.5 * vcoul[igp] * conj(am);
        cdd_add_cprod_dd(&part[tid], creal(c), cimag(c), M);
      }
    }
  }
  cdd_t a;
  memset(&a, 0, sizeof(a));
  for (int t = 0; t < nt; ++t) cdd_add(&a, part[t]);
  achsDtemp[0] = dd_val(a.re);
  achsDtemp[1] = dd_val(a.im);
  free(part);
}

void oracle_asxDtemp(int nvband, int nfreqeval, int ncouls, int ngpown, int nFreq,
                     double freqevalmin, double freqevalstep, double occ, const double* ekq,
                     const double* dFreqGrid, const int* inv_igp_index, const int* indinv,
                     const double* aqsm, const double* aqsn, const double* vcoul,
                     const double* I_epsR, const double* I_epsA, double* asxDtemp, int nthreads) {
  /* FF/ff.cpp:242-281.  wx has no iw dependence (:245), so every asxDtemp[iw] is
   * the same sum; it is computed once and replicated. */
  const int nt = nthreads_of(nthreads);
  cdd_t* part = (cdd_t*)calloc((size_t)nt, sizeof(cdd_t));
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
#pragma omp for schedule(static)
    for (long lp = 0; lp < (long)ngpown; ++lp) {
      const int igp = igp_of(inv_igp_index, indinv, (size_t)lp);
      for (int n1 = 0; n1 < nvband; ++n1) {
        const double wx = freqevalmin - ekq[n1] + freqevalstep;
        double f1 = 0.0, f2 = 0.0;
        int ifreq = 0;
        oracle_compute_fact(wx, nFreq, dFreqGrid, &f1, &f2, &ifreq, 1, 0);
        const double* X = wx > 0 ? I_epsR : I_epsA; /* :254-261 */
        cdd_t M = ff_row_dot(aqsn + 2 * (size_t)n1 * ncouls, X, NULL, ifreq, f1, f2, (size_t)lp,
                             (size_t)ngpown, (size_t)ncouls);
        const double complex am = aqsm[2 * ((size_t)n1 * ncouls + igp)] + I * aqsm[2 * ((size_t)n1 * ncouls + igp) + 1];
        const double complex c = conj(am) * vcoul[igp] * occ;
        cdd_add_cprod_dd(&part[tid], creal(c), cimag(c), M);
      }
    }
  }
  cdd_t a;
  memset(&a, 0, sizeof(a));
  for (int t = 0; t < nt; ++t) cdd_add(&a, part[t]);
  for (int iw = 0; iw < nfreqeval; ++iw) {
    asxDtemp[2 * iw] = dd_val(a.re);
    asxDtemp[2 * iw + 1] = dd_val(a.im);
  }
  free(part);
}

void oracle_achDtemp_cor(int nb, int nvband, int nfreqeval, int ncouls, int ngpown, int nFreq,
                         double freqevalmin, double freqevalstep, const double* ekq,
                         const double* dFreqGrid, const int* inv_igp_index, const int* indinv,
                         const double* aqsm, const double* aqsn, const double* vcoul,
                         const double* I_epsR, const double* I_epsA, double* achDtemp_cor,
                         int nthreads) {
  /* FF/ff.cpp:332-372 with schDttt_corKernel1 (:381-416, wx>0 and unoccupied) and
   * schDttt_corKernel2 (:418-451, wx<=0 and occupied, extra factor -1/2).  wx has
   * no iw dependence (:337): one sum, replicated over iw. */
  const int nt = nthreads_of(nthreads);
  int* kind = (int*)calloc((size_t)nb, sizeof(int));
  int* fr = (int*)calloc((size_t)nb, sizeof(int));
  double* w1 = (double*)calloc((size_t)nb, sizeof(double));
  double* w2 = (double*)calloc((size_t)nb, sizeof(double));
  for (int n1 = 0; n1 < nb; ++n1) {
    const int flag_occ = n1 < nvband;
    const double wx = freqevalmin - ekq[n1] + freqevalstep;
    double f1 = 0.0, f2 = 0.0;
    int ifreq = 0;
    oracle_compute_fact(wx, nFreq, dFreqGrid, &f1, &f2, &ifreq, 2, flag_occ);
    kind[n1] = (wx > 0) ? (!flag_occ ? 1 : 0) : (flag_occ ? 2 : 0);
    fr[n1] = ifreq;
    w1[n1] = f1;
    w2[n1] = f2;
  }
  cdd_t* part = (cdd_t*)calloc((size_t)nt, sizeof(cdd_t));
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
#pragma omp for schedule(static)
    for (long lp = 0; lp < (long)ngpown; ++lp) {
      const int igp = igp_of(inv_igp_index, indinv, (size_t)lp);
      for (int n1 = 0; n1 < nb; ++n1) {
        if (!kind[n1]) continue;
        cdd_t M = ff_row_dot(aqsn + 2 * (size_t)n1 * ncouls, I_epsR, I_epsA, fr[n1], w1[n1],
                             w2[n1], (size_t)lp, (size_t)ngpown, (size_t)ncouls);
        const double complex am = aqsm[2 * ((size_t)n1 * ncouls + igp)] + I * aqsm[2 * ((size_t)n1 * ncouls + igp) + 1];
        double complex c = conj(am) * vcoul[igp];
        if (kind[n1] == 2) c = c * -0.5;
        cdd_add_cprod_dd(&part[tid], creal(c), cimag(c), M);
      }
    }
  }
  cdd_t a;
  memset(&a, 0, sizeof(a));
  for (int t = 0; t < nt; ++t) cdd_add(&a, part[t]);
  for (int iw = 0; iw < nfreqeval; ++iw) {
    achDtemp_cor[2 * iw] = dd_val(a.re);
    achDtemp_cor[2 * iw + 1] = dd_val(a.im);
  }
  free(part);
  free(kind);
  free(fr);
  free(w1);
  free(w2);
}

void oracle_schDttt_cor(int kernel, int ncouls, int ngpown, int ifreq, int n1, double fact1,
                        double fact2, const int* inv_igp_index, const int* indinv,
                        const double* I_epsR, const double* I_epsA, const double* aqsm,
                        const double* aqsn, const double* vcoul, double* schDttt_cor,
                        int nthreads) {
  /* FF/ff.cpp:381-451: sum over my_igp and ig of aqsn(n1,ig) conj(aqsm(n1,igp)) vcoul(igp)
   * times (R-A)(ifreq)*fact1 + (R-A)(ifreq+1)*fact2, times -1/2 in Kernel2 (:443).  The
   * G-vector sum per column is ff_row_dot (compensated), the column sum double-double. */
  const int nt = nthreads_of(nthreads);
  cdd_t* part = (cdd_t*)calloc((size_t)nt, sizeof(cdd_t));
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
#pragma omp for schedule(static)
    for (long lp = 0; lp < (long)ngpown; ++lp) {
      const int igp = igp_of(inv_igp_index, indinv, (size_t)lp);
      cdd_t M = ff_row_dot(aqsn + 2 * (size_t)n1 * ncouls, I_epsR, I_epsA, ifreq, fact1, fact2, (size_t)lp,
                           (size_t)ngpown, (size_t)ncouls);
      const double complex am = aqsm[2 * ((size_t)n1 * ncouls + igp)] + I * aqsm[2 * ((size_t)n1 * ncouls + igp) + 1];
      double complex c = conj(am) * vcoul[igp];
      if (kernel == 2) c = c * -0.5;
      cdd_add_cprod_dd(&part[tid], creal(c), cimag(c), M);
    }
  }
  cdd_t a;
  memset(&a, 0, sizeof(a));
  for (int t = 0; t < nt; ++t) cdd_add(&a, part[t]);
  schDttt_cor[0] = dd_val(a.re);
  schDttt_cor[1] = dd_val(a.im);
  free(part);
}

/* sum_g aqsn(n1,g) (X - Y)(f,p,g) over one frequency slice, compensated. */
static cdd_t ff_slice_dot(const double* an_row, const double* X, const double* Y, int f, size_t p,
                          size_t ngpown, size_t ncouls) {
  const double* x = X + 2 * (((size_t)f * ngpown + p) * ncouls);
  const double* y = Y + 2 * (((size_t)f * ngpown + p) * ncouls);
  lanes_t ar, ai;
  lanes_zero(&ar);
  lanes_zero(&ai);
  size_t g0 = 0;
  for (; g0 + LANES <= ncouls; g0 += LANES) {
    for (int l = 0; l < LANES; ++l) {
      const size_t g = g0 + l;
      const double er = x[2 * g] - y[2 * g], ei = x[2 * g + 1] - y[2 * g + 1];
      const double nr = an_row[2 * g], ni = an_row[2 * g + 1];
      const double tr = nr * er - ni * ei, ti = nr * ei + ni * er;
      double s1, e1, s2, e2;
      two_sum(ar.h[l], tr, &s1, &e1);
      ar.h[l] = s1;
      ar.l[l] += e1;
      two_sum(ai.h[l], ti, &s2, &e2);
      ai.h[l] = s2;
      ai.l[l] += e2;
    }
  }
  cdd_t r = {lanes_fold(&ar), lanes_fold(&ai)};
  for (; g0 < ncouls; ++g0) {
    const double er = x[2 * g0] - y[2 * g0], ei = x[2 * g0 + 1] - y[2 * g0 + 1];
    const double nr = an_row[2 * g0], ni = an_row[2 * g0 + 1];
    dd_add(&r.re, nr * er - ni * ei);
    dd_add(&r.im, nr * ei + ni * er);
  }
  return r;
}

void oracle_achDtemp(int nb, int nvband, int nfreqeval, int ncouls, int ngpown, int nFreq,
                     double freqevalmin, double freqevalstep, const double* ekq,
                     const double* dFreqGrid, const double* pref, const int* inv_igp_index,
                     const int* indinv, const double* aqsm, const double* aqsn, const double* vcoul,
                     const double* I_epsR, const double* I_epsA, double* achDtemp, int nthreads) {
  /* The real-axis frequency sum that the reference prepares and does not apply: the
   * trapezoid weights pref (FF/ff.cpp:619-626) and the spectral normalisation pref_zb =
   * 0.5/3.14 (:468).  achDtemp = sum_{n1,p,g} aqsn(n1,g) conj(aqsm(n1,igp)) vcoul(igp)
   * sum_f w(n1,f) (R - A)(f,p,g), with w(n1,f) = -i/2 pref(f) / (wx(n1) - s(n1) dFreqGrid(f)),
   * wx(n1) = freqevalmin - ekq(n1) + freqevalstep as in the reference's kernels (:245,
   * :337), s = -1 for occupied bands (n1 < nvband), +1 otherwise, and w = 0 where the
   * denominator is exactly zero.  No iw dependence: one sum, replicated over iw. */
  const int nt = nthreads_of(nthreads);
  double* wim = (double*)calloc((size_t)nb * nFreq, sizeof(double));
  for (int n1 = 0; n1 < nb; ++n1) {
    const double wx = freqevalmin - ekq[n1] + freqevalstep;
    for (int f = 0; f < nFreq; ++f) {
      const double den = n1 < nvband ? wx + dFreqGrid[f] : wx - dFreqGrid[f];
      wim[(size_t)n1 * nFreq + f] = den == 0.0 ? 0.0 : -0.5 * (pref[f] / den);
    }
  }
  cdd_t* part = (cdd_t*)calloc((size_t)nt, sizeof(cdd_t));
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
#pragma omp for schedule(static)
    for (long lp = 0; lp < (long)ngpown; ++lp) {
      const int igp = igp_of(inv_igp_index, indinv, (size_t)lp);
      for (int n1 = 0; n1 < nb; ++n1) {
        const double complex am = aqsm[2 * ((size_t)n1 * ncouls + igp)] + I * aqsm[2 * ((size_t)n1 * ncouls + igp) + 1];
        const double complex c = conj(am) * vcoul[igp];
        for (int f = 0; f < nFreq; ++f) {
          const double wf = wim[(size_t)n1 * nFreq + f];
          if (wf == 0.0) continue;
          cdd_t M = ff_slice_dot(aqsn + 2 * (size_t)n1 * ncouls, I_epsR, I_epsA, f, (size_t)lp,
                                 (size_t)ngpown, (size_t)ncouls);
          const double complex cw = c * (I * wf);
          cdd_add_cprod_dd(&part[tid], creal(cw), cimag(cw), M);
        }
      }
    }
  }
  cdd_t a;
  memset(&a, 0, sizeof(a));
  for (int t = 0; t < nt; ++t) cdd_add(&a, part[t]);
  for (int iw = 0; iw < nfreqeval; ++iw) {
    achDtemp[2 * iw] = dd_val(a.re);
    achDtemp[2 * iw + 1] = dd_val(a.im);
  }
  free(part);
  free(wim);
}
