/*
 * bgw_oracle.h -- CPU restatement of the reference GPP and FF kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker.
 * The product (libbgw) never links or calls it.
 *
 * Arithmetic: FP64 per term exactly as the reference writes it, sums
 * compensated (TwoSum lanes over G-vectors, double-double across bands and
 * G'-columns), so the oracle's own rounding error is ~1e-16 relative and the
 * 1e-12 parity bar measures the GPU path, not the oracle.
 *
 * Layouts are the reference's (row-major ArrayMD, interleaved complex).
 * Every function cites the reference lines it restates; the checks that pin
 * each function are listed in tests/test_oracle.py.
 */
#ifndef BGW_ORACLE_H
#define BGW_ORACLE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* GPP forward, GPP/gpp.cpp:339-401 generalised to wx(n1,iw) = wx[n1*3+iw].
 * Outputs as in bgw_gpp_solver (any of acht_n1_loc/achstemp/asxtemp may be NULL). */
void oracle_gpp_forward(size_t nb, size_t nvband, size_t ngpown, size_t ncouls,
                        const int* inv_igp_index, const int* indinv, const double* wx,
                        const double* wtilde, const double* aqsm, const double* aqsn,
                        const double* I_eps, const double* vcoul, double occ,
                        double* achtemp, double* acht_n1_loc, double* achstemp,
                        double* asxtemp, int nthreads);

/* GPP adjoint at selected entries e = p*ncouls+g (entries[k]); outputs
 * dL/dwtilde and dL/dI_eps (c128[n_entries]) for L = Re sum_iw(asxtemp+achtemp). */
void oracle_gpp_grad_entries(size_t nb, size_t nvband, size_t ngpown, size_t ncouls,
                             const int* inv_igp_index, const int* indinv, const double* wx,
                             const double* wtilde, const double* aqsm, const double* aqsn,
                             const double* I_eps, const double* vcoul, double occ,
                             const long* entries, size_t n_entries,
                             double* dL_dwtilde, double* dL_dI_eps, int nthreads);

/* d(achtemp[iw]+asxtemp[iw])/d omega, omega added to every wx(n1,iw); c128[3]. */
void oracle_gpp_dsigma_domega(size_t nb, size_t nvband, size_t ngpown, size_t ncouls,
                              const int* inv_igp_index, const int* indinv, const double* wx,
                              const double* wtilde, const double* aqsm, const double* aqsn,
                              const double* I_eps, const double* vcoul, double occ,
                              double* dsigma_domega, int nthreads);

/* FF/ff.cpp:103-152 */
int oracle_compute_fact(double wx, int nFreq, const double* dFreqGrid, double* fact1,
                         double* fact2, int* ifreq, int loop, int flag_occ);
/* FF/ff.cpp:171-209 */
void oracle_achsDtemp(int nb, int ngpown, int ncouls, const int* inv_igp_index,
                      const int* indinv, const double* aqsn, const double* aqsm,
                      const double* I_epsR, const double* vcoul, double* achsDtemp,
                      int nthreads);
/* FF/ff.cpp:211-293 */
void oracle_asxDtemp(int nvband, int nfreqeval, int ncouls, int ngpown, int nFreq,
                     double freqevalmin, double freqevalstep, double occ, const double* ekq,
                     const double* dFreqGrid, const int* inv_igp_index, const int* indinv,
                     const double* aqsm, const double* aqsn, const double* vcoul,
                     const double* I_epsR, const double* I_epsA, double* asxDtemp,
                     int nthreads);
/* FF/ff.cpp:295-379 with schDttt_corKernel1/2 :381-451 */
void oracle_achDtemp_cor(int nb, int nvband, int nfreqeval, int ncouls, int ngpown, int nFreq,
                         double freqevalmin, double freqevalstep, const double* ekq,
                         const double* dFreqGrid, const int* inv_igp_index, const int* indinv,
                         const double* aqsm, const double* aqsn, const double* vcoul,
                         const double* I_epsR, const double* I_epsA, double* achDtemp_cor,
                         int nthreads);
/* The real-axis frequency sum with the reference's unused weights pref (FF/ff.cpp:619-626)
 * and pref_zb (:468): achDtemp = sum_{n1,p,g} aqsn conj(aqsm(igp)) vcoul(igp)
 * sum_f -i/2 pref(f) (R - A)(f,p,g) / (wx(n1) - s(n1) dFreqGrid(f)), s = -1 occupied. */
void oracle_achDtemp(int nb, int nvband, int nfreqeval, int ncouls, int ngpown, int nFreq,
                     double freqevalmin, double freqevalstep, const double* ekq,
                     const double* dFreqGrid, const double* pref, const int* inv_igp_index,
                     const int* indinv, const double* aqsm, const double* aqsn, const double* vcoul,
                     const double* I_epsR, const double* I_epsA, double* achDtemp, int nthreads);
/* schDttt_corKernel1 (kernel 1, FF/ff.cpp:381-416) / schDttt_corKernel2 (kernel 2,
 * :418-451, factor -1/2): band n1 on slices ifreq, ifreq+1 of I_epsR - I_epsA. */
void oracle_schDttt_cor(int kernel, int ncouls, int ngpown, int ifreq, int n1, double fact1,
                        double fact2, const int* inv_igp_index, const int* indinv,
                        const double* I_epsR, const double* I_epsA, const double* aqsm,
                        const double* aqsn, const double* vcoul, double* schDttt_cor,
                        int nthreads);

#ifdef __cplusplus
}
#endif
#endif
