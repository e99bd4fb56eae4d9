// ref_adapter.cpp -- C entry points onto the REFERENCE's own kernels (TEST INFRASTRUCTURE).
//
// Built by oracle/Makefile together with the reference's unmodified translation
// units GPP/gpp.cpp and FF/ff.cpp (compiled in place under the reference tree,
// their `main` renamed by a -D flag) into oracle/_ref/libbgwref_*.so.  Used only by
// tests (to pin the oracle) and by bench.py's cpu_baseline leg (to time the
// reference CPU path).  Nothing of the reference is copied here: this file only
// wraps caller buffers into the reference's ArrayMD containers and calls the
// reference functions.
#include "common/Defines.h"

#include <chrono>
#include <cstring>

// Prototypes of the reference FF kernels, whose definitions (external linkage)
// are in FF/ff.cpp:171-175, :211-217 and :295-302.  Same signatures, because the
// linker binds them to those definitions.
void achsDtemp_Kernel(int number_bands, int ngpown, int ncouls, ARRAY1D_int &inv_igp_index,
                      ARRAY1D_int &indinv, ARRAY2D &aqsntemp, ARRAY2D &aqsmtemp,
                      ARRAY3D &I_epsR_array, ARRAY1D_DataType &vcoul, ComplexType &achsDtemp);
void asxDtemp_Kernel(int nvband, int nfreqeval, int ncouls, int ngpown, int nFreq,
                     double freqevalmin, double freqevalstep, double occ, ARRAY1D_DataType &ekq,
                     ARRAY1D_DataType &dFreqGrid, ARRAY1D_int &inv_igp_index, ARRAY1D_int &indinv,
                     ARRAY2D &aqsmtemp, ARRAY2D &aqsntemp, ARRAY1D_DataType &vcoul,
                     ARRAY3D &I_epsR_array, ARRAY3D &I_epsA_array, ARRAY1D &asxDtemp);
void achDtemp_cor_Kernel(int number_bands, int nvband, int nfreqeval, int ncouls, int ngpown,
                         int nFreq, double freqevalmin, double freqevalstep,
                         ARRAY1D_DataType &ekq, ARRAY1D_DataType &dFreqGrid,
                         ARRAY1D_int &inv_igp_index, ARRAY1D_int &indinv, ARRAY2D &aqsmtemp,
                         ARRAY2D &aqsntemp, ARRAY1D_DataType &vcoul, ARRAY3D &I_epsR_array,
                         ARRAY3D &I_epsA_array, ARRAY1D &achDtemp_cor);

namespace {
template <typename A, typename T>
void fill(A &arr, const T *src) {
  std::memcpy(arr.dptr, src, arr.size * sizeof(*arr.dptr));
}
template <typename A>
void fill_c(A &arr, const double *src) {
  std::memcpy(static_cast<void *>(arr.dptr), src, arr.size * sizeof(*arr.dptr));
}

// wall time of the last reference kernel call alone (without the copies into ArrayMD)
double g_last_seconds = 0.0;
template <typename F>
void timed(F &&f) {
  const auto t0 = std::chrono::steady_clock::now();
  f();
  g_last_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
}  // namespace

extern "C" {

double ref_last_seconds(void) { return g_last_seconds; }

// noflagOCC_solver (GPP/gpp.cpp:307) on caller data; wx3[3] as the reference takes it.
void ref_noflagOCC_solver(size_t nb, size_t ngpown, size_t ncouls, const int *inv_igp_index,
                          const int *indinv, const double *wx3, const double *wtilde,
                          const double *aqsm, const double *aqsn, const double *I_eps,
                          const double *vcoul, double *achtemp) {
  ARRAY1D achtemp_a(nend - nstart);
  ARRAY2D aqsmtemp(nb, ncouls), aqsntemp(nb, ncouls);
  ARRAY2D I_eps_array(ngpown, ncouls), wtilde_array(ngpown, ncouls);
  ARRAY1D_DataType vcoul_a(ncouls);
  ARRAY1D_int inv_a(ngpown), indinv_a(ncouls + 1);
  ARRAY1D_DataType wx_a(nend - nstart);
  fill_c(aqsmtemp, aqsm);
  fill_c(aqsntemp, aqsn);
  fill_c(I_eps_array, I_eps);
  fill_c(wtilde_array, wtilde);
  fill(vcoul_a, vcoul);
  fill(inv_a, inv_igp_index);
  fill(indinv_a, indinv);
  fill(wx_a, wx3);
  noflagOCC_solver(nb, ngpown, ncouls, inv_a, indinv_a, wx_a, wtilde_array, aqsmtemp, aqsntemp,
                   I_eps_array, vcoul_a, achtemp_a);
  std::memcpy(achtemp, static_cast<void *>(achtemp_a.dptr), 3 * 2 * sizeof(double));
}

// Band-dependent frequencies wx[nb*3]: the reference solver called once per band
// on a one-band slice; per-band results go to acht_n1_loc[nb*3] (c128).
void ref_gpp_per_band(size_t nb, size_t ngpown, size_t ncouls, const int *inv_igp_index,
                      const int *indinv, const double *wx, const double *wtilde,
                      const double *aqsm, const double *aqsn, const double *I_eps,
                      const double *vcoul, double *acht_n1_loc) {
  ARRAY1D achtemp_a(nend - nstart);
  ARRAY2D aqsmtemp(1, ncouls), aqsntemp(1, ncouls);
  ARRAY2D I_eps_array(ngpown, ncouls), wtilde_array(ngpown, ncouls);
  ARRAY1D_DataType vcoul_a(ncouls);
  ARRAY1D_int inv_a(ngpown), indinv_a(ncouls + 1);
  ARRAY1D_DataType wx_a(nend - nstart);
  fill_c(I_eps_array, I_eps);
  fill_c(wtilde_array, wtilde);
  fill(vcoul_a, vcoul);
  fill(inv_a, inv_igp_index);
  fill(indinv_a, indinv);
  // timed: the band loop (its per-band row copies are ~1 MB against the 1.7 GB each
  // band streams), not the copies of wtilde / I_eps above
  timed([&] {
    for (size_t n1 = 0; n1 < nb; ++n1) {
      fill_c(aqsmtemp, aqsm + 2 * n1 * ncouls);
      fill_c(aqsntemp, aqsn + 2 * n1 * ncouls);
      fill(wx_a, wx + 3 * n1);
      noflagOCC_solver(1, ngpown, ncouls, inv_a, indinv_a, wx_a, wtilde_array, aqsmtemp,
                       aqsntemp, I_eps_array, vcoul_a, achtemp_a);
      std::memcpy(acht_n1_loc + 6 * n1, static_cast<void *>(achtemp_a.dptr), 6 * sizeof(double));
    }
  });
}

void ref_achsDtemp(int nb, int ngpown, int ncouls, int nFreq, const int *inv_igp_index,
                   const int *indinv, const double *aqsn, const double *aqsm,
                   const double *I_epsR, const double *vcoul, double *out) {
  ARRAY1D_int inv_a(ngpown), indinv_a(ncouls);
  ARRAY2D aqsntemp(nb, ncouls), aqsmtemp(nb, ncouls);
  ARRAY3D I_epsR_array(nFreq, ngpown, ncouls);
  ARRAY1D_DataType vcoul_a(ncouls);
  fill(inv_a, inv_igp_index);
  fill(indinv_a, indinv);
  fill_c(aqsntemp, aqsn);
  fill_c(aqsmtemp, aqsm);
  fill_c(I_epsR_array, I_epsR);
  fill(vcoul_a, vcoul);
  ComplexType res(0.0, 0.0);
  timed([&] {
    achsDtemp_Kernel(nb, ngpown, ncouls, inv_a, indinv_a, aqsntemp, aqsmtemp, I_epsR_array, vcoul_a,
                     res);
  });
  out[0] = res.real();
  out[1] = res.imag();
}

void ref_asxDtemp(int nb, int nvband, int nfreqeval, int ncouls, int ngpown, int nFreq,
                  double freqevalmin, double freqevalstep, double occ, const double *ekq,
                  const double *dFreqGrid, const int *inv_igp_index, const int *indinv,
                  const double *aqsm, const double *aqsn, const double *vcoul,
                  const double *I_epsR, const double *I_epsA, double *out) {
  ARRAY1D_DataType ekq_a(nb), grid_a(nFreq), vcoul_a(ncouls);
  ARRAY1D_int inv_a(ngpown), indinv_a(ncouls);
  ARRAY2D aqsmtemp(nb, ncouls), aqsntemp(nb, ncouls);
  ARRAY3D R(nFreq, ngpown, ncouls), A(nFreq, ngpown, ncouls);
  ARRAY1D res(nfreqeval);
  fill(ekq_a, ekq);
  fill(grid_a, dFreqGrid);
  fill(vcoul_a, vcoul);
  fill(inv_a, inv_igp_index);
  fill(indinv_a, indinv);
  fill_c(aqsmtemp, aqsm);
  fill_c(aqsntemp, aqsn);
  fill_c(R, I_epsR);
  fill_c(A, I_epsA);
  timed([&] {
    asxDtemp_Kernel(nvband, nfreqeval, ncouls, ngpown, nFreq, freqevalmin, freqevalstep, occ, ekq_a,
                    grid_a, inv_a, indinv_a, aqsmtemp, aqsntemp, vcoul_a, R, A, res);
  });
  std::memcpy(out, static_cast<void *>(res.dptr), nfreqeval * 2 * sizeof(double));
}

void ref_achDtemp_cor(int nb, int nvband, int nfreqeval, int ncouls, int ngpown, int nFreq,
                      double freqevalmin, double freqevalstep, const double *ekq,
                      const double *dFreqGrid, const int *inv_igp_index, const int *indinv,
                      const double *aqsm, const double *aqsn, const double *vcoul,
                      const double *I_epsR, const double *I_epsA, double *out) {
  ARRAY1D_DataType ekq_a(nb), grid_a(nFreq), vcoul_a(ncouls);
  ARRAY1D_int inv_a(ngpown), indinv_a(ncouls);
  ARRAY2D aqsmtemp(nb, ncouls), aqsntemp(nb, ncouls);
  ARRAY3D R(nFreq, ngpown, ncouls), A(nFreq, ngpown, ncouls);
  ARRAY1D res(nfreqeval);
  fill(ekq_a, ekq);
  fill(grid_a, dFreqGrid);
  fill(vcoul_a, vcoul);
  fill(inv_a, inv_igp_index);
  fill(indinv_a, indinv);
  fill_c(aqsmtemp, aqsm);
  fill_c(aqsntemp, aqsn);
  fill_c(R, I_epsR);
  fill_c(A, I_epsA);
  timed([&] {
    achDtemp_cor_Kernel(nb, nvband, nfreqeval, ncouls, ngpown, nFreq, freqevalmin, freqevalstep,
                        ekq_a, grid_a, inv_a, indinv_a, aqsmtemp, aqsntemp, vcoul_a, R, A, res);
  });
  std::memcpy(out, static_cast<void *>(res.dptr), nfreqeval * 2 * sizeof(double));
}

}  // extern "C"
