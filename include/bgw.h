/*
 * bgw.h -- C ABI of the MI355X (gfx950) BerkeleyGW self-energy kernels.
 *
 * Drop-in boundary for the reference mini-apps' kernel API:
 *   GPP  noflagOCC_solver            GPP/gpp.cpp:307-402   (decl common/Defines.h:50-62)
 *   FF   achsDtemp_Kernel            FF/ff.cpp:171-209
 *        asxDtemp_Kernel             FF/ff.cpp:211-293
 *        achDtemp_cor_Kernel         FF/ff.cpp:295-379     (+ schDttt_corKernel1/2 :381-451)
 *        schDttt_corKernel1/2        FF/ff.cpp:381-451     (decl common/Defines.h:36-48)
 *        compute_fact                FF/ff.cpp:103-152
 *
 * Conventions
 *  - Every reference `ArrayMD&` argument becomes a plain pointer to its `.dptr`
 *    (row-major, last index fastest; arrayMD/arrayMD.h:58-81).  Complex arrays are
 *    `double*` to interleaved (re,im) pairs, layout-compatible with
 *    `std::complex<double>[]` (common/Defines.h:21-22 with -Dstd_complex).
 *  - Argument order is the reference's, with a leading `bgw_ctx*`; `void` becomes
 *    an `int` status (BGW_OK or a negative BGW_E* code).  Nothing throws across
 *    the ABI; there is no CPU fallback: without a usable GPU every compute entry
 *    point returns BGW_ENODEV.
 *  - Functions without suffix take HOST pointers and are synchronous: results are
 *    valid on return, exactly like the reference.  `_device` variants take DEVICE
 *    pointers for every array, enqueue on the context's stream and return at once;
 *    call bgw_synchronize() before reading results.  Index arrays passed to a
 *    `_device` variant are checked on the GPU; a bad index is clamped (no fault)
 *    and reported as BGW_EINVAL by the next bgw_synchronize().
 *  - One context = one GPU + one stream.  Distinct contexts may be used from
 *    distinct threads; one context must not be used concurrently.
 */
#ifndef BGW_H
#define BGW_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define BGW_OK      0
#define BGW_EINVAL  (-22) /* null pointer, zero size, bad index, ngpown>ncouls, overflow */
#define BGW_ENOMEM  (-12) /* device allocation failed                                   */
#define BGW_EHIP    (-5)  /* any other HIP runtime error (text in bgw_last_error)       */
#define BGW_ENODEV  (-19) /* no GPU / HIP runtime not usable                            */
#define BGW_EIO     (-74) /* input-set file missing, unreadable or malformed
                             (text in bgw_io_last_error)                              */

/* ---- arithmetic forms (bgw_set_precision) ------------------------------- */
#define BGW_FP64       0  /* plain FP64 throughout (default)                           */
#define BGW_FP32_BULK  1  /* bulk complex products in FP32 (GPP) / FP32-operand MFMA with
                             FP64 stage folds (FF contraction: operands rounded once to
                             FP32, FP32 sums over 16 G-vectors folded into FP64),
                             FP64 pole denominators and FP64 compensated reductions   */
#define BGW_FP64_MOMENTS 2  /* GPP achtemp / acht_n1_loc sum: the pole sum
                             sum_g I_eps(p,g)/(wx - wtilde(p,g)) of each 2048-G chunk from
                             16 moments of the Moebius variable phi = (w - c)/(w - e), e the
                             middle of the band-frequency range, FP64, over the poles with
                             |phi| h/D <= 0.095 (h the range's half-width, D the distance from
                             c to it): truncation below 2^-54 x the sum of the terms'
                             magnitudes; up to 128 other poles per chunk summed directly, a
                             chunk with more by the direct FP64 sum as a whole.  The
                             adjoint's band sums likewise, from 16 moments over a row's band
                             frequencies per pole (waves of 64 poles with a pole near a band
                             frequency take the band loop).  Every other kernel runs as in
                             BGW_FP64.                                                    */

typedef struct bgw_ctx bgw_ctx;

/* ---- context ------------------------------------------------------------ */
/* None of the context, stream, precision and timing calls below has a reference
 * counterpart: the reference has no handle and maps data per call inside the
 * solver (GPP/gpp.cpp:320-334).  Creates a context on HIP device `device`. */
int         bgw_create(bgw_ctx** ctx, int device);
void        bgw_destroy(bgw_ctx* ctx);
const char* bgw_last_error(const bgw_ctx* ctx);
/* Launch on a caller-owned hipStream_t (NULL = the context's own stream).  Switching
 * to a different stream first waits for the work queued on the current one (it may
 * still use the context's buffers). */
int         bgw_set_stream(bgw_ctx* ctx, void* hip_stream);
void*       bgw_get_stream(bgw_ctx* ctx);
int         bgw_set_precision(bgw_ctx* ctx, int precision);
int         bgw_synchronize(bgw_ctx* ctx);
/* Per-kernel HIP-event timing of the dominant kernels (off by default). */
int         bgw_set_timing(bgw_ctx* ctx, int on);
int         bgw_kernel_time(bgw_ctx* ctx, const char* kernel, double* total_ms, long* launches);
int         bgw_reset_timing(bgw_ctx* ctx);
/* Number of compute units of the context's device (for launch sizing reports). */
int         bgw_device_cus(bgw_ctx* ctx);

/* ---- GPP ---------------------------------------------------------------- */
/* Replaces noflagOCC_solver (GPP/gpp.cpp:307-312, common/Defines.h:50-62):
 *   achtemp[iw] = sum_{n1<number_bands} sum_{p<ngpown} sum_{g<ncouls}
 *                 sch_store1(n1,p) * I_eps(p,g) * conj(wdiff)/|wdiff|^2,
 *   wdiff = wx_array[iw] - wtilde(p,g),  iw = 0..2  (nstart=0, nend=3, common/Defines.h:16-17)
 *   sch_store1 = conj(aqsm(n1,igp))*aqsn(n1,igp)*0.5*vcoul(igp)*wtilde(p,igp),
 *   igp = indinv[inv_igp_index[p]]                               (GPP/gpp.cpp:362-371)
 * Sizes: inv_igp_index[ngpown] in [0,ncouls]; indinv[ncouls+1] in [0,ncouls);
 * wx_array[3]; wtilde_array, I_eps_array c128[ngpown*ncouls]; aqsmtemp,
 * aqsntemp c128[number_bands*ncouls]; vcoul[ncouls]; achtemp c128[3] (out). */
int bgw_noflagOCC_solver(bgw_ctx* ctx, size_t number_bands, size_t ngpown, size_t ncouls,
                         const int* inv_igp_index, const int* indinv,
                         const double* wx_array, const double* wtilde_array,
                         const double* aqsmtemp, const double* aqsntemp,
                         const double* I_eps_array, const double* vcoul,
                         double* achtemp);

/* Extended GPP solver (same loop nest, GPP/gpp.cpp:339-391) with band-dependent
 * frequencies wx_array[number_bands*3] (wx(n1,iw) = e_lk - ek[n1] + dw*(iw-1),
 * the reference's GPP/gpp.cpp:271-275 with e_n1kq -> ek[n1]) and the optional
 * outputs of the full BerkeleyGW GPP kernel (NULL = not computed):
 *   achtemp      c128[3]                 required
 *   acht_n1_loc  c128[number_bands*3]    per-band split: sum_n1 acht_n1_loc[n1][iw] = achtemp[iw]
 *   achstemp     c128[1]                 static COH term
 *                  -1/2 sum_{n1,p} vcoul(igp) conj(aqsm(n1,igp)) sum_g aqsn(n1,g) I_eps(p,g) [|I_eps|>to1]
 *   asxtemp      c128[3]                 screened exchange over occupied bands n1<nvband:
 *                  sum occ vcoul(igp) sum_g aqsn(n1,g) conj(aqsm(n1,igp)) ssx(p,g,iw),
 *                  ssx = Omega2/(wx^2 - wtilde^2) with Omega2 = wtilde^2 I_eps in the
 *                  regular region; near-pole and cut-off branches as in BerkeleyGW
 *                  (limittwo = 0.25 of GPP/gpp.cpp:168, limitone = 1/(4 to1), sexcut = 4).
 * No reference counterpart beyond noflagOCC_solver (the reference keeps only achtemp). */
int bgw_gpp_solver(bgw_ctx* ctx, size_t number_bands, size_t nvband, size_t ngpown, size_t ncouls,
                   const int* inv_igp_index, const int* indinv,
                   const double* wx_array, const double* wtilde_array,
                   const double* aqsmtemp, const double* aqsntemp,
                   const double* I_eps_array, const double* vcoul, double occ,
                   double* achtemp, double* acht_n1_loc, double* achstemp, double* asxtemp);
/* Device-pointer form of bgw_gpp_solver (all arrays on the device, results after
 * bgw_synchronize); the same loop nest, GPP/gpp.cpp:339-391. */
int bgw_gpp_solver_device(bgw_ctx* ctx, size_t number_bands, size_t nvband, size_t ngpown, size_t ncouls,
                          const int* inv_igp_index, const int* indinv,
                          const double* wx_array, const double* wtilde_array,
                          const double* aqsmtemp, const double* aqsntemp,
                          const double* I_eps_array, const double* vcoul, double occ,
                          double* achtemp, double* acht_n1_loc, double* achstemp, double* asxtemp);

/* Allocation-free form of bgw_gpp_solver_device for the noflagOCC_solver outputs
 * (achtemp required, acht_n1_loc optional; wx_array[number_bands*3]), in the context's
 * precision and on its stream.  No reference counterpart: it is for callers that cannot
 * allocate device memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  It calls no allocating routine: all workspace comes from a
 * 64 MiB device array of the library (one per GPU, for the life of the process), and a
 * problem whose workspace does not fit returns BGW_EINVAL without launching anything (the
 * benchmark workload, 512 bands x 1638 G', needs 37-51 MiB, depending on the arithmetic form).  Every call on a GPU
 * shares that GPU's array, so calls that overlap in time on one GPU (two streams, or
 * two contexts on one device) race; calls queued on one stream do not.  Results equal
 * bgw_gpp_solver_device's bit for bit.  Bad indices are clamped and reported by the
 * next bgw_synchronize, as for the other _device calls. */
int bgw_gpp_solver_device_noalloc(bgw_ctx* ctx, size_t number_bands, size_t ngpown, size_t ncouls,
                                  const int* inv_igp_index, const int* indinv,
                                  const double* wx_array, const double* wtilde_array,
                                  const double* aqsmtemp, const double* aqsntemp,
                                  const double* I_eps_array, const double* vcoul,
                                  double* achtemp, double* acht_n1_loc);

/* Adjoint of bgw_gpp_solver for L = Re sum_iw (asxtemp[iw] + sum_n1 acht_n1_loc[n1][iw]).
 * Outputs (NULL = not computed):
 *   dsigma_domega c128[3]  d(achtemp[iw]+asxtemp[iw])/d omega, omega shifting every wx(n1,iw)
 *   dL_dwtilde    c128[ngpown*ncouls]   dL/dRe + i dL/dIm of wtilde_array
 *   dL_dI_eps     c128[ngpown*ncouls]   dL/dRe + i dL/dIm of I_eps_array
 * Pole terms are recomputed, not stored.  No reference counterpart. */
int bgw_gpp_backward(bgw_ctx* ctx, size_t number_bands, size_t nvband, size_t ngpown, size_t ncouls,
                     const int* inv_igp_index, const int* indinv,
                     const double* wx_array, const double* wtilde_array,
                     const double* aqsmtemp, const double* aqsntemp,
                     const double* I_eps_array, const double* vcoul, double occ,
                     double* dsigma_domega, double* dL_dwtilde, double* dL_dI_eps);
int bgw_gpp_backward_device(bgw_ctx* ctx, size_t number_bands, size_t nvband, size_t ngpown, size_t ncouls,
                            const int* inv_igp_index, const int* indinv,
                            const double* wx_array, const double* wtilde_array,
                            const double* aqsmtemp, const double* aqsntemp,
                            const double* I_eps_array, const double* vcoul, double occ,
                            double* dsigma_domega, double* dL_dwtilde, double* dL_dI_eps);

/* ---- FF ----------------------------------------------------------------- */
/* Replaces compute_fact (FF/ff.cpp:103-152), including its quirks: strict '<'
 * bracketing, ifreq==0 -> nFreq-2 remap for loop 1 and loop-2 occupied, no remap
 * for loop 2 with wx>0, nothing set for loop 2 unoccupied with wx<=0.  `ifreq`,
 * `fact1`, `fact2` are in/out exactly like the reference's references.  Host only.
 * Returns BGW_EINVAL, without reading the grid, when no interval brackets wx and
 * the incoming ifreq lies outside [0, nFreq-2] (the reference reads out of bounds). */
int bgw_compute_fact(double wx, int nFreq, const double* dFreqGrid, double* fact1,
                     double* fact2, int* ifreq, int loop, int flag_occ);

/* Replaces achsDtemp_Kernel (FF/ff.cpp:171-175; note aqsntemp before aqsmtemp):
 *   achsDtemp = sum_{n1,p} 0.5 vcoul(igp) * (-sum_g aqsn(n1,g) conj(aqsm(n1,igp)) I_epsR(1,p,g))
 * I_epsR_array c128[nFreq*ngpown*ncouls] (slice 1 is read; nFreq>=2), indinv[ncouls]. */
int bgw_achsDtemp_Kernel(bgw_ctx* ctx, int number_bands, int ngpown, int ncouls,
                         const int* inv_igp_index, const int* indinv,
                         const double* aqsntemp, const double* aqsmtemp,
                         const double* I_epsR_array, const double* vcoul,
                         double* achsDtemp);
/* Replaces asxDtemp_Kernel (FF/ff.cpp:211-217).  asxDtemp c128[nfreqeval];
 * ekq[nvband..], dFreqGrid[nFreq], I_eps{R,A}_array c128[nFreq*ngpown*ncouls]. */
int bgw_asxDtemp_Kernel(bgw_ctx* ctx, int nvband, int nfreqeval, int ncouls, int ngpown,
                        int nFreq, double freqevalmin, double freqevalstep, double occ,
                        const double* ekq, const double* dFreqGrid,
                        const int* inv_igp_index, const int* indinv,
                        const double* aqsmtemp, const double* aqsntemp,
                        const double* vcoul, const double* I_epsR_array,
                        const double* I_epsA_array, double* asxDtemp);
/* Replaces achDtemp_cor_Kernel (FF/ff.cpp:295-302) with schDttt_corKernel1/2
 * (FF/ff.cpp:381-451) folded in.  achDtemp_cor c128[nfreqeval]; ekq[number_bands]. */
int bgw_achDtemp_cor_Kernel(bgw_ctx* ctx, int number_bands, int nvband, int nfreqeval,
                            int ncouls, int ngpown, int nFreq, double freqevalmin,
                            double freqevalstep, const double* ekq, const double* dFreqGrid,
                            const int* inv_igp_index, const int* indinv,
                            const double* aqsmtemp, const double* aqsntemp,
                            const double* vcoul, const double* I_epsR_array,
                            const double* I_epsA_array, double* achDtemp_cor);

/* Device-pointer forms of the three FF kernels above (all arrays on the device, results
 * after bgw_synchronize); they replace the same reference functions, FF/ff.cpp:171-302. */
int bgw_achsDtemp_Kernel_device(bgw_ctx* ctx, int number_bands, int ngpown, int ncouls,
                                const int* inv_igp_index, const int* indinv,
                                const double* aqsntemp, const double* aqsmtemp,
                                const double* I_epsR_array, const double* vcoul,
                                double* achsDtemp);
int bgw_asxDtemp_Kernel_device(bgw_ctx* ctx, int nvband, int nfreqeval, int ncouls, int ngpown,
                               int nFreq, double freqevalmin, double freqevalstep, double occ,
                               const double* ekq, const double* dFreqGrid,
                               const int* inv_igp_index, const int* indinv,
                               const double* aqsmtemp, const double* aqsntemp,
                               const double* vcoul, const double* I_epsR_array,
                               const double* I_epsA_array, double* asxDtemp);
int bgw_achDtemp_cor_Kernel_device(bgw_ctx* ctx, int number_bands, int nvband, int nfreqeval,
                                   int ncouls, int ngpown, int nFreq, double freqevalmin,
                                   double freqevalstep, const double* ekq, const double* dFreqGrid,
                                   const int* inv_igp_index, const int* indinv,
                                   const double* aqsmtemp, const double* aqsntemp,
                                   const double* vcoul, const double* I_epsR_array,
                                   const double* I_epsA_array, double* achDtemp_cor);

/* The real-axis frequency sum of the FF self-energy, the term the reference prepares in its
 * driver and never applies: its trapezoid weights pref (FF/ff.cpp:564-565,619-626), the
 * spectral normalisation pref_zb = 0.5/3.14 (:468) and the orphan achDtemp timers (:503).
 * No reference counterpart; defined here as
 *   achDtemp = sum_{n1<number_bands} sum_{p<ngpown} sum_{g<ncouls} aqsn(n1,g) conj(aqsm(n1,igp))
 *              vcoul(igp) sum_{f<nFreq} -i/2 pref(f) (R - A)(f,p,g) / (wx(n1) - s(n1) dFreqGrid(f)),
 *   wx(n1) = freqevalmin - ekq(n1) + freqevalstep  (as in the reference's FF kernels, :245,:337),
 *   s(n1) = -1 for occupied bands n1 < nvband (pole at -dFreqGrid), +1 otherwise,
 * R - A = I_epsR_array - I_epsA_array (the spectral function times 2i, so -i/2 pref(f) is
 * -i pref_zb h_f with h_f the trapezoid step), a term with an exactly zero denominator left
 * out.  pref[nFreq] is the reference driver's array: (dFreqGrid(f+1) - dFreqGrid(f))/3.14,
 * the last entry repeating the one before, both ends halved.  achDtemp c128[nfreqeval], all
 * entries equal (wx has no iw term); the other arguments as bgw_achDtemp_cor_Kernel's. */
int bgw_achDtemp_Kernel(bgw_ctx* ctx, int number_bands, int nvband, int nfreqeval, int ncouls,
                        int ngpown, int nFreq, double freqevalmin, double freqevalstep,
                        const double* ekq, const double* dFreqGrid, const double* pref,
                        const int* inv_igp_index, const int* indinv,
                        const double* aqsmtemp, const double* aqsntemp,
                        const double* vcoul, const double* I_epsR_array,
                        const double* I_epsA_array, double* achDtemp);
int bgw_achDtemp_Kernel_device(bgw_ctx* ctx, int number_bands, int nvband, int nfreqeval,
                               int ncouls, int ngpown, int nFreq, double freqevalmin,
                               double freqevalstep, const double* ekq, const double* dFreqGrid,
                               const double* pref, const int* inv_igp_index, const int* indinv,
                               const double* aqsmtemp, const double* aqsntemp,
                               const double* vcoul, const double* I_epsR_array,
                               const double* I_epsA_array, double* achDtemp);

/* Replace schDttt_corKernel1 (FF/ff.cpp:381-416) and schDttt_corKernel2 (:418-451),
 * declared in common/Defines.h:36-48, with the reference's argument order:
 *   schDttt_cor = s * sum_{p<ngpown} sum_{g<ncouls} aqsn(n1,g) conj(aqsm(n1,igp)) vcoul(igp)
 *                 * ((R-A)(ifreq,p,g) fact1 + (R-A)(ifreq+1,p,g) fact2)
 * with s = 1 (Kernel1) or -1/2 (Kernel2), R/A = I_eps{R,A}_array c128[..][ngpown][ncouls],
 * igp = indinv[inv_igp_index[p]], indinv[ncouls].  schDttt_cor is c128[1].  As in the
 * reference, the arrays must hold slices ifreq and ifreq + 1 and band row n1; ifreq < 0,
 * n1 < 0 or ngpown > ncouls return BGW_EINVAL.  The host forms transfer only those two
 * slices of each array and that row. */
int bgw_schDttt_corKernel1(bgw_ctx* ctx, double* schDttt_cor, const int* inv_igp_index,
                           const int* indinv, const double* I_epsR_array,
                           const double* I_epsA_array, const double* aqsmtemp,
                           const double* aqsntemp, const double* vcoul, int ncouls, int ifreq,
                           int ngpown, int n1, double fact1, double fact2);
int bgw_schDttt_corKernel2(bgw_ctx* ctx, double* schDttt_cor, const int* inv_igp_index,
                           const int* indinv, const double* I_epsR_array,
                           const double* I_epsA_array, const double* aqsmtemp,
                           const double* aqsntemp, const double* vcoul, int ncouls, int ifreq,
                           int ngpown, int n1, double fact1, double fact2);
/* Device-pointer forms of the two above (all arrays on the device, result after
 * bgw_synchronize); same reference functions, FF/ff.cpp:381-451. */
int bgw_schDttt_corKernel1_device(bgw_ctx* ctx, double* schDttt_cor, const int* inv_igp_index,
                                  const int* indinv, const double* I_epsR_array,
                                  const double* I_epsA_array, const double* aqsmtemp,
                                  const double* aqsntemp, const double* vcoul, int ncouls,
                                  int ifreq, int ngpown, int n1, double fact1, double fact2);
int bgw_schDttt_corKernel2_device(bgw_ctx* ctx, double* schDttt_cor, const int* inv_igp_index,
                                  const int* indinv, const double* I_epsR_array,
                                  const double* I_epsA_array, const double* aqsmtemp,
                                  const double* aqsntemp, const double* vcoul, int ncouls,
                                  int ifreq, int ngpown, int n1, double fact1, double fact2);

/* ---- seeded synthetic inputs (host only) ---------------------------------
 * SplitMix64 counter stream: x_i = mix64(seed + (i+1)*0x9E3779B97F4A7C15),
 * double = lo + (hi-lo) * (x>>11) * 2^-53, int in [0,n) = ((x>>32)*n)>>32,
 * permutation = Fisher-Yates driven by the same stream.  Replaces the
 * reference drivers' constant initialisation (GPP/gpp.cpp:246-276,
 * FF/ff.cpp:584-633) by the distributions of the benchmark configurations. */
void bgw_synth_uniform(unsigned long long seed, double lo, double hi, double* out, size_t n);
void bgw_synth_complex(unsigned long long seed, double re_lo, double re_hi,
                       double im_lo, double im_hi, double* out, size_t n);
void bgw_synth_index(unsigned long long seed, int n_max, int* out, size_t n);
void bgw_synth_permutation(unsigned long long seed, int* out, size_t n);

/* ---- input sets on disk (host only) --------------------------------------
 * No reference counterpart: the reference drivers build their inputs in memory
 * (GPP/gpp.cpp:220-276, FF/ff.cpp:552-633) and read no files; these calls store and load the same
 * arrays as a self-describing container, memory-mapped on read so the pointers
 * of a loaded set go to the solver entry points above without a copy.
 * Layout (little-endian):
 *   bytes [0,32)  "BGWINPUT", u32 version = 1, u32 kind, u32 count, u32 0,
 *                 u64 directory offset = 32
 *   directory     count entries of 96 bytes: char name[32] (1-31 printable
 *                 ASCII characters other than space, NUL-padded),
 *                 u32 dtype, u32 ndim (0..4, 0 = one-element scalar),
 *                 u64 dims[4] (unused = 1), u64 offset (multiple of 4096),
 *                 u64 nbytes, u64 checksum
 *   payload       each array at its offset, row-major, complex as (re,im) pairs
 *   checksum      sum_i mix64(w_i ^ (i * 0x9E3779B97F4A7C15)) mod 2^64 over the
 *                 payload's 8-byte little-endian words (zero-padded), mix64 the
 *                 SplitMix64 finalizer of bgw_synth_*.
 * Kinds and the entries bgw_inputs_open checks for them (names = the entry
 * points' argument names; int64 scalars number_bands, nvband, ncouls, ngpown and
 * f64 scalar occ in both):
 *   GPP: inv_igp_index i32[ngpown], indinv i32[ncouls+1], wx_array f64[3] or
 *        [number_bands,3], wtilde_array, I_eps_array c128[ngpown,ncouls],
 *        aqsmtemp, aqsntemp c128[number_bands,ncouls], vcoul f64[ncouls]
 *   FF:  int64 nFreq, nfreqeval; f64 freqevalmin, freqevalstep; inv_igp_index
 *        i32[ngpown], indinv i32[ncouls], ekq f64[number_bands], dFreqGrid
 *        f64[nFreq], aqsmtemp, aqsntemp c128[number_bands,ncouls], vcoul
 *        f64[ncouls], I_epsR_array, I_epsA_array c128[nFreq,ngpown,ncouls]
 *   OTHER: any entries, no check.
 * Errors: BGW_EINVAL for bad arguments, BGW_EIO for files (bgw_io_last_error,
 * per thread, has the text).  No reference counterpart. */
#define BGW_INPUTS_GPP    1
#define BGW_INPUTS_FF     2
#define BGW_INPUTS_OTHER  3
#define BGW_DT_F64   1
#define BGW_DT_C128  2
#define BGW_DT_I32   3
#define BGW_DT_I64   4

typedef struct bgw_array_desc {
  const char* name;
  int         dtype;
  int         ndim;
  size_t      dims[4];
  const void* data;
} bgw_array_desc;
typedef struct bgw_inputs bgw_inputs;

int         bgw_inputs_write(const char* path, int kind, const bgw_array_desc* arrays, int count);
/* Maps `path` read-only and checks its structure and, for GPP/FF, its entries;
 * verify != 0 also checks every checksum.  Descriptors returned by the getters
 * point into the mapping and stay valid until bgw_inputs_close. */
int         bgw_inputs_open(const char* path, int verify, bgw_inputs** out);
int         bgw_inputs_kind(const bgw_inputs* in);
int         bgw_inputs_count(const bgw_inputs* in);
int         bgw_inputs_entry(const bgw_inputs* in, int i, bgw_array_desc* desc);
int         bgw_inputs_get(const bgw_inputs* in, const char* name, bgw_array_desc* desc);
void        bgw_inputs_close(bgw_inputs* in);
const char* bgw_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* BGW_H */
