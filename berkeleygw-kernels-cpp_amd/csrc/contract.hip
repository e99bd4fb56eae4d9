// contract.hip -- FP64-MFMA "trace contraction" shared by the GPP static-COH term
// and the three FF kernels:
//
//   out = sum_job sum_{r < cnt[job]} sum_p  W(job,r,p) * sum_g A(rows[job][r], g) * B_job(p, g)
//   W(job,r,p) = wts[job][r] * vcoul(igp(p)) * conj(aqsm(rows[job][r], igp(p)))
//
// B_job is one [ngpown x ncouls] frequency slice of X (optionally minus the same
// slice of Y, or masked to |X| > to1).  This is the common form of
//   achsDtemp     FF/ff.cpp:191-208   (one job: I_epsR slice 1, w = -1/2)
//   asxDtemp      FF/ff.cpp:242-281   (jobs = slices of I_epsR / I_epsA, w = occ*fact1|fact2)
//   achDtemp_cor  FF/ff.cpp:332-372 + :381-451 (jobs = slices of R - A, w = fact1|fact2, *-1/2 in Kernel2)
//   achstemp      static COH term of the GPP kernel (one job: I_eps masked, w = -1/2)
// The inner sum_g A*B^T is a complex NT-GEMM tile on v_mfma_f64_16x16x4f64 (three
// real products per complex block, 3M scheme); the weights are applied in the epilogue and
// every workgroup writes one double-double partial, summed in fixed order.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "bgw_dd.h"
#include "bgw_fact.h"
#include "bgw_internal.h"
#include "contract.h"
#include "gpp_common.h"

namespace bgw {
namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// 16-byte load with the non-temporal hint (global_load_dwordx4 ... nt)
__device__ inline double2 ntload(const double2* p) {
  const f64x2 v = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(p));
  return make_double2(v.x, v.y);
}

// Workgroup tile: BM rows (bands) x BN columns (G'), 4 waves as WR x WC, each wave
// RB x CB 16x16 MFMA blocks.  Square form (dense row lists, and FP32 sparse plans of
// short band lists): 64 x 64 as 2 x 2 waves of 32 x 32, the fewest LDS operand reads
// per MFMA.  Tall form (other sparse plans): 80 x 64 with each wave owning one
// 16-column block of all 80 rows -- a frequency slice's band list (two neighbouring
// frequency groups, ~60-80 bands) then fits one tile and its B tile is streamed once.
// contract_row_tile picks the form.
template <bool TALL>
struct TileShape {
  static constexpr int WR = TALL ? 1 : 2, WC = TALL ? 4 : 2;
  static constexpr int RB = TALL ? 5 : 2, CB = TALL ? 1 : 2;
  static constexpr int BM = WR * RB * 16;
  static constexpr int LW = (BM + 31) / 32 * 32;  // FP32 LDS row width: its swizzle permutes within 32
};
constexpr int BN = 64;   // columns (G') per workgroup tile
constexpr int BK = 16;   // G-vectors per LDS stage (double-buffered)
constexpr int NT = 256;  // 4 waves
constexpr int kXcds = 8;  // MI355X: 8 XCDs, workgroups dealt round-robin by linear id

// F32 = false: FP64 MFMA throughout.
// F32 = true (precision form): operands rounded to FP32 in LDS, products on
// v_mfma_f32_16x16x4f32 with FP32 accumulation over one LDS stage (16 G-vectors),
// folded into the FP64 accumulators after every stage.
template <int MODE, bool F32, bool TALL>
__global__ __launch_bounds__(NT, 2) void k_contract(ContractArgs a) {
  typedef typename std::conditional<F32, float, double>::type T;
  typedef TileShape<TALL> TS;
  constexpr int BM = TS::BM, RB = TS::RB, CB = TS::CB;
  constexpr int AE = BM * BK / NT;  // A elements staged per thread and stage
  constexpr int BE = BN * BK / NT;  // B elements staged per thread and stage
  // [buffer][k][row ^ swz(k)] of (re, im) pairs: two stages in flight, one barrier per
  // stage; a stage element is one 16-byte (FP64) or 8-byte (FP32) write and every
  // operand pair one read.  The XOR swizzle keeps both the staging writes (16 k of a
  // row per 16 lanes) and the MFMA operand reads (16 rows of a k per 16 lanes) on
  // distinct LDS banks: FP64 pairs fill the 64 banks 16 at a time, so swz(k) = k
  // permutes within rows of 16; FP32 pairs take two 16-lane groups per pass, so
  // swz(k) = 16*(k&1) + k permutes within 32 and alternates the halves.
  typedef typename std::conditional<F32, float2, double2>::type T2;
  constexpr int LWP = F32 ? TS::LW : (BM + 15) / 16 * 16;  // LDS row width
  __shared__ __align__(16) T2 As[2][BK][LWP];
  __shared__ __align__(16) T2 Bs[2][BK][BN];

  // Work items are (G' tile ct, live row tile t, K split ks), ct fastest.  The live
  // row tiles of all jobs are numbered densely by tile_start (prefix of
  // ceil(cnt/BM)), so empty jobs and empty row-tile tails take no workgroup.  The
  // hardware deals workgroup L to XCD L % 8; the order of the items over the XCDs
  // decides what meets in each XCD's L2.
  //   Dense row lists (a.job_tiles row tiles per job, K splits a multiple of 8):
  //   XCD x takes the K splits ks = x (mod 8), its items run row tile fastest, then
  //   G' tile, then job.  The row tiles that share a B tile and the G' tiles that
  //   share an A row block then run side by side on one XCD, and each XCD reads only
  //   its K range of A and of the slices.
  //   Sparse plans (a.xcd_ks, K splits a multiple of 8): XCD x takes the K splits
  //   ks = x (mod 8), its items G' tile fastest, then row tile, so the G' tiles of a row
  //   tile (sharing its A rows) meet in one L2 and each XCD reads its K range only.
  //   Otherwise chunks of a.chunk consecutive items (the G' tiles of two row tiles) are
  //   dealt round-robin over the XCDs.
  const int ntn = (a.N + BN - 1) / BN;
  int item;
  if (a.job_tiles) {
    const int x = (int)(blockIdx.x % kXcds), q = (int)(blockIdx.x / kXcds), per_ks = a.tiles * ntn;
    const int ks_ = (q / per_ks) * kXcds + x, r = q % per_ks;
    const int rt_ = r % a.job_tiles, ct_ = (r / a.job_tiles) % ntn, job_ = r / a.job_tiles / ntn;
    item = ct_ + ntn * (job_ * a.job_tiles + rt_ + a.tiles * ks_);
  } else if (a.xcd_ks) {
    const int x = (int)(blockIdx.x % kXcds), q = (int)(blockIdx.x / kXcds), per_ks = a.tiles * ntn;
    item = q % per_ks + per_ks * ((q / per_ks) * kXcds + x);
  } else {
    const int slot = (int)(blockIdx.x / kXcds);
    item = ((slot / a.chunk) * kXcds + (int)(blockIdx.x % kXcds)) * a.chunk + slot % a.chunk;
  }
  if (item >= a.items) return;  // grid padding to whole chunks
  const size_t pid = (size_t)item;
  const int ct = item % ntn;
  const int t = (item / ntn) % a.tiles;
  const int ks = item / ntn / a.tiles;
  if (t >= a.tile_start[a.njobs]) {  // slot beyond the live tiles: a zero partial keeps the order fixed
    if (threadIdx.x == 0) cdd_store(&a.partial[pid], cdd_zero());
    return;
  }
  // job = the largest j with tile_start[j] <= t (jobs without tiles repeat a value)
  int lo = 0, hi = a.njobs;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.tile_start[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int job = lo;
  const int rt = t - a.tile_start[job];
  const int nrows = a.cnt[job];
  const int slice = a.slice_of_job[job];
  const double2* __restrict__ X = (a.Xb && job >= a.jobs_a ? a.Xb : a.Xa) + (size_t)slice * a.slice_stride;
  const double2* __restrict__ Y = a.Y ? a.Y + (size_t)slice * a.slice_stride : nullptr;
  const int* __restrict__ rows = a.rows + (size_t)job * a.max_rows;
  const int K = a.K;
  const int kper = ((K + a.ksplit - 1) / a.ksplit + BK - 1) / BK * BK;
  const int k0 = ks * kper, k1 = min(K, k0 + kper);

  // the wave index is wave-uniform: held in a scalar register, the per-wave block
  // liveness below branches on scalars instead of masking lanes
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wr = (wave / TS::WC) * (RB * 16), wc = (wave % TS::WC) * (CB * 16);

  // complex blocks by the 3M scheme: e = Ar*Br, f = Ai*Bi, h = (Ar+Ai)*(Br+Bi), so
  // re = e - f and im = h - e - f take three real MFMAs instead of four.  FP64
  // form: the MFMAs accumulate e, f, h in de, df, dh.  FP32 form: the FP32 stage sums
  // are folded as re into de and im into df, so dh is not needed (40 VGPRs fewer
  // in the tall tile)
  f64x4 de[RB][CB], df[RB][CB], dh[F32 ? 1 : RB][CB];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < CB; ++j) {
      de[i][j] = (f64x4){0.0, 0.0, 0.0, 0.0};
      df[i][j] = (f64x4){0.0, 0.0, 0.0, 0.0};
      if constexpr (!F32) dh[i][j] = (f64x4){0.0, 0.0, 0.0, 0.0};
    }

  // each thread stages AE elements of A and BE of B per stage: idx = tid + NT*i -> (row, k)
  int arow[AE];
  bool aok[AE];
  int bcol[BE];  // 32-bit: the 64-bit row offset is formed at fetch time (fewer live VGPRs)
  bool bok[BE];
#pragma unroll
  for (int i = 0; i < AE; ++i) {
    const int r = (tid + NT * i) / BK;
    const int gr = rt * BM + r;
    aok[i] = gr < nrows;
    arow[i] = aok[i] ? rows[gr] : 0;
  }
#pragma unroll
  for (int i = 0; i < BE; ++i) {
    const int gc = ct * BN + (tid + NT * i) / BK;
    bok[i] = gc < a.N;
    bcol[i] = bok[i] ? gc : 0;
  }

  // 16x16 blocks of this wave that hold at least one live row / column.  The others
  // (row-tile tails of short band lists) skip their MFMAs, except in the FP64 square
  // tile, which only dense lists use: there every block runs (a dead block multiplies
  // the zeros staged for out-of-range elements), which keeps the MFMA loop free of
  // branches (measured: achsDtemp 5% faster; achDtemp_cor, a sparse plan, 7% slower
  // without the skip)
  bool live_r[RB], live_c[CB];
#pragma unroll
  for (int i = 0; i < RB; ++i) live_r[i] = rt * BM + wr + i * 16 < nrows;
#pragma unroll
  for (int j = 0; j < CB; ++j) live_c[j] = ct * BN + wc + j * 16 < a.N;

  // software pipeline: the global loads of stage kb+BK are in flight while the
  // MFMAs of stage kb run.  The loads are unconditional, from clamped in-range
  // addresses, and land in registers as loaded; the R - A difference, the mask
  // and the zeroing of out-of-range elements are applied when the stage is
  // stored to LDS, so no wave waits on its loads before it has to.
  double2 av[AE], bx[BE], by[MODE == kModeDiff ? BE : 1];
  auto fetch = [&](int kb) {
#pragma unroll
    for (int i = 0; i < AE; ++i) {
      const int k = min(kb + (tid + NT * i) % BK, K - 1);
      av[i] = a.A[(size_t)arow[i] * K + k];
    }
#pragma unroll
    for (int i = 0; i < BE; ++i) {
      const int k = min(kb + (tid + NT * i) % BK, K - 1);
      const size_t off = (size_t)bcol[i] * K + k;
      // the tall tile reads each slice once: stream it past L2 (non-temporal), which
      // leaves the aqsntemp rows that all G' tiles of a row tile share in L2 longer
      // (achDtemp_cor 2-3% faster); square tiles of dense lists re-read B across
      // row tiles
      if constexpr (TALL) {
        bx[i] = ntload(X + off);
        if constexpr (MODE == kModeDiff) by[i] = ntload(Y + off);
      } else {
        bx[i] = X[off];
        if constexpr (MODE == kModeDiff) by[i] = Y[off];
      }
    }
  };
  auto swz = [](int k) { return F32 ? 16 * (k & 1) + k : k; };
  auto stage_store = [&](int buf, int kb) {
#pragma unroll
    for (int i = 0; i < AE; ++i) {
      const int idx = tid + NT * i, k = idx % BK, col = (idx / BK) ^ swz(k);
      const bool ok = aok[i] && kb + k < k1;
      T2 w;
      w.x = ok ? (T)av[i].x : (T)0;
      w.y = ok ? (T)av[i].y : (T)0;
      As[buf][k][col] = w;
    }
#pragma unroll
    for (int i = 0; i < BE; ++i) {
      const int idx = tid + NT * i, k = idx % BK, col = (idx / BK) ^ swz(k);
      double2 v = bx[i];
      if constexpr (MODE == kModeDiff) {
        v.x -= by[i].x;
        v.y -= by[i].y;
      } else if constexpr (MODE == kModeMask) {
        if (!(v.x * v.x + v.y * v.y > a.thr2)) v = make_double2(0.0, 0.0);
      }
      const bool ok = bok[i] && kb + k < k1;
      T2 w;
      w.x = ok ? (T)v.x : (T)0;
      w.y = ok ? (T)v.y : (T)0;
      Bs[buf][k][col] = w;
    }
  };
  if (k0 < k1) {
    fetch(k0);
    stage_store(0, k0);
  }
  __syncthreads();
  int buf = 0;
  for (int kb = k0; kb < k1; kb += BK) {
    // stage kb+BK is loaded into registers while stage kb is multiplied, then
    // stored to the other buffer; the barrier at the end of the iteration both
    // publishes it and retires the reads of this one
    const bool more = kb + BK < k1;
    if (more) fetch(kb + BK);
    if constexpr (!F32) {
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        const int kr = kk + (lane >> 4), cl = lane & 15, sk = swz(kr);
        double ar[RB], ai[RB], sa[RB], br[CB], bi[CB], sb[CB];
#pragma unroll
        for (int i = 0; i < RB; ++i) {
          const T2 av2 = As[buf][kr][(wr + i * 16 + cl) ^ sk];
          ar[i] = av2.x;
          ai[i] = av2.y;
          sa[i] = ar[i] + ai[i];
        }
#pragma unroll
        for (int j = 0; j < CB; ++j) {
          const T2 bv2 = Bs[buf][kr][(wc + j * 16 + cl) ^ sk];
          br[j] = bv2.x;
          bi[j] = bv2.y;
          sb[j] = br[j] + bi[j];
        }
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
          for (int j = 0; j < CB; ++j) {
            if (TALL && (!live_r[i] || !live_c[j])) continue;  // FP64 sparse plans: tall tiles only
            de[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], de[i][j], 0, 0, 0);
            df[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], bi[j], df[i][j], 0, 0, 0);
            dh[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(sa[i], sb[j], dh[i][j], 0, 0, 0);
          }
      }
    } else {
      f32x4 se[RB][CB], sf[RB][CB], sh[RB][CB];
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
          se[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
          sf[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
          sh[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
        }
      // the tall FP32 tile fits 256 VGPRs only with the operand reads of at most two
      // k-steps in flight (fully unrolled, the compiler hoists all four and spills)
      constexpr int kSteps = TALL ? 2 : 4;
#pragma unroll kSteps
      for (int kk = 0; kk < BK; kk += 4) {
        const int kr = kk + (lane >> 4), cl = lane & 15, sk = swz(kr);
        float ar[RB], ai[RB], sa[RB], br[CB], bi[CB], sb[CB];
#pragma unroll
        for (int i = 0; i < RB; ++i) {
          const T2 av2 = As[buf][kr][(wr + i * 16 + cl) ^ sk];
          ar[i] = av2.x;
          ai[i] = av2.y;
          sa[i] = ar[i] + ai[i];
        }
#pragma unroll
        for (int j = 0; j < CB; ++j) {
          const T2 bv2 = Bs[buf][kr][(wc + j * 16 + cl) ^ sk];
          br[j] = bv2.x;
          bi[j] = bv2.y;
          sb[j] = br[j] + bi[j];
        }
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
          for (int j = 0; j < CB; ++j) {
            if (!live_r[i] || !live_c[j]) continue;
            se[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[i], br[j], se[i][j], 0, 0, 0);
            sf[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[i], bi[j], sf[i][j], 0, 0, 0);
            sh[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(sa[i], sb[j], sh[i][j], 0, 0, 0);
          }
      }
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const double e = se[i][j][v], f = sf[i][j][v];
            de[i][j][v] += e - f;
            df[i][j][v] += ((double)sh[i][j][v] - e) - f;
          }
    }
    if (more) stage_store(buf ^ 1, kb + BK);
    __syncthreads();
    buf ^= 1;
  }

  // epilogue.  Output layouts on gfx950 (measured): v_mfma_f64_16x16x4f64 leaves
  // D[row = lane/16 + 4v][col = lane%16] in element v of lane `lane`;
  // v_mfma_f32_16x16x4f32 leaves D[row = 4(lane/16) + v][col = lane%16].  The FP32
  // path folds its accumulators element-wise, so it keeps the FP32 layout.
  cdd acc = cdd_zero();
#pragma unroll
  for (int j = 0; j < CB; ++j) {
    const int p = ct * BN + wc + j * 16 + (lane & 15);
    if (p >= a.N) continue;
    const int igp = a.gpp_index ? gpp_igp(a.inv, a.indinv, p, a.ncouls) : ff_igp(a.inv, a.indinv, p, a.ncouls);
    const double vc = a.vcoul[igp];
#pragma unroll
    for (int i = 0; i < RB; ++i) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = rt * BM + wr + i * 16 + (F32 ? 4 * (lane >> 4) + v : (lane >> 4) + 4 * v);
        if (r >= nrows) continue;
        const int n1 = rows[r];
        const double2 am = a.aqsm[(size_t)n1 * a.ncouls + igp];
        const double2 w = cmul(a.wts[(size_t)job * a.max_rows + r], make_double2(vc * am.x, -vc * am.y));
        double2 s;
        if constexpr (F32) {
          s = make_double2(de[i][j][v], df[i][j][v]);
        } else {
          const double e = de[i][j][v], f = df[i][j][v];
          s = make_double2(e - f, dh[i][j][v] - e - f);
        }
        const double2 t = cmul(w, s);
        acc = cdd_add(acc, t.x, t.y);
      }
    }
  }
  // the stage buffers are free after the loop's last barrier: the reduction
  // scratch lives in them (two FP64 80-row workgroups take 144 of the CU's 160 KB of LDS)
  cdd tot = cdd_block_reduce(acc, reinterpret_cast<cdd*>(&As[0][0][0]));
  if (tid == 0) cdd_store(&a.partial[pid], tot);
}

// tile_start[j] = sum_{i<j} ceil(cnt[i]/bm), j <= njobs: each thread sums a run
// of consecutive jobs, then one LDS scan over the runs.
constexpr int kPrefixThreads = 1024;
__global__ __launch_bounds__(kPrefixThreads) void k_tile_prefix(const int* __restrict__ cnt, int njobs, int bm,
                                                               int* __restrict__ tile_start) {
  __shared__ int part[kPrefixThreads];
  const int per = (njobs + kPrefixThreads - 1) / kPrefixThreads;
  const int j0 = min(njobs, (int)threadIdx.x * per), j1 = min(njobs, j0 + per);
  int s = 0;
  for (int j = j0; j < j1; ++j) s += (cnt[j] + bm - 1) / bm;
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < kPrefixThreads; off <<= 1) {
    const int v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int run = part[threadIdx.x] - s;  // exclusive prefix of this thread's run
  for (int j = j0; j < j1; ++j) {
    tile_start[j] = run;
    run += (cnt[j] + bm - 1) / bm;
  }
  if (threadIdx.x == kPrefixThreads - 1) tile_start[njobs] = part[threadIdx.x];
}

// Workgroups of `kern` resident on the whole device at once.
int resident_slots(bgw_ctx* ctx, const void* kern) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, NT, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  return per_cu * ctx->cus;
}

// K split minimising the estimated makespan: whole rounds of `slots` resident
// workgroups, each item costing its bk-wide stages plus ~2 stages of epilogue.
int pick_ksplit(int K, int BK, size_t tiles_x_ntn, int slots, int step) {
  int best = step;
  double best_cost = 0.0;
  const int max_ks = std::max(step, std::min(64, K / BK));
  for (int ks = step; ks <= max_ks; ks += step) {
    const double items = (double)tiles_x_ntn * ks;
    const double rounds = std::ceil(items / slots);
    const double stages = std::ceil(std::ceil((double)K / ks) / BK);
    const double cost = rounds * (stages + 2.0);
    if (ks == step || cost < best_cost * 0.999) {
      best = ks;
      best_cost = cost;
    }
  }
  return best;
}


// Streaming form for sparse plans whose jobs hold at most kStreamRows rows
// (asxDtemp's valence bands, the one-band schDttt terms).  There each slice element
// meets at most a few rows, about 1 FLOP per byte, so the work is reading the slices
// once: workgroup (job, pair of G' rows) streams the pair's rows of X (minus Y) with
// unrolled coalesced loads, forms sum_g A[n1][g] x[g] per row in FP64 (the few A rows
// stay in L2) and weights it with the row's w conj(aqsm[n1][igp]) vcoul[igp], as the
// MFMA epilogue does.  One double-double partial per workgroup, zero for empty jobs.
constexpr int kStreamNT = 256, kStreamUnroll = 4, kStreamPPer = 2;

//
// Workgroup (live job li, pair pb): the live jobs (cnt > 0) are numbered densely by
// a.tile_start (the prefix of [cnt > 0]), so a plan of many frequency slices with a few
// live ones launches workgroups for the live ones only; slots beyond the live count write
// a zero partial, which keeps the reduction's order fixed.
__global__ __launch_bounds__(kStreamNT) void k_contract_stream(ContractArgs a, int pblocks) {
  const int li = blockIdx.x / pblocks, pb = blockIdx.x - li * pblocks;
  const int tid = threadIdx.x;
  __shared__ cdd red[kStreamNT / 64];
  int job = 0, nrows = 0;
  if (li < a.tile_start[a.njobs]) {
    // job = the largest j with tile_start[j] <= li (jobs without rows repeat a value)
    int lo = 0, hi = a.njobs;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.tile_start[mid] <= li)
        lo = mid;
      else
        hi = mid - 1;
    }
    job = lo;
    nrows = a.cnt[job];
  }
  cdd acc = cdd_zero();
  if (nrows > 0) {
    const int slice = a.slice_of_job[job];
    const int* rows = a.rows + (size_t)job * a.max_rows;
    int n1[kStreamRows];
#pragma unroll
    for (int r = 0; r < kStreamRows; ++r) n1[r] = rows[r < nrows ? r : 0];
    // the workgroup's G' rows p0 .. p0 + kStreamPPer - 1 in one pass, so each A value
    // loaded serves all of them; rows beyond N read row p0 and are left out below
    const int p0 = pb * kStreamPPer;
    const double2* __restrict__ X[kStreamPPer];
    const double2* __restrict__ Y[kStreamPPer];
    const double2* Xs = a.Xb && job >= a.jobs_a ? a.Xb : a.Xa;
#pragma unroll
    for (int q = 0; q < kStreamPPer; ++q) {
      const int p = p0 + q < a.N ? p0 + q : p0;
      const size_t off = (size_t)slice * a.slice_stride + (size_t)p * a.K;
      X[q] = Xs + off;
      Y[q] = a.Y ? a.Y + off : nullptr;
    }
    double sr[kStreamPPer][kStreamRows], si[kStreamPPer][kStreamRows];
#pragma unroll
    for (int q = 0; q < kStreamPPer; ++q)
#pragma unroll
      for (int r = 0; r < kStreamRows; ++r) sr[q][r] = si[q][r] = 0.0;
    for (int g0 = tid; g0 < a.K; g0 += kStreamNT * kStreamUnroll) {
      double2 x[kStreamPPer][kStreamUnroll];
#pragma unroll
      for (int q = 0; q < kStreamPPer; ++q)
#pragma unroll
        for (int u = 0; u < kStreamUnroll; ++u) {
          const int g = g0 + u * kStreamNT;
          x[q][u] = g < a.K ? X[q][g] : make_double2(0.0, 0.0);
        }
      if (a.Y) {
#pragma unroll
        for (int q = 0; q < kStreamPPer; ++q)
#pragma unroll
          for (int u = 0; u < kStreamUnroll; ++u) {
            const int g = g0 + u * kStreamNT;
            const double2 y = g < a.K ? Y[q][g] : make_double2(0.0, 0.0);
            x[q][u].x -= y.x;
            x[q][u].y -= y.y;
          }
      }
#pragma unroll
      for (int r = 0; r < kStreamRows; ++r) {
        if (r >= nrows) break;
        const double2* __restrict__ Ar = a.A + (size_t)n1[r] * a.K;
#pragma unroll
        for (int u = 0; u < kStreamUnroll; ++u) {
          const int g = g0 + u * kStreamNT;
          const double2 av = g < a.K ? Ar[g] : make_double2(0.0, 0.0);
#pragma unroll
          for (int q = 0; q < kStreamPPer; ++q) {
            sr[q][r] = fma(av.x, x[q][u].x, fma(-av.y, x[q][u].y, sr[q][r]));
            si[q][r] = fma(av.x, x[q][u].y, fma(av.y, x[q][u].x, si[q][r]));
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kStreamPPer; ++q) {
      const int p = p0 + q;
      if (p >= a.N) break;
      const int igp = ff_igp(a.inv, a.indinv, p, a.ncouls);
      const double vc = a.vcoul[igp];
#pragma unroll
      for (int r = 0; r < kStreamRows; ++r) {
        if (r >= nrows) break;
        const double2 am = a.aqsm[(size_t)n1[r] * a.ncouls + igp];
        const double2 w = cmul(a.wts[(size_t)job * a.max_rows + r], make_double2(vc * am.x, -vc * am.y));
        const double2 t = cmul(w, make_double2(sr[q][r], si[q][r]));
        acc = cdd_add(acc, t.x, t.y);
      }
    }
  }
  const cdd tot = cdd_block_reduce(acc, red);
  if (tid == 0) cdd_store(&a.partial[blockIdx.x], tot);
}

}  // namespace

int contract_launch(bgw_ctx* ctx, ContractArgs a, int njobs, int bm, int max_tiles, bool exact_tiles,
                    int est_tiles, const char* tname) {
  const int ntn = (a.N + BN - 1) / BN;
  const bool f32 = ctx->precision == BGW_FP32_BULK;
  const int mode = a.Y ? kModeDiff : (a.thr2 > 0.0 ? kModeMask : kModePlain);
  if (bm != TileShape<true>::BM && bm != TileShape<false>::BM) return fail(ctx, BGW_EINVAL, "unsupported row tile");
  const bool tall = bm == TileShape<true>::BM;
  void (*kern)(ContractArgs) = nullptr;
  if (mode == kModeDiff)
    kern = f32 ? (tall ? k_contract<kModeDiff, true, true> : k_contract<kModeDiff, true, false>)
               : (tall ? k_contract<kModeDiff, false, true> : k_contract<kModeDiff, false, false>);
  else if (mode == kModeMask)
    kern = f32 ? (tall ? k_contract<kModeMask, true, true> : k_contract<kModeMask, true, false>)
               : (tall ? k_contract<kModeMask, false, true> : k_contract<kModeMask, false, false>);
  else
    kern = f32 ? (tall ? k_contract<kModePlain, true, true> : k_contract<kModePlain, true, false>)
               : (tall ? k_contract<kModePlain, false, true> : k_contract<kModePlain, false, false>);
  // Row counts per job are only known on the device (plan kernel), so the grid
  // covers max_tiles row tiles, a host bound on sum_j ceil(cnt[j]/bm); slots
  // beyond the live tiles exit after a zero partial.  When max_tiles is exact the
  // K split is chosen for whole rounds of resident workgroups.  Otherwise each
  // workgroup takes at most ~1024 G-vectors, and plans with few live tiles
  // (est_tiles, a host bound) are split further so that about four workgroups per
  // resident slot keep enough loads in flight.
  if (max_tiles < 1) max_tiles = 1;
  const int bk = BK;
  const int slots = resident_slots(ctx, (const void*)kern);
  int ksplit;
  a.job_tiles = 0;
  a.xcd_ks = 0;
  if (exact_tiles) {
    // dense lists: K splits in multiples of the XCD count when K allows (one split
    // per XCD in turn, see k_contract's work order)
    const bool per_xcd = a.K / bk >= kXcds && max_tiles % njobs == 0;
    ksplit = pick_ksplit(a.K, bk, (size_t)ntn * max_tiles, slots, per_xcd ? kXcds : 1);
    if (per_xcd) a.job_tiles = max_tiles / njobs;
  } else {
    const long long est = std::max(1, std::min(max_tiles, est_tiles));
    const long long want = (4LL * slots + (long long)ntn * est - 1) / ((long long)ntn * est);
    ksplit = (int)std::max<long long>(a.K / 1024, want);
    ksplit = std::max(1, std::min({ksplit, 64, std::max(1, a.K / bk)}));
    // one K split per XCD in turn when K allows (achDtemp_cor at config 4: 4% faster,
    // 6.39e9 B fetched instead of 6.51e9)
    if (a.K / bk >= kXcds) {
      ksplit = std::min(64, (ksplit + kXcds - 1) / kXcds * kXcds);
      a.xcd_ks = 1;
    }
  }
  a.ksplit = ksplit;
  a.tiles = max_tiles;
  a.njobs = njobs;
  const size_t W = (size_t)ntn * ksplit * max_tiles;
  // the chunked order (K below 8 stages): one row tile's G' tiles per chunk for dense
  // lists (measured 6-8% faster for achsDtemp than two), two row tiles (one or two
  // frequency slices) for sparse plans.  Under the per-XCD orders the grid is rounded
  // the same way; blocks past the items return at once.
  a.chunk = (exact_tiles ? 1 : 2) * ntn;
  const size_t per_round = (size_t)kXcds * a.chunk;
  const size_t G = (W + per_round - 1) / per_round * per_round;
  if (G > (size_t)INT32_MAX) return fail(ctx, BGW_EINVAL, "contraction grid too large");
  a.items = (int)W;
  dim3 grid((unsigned)G);
  void* pts;
  int rc = pool_get(ctx, "ws_tile_start", (size_t)(njobs + 1) * sizeof(int), &pts);
  if (rc) return rc;
  a.tile_start = static_cast<const int*>(pts);
  hipLaunchKernelGGL(k_tile_prefix, dim3(1), dim3(kPrefixThreads), 0, ctx->stream, a.cnt, njobs, bm,
                     static_cast<int*>(pts));
  if ((rc = check_launch(ctx, "k_tile_prefix"))) return rc;
  void* pp;
  rc = pool_get(ctx, "ws_contract", W * sizeof(double4), &pp);
  if (rc) return rc;
  a.partial = static_cast<double4*>(pp);
  {
    KernelTimer kt(ctx, tname);
    hipLaunchKernelGGL(kern, grid, dim3(NT), 0, ctx->stream, a);
  }
  if ((rc = check_launch(ctx, "k_contract"))) return rc;
  launch_reduce_cdd(static_cast<const double4*>(pp), (int)W, 1, a.out, ctx->stream);
  return check_launch(ctx, "reduce");
}

int contract_stream(bgw_ctx* ctx, ContractArgs a, int njobs, int max_job_rows, int live_bound, const char* tname) {
  if (max_job_rows > kStreamRows || a.thr2 > 0.0)
    return fail(ctx, BGW_EINVAL, "streaming contraction: a job has more rows than it handles");
  const int pblocks = (a.N + kStreamPPer - 1) / kStreamPPer;
  const int live = std::max(0, std::min(njobs, live_bound));
  const long long W = (long long)live * pblocks;
  if (W < 1) {
    fill_zero(a.out, sizeof(double2), ctx->stream);
    return check_launch(ctx, "fill");
  }
  if (W > INT32_MAX) return fail(ctx, BGW_EINVAL, "contraction grid too large");
  void *pts, *pp;
  int rc = pool_get(ctx, "ws_tile_start", (size_t)(njobs + 1) * sizeof(int), &pts);
  if (!rc) rc = pool_get(ctx, "ws_contract", (size_t)W * sizeof(double4), &pp);
  if (rc) return rc;
  // live-job numbering: the tile prefix with one tile per job that has rows
  hipLaunchKernelGGL(k_tile_prefix, dim3(1), dim3(kPrefixThreads), 0, ctx->stream, a.cnt, njobs, 1 << 30,
                     static_cast<int*>(pts));
  if ((rc = check_launch(ctx, "k_tile_prefix"))) return rc;
  a.tile_start = static_cast<const int*>(pts);
  a.njobs = njobs;
  a.partial = static_cast<double4*>(pp);
  {
    KernelTimer kt(ctx, tname);
    hipLaunchKernelGGL(k_contract_stream, dim3((unsigned)W), dim3(kStreamNT), 0, ctx->stream, a, pblocks);
  }
  if ((rc = check_launch(ctx, "k_contract_stream"))) return rc;
  launch_reduce_cdd(static_cast<const double4*>(pp), (int)W, 1, a.out, ctx->stream);
  return check_launch(ctx, "reduce");
}

int contract_achstemp(bgw_ctx* ctx, int nb, int ngpown, int ncouls, const int* inv, const int* indinv,
                      const double2* aqsn, const double2* aqsm, const double2* Ieps, const double* vcoul,
                      double2* achstemp) {
  void *prow, *pw, *pcnt, *pslice;
  int rc = pool_get(ctx, "ws_rows_gpp", (size_t)nb * sizeof(int), &prow);
  if (!rc) rc = pool_get(ctx, "ws_wts_gpp", (size_t)nb * 16, &pw);
  if (!rc) rc = pool_get(ctx, "ws_cnt_gpp", 64, &pcnt);
  if (!rc) rc = pool_get(ctx, "ws_slice_gpp", 64, &pslice);
  if (rc) return rc;
  launch_identity_rows(static_cast<int*>(prow), static_cast<double2*>(pw), static_cast<int*>(pcnt),
                       static_cast<int*>(pslice), nb, make_double2(-0.5, 0.0), 0, ctx->stream);
  ContractArgs a{};
  a.A = aqsn;
  a.K = ncouls;
  a.Xa = Ieps;
  a.slice_stride = 0;
  a.jobs_a = 1;
  a.N = ngpown;
  a.slice_of_job = static_cast<const int*>(pslice);
  a.rows = static_cast<const int*>(prow);
  a.wts = static_cast<const double2*>(pw);
  a.cnt = static_cast<const int*>(pcnt);
  a.max_rows = nb;
  a.inv = inv;
  a.indinv = indinv;
  a.gpp_index = 1;
  a.aqsm = aqsm;
  a.vcoul = vcoul;
  a.ncouls = ncouls;
  a.thr2 = kTo1 * kTo1;  // |schs| > to1 cut of the static COH term
  a.out = achstemp;
  const int bm = contract_row_tile(ctx->precision == BGW_FP32_BULK, false, nb);
  const int tiles = (nb + bm - 1) / bm;
  return contract_launch(ctx, a, 1, bm, tiles, true, tiles, "contract_achs");
}

int contract_row_tile(bool f32, bool sparse, int max_job_rows) {
  // measured at config 4: dense achsDtemp 10% slower in 80-row tiles; asxDtemp (2 rows
  // per job) 10% faster tall in FP64 and 11% faster square in FP32; achDtemp_cor (up to
  // 79 rows per slice) 13% faster tall in FP32
  const bool tall = sparse && (!f32 || max_job_rows > TileShape<false>::BM);
  return tall ? TileShape<true>::BM : TileShape<false>::BM;
}

}  // namespace bgw
