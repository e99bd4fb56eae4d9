// ceilings.hip -- measured ceilings of the instruction forms and memory streams the
// GPP and FF kernels are built from.  Not part of the library: a measurement tool.
//
//   valu   FP64/FP32 FMA and MUL, v_rcp_f64/f32: 8 independent dependency chains per
//          lane (register-only work), 256-thread blocks, 8 blocks per CU; printed in
//          wave instructions per CU per ns and, for the arithmetic forms, TFLOP/s.
//   fixed  operands pinned to fixed registers with inline asm (VALU and MFMA only, no
//          memory access): FP32 FMAs whose source VGPRs share a register bank or include
//          an SGPR, and VALU FMAs interleaved with MFMAs of the same precision.
//   mfma   v_mfma_f64_4x4x4f64 and v_mfma_f64_16x16x4f64 streams: 1..8 waves per SIMD
//          x 4, 8 or 16 independent accumulators per wave, up to saturation.
//   hbm    read stream of 4 GiB (16-byte loads, plain and non-temporal, several grid
//          sizes): the attainable HBM read bandwidth.
//   dual   FP32 VALU dual issue: independent vs dependent streams, transcendental and
//          FP64 instructions mixed in, at 1..8 waves per SIMD.
//   mfmalds v_mfma_f64_16x16x4f64 fed by LDS operand reads in the FF contraction's
//          pattern (re+im sums from VALU or from a third plane, with or without a
//          barrier per stage) at 1 and 2 waves per SIMD.
//   copy   one 1 GiB device-to-device copy by a plain kernel (16-byte loads and
//          stores): the reference dispatch for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE.
//
//   make -C berkeleygw-kernels-cpp_amd ceilings
//   build/ceilings [valu|fixed|mfma|dual|hbm|copy ...]      (no argument: all but copy)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

#define HIP_OK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) {                                                                \
      std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));                 \
      std::exit(1);                                                                        \
    }                                                                                      \
  } while (0)

enum Op { kFmaF64, kMulF64, kRcpF64, kFmaF32Vgpr3, kFmaF32Vgpr2, kMulF32, kRcpF32 };

// `in` holds zeros; values read from it are runtime constants the compiler
// cannot fold, so every chain stays a register dependency chain.
template <Op OP>
__global__ __launch_bounds__(256) void k_valu(const double* __restrict__ in, double* __restrict__ out, int iters) {
  if (OP <= kRcpF64) {
    double x[8], y[8], z[8];
    for (int c = 0; c < 8; ++c) {
      x[c] = 1.0 + threadIdx.x * 1e-6 + c;
      y[c] = in[c] + 0.999999;
      z[c] = in[8 + c] + 1e-7;
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        if (OP == kFmaF64) x[c] = fma(y[c], z[c], x[c]);
        if (OP == kMulF64) x[c] = x[c] * y[c];
        if (OP == kRcpF64) x[c] = __builtin_amdgcn_rcp(x[c]);
      }
    }
    double s = 0.0;
    for (int c = 0; c < 8; ++c) s += x[c];
    out[blockIdx.x * 256 + threadIdx.x] = s;
  } else {
    float x[8], y[8], z[8];
    for (int c = 0; c < 8; ++c) {
      x[c] = 1.0f + threadIdx.x * 1e-6f + c;
      y[c] = (float)in[c] + 0.999999f;
      z[c] = (float)in[8 + c] + 1e-7f;
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        if (OP == kFmaF32Vgpr3) x[c] = fmaf(y[c], z[c], x[c]);  // three VGPR operands
        if (OP == kFmaF32Vgpr2) x[c] = fmaf(x[c], 0.999999f, z[c]);  // two VGPR operands + literal
        if (OP == kMulF32) x[c] = x[c] * y[c];
        if (OP == kRcpF32) x[c] = __builtin_amdgcn_rcpf(x[c]);
      }
    }
    float s = 0.f;
    for (int c = 0; c < 8; ++c) s += x[c];
    out[blockIdx.x * 256 + threadIdx.x] = s;
  }
}

template <int NACC>
__global__ __launch_bounds__(256) void k_mfma_f64(const double* __restrict__ in, double* __restrict__ out, int iters) {
  const double a = 1.0 + threadIdx.x * 1e-9 + in[0], b = 0.999999 + in[1];
  f64x4 acc[NACC];
#pragma unroll
  for (int j = 0; j < NACC; ++j) acc[j] = (f64x4){0.0, 0.0, 0.0, 0.0};
  // 16 rounds per loop iteration: the compiler moves the accumulators between AGPRs and
  // VGPRs at the loop's back edge (2 x 8 NACC v_accvgpr moves), which a short body would
  // measure instead of the MFMAs
  for (int i = 0; i < iters; i += 16) {
#pragma unroll
    for (int u = 0; u < 16; ++u)
#pragma unroll
      for (int j = 0; j < NACC; ++j) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < NACC; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

// v_mfma_f64_4x4x4f64 (four 4x4x4 blocks per wave, 512 FLOP) stream
template <int NACC>
__global__ __launch_bounds__(256) void k_mfma4_f64(const double* __restrict__ in, double* __restrict__ out, int iters) {
  const double a = 1.0 + threadIdx.x * 1e-9 + in[0], b = 0.999999 + in[1];
  double acc[NACC];
#pragma unroll
  for (int j = 0; j < NACC; ++j) acc[j] = 0.0;
  for (int i = 0; i < iters; i += 16) {  // 16 rounds per iteration, as in k_mfma_f64
#pragma unroll
    for (int u = 0; u < 16; ++u)
#pragma unroll
      for (int j = 0; j < NACC; ++j) acc[j] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, acc[j], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < NACC; ++j) s += acc[j];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

// MFMA fed from LDS in the FF contraction's pattern: a wave owns one 16-column block of
// RB 16-row blocks and per k-step reads its operands from LDS (8-byte reads, the
// contraction's XOR-swizzled layout: A[k][row ^ swz(k)], B[k][col ^ swz(k)]) and issues
// the 3M scheme's three v_mfma_f64_16x16x4f64 per row block; 4 k-steps per 16-G stage.
// SUMS: re and im planes only, the sums re + im formed by VALU adds as the contraction
// does (else read from a third plane).  BARRIER: a workgroup barrier after every stage.
// Waves per SIMD are set by the LDS each workgroup reserves (registers capped for two
// workgroups per CU, as in the contraction).  The LDS holds constants:
// the operands never change, only the issue pattern is measured.
template <int RB, bool SUMS, bool BARRIER>
__global__ __launch_bounds__(256, 2) void k_mfma_lds(const double* __restrict__ in, double* __restrict__ out,
                                                  int stages) {
  constexpr int LW = (RB * 16 + 31) / 32 * 32, BK = 16, BN = 64;
  extern __shared__ double lds[];
  double* As = lds;                 // [3][BK][LW]
  double* Bs = lds + 3 * BK * LW;   // [3][BK][BN]
  for (int i = threadIdx.x; i < 3 * BK * (LW + BN); i += 256) lds[i] = in[i & 15] + 1e-3 * (i & 7);
  __syncthreads();
  const int lane = threadIdx.x & 63, wc = (threadIdx.x >> 6) * 16, cl = lane & 15;
  auto swz = [](int k) { return 16 * (k & 1) + k; };
  f64x4 de[RB], df[RB], dh[RB];
#pragma unroll
  for (int i = 0; i < RB; ++i) de[i] = df[i] = dh[i] = (f64x4){0.0, 0.0, 0.0, 0.0};
  for (int s = 0; s < stages; ++s) {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4), sk = swz(kr);
      double ar[RB], ai[RB], sa[RB];
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        ar[i] = As[(0 * BK + kr) * LW + ((i * 16 + cl) ^ sk)];
        ai[i] = As[(1 * BK + kr) * LW + ((i * 16 + cl) ^ sk)];
        sa[i] = SUMS ? ar[i] + ai[i] : As[(2 * BK + kr) * LW + ((i * 16 + cl) ^ sk)];
      }
      const double br = Bs[(0 * BK + kr) * BN + ((wc + cl) ^ sk)];
      const double bi = Bs[(1 * BK + kr) * BN + ((wc + cl) ^ sk)];
      const double sb = SUMS ? br + bi : Bs[(2 * BK + kr) * BN + ((wc + cl) ^ sk)];
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        de[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br, de[i], 0, 0, 0);
        df[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], bi, df[i], 0, 0, 0);
        dh[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(sa[i], sb, dh[i], 0, 0, 0);
      }
    }
    if (BARRIER) __syncthreads();
  }
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < RB; ++i) t += de[i][0] + df[i][1] + dh[i][2] + de[i][3];
  out[blockIdx.x * 256 + threadIdx.x] = t;
}

// Read stream: every element of p[0, n) is read once (grid-stride, 4 loads of 16 B
// in flight per lane); one sum per thread is written so the loads stay live.
template <bool NONTEMPORAL>
__global__ __launch_bounds__(256) void k_read(const double2* __restrict__ p, size_t n, double* __restrict__ out) {
  double s0 = 0.0, s1 = 0.0;
  const size_t stride = (size_t)gridDim.x * 1024;
  for (size_t i = (size_t)blockIdx.x * 1024 + threadIdx.x; i < n; i += stride) {
    double2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const size_t k = i + (size_t)u * 256;
      if (k < n) {
        if (NONTEMPORAL) {
          v[u].x = __builtin_nontemporal_load(&p[k].x);
          v[u].y = __builtin_nontemporal_load(&p[k].y);
        } else {
          v[u] = p[k];
        }
      } else {
        v[u] = make_double2(0.0, 0.0);
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      s0 += v[u].x;
      s1 += v[u].y;
    }
  }
  out[(size_t)blockIdx.x * 256 + threadIdx.x] = s0 + s1;
}

// Plain copy dst[0, n) = src[0, n), grid-stride, 16-byte loads and stores.
__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ src, double2* __restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// ---- fixed-register forms.  Chains live in v16..v31, operands in v40..v47 and s20,
// MFMA accumulators in a0..a15; none of the operand registers is written in the loop.
#define FIXED_CLOBBERS                                                                                  \
  "v16", "v17", "v18", "v19", "v20", "v21", "v22", "v23", "v24", "v25", "v26", "v27", "v28", "v29", "v30", \
      "v31", "v40", "v41", "v42", "v43", "v44", "v45", "v46", "v47", "s20", "a0", "a1", "a2", "a3", "a4",   \
      "a5", "a6", "a7", "a8", "a9", "a10", "a11", "a12", "a13", "a14", "a15"

enum Fixed {
  kF32Distinct,   // v_fma_f32 acc, m1, m2, acc with acc, m1, m2 in three different banks
  kF32SameBank,   // m1 in the bank of acc
  kF32SrcPair,    // m1 and m2 in one bank
  kF32Sgpr,       // m2 an SGPR
  kF64Valu,       // 8 v_fma_f64
  kF64Mfma,       // 2 v_mfma_f64_4x4x4_4b_f64
  kF64Mix,        // the 8 FMAs and the 2 MFMAs interleaved
  kF32Valu,       // 8 v_fma_f32
  kF32Mfma,       // 2 v_mfma_f32_4x4x1_16b_f32
  kF32Mix,        // interleaved
  kF32Chain1,     // 32 v_fma_f32 in one dependency chain
  kF32Chain2,     // two chains, alternating
  kF32Rcp,        // per group of 8: 7 independent v_fma_f32 + 1 v_rcp_f32
  kF32F64,        // per group of 8: 7 independent v_fma_f32 + 1 v_fma_f64
};

#define F32_DISTINCT                                                                   \
  "v_fma_f32 v16, v41, v46, v16\n v_fma_f32 v17, v42, v47, v17\n"                     \
  "v_fma_f32 v18, v43, v44, v18\n v_fma_f32 v19, v40, v45, v19\n"                     \
  "v_fma_f32 v20, v41, v46, v20\n v_fma_f32 v21, v42, v47, v21\n"                     \
  "v_fma_f32 v22, v43, v44, v22\n v_fma_f32 v23, v40, v45, v23\n"
#define F32_SAMEBANK                                                                   \
  "v_fma_f32 v16, v40, v46, v16\n v_fma_f32 v17, v41, v47, v17\n"                     \
  "v_fma_f32 v18, v42, v44, v18\n v_fma_f32 v19, v43, v45, v19\n"                     \
  "v_fma_f32 v20, v40, v46, v20\n v_fma_f32 v21, v41, v47, v21\n"                     \
  "v_fma_f32 v22, v42, v44, v22\n v_fma_f32 v23, v43, v45, v23\n"
#define F32_SRCPAIR                                                                    \
  "v_fma_f32 v16, v41, v45, v16\n v_fma_f32 v17, v42, v46, v17\n"                     \
  "v_fma_f32 v18, v43, v47, v18\n v_fma_f32 v19, v40, v44, v19\n"                     \
  "v_fma_f32 v20, v41, v45, v20\n v_fma_f32 v21, v42, v46, v21\n"                     \
  "v_fma_f32 v22, v43, v47, v22\n v_fma_f32 v23, v40, v44, v23\n"
#define F32_SGPR                                                                       \
  "v_fma_f32 v16, v41, s20, v16\n v_fma_f32 v17, v42, s20, v17\n"                     \
  "v_fma_f32 v18, v43, s20, v18\n v_fma_f32 v19, v40, s20, v19\n"                     \
  "v_fma_f32 v20, v41, s20, v20\n v_fma_f32 v21, v42, s20, v21\n"                     \
  "v_fma_f32 v22, v43, s20, v22\n v_fma_f32 v23, v40, s20, v23\n"
#define F64_A "v_fma_f64 v[16:17], v[40:41], v[16:17], v[42:43]\n v_fma_f64 v[18:19], v[40:41], v[18:19], v[42:43]\n"
#define F64_B "v_fma_f64 v[20:21], v[40:41], v[20:21], v[42:43]\n v_fma_f64 v[22:23], v[40:41], v[22:23], v[42:43]\n"
#define F64_C "v_fma_f64 v[24:25], v[40:41], v[24:25], v[42:43]\n v_fma_f64 v[26:27], v[40:41], v[26:27], v[42:43]\n"
#define F64_D "v_fma_f64 v[28:29], v[40:41], v[28:29], v[42:43]\n v_fma_f64 v[30:31], v[40:41], v[30:31], v[42:43]\n"
#define M64_0 "v_mfma_f64_4x4x4_4b_f64 a[0:1], v[44:45], v[46:47], a[0:1]\n"
#define M64_1 "v_mfma_f64_4x4x4_4b_f64 a[2:3], v[44:45], v[46:47], a[2:3]\n"
#define F32_A "v_fma_f32 v16, v40, v16, v41\n v_fma_f32 v17, v42, v17, v43\n"
#define F32_B "v_fma_f32 v18, v40, v18, v41\n v_fma_f32 v19, v42, v19, v43\n"
#define F32_C "v_fma_f32 v20, v40, v20, v41\n v_fma_f32 v21, v42, v21, v43\n"
#define F32_D "v_fma_f32 v22, v40, v22, v41\n v_fma_f32 v23, v42, v23, v43\n"
#define M32_0 "v_mfma_f32_4x4x1_16b_f32 a[0:3], v44, v45, a[0:3]\n"
#define M32_1 "v_mfma_f32_4x4x1_16b_f32 a[4:7], v44, v45, a[4:7]\n"
#define F32_CHAIN1_1 "v_fma_f32 v16, v40, v16, v41\n"
#define F32_CHAIN1 F32_CHAIN1_1 F32_CHAIN1_1 F32_CHAIN1_1 F32_CHAIN1_1 F32_CHAIN1_1 F32_CHAIN1_1 F32_CHAIN1_1 F32_CHAIN1_1
#define F32_CHAIN2_2 "v_fma_f32 v16, v40, v16, v41\n v_fma_f32 v17, v42, v17, v43\n"
#define F32_CHAIN2 F32_CHAIN2_2 F32_CHAIN2_2 F32_CHAIN2_2 F32_CHAIN2_2
#define F32_SEVEN                                                                      \
  "v_fma_f32 v16, v41, v46, v16\n v_fma_f32 v17, v42, v47, v17\n"                     \
  "v_fma_f32 v18, v43, v44, v18\n v_fma_f32 v19, v40, v45, v19\n"                     \
  "v_fma_f32 v20, v41, v46, v20\n v_fma_f32 v21, v42, v47, v21\n"                     \
  "v_fma_f32 v22, v43, v44, v22\n"
// v24 stays 1.0 under v_rcp_f32; its next use is 8 instructions later (past the
// transcendental-use hazard)
#define F32_RCP F32_SEVEN "v_rcp_f32 v24, v24\n"
#define F32_F64 F32_SEVEN "v_fma_f64 v[28:29], v[40:41], v[28:29], v[42:43]\n"
#define X4(G) G G G G

// Operand values come from `seed` (1.0f - 2^-24 as bits of each 32-bit half), so
// the chains neither overflow nor go subnormal.
template <Fixed F>
__global__ __launch_bounds__(256) void k_fixed(float* __restrict__ out, int iters, float seed) {
  asm volatile(
      "v_mov_b32 v16, 1.0\n v_mov_b32 v17, 1.0\n v_mov_b32 v18, 1.0\n v_mov_b32 v19, 1.0\n"
      "v_mov_b32 v20, 1.0\n v_mov_b32 v21, 1.0\n v_mov_b32 v22, 1.0\n v_mov_b32 v23, 1.0\n"
      "v_mov_b32 v24, 1.0\n v_mov_b32 v25, 1.0\n v_mov_b32 v26, 1.0\n v_mov_b32 v27, 1.0\n"
      "v_mov_b32 v28, 1.0\n v_mov_b32 v29, 1.0\n v_mov_b32 v30, 1.0\n v_mov_b32 v31, 1.0\n"
      "v_mov_b32 v40, %0\n v_mov_b32 v41, %0\n v_mov_b32 v42, %0\n v_mov_b32 v43, %0\n"
      "v_mov_b32 v44, %0\n v_mov_b32 v45, %0\n v_mov_b32 v46, %0\n v_mov_b32 v47, %0\n"
      "s_mov_b32 s20, 0x3f7fffff\n"
      "v_accvgpr_write_b32 a0, 0\n v_accvgpr_write_b32 a1, 0\n v_accvgpr_write_b32 a2, 0\n"
      "v_accvgpr_write_b32 a3, 0\n v_accvgpr_write_b32 a4, 0\n v_accvgpr_write_b32 a5, 0\n"
      "v_accvgpr_write_b32 a6, 0\n v_accvgpr_write_b32 a7, 0\n s_nop 7\n" ::"v"(seed)
      : FIXED_CLOBBERS);
  // four copies of the group per iteration keep the loop branch and the code
  // alignment of the short body out of the measurement
  for (int i = 0; i < iters; ++i) {
    if (F == kF32Distinct) asm volatile(X4(F32_DISTINCT) ::: FIXED_CLOBBERS);
    if (F == kF32SameBank) asm volatile(X4(F32_SAMEBANK) ::: FIXED_CLOBBERS);
    if (F == kF32SrcPair) asm volatile(X4(F32_SRCPAIR) ::: FIXED_CLOBBERS);
    if (F == kF32Sgpr) asm volatile(X4(F32_SGPR) ::: FIXED_CLOBBERS);
    if (F == kF64Valu) asm volatile(X4(F64_A F64_B F64_C F64_D) ::: FIXED_CLOBBERS);
    if (F == kF64Mfma) asm volatile(X4(M64_0 M64_1) ::: FIXED_CLOBBERS);
    if (F == kF64Mix) asm volatile(X4(F64_A M64_0 F64_B F64_C M64_1 F64_D) ::: FIXED_CLOBBERS);
    if (F == kF32Valu) asm volatile(X4(F32_A F32_B F32_C F32_D) ::: FIXED_CLOBBERS);
    if (F == kF32Mfma) asm volatile(X4(M32_0 M32_1) ::: FIXED_CLOBBERS);
    if (F == kF32Mix) asm volatile(X4(F32_A M32_0 F32_B F32_C M32_1 F32_D) ::: FIXED_CLOBBERS);
    if (F == kF32Chain1) asm volatile(X4(F32_CHAIN1) ::: FIXED_CLOBBERS);
    if (F == kF32Chain2) asm volatile(X4(F32_CHAIN2) ::: FIXED_CLOBBERS);
    if (F == kF32Rcp) asm volatile(X4(F32_RCP) ::: FIXED_CLOBBERS);
    if (F == kF32F64) asm volatile(X4(F32_F64) ::: FIXED_CLOBBERS);
  }
  float r;
  asm volatile("s_nop 7\n s_nop 7\n v_mov_b32 %0, v16" : "=v"(r) : : FIXED_CLOBBERS);
  out[blockIdx.x * 256 + threadIdx.x] = r;
}

template <typename F>
float time_ms(F launch) {
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  float ms = 0.f;
  for (int rep = 0; rep < 2; ++rep) {  // the first run warms up
    HIP_OK(hipEventRecord(e0));
    launch();
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(e1));
    HIP_OK(hipEventSynchronize(e1));
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
  }
  HIP_OK(hipEventDestroy(e0));
  HIP_OK(hipEventDestroy(e1));
  return ms;
}

bool want(int argc, char** argv, const char* mode) {
  if (argc < 2) return std::strcmp(mode, "copy") != 0;
  for (int i = 1; i < argc; ++i)
    if (std::strcmp(argv[i], mode) == 0) return true;
  return false;
}

template <int NACC>
void mfma_sweep(int cus, const double* in, double* out) {
  const int miters = 4000;
  int max_per_cu = 0;
  HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&max_per_cu, k_mfma_f64<NACC>, 256, 0));
  for (int per_cu = 1; per_cu <= 8; ++per_cu) {
    if (per_cu > max_per_cu) {
      std::printf("v_mfma_f64_16x16x4f64, %2d accumulators, %d wave(s)/SIMD: not resident (max %d)\n", NACC,
                  per_cu, max_per_cu);
      continue;
    }
    const int blocks = cus * per_cu;
    const float ms =
        time_ms([&] { hipLaunchKernelGGL(k_mfma_f64<NACC>, dim3(blocks), dim3(256), 0, 0, in, out, miters); });
    const double flop = (double)blocks * 4 * miters * NACC * 2048.0;  // 4 waves per block, 16x16x4x2 per MFMA
    std::printf("v_mfma_f64_16x16x4f64, %2d accumulators, %d wave(s)/SIMD %8.3f ms  %6.2f TFLOP/s\n", NACC, per_cu,
                ms, flop / ms / 1e9);
  }
}

template <int NACC>
void mfma4_sweep(int cus, const double* in, double* out) {
  const int miters = 16000;
  int max_per_cu = 0;
  HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&max_per_cu, k_mfma4_f64<NACC>, 256, 0));
  for (int per_cu = 1; per_cu <= 8; per_cu *= 2) {
    if (per_cu > max_per_cu) continue;
    const int blocks = cus * per_cu;
    const float ms =
        time_ms([&] { hipLaunchKernelGGL(k_mfma4_f64<NACC>, dim3(blocks), dim3(256), 0, 0, in, out, miters); });
    const double flop = (double)blocks * 4 * miters * NACC * 512.0;
    std::printf("v_mfma_f64_4x4x4f64,   %2d accumulators, %d wave(s)/SIMD %8.3f ms  %6.2f TFLOP/s\n", NACC, per_cu,
                ms, flop / ms / 1e9);
  }
}

template <int RB, bool SUMS, bool BARRIER>
void mfma_lds_case(int cus, const double* in, double* out) {
  constexpr int LW = (RB * 16 + 31) / 32 * 32;
  const size_t need = (size_t)3 * 16 * (LW + 64) * sizeof(double);
  const int stages = 2000;
  for (int per_cu = 1; per_cu <= 2; ++per_cu) {
    // reserve LDS so that exactly per_cu workgroups (per_cu waves per SIMD) fit a CU's 160 KiB
    const size_t lds = per_cu == 1 ? std::max(need, (size_t)100 << 10) : std::max(need, (size_t)64 << 10);
    auto k = k_mfma_lds<RB, SUMS, BARRIER>;
    HIP_OK(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int resident = 0;
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, k, 256, lds));
    if (resident != per_cu) {
      std::printf("mfma from LDS, %d row blocks: %d workgroups/CU resident, wanted %d\n", RB, resident, per_cu);
      continue;
    }
    const int blocks = cus * per_cu;
    const float ms = time_ms([&] { hipLaunchKernelGGL(k, dim3(blocks), dim3(256), lds, 0, in, out, stages); });
    const double flop = (double)blocks * 4 * stages * 4 * 3 * RB * 2048.0;  // 4 waves, 4 k-steps, 3 MFMAs per block
    std::printf("mfma from LDS, %d row blocks, %s, %s barrier/stage, %d wave(s)/SIMD %8.3f ms  %6.2f TFLOP/s\n",
                RB, SUMS ? "VALU sums " : "sum plane ", BARRIER ? "   " : "no ", per_cu, ms, flop / ms / 1e9);
  }
}

}  // namespace

int main(int argc, char** argv) {
  hipDeviceProp_t pr;
  HIP_OK(hipGetDeviceProperties(&pr, 0));
  const int cus = pr.multiProcessorCount;
  const int max_blocks = cus * 16;
  double *in, *out;
  HIP_OK(hipMalloc(&in, 32 * sizeof(double)));
  HIP_OK(hipMemset(in, 0, 32 * sizeof(double)));
  HIP_OK(hipMalloc(&out, (size_t)max_blocks * 256 * sizeof(double)));
  std::printf("device %s, %d CUs\n", pr.name, cus);

  if (want(argc, argv, "valu")) {
    // VALU forms: 8 blocks per CU (8 waves per SIMD), 8 chains per lane
    struct V {
      const char* name;
      void (*k)(const double*, double*, int);
      double flop_per_lane_instr;
    };
    const V valu[] = {
        {"v_fma_f64  (3 VGPR operands)", k_valu<kFmaF64>, 2.0},
        {"v_mul_f64", k_valu<kMulF64>, 1.0},
        {"v_rcp_f64", k_valu<kRcpF64>, 0.0},
        {"v_fma_f32  (3 VGPR operands)", k_valu<kFmaF32Vgpr3>, 2.0},
        {"v_fma_f32  (2 VGPR + literal)", k_valu<kFmaF32Vgpr2>, 2.0},
        {"v_mul_f32", k_valu<kMulF32>, 1.0},
        {"v_rcp_f32", k_valu<kRcpF32>, 0.0},
    };
    const int iters = 20000;
    const int blocks = cus * 8;
    for (const V& v : valu) {
      const float ms = time_ms([&] { hipLaunchKernelGGL(v.k, dim3(blocks), dim3(256), 0, 0, in, out, iters); });
      const double lane_instr = (double)blocks * 256 * iters * 8;
      const double wave_per_cu_ns = lane_instr / 64 / ((double)ms * 1e6) / cus;
      std::printf("%-30s %8.3f ms  %6.3f wave-instr/CU/ns", v.name, ms, wave_per_cu_ns);
      if (v.flop_per_lane_instr > 0) std::printf("  %6.2f TFLOP/s", lane_instr * v.flop_per_lane_instr / ms / 1e9);
      std::printf("\n");
    }
  }
  if (want(argc, argv, "mfma")) {
    mfma4_sweep<4>(cus, in, out);
    mfma4_sweep<8>(cus, in, out);
    mfma4_sweep<16>(cus, in, out);
    mfma_sweep<4>(cus, in, out);
    mfma_sweep<8>(cus, in, out);
    mfma_sweep<16>(cus, in, out);
  }
  if (want(argc, argv, "mfmalds")) {
    mfma_lds_case<4, true, true>(cus, in, out);
    mfma_lds_case<4, false, true>(cus, in, out);
    mfma_lds_case<4, false, false>(cus, in, out);
    mfma_lds_case<5, true, true>(cus, in, out);
    mfma_lds_case<5, false, true>(cus, in, out);
    mfma_lds_case<5, false, false>(cus, in, out);
  }
  if (want(argc, argv, "fixed")) {
    // fixed-register forms, 8 blocks per CU (8 waves per SIMD); per iteration 32 VALU
    // instructions, or 8 MFMAs, or both interleaved
    struct FX {
      const char* name;
      void (*k)(float*, int, float);
    };
    const FX fixed[] = {
        {"v_fma_f32  fixed regs, 3 banks", k_fixed<kF32Distinct>},
        {"v_fma_f32  acc + src one bank", k_fixed<kF32SameBank>},
        {"v_fma_f32  two srcs one bank", k_fixed<kF32SrcPair>},
        {"v_fma_f32  VGPR * SGPR + acc", k_fixed<kF32Sgpr>},
        {"8 v_fma_f64", k_fixed<kF64Valu>},
        {"2 v_mfma_f64_4x4x4_4b", k_fixed<kF64Mfma>},
        {"8 v_fma_f64 + 2 mfma_f64 mixed", k_fixed<kF64Mix>},
        {"8 v_fma_f32", k_fixed<kF32Valu>},
        {"2 v_mfma_f32_4x4x1_16b", k_fixed<kF32Mfma>},
        {"8 v_fma_f32 + 2 mfma_f32 mixed", k_fixed<kF32Mix>},
    };
    const int fiters = 5000;
    const int blocks = cus * 8;
    float* fout = reinterpret_cast<float*>(out);
    for (int f = 0; f < 10; ++f) {
      const FX& v = fixed[f];
      const float ms =
          time_ms([&] { hipLaunchKernelGGL(v.k, dim3(blocks), dim3(256), 0, 0, fout, fiters, 0.99999994f); });
      if (f < 4) {
        const double winstr = (double)blocks * 4 * fiters * 32;
        std::printf("%-32s %8.3f ms  %6.3f wave-instr/CU/ns\n", v.name, ms, winstr / ((double)ms * 1e6) / cus);
      } else {
        std::printf("%-32s %8.3f ms per %d iterations\n", v.name, ms, fiters);
      }
    }
  }
  if (want(argc, argv, "dual")) {
    // VALU dual issue: the same fixed-register groups at 1, 2, 4 and 8 waves per SIMD.
    // Two VALU instructions issue in one quad-cycle only if they are independent; the
    // rates at 1 wave/SIMD tell whether the pair may come from one wave.
    struct FX {
      const char* name;
      void (*k)(float*, int, float);
    };
    const FX forms[] = {
        {"8 independent v_fma_f32", k_fixed<kF32Distinct>},
        {"1 chain of v_fma_f32", k_fixed<kF32Chain1>},
        {"2 alternating chains", k_fixed<kF32Chain2>},
        {"7 v_fma_f32 + 1 v_rcp_f32", k_fixed<kF32Rcp>},
        {"7 v_fma_f32 + 1 v_fma_f64", k_fixed<kF32F64>},
    };
    const int fiters = 5000;
    float* fout = reinterpret_cast<float*>(out);
    for (const FX& v : forms)
      for (int per_cu = 1; per_cu <= 8; per_cu *= 2) {
        const int blocks = cus * per_cu;
        const float ms =
            time_ms([&] { hipLaunchKernelGGL(v.k, dim3(blocks), dim3(256), 0, 0, fout, fiters, 0.99999994f); });
        const double winstr = (double)blocks * 4 * fiters * 32;
        std::printf("dual: %-28s %d wave(s)/SIMD %8.3f ms  %6.3f wave-instr/CU/ns\n", v.name, per_cu, ms,
                    winstr / ((double)ms * 1e6) / cus);
      }
  }
  if (want(argc, argv, "hbm")) {
    const size_t bytes = size_t(4) << 30;
    const size_t n = bytes / sizeof(double2);
    double2* buf;
    HIP_OK(hipMalloc(&buf, bytes));
    HIP_OK(hipMemset(buf, 0, bytes));
    for (int nt = 0; nt < 2; ++nt)
      for (int per_cu : {4, 8, 16}) {
        const int blocks = cus * per_cu;
        const float ms = time_ms([&] {
          if (nt)
            hipLaunchKernelGGL(k_read<true>, dim3(blocks), dim3(256), 0, 0, buf, n, out);
          else
            hipLaunchKernelGGL(k_read<false>, dim3(blocks), dim3(256), 0, 0, buf, n, out);
        });
        std::printf("read stream %s, %2d blocks/CU: 4 GiB in %8.3f ms  %6.3f TB/s\n", nt ? "non-temporal" : "plain",
                    per_cu, ms, bytes / (ms * 1e-3) / 1e12);
      }
    HIP_OK(hipFree(buf));
  }
  if (want(argc, argv, "copy")) {
    const size_t bytes = size_t(1) << 30;
    const size_t n = bytes / sizeof(double2);
    double2 *src, *dst;
    HIP_OK(hipMalloc(&src, bytes));
    HIP_OK(hipMalloc(&dst, bytes));
    HIP_OK(hipMemset(src, 0, bytes));
    HIP_OK(hipDeviceSynchronize());
    hipLaunchKernelGGL(k_copy, dim3(cus * 8), dim3(256), 0, 0, src, dst, n);
    HIP_OK(hipGetLastError());
    HIP_OK(hipDeviceSynchronize());
    std::printf("copy: %zu bytes read and %zu written by k_copy\n", bytes, bytes);
    HIP_OK(hipFree(src));
    HIP_OK(hipFree(dst));
  }
  HIP_OK(hipDeviceSynchronize());
  HIP_OK(hipFree(in));
  HIP_OK(hipFree(out));
  return 0;
}
