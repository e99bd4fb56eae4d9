// ceilings.hip -- issue-rate ceilings of the instruction forms the GPP and FF
// kernels are built from, measured on the device: FP64/FP32 FMA and MUL,
// v_rcp_f64/f32, and v_mfma_f64_16x16x4f64 at one and two waves per SIMD.
// Each kernel runs 8 independent dependency chains per lane (register-only work),
// 256-thread blocks; the rate is printed in wave instructions per CU per ns and,
// for the arithmetic forms, in TFLOP/s.
// A second group pins the operands to fixed registers with inline asm (VALU and
// MFMA instructions only, no memory access) to isolate operand forms: FP32 FMAs
// whose source VGPRs share a register bank (index mod 4) or include an SGPR, and
// VALU FMAs interleaved with MFMAs of the same precision (do the two overlap?).
//
//   build/ceilings            (make -C berkeleygw-kernels-cpp_amd build/ceilings)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

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

__global__ __launch_bounds__(256) void k_mfma_f64(const double* __restrict__ in, double* __restrict__ out, int iters) {
  const double a = 1.0 + threadIdx.x * 1e-9 + in[0], b = 0.999999 + in[1];
  f64x4 acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = (f64x4){0.0, 0.0, 0.0, 0.0};
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  out[blockIdx.x * 256 + threadIdx.x] = s;
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

}  // namespace

int main() {
  hipDeviceProp_t pr;
  HIP_OK(hipGetDeviceProperties(&pr, 0));
  const int cus = pr.multiProcessorCount;
  const int max_blocks = cus * 8;
  double *in, *out;
  HIP_OK(hipMalloc(&in, 32 * sizeof(double)));
  HIP_OK(hipMemset(in, 0, 32 * sizeof(double)));
  HIP_OK(hipMalloc(&out, (size_t)max_blocks * 256 * sizeof(double)));
  std::printf("device %s, %d CUs\n", pr.name, cus);

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
  for (const V& v : valu) {
    const float ms = time_ms([&] { hipLaunchKernelGGL(v.k, dim3(max_blocks), dim3(256), 0, 0, in, out, iters); });
    const double lane_instr = (double)max_blocks * 256 * iters * 8;
    const double wave_per_cu_ns = lane_instr / 64 / ((double)ms * 1e6) / cus;
    std::printf("%-30s %8.3f ms  %6.3f wave-instr/CU/ns", v.name, ms, wave_per_cu_ns);
    if (v.flop_per_lane_instr > 0) std::printf("  %6.2f TFLOP/s", lane_instr * v.flop_per_lane_instr / ms / 1e9);
    std::printf("\n");
  }
  // FP64 MFMA: 1 and 2 blocks per CU (1 and 2 waves per SIMD), 8 accumulators per wave
  const int miters = 4000;
  for (int per_cu = 1; per_cu <= 2; ++per_cu) {
    const int blocks = cus * per_cu;
    const float ms = time_ms([&] { hipLaunchKernelGGL(k_mfma_f64, dim3(blocks), dim3(256), 0, 0, in, out, miters); });
    const double flop = (double)blocks * 4 * miters * 8 * 2048.0;  // 4 waves per block, 16x16x4x2 per MFMA
    std::printf("v_mfma_f64_16x16x4f64, %d wave(s)/SIMD %8.3f ms  %6.2f TFLOP/s\n", per_cu, ms, flop / ms / 1e9);
  }
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
  float* fout = reinterpret_cast<float*>(out);
  for (int f = 0; f < 10; ++f) {
    const FX& v = fixed[f];
    const float ms = time_ms([&] { hipLaunchKernelGGL(v.k, dim3(max_blocks), dim3(256), 0, 0, fout, fiters, 0.99999994f); });
    if (f < 4) {
      const double winstr = (double)max_blocks * 4 * fiters * 32;
      std::printf("%-32s %8.3f ms  %6.3f wave-instr/CU/ns\n", v.name, ms, winstr / ((double)ms * 1e6) / cus);
    } else {
      std::printf("%-32s %8.3f ms per %d iterations\n", v.name, ms, fiters);
    }
  }
  HIP_OK(hipDeviceSynchronize());
  HIP_OK(hipFree(in));
  HIP_OK(hipFree(out));
  return 0;
}
