// ceilings.hip -- issue-rate ceilings of the instruction forms the GPP and FF
// kernels are built from, measured on the device: FP64/FP32 FMA and MUL,
// v_rcp_f64/f32, and v_mfma_f64_16x16x4f64 at one and two waves per SIMD.
// Each kernel runs 8 independent dependency chains per lane (register-only work),
// 256-thread blocks; the rate is printed in wave instructions per CU per ns and,
// for the arithmetic forms, in TFLOP/s.
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
  HIP_OK(hipDeviceSynchronize());
  HIP_OK(hipFree(in));
  HIP_OK(hipFree(out));
  return 0;
}
