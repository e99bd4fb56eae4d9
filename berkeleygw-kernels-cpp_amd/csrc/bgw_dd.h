// bgw_dd.h -- device double-double helpers for the fixed-order reductions.
#pragma once

#include <hip/hip_runtime.h>

namespace bgw {

struct dd {
  double hi, lo;
};

__host__ __device__ inline void two_sum(double a, double b, double& s, double& e) {
  double x = a + b;
  double bb = x - a;
  e = (a - (x - bb)) + (b - bb);
  s = x;
}

__host__ __device__ inline dd dd_add(dd a, double b) {
  double s, e;
  two_sum(a.hi, b, s, e);
  e += a.lo;
  dd r;
  r.hi = s + e;
  r.lo = e - (r.hi - s);
  return r;
}

__host__ __device__ inline dd dd_add(dd a, dd b) {
  double s, e;
  two_sum(a.hi, b.hi, s, e);
  e += a.lo + b.lo;
  dd r;
  r.hi = s + e;
  r.lo = e - (r.hi - s);
  return r;
}

// complex double-double
struct cdd {
  dd re, im;
};

__host__ __device__ inline cdd cdd_zero() {
  cdd z;
  z.re.hi = z.re.lo = z.im.hi = z.im.lo = 0.0;
  return z;
}

__host__ __device__ inline cdd cdd_add(cdd a, cdd b) {
  cdd r;
  r.re = dd_add(a.re, b.re);
  r.im = dd_add(a.im, b.im);
  return r;
}

__host__ __device__ inline cdd cdd_add(cdd a, double re, double im) {
  cdd r;
  r.re = dd_add(a.re, re);
  r.im = dd_add(a.im, im);
  return r;
}

// Stored form of a cdd partial: (re.hi, re.lo, im.hi, im.lo).
__device__ inline void cdd_store(double4* dst, cdd v) { *dst = make_double4(v.re.hi, v.re.lo, v.im.hi, v.im.lo); }
__device__ inline cdd cdd_load(const double4* src) {
  double4 v = *src;
  cdd r;
  r.re.hi = v.x;
  r.re.lo = v.y;
  r.im.hi = v.z;
  r.im.lo = v.w;
  return r;
}

// Wave-level (64 lanes) fixed-pattern butterfly reduction of a cdd.
__device__ inline cdd cdd_wave_reduce(cdd v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    cdd o;
    o.re.hi = __shfl_xor(v.re.hi, off, 64);
    o.re.lo = __shfl_xor(v.re.lo, off, 64);
    o.im.hi = __shfl_xor(v.im.hi, off, 64);
    o.im.lo = __shfl_xor(v.im.lo, off, 64);
    // order the operands by lane bit so both partners compute the identical sum
    v = (threadIdx.x & off) ? cdd_add(o, v) : cdd_add(v, o);
  }
  return v;
}

// Block reduction (blockDim.x a multiple of 64, <= 1024) of one cdd per thread;
// the result is valid in thread 0.  `sh` needs blockDim.x/64 entries.
__device__ inline cdd cdd_block_reduce(cdd v, cdd* sh) {
  v = cdd_wave_reduce(v);
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[wave] = v;
  __syncthreads();
  cdd r = cdd_zero();
  if (threadIdx.x == 0)
    for (int w = 0; w < nw; ++w) r = cdd_add(r, sh[w]);
  return r;
}

// Wave sums of N double-doubles v[] (N a power of two, at most 64), reduce-scatter style:
// each of the first log2(N) exchanges halves the values a lane carries (the lane keeps the
// lower or upper half by its lane bit and adds its partner's copy of that half), the
// remaining lane bits sum the one value left.  On return v[0] of a lane is the wave total
// of component `comp`; `writer` is set in one lane per component.  That takes
// N - 1 + 6 - log2(N) additions per lane, against 6 N for N butterflies; every addition has
// a fixed operand order (lower lane first), so the sums are bitwise reproducible.
template <int N, int MASK = 32>
__device__ __forceinline__ void dd_wave_sum_scatter(dd* v, int lane, int& comp, bool& writer) {
  static_assert(N >= 1 && (N & (N - 1)) == 0 && (MASK == 0 ? N == 1 : N <= 2 * MASK), "N: a power of two up to 64");
  if constexpr (MASK > 0) {
    const bool up = (lane & MASK) != 0;
    if constexpr (N > 1) {
#pragma unroll
      for (int j = 0; j < N / 2; ++j) {
        const dd send = up ? v[j] : v[j + N / 2];
        const dd keep = up ? v[j + N / 2] : v[j];
        dd o;
        o.hi = __shfl_xor(send.hi, MASK, 64);
        o.lo = __shfl_xor(send.lo, MASK, 64);
        v[j] = up ? dd_add(o, keep) : dd_add(keep, o);
      }
      if (up) comp += N / 2;
      dd_wave_sum_scatter<N / 2, MASK / 2>(v, lane, comp, writer);
    } else {
      dd o;
      o.hi = __shfl_xor(v[0].hi, MASK, 64);
      o.lo = __shfl_xor(v[0].lo, MASK, 64);
      v[0] = up ? dd_add(o, v[0]) : dd_add(v[0], o);
      if (up) writer = false;
      dd_wave_sum_scatter<1, MASK / 2>(v, lane, comp, writer);
    }
  }
}

}  // namespace bgw
