// bgw_synth.cpp -- seeded synthetic inputs (host only).
//
// Counter-based SplitMix64 so that any element can be produced independently
// (OpenMP-parallel, order independent) and the CPU oracle, the GPU path and the
// drivers all see bit-identical inputs.  Replaces the reference drivers'
// constant initialisation (GPP/gpp.cpp:246-276, FF/ff.cpp:584-633).
#include <cstdint>

#include "bgw.h"

namespace {

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t stream_at(uint64_t seed, uint64_t i) { return mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull); }
inline double unit(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }
inline uint32_t bounded(uint64_t x, uint32_t n) { return (uint32_t)(((x >> 32) * (uint64_t)n) >> 32); }

}  // namespace

extern "C" {

void bgw_synth_uniform(unsigned long long seed, double lo, double hi, double* out, size_t n) {
  const double w = hi - lo;
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) out[i] = lo + w * unit(stream_at(seed, (uint64_t)i));
}

void bgw_synth_complex(unsigned long long seed, double re_lo, double re_hi, double im_lo,
                       double im_hi, double* out, size_t n) {
  const double wr = re_hi - re_lo, wi = im_hi - im_lo;
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    out[2 * i] = re_lo + wr * unit(stream_at(seed, 2 * (uint64_t)i));
    out[2 * i + 1] = im_lo + wi * unit(stream_at(seed, 2 * (uint64_t)i + 1));
  }
}

void bgw_synth_index(unsigned long long seed, int n_max, int* out, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = (int)bounded(stream_at(seed, i), (uint32_t)n_max);
}

void bgw_synth_permutation(unsigned long long seed, int* out, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = (int)i;
  if (n < 2) return;
  uint64_t k = 0;
  for (size_t i = n - 1; i >= 1; --i) {
    size_t j = bounded(stream_at(seed, k++), (uint32_t)(i + 1));
    int t = out[i];
    out[i] = out[j];
    out[j] = t;
  }
}

}  // extern "C"
