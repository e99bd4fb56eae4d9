// bgw_fact.h -- frequency-grid bracketing of the FF kernels (host and device).
#pragma once

#include <hip/hip_runtime.h>

namespace bgw {

// compute_fact, FF/ff.cpp:103-152, with the reference's behaviour kept exactly:
//  * strict '<' on both sides: a wx on a grid point is "not found" (ifreq unchanged);
//  * loop 1 (and loop 2 for occupied bands with wx<=0, on -wx): ifreq==0 after the
//    search is remapped to nFreq-2, so the first interval extrapolates from the top;
//  * loop 2 with wx>0: weights times -1/2 and no remap (its test is ifreq==-1);
//  * loop 2, unoccupied, wx<=0: nothing is written.
// fact1, fact2, ifreq are in/out, like the reference's reference arguments.  When no
// interval brackets wx, the incoming ifreq survives the search; the reference then
// reads the grid at it even when it lies outside [0, nFreq-2].  Here that case
// returns false before any read (fact1, fact2 unchanged); the device plan and the
// host slice selection always start from ifreq = 0 and never meet it.
__host__ __device__ inline bool compute_fact(double wx, int nFreq, const double* grid, double& fact1,
                                             double& fact2, int& ifreq, int loop, bool flag_occ) {
  auto in_grid = [&]() { return ifreq >= 0 && ifreq <= nFreq - 2; };
  if (loop == 1) {
    const double w = wx > 0.0 ? wx : -wx;
    for (int k = 0; k < nFreq - 1; ++k)
      if (w > grid[k] && w < grid[k + 1]) ifreq = k;
    if (ifreq == 0) ifreq = nFreq - 2;
    if (!in_grid()) return false;
    const double den = grid[ifreq + 1] - grid[ifreq];
    if (wx > 0.0) {
      fact1 = (grid[ifreq + 1] - wx) / den;
      fact2 = (wx - grid[ifreq]) / den;
    } else {
      fact1 = (grid[ifreq + 1] + wx) / den;
      fact2 = (-grid[ifreq] - wx) / den;
    }
    return true;
  }
  if (loop == 2 && wx > 0.0) {
    for (int k = 0; k < nFreq - 1; ++k)
      if (wx > grid[k] && wx < grid[k + 1]) ifreq = k;
    if (ifreq == -1) ifreq = nFreq - 2;
    if (!in_grid()) return false;
    const double den = grid[ifreq + 1] - grid[ifreq];
    fact1 = -0.5 * (grid[ifreq + 1] - wx) / den;
    fact2 = -0.5 * (wx - grid[ifreq]) / den;
  } else if (loop == 2 && flag_occ) {
    const double w = -wx;
    ifreq = 0;
    for (int k = 0; k < nFreq - 1; ++k)
      if (w > grid[k] && w < grid[k + 1]) ifreq = k;
    if (ifreq == 0) ifreq = nFreq - 2;
    const double den = grid[ifreq + 1] - grid[ifreq];
    fact1 = (grid[ifreq + 1] - w) / den;
    fact2 = (w - grid[ifreq]) / den;
  }
  return true;
}

}  // namespace bgw
