// driver_args.hpp -- command-line checks shared by the gpp.ex and ff.ex drivers.
#pragma once

#include <unistd.h>

#include <cerrno>
#include <cstdlib>
#include <string>

namespace bgw {

// The whole of `s` as a decimal integer in [lo, hi].  std::atoi, which the reference's
// drivers use, would read "12abc" as 12 and overflow on large values.
inline bool parse_int(const std::string& s, long long lo, long long hi, long long* out) {
  if (s.empty()) return false;
  errno = 0;
  char* end = nullptr;
  const long long v = std::strtoll(s.c_str(), &end, 10);
  if (errno != 0 || end != s.c_str() + s.size() || v < lo || v > hi) return false;
  *out = v;
  return true;
}

// Physical memory of the host in bytes (0 if unknown): a run whose arrays need more is
// refused before it allocates, instead of being killed for want of memory part-way.
inline double host_memory_bytes() {
  const long pages = sysconf(_SC_PHYS_PAGES), page = sysconf(_SC_PAGE_SIZE);
  return pages > 0 && page > 0 ? (double)pages * (double)page : 0.0;
}

}  // namespace bgw
