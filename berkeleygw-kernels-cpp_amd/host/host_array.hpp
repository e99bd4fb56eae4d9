// host_array.hpp -- minimal row-major host container with the `dptr`/`size`
// fields the reference's ArrayMD exposes (arrayMD/arrayMD.h:156-163), so the
// drivers can call the reference-signature API unchanged.
#pragma once

#include <array>
#include <cstddef>
#include <vector>

template <typename T, size_t Dim>
struct HostArray {
  std::array<size_t, Dim> dims{};
  size_t size = 0;
  std::vector<T> store;
  T* dptr = nullptr;

  template <typename... Idx>
  explicit HostArray(Idx... n) : dims{{static_cast<size_t>(n)...}} {
    static_assert(sizeof...(Idx) == Dim, "dimension mismatch");
    size = 1;
    for (size_t d : dims) size *= d;
    store.resize(size);
    dptr = store.data();
  }
  template <typename... Idx>
  T& operator()(Idx... i) {
    static_assert(sizeof...(Idx) == Dim, "dimension mismatch");
    const size_t idx[] = {static_cast<size_t>(i)...};
    size_t off = 0;
    for (size_t d = 0; d < Dim; ++d) off = off * dims[d] + idx[d];
    return dptr[off];
  }
};
