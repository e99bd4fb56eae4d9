// input_set.hpp -- driver-side helpers over the input-set calls of bgw.h
// (bgw_inputs_open / _get / _write): a RAII handle whose getters return the
// mapped arrays, and a builder for the descriptor list of a dump.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "bgw.h"

namespace bgw {

class InputSetFile {
 public:
  // Opens and checks `path` (checksums included); exits with status 1 and the
  // library's message if it is not a readable input set of `kind`.
  InputSetFile(const char* path, int kind) {
    if (bgw_inputs_open(path, 1, &h_) != BGW_OK) {
      std::printf("cannot read input set: %s\n", bgw_io_last_error());
      std::exit(1);
    }
    if (bgw_inputs_kind(h_) != kind) {
      std::printf("%s: not a %s input set\n", path, kind == BGW_INPUTS_GPP ? "GPP" : "FF");
      std::exit(1);
    }
  }
  ~InputSetFile() { bgw_inputs_close(h_); }
  InputSetFile(const InputSetFile&) = delete;
  InputSetFile& operator=(const InputSetFile&) = delete;

  bgw_array_desc entry(const char* name) const {
    bgw_array_desc d{};
    if (bgw_inputs_get(h_, name, &d) != BGW_OK) {
      std::printf("input set has no entry '%s'\n", name);
      std::exit(1);
    }
    return d;
  }
  long long i64(const char* name) const { return *static_cast<const long long*>(entry(name).data); }
  double f64(const char* name) const { return *static_cast<const double*>(entry(name).data); }
  template <typename T>
  const T* array(const char* name) const {
    return static_cast<const T*>(entry(name).data);
  }

 private:
  bgw_inputs* h_ = nullptr;
};

// Descriptor list for bgw_inputs_write; the scalars are kept in the builder.
class InputSetWriter {
 public:
  void scalar_i64(const char* name, long long v) { add_scalar(name, BGW_DT_I64, &v, sizeof v); }
  void scalar_f64(const char* name, double v) { add_scalar(name, BGW_DT_F64, &v, sizeof v); }
  void array(const char* name, int dtype, const void* data, std::vector<size_t> dims) {
    bgw_array_desc d{};
    d.name = name;
    d.dtype = dtype;
    d.ndim = (int)dims.size();
    for (size_t k = 0; k < 4; ++k) d.dims[k] = k < dims.size() ? dims[k] : 1;
    d.data = data;
    descs_.push_back(d);
  }
  // Writes the set; exits with status 1 and the library's message on failure.
  void write(const char* path, int kind) {
    for (size_t i = 0, s = 0; i < descs_.size(); ++i)
      if (descs_[i].ndim == 0 && descs_[i].data == nullptr) descs_[i].data = &scalars_[s++];
    if (bgw_inputs_write(path, kind, descs_.data(), (int)descs_.size()) != BGW_OK) {
      std::printf("cannot write input set: %s\n", bgw_io_last_error());
      std::exit(1);
    }
  }

 private:
  void add_scalar(const char* name, int dtype, const void* v, size_t n) {
    long long bits = 0;
    __builtin_memcpy(&bits, v, n);
    scalars_.push_back(bits);
    bgw_array_desc d{};
    d.name = name;
    d.dtype = dtype;
    d.ndim = 0;
    for (int k = 0; k < 4; ++k) d.dims[k] = 1;
    d.data = nullptr;  // bound to its slot in write(), once scalars_ stops growing
    descs_.push_back(d);
  }
  std::vector<bgw_array_desc> descs_;
  std::vector<long long> scalars_;
};

}  // namespace bgw
