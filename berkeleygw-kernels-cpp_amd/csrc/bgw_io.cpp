// bgw_io.cpp -- input sets on disk (host only).
//
// The reference mini-apps build their inputs in memory from constants
// (GPP/gpp.cpp:220-276, FF/ff.cpp:552-633) and have no file input; a production
// caller hands over dielectric-matrix and matrix-element dumps instead.  This file
// defines a small self-describing container for exactly the arrays and scalars the
// GPP and FF entry points take, memory-mapped on read so that a multi-GB set goes
// to the solver without a copy.
//
// Layout (little-endian):
//   [0,32)     "BGWINPUT", u32 version (=1), u32 kind, u32 count, u32 0, u64 directory offset (=32)
//   directory  count entries of 96 bytes:
//                char name[32] (1-31 printable ASCII characters other than space,
//                NUL-padded), u32 dtype, u32 ndim, u64 dims[4] (unused = 1),
//                u64 offset (multiple of 4096), u64 nbytes, u64 checksum
//   payload    each array at its offset, row-major, complex as interleaved (re,im)
// checksum = sum_i mix64(w_i ^ (i * 0x9E3779B97F4A7C15)) mod 2^64 over the payload's
// 8-byte little-endian words w_i (zero-padded to a multiple of 8 bytes), mix64 being
// the SplitMix64 finalizer.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bgw.h"

namespace {

constexpr char kMagic[8] = {'B', 'G', 'W', 'I', 'N', 'P', 'U', 'T'};
constexpr uint32_t kVersion = 1;
constexpr size_t kHeader = 32, kEntry = 96, kAlign = 4096, kMaxEntries = 256;

thread_local std::string g_err;

int fail(const std::string& msg) {
  g_err = msg;
  return BGW_EIO;
}

size_t elem_size(int dtype) {
  switch (dtype) {
    case BGW_DT_F64: return 8;
    case BGW_DT_C128: return 16;
    case BGW_DT_I32: return 4;
    case BGW_DT_I64: return 8;
    default: return 0;
  }
}

// Entry names: 1 to 31 printable ASCII characters other than space.
bool valid_name(const char* s) {
  const size_t n = std::strlen(s);
  if (n == 0 || n >= 32) return false;
  for (size_t i = 0; i < n; ++i)
    if ((unsigned char)s[i] < 0x21 || (unsigned char)s[i] > 0x7E) return false;
  return true;
}

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t checksum(const unsigned char* p, size_t nbytes) {
  const size_t nfull = nbytes / 8;
  uint64_t h = 0;
#pragma omp parallel for reduction(+ : h) schedule(static)
  for (long long i = 0; i < (long long)nfull; ++i) {
    uint64_t w;
    std::memcpy(&w, p + 8 * (size_t)i, 8);
    h += mix64(w ^ ((uint64_t)i * 0x9E3779B97F4A7C15ull));
  }
  if (nbytes % 8) {
    uint64_t w = 0;
    std::memcpy(&w, p + 8 * nfull, nbytes % 8);
    h += mix64(w ^ ((uint64_t)nfull * 0x9E3779B97F4A7C15ull));
  }
  return h;
}

template <typename T>
void put(std::vector<unsigned char>& b, size_t off, T v) {
  std::memcpy(b.data() + off, &v, sizeof(T));
}
template <typename T>
T get(const unsigned char* b, size_t off) {
  T v;
  std::memcpy(&v, b + off, sizeof(T));
  return v;
}

// element count of a descriptor, 0 on overflow or bad ndim
bool desc_elems(const bgw_array_desc& d, size_t* n) {
  if (d.ndim < 0 || d.ndim > 4) return false;
  size_t e = 1;
  for (int k = 0; k < d.ndim; ++k) {
    if (d.dims[k] != 0 && e > SIZE_MAX / d.dims[k]) return false;
    e *= d.dims[k];
  }
  *n = e;
  return true;
}

}  // namespace

struct bgw_inputs {
  int fd = -1;
  void* map = nullptr;
  size_t size = 0;
  int kind = 0;
  std::vector<bgw_array_desc> entries;
  std::vector<std::string> names;
};

namespace {

const bgw_array_desc* find(const bgw_inputs* in, const char* name) {
  for (size_t i = 0; i < in->entries.size(); ++i)
    if (in->names[i] == name) return &in->entries[i];
  return nullptr;
}

bool scalar_i(const bgw_inputs* in, const char* name, long long* v) {
  const bgw_array_desc* d = find(in, name);
  if (!d || d->dtype != BGW_DT_I64 || d->ndim != 0) return false;
  *v = *static_cast<const long long*>(d->data);
  return true;
}

// Checks that a GPP or FF set holds every array of its entry points with the
// shapes its scalars imply (GPP: bgw_gpp_solver; FF: the three FF kernels).
int check_kind(const bgw_inputs* in) {
  auto need = [&](const char* name, int dtype, std::initializer_list<size_t> dims) -> int {
    const bgw_array_desc* d = find(in, name);
    if (!d) return fail(std::string("input set lacks '") + name + "'");
    if (d->dtype != dtype) return fail(std::string("'") + name + "' has the wrong dtype");
    if ((size_t)d->ndim != dims.size()) return fail(std::string("'") + name + "' has the wrong rank");
    int k = 0;
    for (size_t n : dims)
      if (d->dims[k++] != n) return fail(std::string("'") + name + "' has the wrong shape");
    return BGW_OK;
  };
  int rc;
  long long nb, nv, nc, ng;
  if (!scalar_i(in, "number_bands", &nb) || !scalar_i(in, "nvband", &nv) || !scalar_i(in, "ncouls", &nc) ||
      !scalar_i(in, "ngpown", &ng))
    return fail("input set lacks one of the int64 scalars number_bands, nvband, ncouls, ngpown");
  if (nb <= 0 || nc <= 0 || ng <= 0 || nv < 0 || nv > nb || ng > nc)
    return fail("input set scalars out of range (need 0 < ngpown <= ncouls, 0 <= nvband <= number_bands)");
  const size_t B = (size_t)nb, C = (size_t)nc, G = (size_t)ng;
  if ((rc = need("occ", BGW_DT_F64, {}))) return rc;
  if ((rc = need("inv_igp_index", BGW_DT_I32, {G}))) return rc;
  if ((rc = need("aqsmtemp", BGW_DT_C128, {B, C}))) return rc;
  if ((rc = need("aqsntemp", BGW_DT_C128, {B, C}))) return rc;
  if ((rc = need("vcoul", BGW_DT_F64, {C}))) return rc;
  if (in->kind == BGW_INPUTS_GPP) {
    if ((rc = need("indinv", BGW_DT_I32, {C + 1}))) return rc;
    if ((rc = need("wtilde_array", BGW_DT_C128, {G, C}))) return rc;
    if ((rc = need("I_eps_array", BGW_DT_C128, {G, C}))) return rc;
    const bgw_array_desc* wx = find(in, "wx_array");
    if (!wx || wx->dtype != BGW_DT_F64) return fail("input set lacks the f64 array 'wx_array'");
    const bool ref_form = wx->ndim == 1 && wx->dims[0] == 3;
    const bool band_form = wx->ndim == 2 && wx->dims[0] == B && wx->dims[1] == 3;
    if (!ref_form && !band_form) return fail("'wx_array' must be [3] or [number_bands, 3]");
    return BGW_OK;
  }
  long long nf, ne;
  if (!scalar_i(in, "nFreq", &nf) || !scalar_i(in, "nfreqeval", &ne))
    return fail("FF input set lacks the int64 scalars nFreq, nfreqeval");
  if (nf < 2 || ne <= 0) return fail("FF input set needs nFreq >= 2 and nfreqeval >= 1");
  const size_t F = (size_t)nf;
  if ((rc = need("freqevalmin", BGW_DT_F64, {}))) return rc;
  if ((rc = need("freqevalstep", BGW_DT_F64, {}))) return rc;
  if ((rc = need("indinv", BGW_DT_I32, {C}))) return rc;
  if ((rc = need("ekq", BGW_DT_F64, {B}))) return rc;
  if ((rc = need("dFreqGrid", BGW_DT_F64, {F}))) return rc;
  if ((rc = need("I_epsR_array", BGW_DT_C128, {F, G, C}))) return rc;
  if ((rc = need("I_epsA_array", BGW_DT_C128, {F, G, C}))) return rc;
  return BGW_OK;
}

}  // namespace

extern "C" {

const char* bgw_io_last_error(void) { return g_err.c_str(); }

int bgw_inputs_write(const char* path, int kind, const bgw_array_desc* arrays, int count) {
  if (!path || !arrays || count <= 0 || (size_t)count > kMaxEntries ||
      (kind != BGW_INPUTS_GPP && kind != BGW_INPUTS_FF && kind != BGW_INPUTS_OTHER)) {
    g_err = "bgw_inputs_write: bad path, kind or array list";
    return BGW_EINVAL;
  }
  std::vector<unsigned char> head(kHeader + kEntry * (size_t)count, 0);
  std::memcpy(head.data(), kMagic, 8);
  put<uint32_t>(head, 8, kVersion);
  put<uint32_t>(head, 12, (uint32_t)kind);
  put<uint32_t>(head, 16, (uint32_t)count);
  put<uint64_t>(head, 24, (uint64_t)kHeader);
  size_t off = (head.size() + kAlign - 1) / kAlign * kAlign;
  std::vector<size_t> offs(count), sizes(count);
  for (int i = 0; i < count; ++i) {
    const bgw_array_desc& d = arrays[i];
    size_t n = 0;
    const size_t es = elem_size(d.dtype);
    if (!d.name || !valid_name(d.name) || !es || !desc_elems(d, &n) || n > SIZE_MAX / es ||
        (!d.data && n)) {
      g_err = "bgw_inputs_write: bad descriptor #" + std::to_string(i);
      return BGW_EINVAL;
    }
    for (int j = 0; j < i; ++j)
      if (!std::strcmp(arrays[j].name, d.name)) {
        g_err = std::string("bgw_inputs_write: duplicate name '") + d.name + "'";
        return BGW_EINVAL;
      }
    const size_t e = kHeader + kEntry * (size_t)i;
    std::memcpy(head.data() + e, d.name, std::strlen(d.name));
    put<uint32_t>(head, e + 32, (uint32_t)d.dtype);
    put<uint32_t>(head, e + 36, (uint32_t)d.ndim);
    for (int k = 0; k < 4; ++k) put<uint64_t>(head, e + 40 + 8 * k, k < d.ndim ? (uint64_t)d.dims[k] : 1);
    offs[i] = off;
    sizes[i] = n * es;
    put<uint64_t>(head, e + 72, off);
    put<uint64_t>(head, e + 80, sizes[i]);
    put<uint64_t>(head, e + 88, checksum(static_cast<const unsigned char*>(d.data), sizes[i]));
    off = (off + sizes[i] + kAlign - 1) / kAlign * kAlign;
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(std::string("cannot create ") + path + ": " + std::strerror(errno));
  bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
  size_t pos = head.size();
  static const unsigned char zeros[kAlign] = {0};
  for (int i = 0; ok && i < count; ++i) {
    while (ok && pos < offs[i]) {
      const size_t z = std::min(kAlign, offs[i] - pos);
      ok = std::fwrite(zeros, 1, z, f) == z;
      pos += z;
    }
    const unsigned char* p = static_cast<const unsigned char*>(arrays[i].data);
    for (size_t done = 0; ok && done < sizes[i];) {
      const size_t chunk = std::min<size_t>(sizes[i] - done, (size_t)1 << 26);
      ok = std::fwrite(p + done, 1, chunk, f) == chunk;
      done += chunk;
    }
    pos += sizes[i];
  }
  if (std::fclose(f) != 0) ok = false;
  if (!ok) return fail(std::string("write error on ") + path);
  return BGW_OK;
}

int bgw_inputs_open(const char* path, int verify, bgw_inputs** out) {
  if (!path || !out) {
    g_err = "bgw_inputs_open: null argument";
    return BGW_EINVAL;
  }
  *out = nullptr;
  auto* in = new bgw_inputs;
  auto bad = [&](const std::string& m) {
    bgw_inputs_close(in);
    return fail(std::string(path) + ": " + m);
  };
  in->fd = ::open(path, O_RDONLY);
  if (in->fd < 0) return bad(std::strerror(errno));
  struct stat st;
  if (fstat(in->fd, &st) != 0) return bad(std::strerror(errno));
  in->size = (size_t)st.st_size;
  if (in->size < kHeader) return bad("too short for an input set");
  in->map = mmap(nullptr, in->size, PROT_READ, MAP_PRIVATE, in->fd, 0);
  if (in->map == MAP_FAILED) {
    in->map = nullptr;
    return bad(std::string("mmap failed: ") + std::strerror(errno));
  }
  const auto* b = static_cast<const unsigned char*>(in->map);
  if (std::memcmp(b, kMagic, 8) != 0) return bad("not an input set (magic)");
  if (get<uint32_t>(b, 8) != kVersion) return bad("unsupported version " + std::to_string(get<uint32_t>(b, 8)));
  in->kind = (int)get<uint32_t>(b, 12);
  const uint32_t count = get<uint32_t>(b, 16);
  const uint64_t dir = get<uint64_t>(b, 24);
  if (in->kind != BGW_INPUTS_GPP && in->kind != BGW_INPUTS_FF && in->kind != BGW_INPUTS_OTHER)
    return bad("unknown kind " + std::to_string(in->kind));
  if (count == 0 || count > kMaxEntries) return bad("bad entry count");
  if (get<uint32_t>(b, 20) != 0) return bad("reserved header word is not zero");
  if (dir != kHeader || dir + (uint64_t)count * kEntry > in->size) return bad("directory outside the file");
  for (uint32_t i = 0; i < count; ++i) {
    const size_t e = dir + (size_t)i * kEntry;
    char name[33] = {0};
    std::memcpy(name, b + e, 32);
    if (!valid_name(name)) return bad("entry #" + std::to_string(i) + " has no valid name");
    for (const auto& prev : in->names)
      if (prev == name) return bad(std::string("duplicate entry '") + name + "'");
    bgw_array_desc d{};
    d.dtype = (int)get<uint32_t>(b, e + 32);
    d.ndim = (int)get<uint32_t>(b, e + 36);
    for (int k = 0; k < 4; ++k) d.dims[k] = (size_t)get<uint64_t>(b, e + 40 + 8 * k);
    for (int k = d.ndim; k >= 0 && k < 4; ++k)
      if (d.dims[k] != 1) return bad(std::string("entry '") + name + "': unused dims must be 1");
    const uint64_t off = get<uint64_t>(b, e + 72), nbytes = get<uint64_t>(b, e + 80);
    size_t n = 0;
    const size_t es = elem_size(d.dtype);
    if (!es || !desc_elems(d, &n) || n > SIZE_MAX / es || n * es != nbytes)
      return bad(std::string("entry '") + name + "': dtype, rank or shape inconsistent with its size");
    if (off % kAlign || off < dir + (uint64_t)count * kEntry || off > in->size || nbytes > in->size - off)
      return bad(std::string("entry '") + name + "' lies outside the file");
    d.data = b + off;
    if (verify && checksum(b + off, nbytes) != get<uint64_t>(b, e + 88))
      return bad(std::string("checksum mismatch in '") + name + "'");
    in->names.push_back(name);
    in->entries.push_back(d);
  }
  for (size_t i = 0; i < in->entries.size(); ++i) in->entries[i].name = in->names[i].c_str();
  if (in->kind != BGW_INPUTS_OTHER) {
    const int rc = check_kind(in);
    if (rc) {
      const std::string m = g_err;
      return bad(m);
    }
  }
  *out = in;
  g_err.clear();
  return BGW_OK;
}

int bgw_inputs_kind(const bgw_inputs* in) { return in ? in->kind : BGW_EINVAL; }

int bgw_inputs_count(const bgw_inputs* in) { return in ? (int)in->entries.size() : BGW_EINVAL; }

int bgw_inputs_entry(const bgw_inputs* in, int i, bgw_array_desc* desc) {
  if (!in || !desc || i < 0 || (size_t)i >= in->entries.size()) return BGW_EINVAL;
  *desc = in->entries[(size_t)i];
  return BGW_OK;
}

int bgw_inputs_get(const bgw_inputs* in, const char* name, bgw_array_desc* desc) {
  if (!in || !name || !desc) return BGW_EINVAL;
  const bgw_array_desc* d = find(in, name);
  if (!d) {
    g_err = std::string("input set has no entry '") + name + "'";
    return BGW_EINVAL;
  }
  *desc = *d;
  return BGW_OK;
}

void bgw_inputs_close(bgw_inputs* in) {
  if (!in) return;
  if (in->map) munmap(in->map, in->size);
  if (in->fd >= 0) ::close(in->fd);
  delete in;
}

}  // extern "C"
