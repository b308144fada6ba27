// Argument conversion and context cache of the MATLAB gateway; see vie_mex_args.hpp.
#include "vie_mex_args.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace vie_mex {

namespace {

template <typename T>
bool convert_int(const T* p, size_t n, int64_t nvox, std::vector<int64_t>& out, std::string& err) {
  for (size_t s = 0; s < n; ++s) {
    // 1-based MATLAB index -> 0-based; the comparison is made before the
    // subtraction so that unsigned values never wrap
    if (!(p[s] >= (T)1) || (uint64_t)p[s] > (uint64_t)nvox) {
      err = "idxS entry " + std::to_string(s + 1) + " is outside 1.." + std::to_string(nvox);
      return false;
    }
    out[s] = (int64_t)p[s] - 1;
  }
  return true;
}

template <typename T>
bool convert_float(const T* p, size_t n, int64_t nvox, std::vector<int64_t>& out, std::string& err) {
  for (size_t s = 0; s < n; ++s) {
    const double v = (double)p[s];
    if (!(v >= 1.0) || v > (double)nvox || std::floor(v) != v) {
      err = "idxS entry " + std::to_string(s + 1) + " is not an integer in 1.." + std::to_string(nvox);
      return false;
    }
    out[s] = (int64_t)v - 1;
  }
  return true;
}

inline uint64_t load64(const unsigned char* p) {
  uint64_t v;
  std::memcpy(&v, p, 8);
  return v;
}

// splitmix64 finaliser
inline uint64_t mix(uint64_t x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}

}  // namespace

bool idx_to_zero_based(NumClass cls, const void* data, size_t n, int64_t nvox, std::vector<int64_t>& sorted,
                       std::vector<int64_t>& perm, std::string& err) {
  sorted.assign(n, 0);
  perm.clear();
  if (n > 0 && !data) {
    err = "idxS has no data";
    return false;
  }
  bool ok = false;
  switch (cls) {
    case NumClass::Double: ok = convert_float(static_cast<const double*>(data), n, nvox, sorted, err); break;
    case NumClass::Single: ok = convert_float(static_cast<const float*>(data), n, nvox, sorted, err); break;
    case NumClass::Int32: ok = convert_int(static_cast<const int32_t*>(data), n, nvox, sorted, err); break;
    case NumClass::UInt32: ok = convert_int(static_cast<const uint32_t*>(data), n, nvox, sorted, err); break;
    case NumClass::Int64: ok = convert_int(static_cast<const int64_t*>(data), n, nvox, sorted, err); break;
    case NumClass::UInt64: ok = convert_int(static_cast<const uint64_t*>(data), n, nvox, sorted, err); break;
    default: err = "idxS must be double, single, int32, uint32, int64 or uint64"; return false;
  }
  if (!ok) return false;
  bool increasing = true;
  for (size_t s = 1; s < n && increasing; ++s) increasing = sorted[s - 1] < sorted[s];
  if (increasing) return true;
  perm.resize(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) { return sorted[a] < sorted[b]; });
  std::vector<int64_t> s2(n);
  for (size_t k = 0; k < n; ++k) s2[k] = sorted[perm[k]];
  for (size_t k = 1; k < n; ++k)
    if (s2[k - 1] == s2[k]) {
      err = "idxS holds voxel " + std::to_string(s2[k] + 1) + " twice";
      perm.clear();
      return false;
    }
  sorted.swap(s2);
  return true;
}

void packed_to_sorted(const double* in, double* out, const std::vector<int64_t>& perm, int ql, int nrhs) {
  const size_t nS = perm.size();
  for (size_t b = 0; b < (size_t)ql * nrhs; ++b)
    for (size_t k = 0; k < nS; ++k) {
      out[2 * (b * nS + k)] = in[2 * (b * nS + perm[k])];
      out[2 * (b * nS + k) + 1] = in[2 * (b * nS + perm[k]) + 1];
    }
}

void packed_from_sorted(const double* in, double* out, const std::vector<int64_t>& perm, int ql, int nrhs) {
  const size_t nS = perm.size();
  for (size_t b = 0; b < (size_t)ql * nrhs; ++b)
    for (size_t k = 0; k < nS; ++k) {
      out[2 * (b * nS + perm[k])] = in[2 * (b * nS + k)];
      out[2 * (b * nS + perm[k]) + 1] = in[2 * (b * nS + k) + 1];
    }
}

uint64_t fingerprint(const void* data, size_t bytes, uint64_t h) {
  const unsigned char* p = static_cast<const unsigned char*>(data);
  // four independent lanes over 8-byte words keep the multiplies pipelined
  uint64_t l[4] = {h ^ 0x243f6a8885a308d3ull, h ^ 0x13198a2e03707344ull, h ^ 0xa4093822299f31d0ull,
                   h ^ 0x082efa98ec4e6c89ull};
  const size_t nw = bytes / 8;
  size_t i = 0;
  for (; i + 4 <= nw; i += 4)
    for (int k = 0; k < 4; ++k) l[k] = mix(l[k] ^ load64(p + 8 * (i + k)));
  for (; i < nw; ++i) l[0] = mix(l[0] ^ load64(p + 8 * i));
  uint64_t tail = 0;
  for (size_t b = nw * 8; b < bytes; ++b) tail = (tail << 8) | p[b];
  uint64_t r = mix(l[0] ^ mix(l[1] ^ mix(l[2] ^ mix(l[3] ^ tail))));
  return mix(r ^ (uint64_t)bytes);
}

bool RawArray::same(const void* data, size_t nbytes, int c, bool cplx) const {
  return c == cls && cplx == complex && nbytes == bytes.size() &&
         (nbytes == 0 || std::memcmp(data, bytes.data(), nbytes) == 0);
}

void RawArray::assign(const void* data, size_t nbytes, int c, bool cplx) {
  const unsigned char* p = static_cast<const unsigned char*>(data);
  bytes.assign(p, p + nbytes);
  cls = c;
  complex = cplx;
}

CtxCache::Entry* CtxCache::find(const TKey& key) {
  for (Entry* e : entries_)
    if (e->key == key) {
      e->last_use = ++tick_;
      return e;
    }
  return nullptr;
}

CtxCache::Entry* CtxCache::insert(const TKey& key, void* ctx) {
  while (max_ > 0 && entries_.size() >= max_) {
    auto lru = std::min_element(entries_.begin(), entries_.end(),
                                [](const Entry* a, const Entry* b) { return a->last_use < b->last_use; });
    destroy_((*lru)->ctx);
    delete *lru;
    entries_.erase(lru);
  }
  Entry* e = new Entry();
  e->key = key;
  e->ctx = ctx;
  e->last_use = ++tick_;
  entries_.push_back(e);
  return e;
}

void CtxCache::clear() {
  for (Entry* e : entries_) {
    destroy_(e->ctx);
    delete e;
  }
  entries_.clear();
}

}  // namespace vie_mex
