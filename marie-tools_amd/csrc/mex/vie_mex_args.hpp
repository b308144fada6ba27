// Argument conversion and context cache of the MATLAB gateway (vie_hip_mex.cpp),
// in plain C++ with no MATLAB header, so that a CPU test (test_vie_mex_args.cpp)
// can drive every conversion the gateway makes.
//
//  * idx_to_zero_based: MARIE's dims.idxS (geo_body_domain.m:13, copied from
//    RHBM.idxS; the shipped body models store it as int32, other callers as
//    double) -> the library's sorted 0-based int64 voxel list, with the
//    permutation that maps the caller's packed order onto the sorted one.
//  * fingerprint: 64-bit content hash of the T-struct factors.  The cache
//    recognises a T-struct by content, not by the address MATLAB happens to
//    give it (a T rebuilt for another frequency can land at a freed address).
//  * CtxCache: contexts keyed by grid, operator, basis and T content, shared
//    by every command that uses the same T (mvp_N, mvp_vie, gmres_vie), with
//    the material of each context compared byte for byte before it is reused.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace vie_mex {

// numeric classes the gateway accepts for index arrays (the gateway maps
// mxGetClassID onto these)
enum class NumClass { Double, Single, Int32, UInt32, Int64, UInt64, Other };

// Converts n MATLAB 1-based voxel indices of class cls to 0-based int64 in
// [0, nvox).  Doubles and singles must hold integers.  On success `sorted`
// holds the indices in increasing order and `perm` is empty when the input
// was already strictly increasing, else perm[k] = position in the input of
// sorted[k].  Duplicates, out-of-range or non-integral entries and
// unsupported classes fail with a message in err.
bool idx_to_zero_based(NumClass cls, const void* data, size_t n, int64_t nvox, std::vector<int64_t>& sorted,
                       std::vector<int64_t>& perm, std::string& err);

// Packed vectors of ql blocks of nS values (component-major, idx_solution
// order).  to_sorted: out[c nS + k] = in[c nS + perm[k]] (caller order ->
// sorted order); from_sorted: the inverse.  Interleaved complex doubles.
void packed_to_sorted(const double* in, double* out, const std::vector<int64_t>& perm, int ql, int nrhs);
void packed_from_sorted(const double* in, double* out, const std::vector<int64_t>& perm, int ql, int nrhs);

// 64-bit content hash of `bytes` bytes, continuing from state h.
uint64_t fingerprint(const void* data, size_t bytes, uint64_t h);
constexpr uint64_t kFingerprintSeed = 0x6a09e667f3bcc909ull;

// Identity of a T-struct context: grid, operator (VIE_OP_N / VIE_OP_K),
// basis, and the fingerprint of every term's ranks and factor data.
struct TKey {
  int n1 = 0, n2 = 0, n3 = 0, op = 0, basis = 0;
  uint64_t fp = 0;
  bool operator==(const TKey& o) const {
    return n1 == o.n1 && n2 == o.n2 && n3 == o.n3 && op == o.op && basis == o.basis && fp == o.fp;
  }
};

// Raw bytes of one material input (EP.Mr, EP.Mc or dims.idxS) with its class
// and complexity, compared exactly before a cached material is reused.
struct RawArray {
  std::vector<unsigned char> bytes;
  int cls = 0;
  bool complex = false;
  bool same(const void* data, size_t nbytes, int c, bool cplx) const;
  void assign(const void* data, size_t nbytes, int c, bool cplx);
};

class CtxCache {
 public:
  using Destroy = void (*)(void*);
  explicit CtxCache(Destroy d, size_t max_entries = 4) : destroy_(d), max_(max_entries) {}
  ~CtxCache() { clear(); }
  CtxCache(const CtxCache&) = delete;
  CtxCache& operator=(const CtxCache&) = delete;

  struct Entry {
    TKey key;
    void* ctx = nullptr;
    double gram_res = 0.0;     // voxel edge of the context's invG (0: not set by mvp_vie yet)
    bool have_material = false;
    RawArray mr, mc, idx;      // the material last given to the context
    std::vector<int64_t> perm; // caller order -> sorted idxS (empty: already sorted)
    uint64_t last_use = 0;
  };

  // cached entry for key (marked most recently used), or null
  Entry* find(const TKey& key);
  // adds a context for key; evicts (destroys) least recently used entries
  // beyond the capacity first.  The cache owns ctx from here on.
  Entry* insert(const TKey& key, void* ctx);
  // destroys every context
  void clear();
  size_t size() const { return entries_.size(); }

 private:
  Destroy destroy_;
  size_t max_;
  uint64_t tick_ = 0;
  std::vector<Entry*> entries_;
};

}  // namespace vie_mex
