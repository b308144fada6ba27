// CPU test of the gateway's argument conversion and context cache
// (vie_mex_args.cpp).  Built with g++ under AddressSanitizer and
// UndefinedBehaviorSanitizer by the Makefile target mexargs-test and run by
// tests/test_mex_args.py; prints one line per check and "ALL OK" at the end.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vie_mex_args.hpp"

using namespace vie_mex;

static int g_fail = 0;

#define CHECK(cond)                                                              \
  do {                                                                           \
    if (!(cond)) {                                                               \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);               \
      ++g_fail;                                                                  \
    }                                                                            \
  } while (0)

static void test_idx_classes() {
  // the shipped body models store idxS as int32, 1-based (geo_body_domain.m:13)
  const int32_t i32[] = {2, 5, 9};
  const uint32_t u32[] = {2, 5, 9};
  const int64_t i64[] = {2, 5, 9};
  const uint64_t u64[] = {2, 5, 9};
  const double f64[] = {2.0, 5.0, 9.0};
  const float f32[] = {2.0f, 5.0f, 9.0f};
  const void* data[] = {i32, u32, i64, u64, f64, f32};
  const NumClass cls[] = {NumClass::Int32, NumClass::UInt32, NumClass::Int64, NumClass::UInt64, NumClass::Double,
                          NumClass::Single};
  for (int k = 0; k < 6; ++k) {
    std::vector<int64_t> s, p;
    std::string err;
    CHECK(idx_to_zero_based(cls[k], data[k], 3, 10, s, p, err));
    CHECK(s.size() == 3 && s[0] == 1 && s[1] == 4 && s[2] == 8);
    CHECK(p.empty());
  }
  std::printf("idx classes ok\n");
}

static void test_idx_errors() {
  std::vector<int64_t> s, p;
  std::string err;
  const int32_t zero[] = {0, 3};
  CHECK(!idx_to_zero_based(NumClass::Int32, zero, 2, 10, s, p, err) && !err.empty());
  const int32_t neg[] = {-4};
  CHECK(!idx_to_zero_based(NumClass::Int32, neg, 1, 10, s, p, err));
  const int64_t big[] = {11};
  CHECK(!idx_to_zero_based(NumClass::Int64, big, 1, 10, s, p, err));
  const uint64_t ubig[] = {~0ull};
  CHECK(!idx_to_zero_based(NumClass::UInt64, ubig, 1, 10, s, p, err));
  const double frac[] = {2.5};
  CHECK(!idx_to_zero_based(NumClass::Double, frac, 1, 10, s, p, err));
  const double nan[] = {std::nan("")};
  CHECK(!idx_to_zero_based(NumClass::Double, nan, 1, 10, s, p, err));
  const int32_t dup[] = {4, 2, 4};
  CHECK(!idx_to_zero_based(NumClass::Int32, dup, 3, 10, s, p, err));
  CHECK(!idx_to_zero_based(NumClass::Other, dup, 3, 10, s, p, err));
  CHECK(idx_to_zero_based(NumClass::Int32, nullptr, 0, 10, s, p, err) && s.empty());
  std::printf("idx errors ok\n");
}

static void test_unsorted_roundtrip() {
  // caller order [7 2 5] (1-based) -> sorted [1 4 6] (0-based), perm = [1 2 0]
  const int32_t idx[] = {7, 2, 5};
  std::vector<int64_t> s, p;
  std::string err;
  CHECK(idx_to_zero_based(NumClass::Int32, idx, 3, 10, s, p, err));
  CHECK(s.size() == 3 && s[0] == 1 && s[1] == 4 && s[2] == 6);
  CHECK(p.size() == 3 && p[0] == 1 && p[1] == 2 && p[2] == 0);
  // two components, two rhs: value = 100 rhs + 10 comp + position in caller order
  const int ql = 2, nrhs = 2, nS = 3;
  std::vector<double> in(2 * ql * nrhs * nS), mid(in.size()), back(in.size());
  for (int r = 0; r < nrhs; ++r)
    for (int c = 0; c < ql; ++c)
      for (int k = 0; k < nS; ++k) {
        const size_t e = (size_t)(r * ql + c) * nS + k;
        in[2 * e] = 100 * r + 10 * c + k;
        in[2 * e + 1] = -in[2 * e];
      }
  packed_to_sorted(in.data(), mid.data(), p, ql, nrhs);
  // sorted slot 0 is caller position 1 (voxel 2), slot 2 is caller position 0 (voxel 7)
  CHECK(mid[0] == 1 && mid[2] == 2 && mid[4] == 0);
  CHECK(mid[2 * (1 * nS + 0)] == 11 && mid[2 * (3 * nS + 2) + 1] == -110);
  packed_from_sorted(mid.data(), back.data(), p, ql, nrhs);
  CHECK(std::memcmp(in.data(), back.data(), in.size() * sizeof(double)) == 0);
  std::printf("unsorted round trip ok\n");
}

static void test_fingerprint() {
  std::vector<std::complex<double>> a(1000);
  for (size_t i = 0; i < a.size(); ++i) a[i] = {std::sin(0.1 * i), std::cos(0.3 * i)};
  const uint64_t h0 = fingerprint(a.data(), a.size() * 16, kFingerprintSeed);
  CHECK(h0 == fingerprint(a.data(), a.size() * 16, kFingerprintSeed));
  // a copy at another address hashes the same
  std::vector<std::complex<double>> b = a;
  CHECK(h0 == fingerprint(b.data(), b.size() * 16, kFingerprintSeed));
  // one changed value in any position changes the hash
  int same = 0;
  for (size_t i = 0; i < a.size(); i += 37) {
    b = a;
    b[i] += std::complex<double>(1e-12, 0);
    same += fingerprint(b.data(), b.size() * 16, kFingerprintSeed) == h0;
  }
  CHECK(same == 0);
  // length and tail bytes count
  CHECK(fingerprint(a.data(), 15999, kFingerprintSeed) != h0);
  const unsigned char t1[5] = {1, 2, 3, 4, 5}, t2[5] = {1, 2, 3, 4, 6};
  CHECK(fingerprint(t1, 5, 0) != fingerprint(t2, 5, 0));
  std::printf("fingerprint ok\n");
}

static int g_destroyed = 0;
static void fake_destroy(void* p) {
  ++g_destroyed;
  delete static_cast<int*>(p);
}

static void test_cache() {
  g_destroyed = 0;
  {
    CtxCache cache(fake_destroy, 2);
    TKey k1{64, 64, 80, 0, 1, 111}, k2{64, 64, 80, 0, 1, 222}, k3{32, 32, 32, 0, 0, 333};
    CHECK(cache.find(k1) == nullptr);
    auto* e1 = cache.insert(k1, new int(1));
    CHECK(cache.find(k1) == e1);
    // the same factors at another address give the same key: a hit; other
    // content at the old address gives another fingerprint: a miss
    TKey k1b = k1;
    CHECK(cache.find(k1b) == e1);
    CHECK(cache.find(k2) == nullptr);
    cache.insert(k2, new int(2));
    CHECK(cache.find(k1) == e1);   // k1 most recently used now
    cache.insert(k3, new int(3));  // evicts k2
    CHECK(g_destroyed == 1 && cache.size() == 2);
    CHECK(cache.find(k2) == nullptr && cache.find(k1) == e1 && cache.find(k3) != nullptr);
    // material compared byte for byte
    const double mr[4] = {1, 0, 2, 0};
    double mr2[4] = {1, 0, 2, 0};
    e1->mr.assign(mr, sizeof(mr), 6, true);
    CHECK(e1->mr.same(mr2, sizeof(mr2), 6, true));
    mr2[2] = 2.0000000001;
    CHECK(!e1->mr.same(mr2, sizeof(mr2), 6, true));
    CHECK(!e1->mr.same(mr, sizeof(mr), 6, false));
    CHECK(!e1->mr.same(mr, sizeof(mr), 12, true));
    CHECK(!e1->mr.same(mr, sizeof(mr) - 16, 6, true));
    cache.clear();
    CHECK(g_destroyed == 3 && cache.size() == 0);
  }
  std::printf("cache ok\n");
}

int main() {
  test_idx_classes();
  test_idx_errors();
  test_unsorted_roundtrip();
  test_fingerprint();
  test_cache();
  if (g_fail) {
    std::printf("%d FAILED\n", g_fail);
    return 1;
  }
  std::printf("ALL OK\n");
  return 0;
}
