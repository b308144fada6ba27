// Compile-time specialised building blocks for the common FFT lengths:
// constexpr Stockham plans (all index arithmetic by constants) and the
// spectral coupling terms of the N/K operators as template constants, so the
// block product of the fused x pass unrolls completely.
#pragma once

#include <utility>

#include "vie_fft.hpp"

namespace vie {

constexpr int kThreads = 256;

// Stockham stage radices, first stage first.  Order 0: the largest power of
// two first, the odd radices last; order 1: the odd radices (7, 5, 3) first,
// where the first stage needs no twiddles, then the powers of two (the same
// stages as order 0 for a power-of-two length).  The autotuner times both.
constexpr int pick_radix(int rem) {
  return rem % 8 == 0 ? 8 : rem % 4 == 0 ? 4 : rem % 2 == 0 ? 2 : rem % 3 == 0 ? 3 : rem % 5 == 0 ? 5 : 7;
}
constexpr int pick_radix_odd_first(int rem) {
  return rem % 7 == 0 ? 7 : rem % 5 == 0 ? 5 : rem % 3 == 0 ? 3 : pick_radix(rem);
}
template <int ORD>
constexpr int pick_radix_ord(int rem) {
  return ORD ? pick_radix_odd_first(rem) : pick_radix(rem);
}

// one Stockham stage, everything but the data and the twiddle index known at compile time
template <int R, int P, int Ns, int NPEN, bool INV, bool PF, typename C>
__device__ __forceinline__ void stage_ct(const C* __restrict__ src, C* __restrict__ dst, const C* __restrict__ tw,
                                         int tid) {
  constexpr int PR = P / R;
  constexpr int twstep = P / (Ns * R);
  constexpr int NITEMS = NPEN * PR;
#pragma unroll
  for (int i0 = 0; i0 < NITEMS; i0 += kThreads) {
    const int i = i0 + tid;
    if (NITEMS % kThreads != 0 && i >= NITEMS) break;
    int pen, j;
    if (PF) { pen = i % NPEN; j = i / NPEN; } else { pen = i / PR; j = i - pen * PR; }
    const int k = j % Ns;
    C x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int a = j + r * PR;
      x[r] = PF ? src[a * NPEN + pen] : src[pen * P + a];
    }
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const C w = tw[r * k * twstep];
        x[r] = INV ? cmulc(x[r], w) : cmul(x[r], w);
      }
    }
    dft<R, INV>(x);
    const int o = (j / Ns) * (Ns * R) + k;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int a = o + r * Ns;
      if (PF) dst[a * NPEN + pen] = x[r]; else dst[pen * P + a] = x[r];
    }
  }
}

template <int P, int Ns, int NPEN, bool INV, bool PF, int ORD = 0, typename C>
__device__ __forceinline__ C* fft_ct(C* src, C* dst, const C* __restrict__ tw, int tid) {
  if constexpr (Ns == P) {
    return src;
  } else {
    constexpr int R = pick_radix_ord<ORD>(P / Ns);
    stage_ct<R, P, Ns, NPEN, INV, PF>(src, dst, tw, tid);
    __syncthreads();
    return fft_ct<P, Ns * R, NPEN, INV, PF, ORD>(dst, src, tw, tid);
  }
}

// ---------------------------------------------------------------------------
// register-resident DFTs of small composite length with constant twiddles
// ---------------------------------------------------------------------------
// cos / sin of 2 pi k / n from their Taylor series (evaluated by the compiler;
// the argument is reduced to [-pi, pi] first, 24 terms give full FP64 accuracy)
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double taylor_sin(double x) {
  double t = x, s = x;
  for (int n = 1; n < 24; ++n) { t *= -x * x / ((2.0 * n) * (2.0 * n + 1.0)); s += t; }
  return s;
}
constexpr double taylor_cos(double x) {
  double t = 1.0, s = 1.0;
  for (int n = 1; n < 24; ++n) { t *= -x * x / ((2.0 * n - 1.0) * (2.0 * n)); s += t; }
  return s;
}
// cos and sin of 2 pi k / n with the angle folded to [-pi/2, pi/2] exactly in k
constexpr double omega_cos(int k, int n) {
  k %= n;
  if (2 * k > n) k -= n;                        // [-n/2, n/2]
  const int a = k < 0 ? -k : k;
  if (4 * a > n) {                              // |angle| > pi/2: cos(t) = -cos(pi - |t|)
    return -taylor_cos(kTwoPi * (double)(n - 2 * a) / (2.0 * n));
  }
  return taylor_cos(kTwoPi * (double)a / (double)n);
}
constexpr double omega_sin(int k, int n) {
  k %= n;
  if (2 * k > n) k -= n;
  const double s = k < 0 ? -1.0 : 1.0;
  const int a = k < 0 ? -k : k;
  if (4 * a > n) return s * taylor_sin(kTwoPi * (double)(n - 2 * a) / (2.0 * n));  // sin(t) = sin(pi - t)
  return s * taylor_sin(kTwoPi * (double)a / (double)n);
}
// exp(-2 pi i k / n) (forward) or its conjugate
template <int N, int K, bool INV> struct Omega {
  static constexpr double re = omega_cos(K, N);
  static constexpr double im = INV ? omega_sin(K, N) : -omega_sin(K, N);
};

template <int I, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < E) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, E>(f);
  }
}

constexpr bool is_base_radix(int n) { return n == 2 || n == 3 || n == 4 || n == 5 || n == 7 || n == 8; }
constexpr int split_radix(int n) {
  return n % 4 == 0 ? 4 : n % 8 == 0 ? 8 : n % 2 == 0 ? 2 : n % 3 == 0 ? 3 : n % 5 == 0 ? 5 : 7;
}

// natural-order in, natural-order out DFT of length N held in registers.
// N = R M decimation in time: X[k1 + M k2] = DFT_R over r of w_N^{r k1} DFT_M(x[r + R m])[k1]
template <int N, bool INV, typename C>
__device__ __forceinline__ void reg_fft(C* x) {
  if constexpr (N == 1) {
  } else if constexpr (is_base_radix(N)) {
    dft<N, INV>(x);
  } else {
    constexpr int R = split_radix(N);
    constexpr int M = N / R;
    using T = typename scalar_of<C>::type;
    C sub[R][M];
    static_for<0, R>([&](auto r) {
      static_for<0, M>([&](auto m) { sub[r.value][m.value] = x[r.value + R * m.value]; });
      reg_fft<M, INV>(sub[r.value]);
    });
    static_for<0, M>([&](auto k1) {
      C t[R];
      static_for<0, R>([&](auto r) {
        constexpr int e = r.value * k1.value;
        if constexpr (e == 0) {
          t[r.value] = sub[r.value][k1.value];
        } else {
          C w;
          w.x = (T)Omega<N, e, INV>::re;
          w.y = (T)Omega<N, e, INV>::im;
          t[r.value] = cmul(sub[r.value][k1.value], w);
        }
      });
      dft<R, INV>(t);
      static_for<0, R>([&](auto k2) { x[k1.value + M * k2.value] = t[k2.value]; });
    });
  }
}

// ---------------------------------------------------------------------------
// coupling terms of the block product as compile-time constants
//   output channel U, term K -> input channel v, unique spectrum g, sign
// Tables: pq (mvp_N_pwc_tucker.m:12), curl pq (mvp_K_pwc_tucker.m:12),
// llp (mvp_N_pwl_tucker.m:13), slp (mvp_N_herm_adj_pwl_tucker.m:14), parity
// (assembly_fft_circ_tucker_pwc.m:9-15, assembly_fft_circ_tucker_pwl.m:45-47).
// ---------------------------------------------------------------------------
struct OpTables {
  static constexpr int pqN[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  static constexpr int pqK[3][3] = {{0, -3, 2}, {3, 0, -1}, {-2, 1, 0}};
  static constexpr int llp[4][4] = {{1, 5, 6, 7}, {-5, 2, 8, 9}, {-6, 8, 3, 10}, {-7, 9, 10, 4}};
  static constexpr int slp[4][4] = {{1, -1, -1, -1}, {-1, 1, 1, 1}, {-1, 1, 1, 1}, {-1, 1, 1, 1}};
  static constexpr int gN[3][6] = {{+1, -1, -1, +1, +1, +1}, {+1, -1, +1, +1, -1, +1}, {+1, +1, -1, +1, -1, +1}};
  static constexpr int gK[3][3] = {{-1, +1, +1}, {+1, -1, +1}, {+1, +1, -1}};
  static constexpr int bt[3][10] = {{+1, +1, +1, +1, -1, +1, +1, -1, -1, +1},
                                    {+1, +1, +1, +1, +1, -1, +1, -1, +1, -1},
                                    {+1, +1, +1, +1, +1, +1, -1, +1, -1, -1}};
};

constexpr int iabs(int v) { return v < 0 ? -v : v; }
constexpr int isgn(int v) { return v < 0 ? -1 : 1; }

// per-axis parity of the unique spectrum (component comp, pair) along axis d
constexpr int spectrum_parity(bool isk, bool pwl, int comp, int pair, int d) {
  return (isk ? OpTables::gK[d][comp] : OpTables::gN[d][comp]) * (pwl ? OpTables::bt[d][pair] : 1);
}

template <bool PWL, bool ISK>
struct OpShape {
  static constexpr int QL = PWL ? 12 : 3;
  static constexpr int NLP = PWL ? 4 : 1;
  static constexpr int NT = (ISK ? 2 : 3) * NLP;       // terms per output channel
  static constexpr int NCOMP = ISK ? 3 : 6;
  static constexpr int NG = NCOMP * (PWL ? 10 : 1);
  static constexpr int UPG = QL / 3;                   // outputs per group (one group per p)
};

template <bool PWL, bool ISK, bool ADJ, int U, int K>
struct TermCT {
  using S = OpShape<PWL, ISK>;
  static constexpr int p = PWL ? U / 4 : U;
  static constexpr int l = PWL ? U % 4 : 0;
  static constexpr int qi = K / S::NLP;
  static constexpr int lp = K % S::NLP;
  static constexpr int q = ISK ? (qi < p ? qi : qi + 1) : qi;
  static constexpr int v = PWL ? 4 * q + lp : q;
  static constexpr int comp = ISK ? iabs(OpTables::pqK[p][q]) - 1 : OpTables::pqN[p][q];
  static constexpr int pair = PWL ? iabs(OpTables::llp[l][lp]) - 1 : 0;
  static constexpr int g = pair * S::NCOMP + comp;
  static constexpr int sign = (ISK ? isgn(OpTables::pqK[p][q]) : 1) * (PWL ? isgn(OpTables::llp[l][lp]) : 1) *
                              ((ADJ && PWL) ? OpTables::slp[l][lp] : 1) * ((ADJ && ISK) ? -1 : 1);
  // parity bits of spectrum g: 1 odd along x, 2 along y, 4 along z
  static constexpr int bits = (spectrum_parity(ISK, PWL, comp, pair, 0) < 0 ? 1 : 0) |
                              (spectrum_parity(ISK, PWL, comp, pair, 1) < 0 ? 2 : 0) |
                              (spectrum_parity(ISK, PWL, comp, pair, 2) < 0 ? 4 : 0);
};

}  // namespace vie
