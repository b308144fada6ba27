// Compile-time-length pass kernels for the FFT lengths of the benchmark and
// configuration grids and of their scatterer boxes (axis passes: 64, 80, 96,
// 112, 128, 160, 192, 224, 256, 320; x pass: 64, 80, 96, 112, 128, 192, 256); other
// lengths run the generic kernels of vie_kernels.hip.
//
// Same data flow as the generic kernels; differences:
//  * axis passes: the Stockham plan (FP64) or the two-step register FFT
//    (FP32), pencil count and tile width are constants, so all index
//    arithmetic folds and every stage loop unrolls;
//  * fused x pass: a two-step FFT held in registers (P = A B: an A-point
//    DFT per thread, one padded LDS transpose, B-point DFTs), so a column
//    crosses LDS four times instead of once per radix stage; the block
//    product runs one frequency per work item with all ql inputs and
//    outputs in registers and its terms as compile-time constants sorted by
//    unique spectrum (each spectrum value loaded once per frequency);
//  * fused x pass: columns are walked grouped by octant column (the up to
//    four mirror columns (+-Y, +-Z) that read the same octant spectrum are
//    adjacent), rhs fastest, in contiguous chunks per XCD, so the spectrum
//    column stays in that XCD's L2 while every mirror and rhs consumes it.

#pragma once

#include <utility>

#include "vie_fast.hpp"
#include "vie_passes.hpp"

namespace vie {

// Number of consecutive x tiles (TX complex values of T each) that share one
// 128-byte cache line.
template <typename T, int TX>
constexpr int line_group() {
  return TX * (int)sizeof(V2<T>) >= 128 ? 1 : 128 / (TX * (int)sizeof(V2<T>));
}

// Block -> (tile, line) work item.  Workgroups go to the 8 XCDs round-robin by
// id and each XCD has its own L2: the G consecutive items whose tiles share
// cache lines run on one XCD (block ids k, k+8, ..., k+8(G-1)), so a line is
// fetched from HBM once, and those groups go round-robin over the XCDs, so a
// scatterer concentrated in a few lines still spreads over all of them.
// Bijective on [0, 8 G R) for any R; the identity for G = 1.
template <int G>
__device__ __forceinline__ int64_t axis_item(unsigned bx) {
  const unsigned k = bx % 8, m = bx / 8;
  return (int64_t)G * (8 * (m / G) + k) + m % G;
}

// LDS of one Stockham axis-pass block: two TX x P buffers
template <typename T, int P, int TX>
constexpr size_t axis_lds_bytes() {
  return 2 * (size_t)TX * P * sizeof(V2<T>);
}

template <typename T, int IN, int OUT, bool INV, int P, int TX, int ORD>
__global__ void __launch_bounds__(kThreads) axis_pass_ct(AxisArgs a, const V2<T>* __restrict__ tw) {
  using C = V2<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* bufA = reinterpret_cast<C*>(smem_raw);
  C* bufB = bufA + TX * P;
  const int64_t q = axis_item<line_group<T, TX>()>(blockIdx.x);
  if (q >= a.nitem) return;
  const int tile = (int)(q % a.ntile);
  const int64_t line = q / a.ntile;
  const int cb = blockIdx.y;
  const int slot = cb / a.ql, ch = cb - slot * a.ql;
  if (skip_slot(a.rmap, a.kdone, a.step, slot)) return;
  const int64_t rhs = a.rmap ? a.rmap[slot] : slot;
  const int x0 = tile * TX;
  if (zline_empty(a, x0, TX, line) || xtile_outside(a, x0, TX)) return;
  const int tid = threadIdx.x;
  constexpr int NLOAD = TX * P;
#pragma unroll
  for (int i0 = 0; i0 < NLOAD; i0 += kThreads) {
    const int i = i0 + tid;
    if (NLOAD % kThreads == 0 || i < NLOAD) {
      const int t = i % TX, aa = i / TX;
      bufA[aa * TX + t] = axis_load<T, IN>(a, cb, rhs, ch, x0 + t, line, aa);
    }
  }
  __syncthreads();
  C* res = fft_ct<P, 1, TX, INV, true, ORD>(bufA, bufB, tw, tid);
  const int nout = TX * a.Lout;
  for (int i = tid; i < nout; i += kThreads) {
    const int t = i % TX, aa = i / TX;
    const int x = x0 + t;
    if (x < a.n1) axis_store<T, OUT>(a, cb, rhs, ch, x, line, aa, res[aa * TX + t]);
  }
}

// ---------------------------------------------------------------------------
// fused x pass: two-step register FFT (P = A B), block product, inverse
// ---------------------------------------------------------------------------
// Length split of the x-axis FFT: step 1 is an A-point DFT per thread over
// the elements j + B m, step 2 a B-point DFT over j after one LDS transpose
// (X[k1 + A k2] = DFT_B over j of w_P^{j k1} DFT_A over m of x[j + B m]).
// P = 192 takes 12 x 16, not 24 x 8: at 24 x 8 a PWC block held 10 columns of
// 3 x 8 lanes (232 values per channel and column: FP32 56 KB, FP64 111 KB of
// LDS, 2 resp. 1 block per CU) and a PWL block 2 columns of 12 x 8 (FP64
// 89 KB, one block of 3 wavefronts per CU), with a 24-point register DFT; at
// 12 x 16 a PWC block holds 5 columns (26 resp. 52 KB), a PWL block one
// column of 192 threads (21 resp. 41 KB), and the register DFT has 12 points.
template <int P> struct XSplit;
template <> struct XSplit<64> { static constexpr int A = 8, B = 8; };
template <> struct XSplit<80> { static constexpr int A = 10, B = 8; };
template <> struct XSplit<96> { static constexpr int A = 12, B = 8; };
template <> struct XSplit<112> { static constexpr int A = 14, B = 8; };
template <> struct XSplit<128> { static constexpr int A = 16, B = 8; };
template <> struct XSplit<192> { static constexpr int A = 12, B = 16; };
template <> struct XSplit<256> { static constexpr int A = 16, B = 16; };

constexpr int kXThreads = 256;

// Block of the fused x pass: CPB columns of ql pencils of B lanes, at most
// kXThreads threads, rounded up to whole wavefronts.  PWL (96 threads per
// column at B = 8, 192 at B = 16) launches 192 threads, not 256: a fourth
// wavefront would have no column and no product item and would only hold a
// wave slot of the CU.
template <bool PWL, bool ISK, int P> struct XBlock {
  static constexpr int TPC = OpShape<PWL, ISK>::QL * XSplit<P>::B;   // threads per column
  static constexpr int CPB = kXThreads / TPC;                         // columns per block
  static constexpr int threads = (CPB * TPC + 63) / 64 * 64;
};

// LDS complex values per channel of the fused x pass: the transpose rows are
// padded to B + 1, and the channel stride is 8 mod 16 so the pencils of a
// wavefront spread over the LDS banks
constexpr int x_lcc(int A, int B) { return (A * (B + 1) + 15) / 16 * 16 + 8; }

// block-product terms sorted by unique spectrum g (each spectrum value is
// loaded once per frequency pair and applied to every (u, v) that uses it)
struct TermList {
  int n;
  int u[144], v[144], g[144], s[144];
};

template <bool PWL, bool ISK, bool ADJ>
constexpr TermList make_terms() {
  using S = OpShape<PWL, ISK>;
  TermList L{};
  L.n = 0;
  for (int u = 0; u < S::QL; ++u) {
    const int p = PWL ? u / 4 : u, l = PWL ? u % 4 : 0;
    for (int q = 0; q < 3; ++q) {
      if (ISK && q == p) continue;
      for (int lp = 0; lp < S::NLP; ++lp) {
        const int comp = ISK ? iabs(OpTables::pqK[p][q]) - 1 : OpTables::pqN[p][q];
        const int pair = PWL ? iabs(OpTables::llp[l][lp]) - 1 : 0;
        int s = (ISK ? isgn(OpTables::pqK[p][q]) : 1) * (PWL ? isgn(OpTables::llp[l][lp]) : 1);
        if (ADJ && PWL) s *= OpTables::slp[l][lp];
        if (ADJ && ISK) s = -s;
        L.u[L.n] = u;
        L.v[L.n] = PWL ? 4 * q + lp : q;
        L.g[L.n] = pair * S::NCOMP + comp;
        L.s[L.n] = s;
        ++L.n;
      }
    }
  }
  // insertion sort by g (stable)
  for (int i = 1; i < L.n; ++i) {
    const int u = L.u[i], v = L.v[i], g = L.g[i], s = L.s[i];
    int k = i - 1;
    while (k >= 0 && L.g[k] > g) {
      L.u[k + 1] = L.u[k]; L.v[k + 1] = L.v[k]; L.g[k + 1] = L.g[k]; L.s[k + 1] = L.s[k];
      --k;
    }
    L.u[k + 1] = u; L.v[k + 1] = v; L.g[k + 1] = g; L.s[k + 1] = s;
  }
  return L;
}

template <bool PWL, bool ISK, bool ADJ> struct Terms {
  static constexpr TermList L = make_terms<PWL, ISK, ADJ>();
};

template <bool ISK, bool PWL>
constexpr int gbits(int g) {
  using S = OpShape<PWL, ISK>;
  const int comp = g % S::NCOMP, pair = g / S::NCOMP;
  return (spectrum_parity(ISK, PWL, comp, pair, 0) < 0 ? 1 : 0) | (spectrum_parity(ISK, PWL, comp, pair, 1) < 0 ? 2 : 0) |
         (spectrum_parity(ISK, PWL, comp, pair, 2) < 0 ? 4 : 0);
}

// Mirror signs.  A unique spectrum g = (pair (l, lp), component of (p, q)) is
// odd along axis d exactly when [p==d] ^ [q==d] ^ [l==d+1] ^ [lp==d+1] is 1
// (the tables of assembly_fft_circ_tucker_*.m; K components add 1 per axis).
// Its sign at a mirrored frequency therefore factorises into a sign of the
// input channel v = (q, lp), a sign of the output channel u = (p, l) and,
// for K, one sign per mirrored axis: the block product runs on the octant
// spectrum with the inputs and outputs sign-flipped instead of every spectrum.
template <bool PWL>
constexpr int channel_axes(int u) {
  const int p = PWL ? u / 4 : u, l = PWL ? u % 4 : 0;
  int m = 0;
  for (int d = 0; d < 3; ++d)
    if ((p == d) != (l == d + 1)) m |= 1 << d;
  return m;
}

// acc += s * x * G (ADJ: x * conj(G)) as four FMAs, s a compile-time sign
template <int S, bool ADJ, typename C>
__device__ __forceinline__ void cfma(C& acc, const C& x, const C& G) {
  using T = typename scalar_of<C>::type;
  const T xr = S > 0 ? x.x : -x.x, xi = S > 0 ? x.y : -x.y;
  if constexpr (!ADJ) {
    acc.x = fma(xr, G.x, acc.x);
    acc.x = fma(-xi, G.y, acc.x);
    acc.y = fma(xr, G.y, acc.y);
    acc.y = fma(xi, G.x, acc.y);
  } else {
    acc.x = fma(xr, G.x, acc.x);
    acc.x = fma(xi, G.y, acc.x);
    acc.y = fma(xi, G.x, acc.y);
    acc.y = fma(-xr, G.y, acc.y);
  }
}

// the factorisation above, checked against the parity tables for every term
template <bool PWL, bool ISK, bool ADJ>
constexpr bool mirror_signs_factorise() {
  const TermList L = make_terms<PWL, ISK, ADJ>();
  for (int i = 0; i < L.n; ++i)
    if (gbits<ISK, PWL>(L.g[i]) != (channel_axes<PWL>(L.u[i]) ^ channel_axes<PWL>(L.v[i]) ^ (ISK ? 7 : 0)))
      return false;
  return true;
}
static_assert(mirror_signs_factorise<false, false, false>() && mirror_signs_factorise<true, false, false>() &&
                  mirror_signs_factorise<false, true, false>() && mirror_signs_factorise<true, true, false>(),
              "mirror parity does not factorise into channel signs");

constexpr int first_term_of(const TermList& L, int g0) {
  int i = 0;
  while (i < L.n && L.g[i] < g0) ++i;
  return i;
}

// block product at one frequency X of one column, all ql outputs, in place in
// the column's natural-order spectrum S[c][k] (stride LCC).  The octant index
// of X is min(X, P - X); mirror = mirrored axes (1: x, 2: y, 4: z).
template <typename T, bool PWL, bool ISK, bool ADJ, int P, bool DB = (sizeof(T) == 4)>
__device__ __forceinline__ void product_freq(V2<T>* S, int LCC, const V2<T>* __restrict__ sp, int O1, int X,
                                             int colbits) {
  using C = V2<T>;
  using Sh = OpShape<PWL, ISK>;
  constexpr int QL = Sh::QL, H = P / 2;
  constexpr int GC = sizeof(T) == 8 ? (DB ? 8 : 12) : 6;   // spectrum values in flight per chunk
  const bool mx = X > H;
  const int xo = mx ? P - X : X;
  const int mirror = colbits | (mx ? 1 : 0);
  C in[QL], acc[QL];
  static_for<0, QL>([&](auto V) {
    constexpr int v = V.value;
    C t = S[v * LCC + X];
    if (__popc(mirror & channel_axes<PWL>(v)) & 1) { t.x = -t.x; t.y = -t.y; }
    in[v] = t;
    acc[v] = mk<T>(0, 0);
  });
  constexpr TermList L = Terms<PWL, ISK, ADJ>::L;
  const C* spx = sp + xo;
  constexpr int NCH = (Sh::NG + GC - 1) / GC;
  if constexpr (DB) {
    // double-buffered chunks, the loads of chunk c+1 are issued before the
    // FMAs of chunk c (FP32: the register room exists at 4 waves per SIMD;
    // FP64 with the spectrum staged in LDS: 2 waves per SIMD)
    C Gb[2][GC];
    static_for<0, (GC < Sh::NG ? GC : Sh::NG)>([&](auto Gi) { Gb[0][Gi.value] = spx[Gi.value * O1]; });
    static_for<0, NCH>([&](auto CH) {
      constexpr int g0 = CH.value * GC;
      constexpr int g1 = g0 + GC < Sh::NG ? g0 + GC : Sh::NG;
      constexpr int cur = CH.value & 1;
      if constexpr (CH.value + 1 < NCH) {
        constexpr int h0 = g1, h1 = h0 + GC < Sh::NG ? h0 + GC : Sh::NG;
        static_for<h0, h1>([&](auto Gi) { Gb[cur ^ 1][Gi.value - h0] = spx[Gi.value * O1]; });
      }
      static_for<first_term_of(L, g0), first_term_of(L, g1)>([&](auto I) {
        constexpr int i = I.value;
        cfma<L.s[i], ADJ>(acc[L.u[i]], in[L.v[i]], Gb[cur][L.g[i] - g0]);
      });
    });
  } else {
    static_for<0, NCH>([&](auto CH) {
      constexpr int g0 = CH.value * GC;
      constexpr int g1 = g0 + GC < Sh::NG ? g0 + GC : Sh::NG;
      C G[GC];
      static_for<g0, g1>([&](auto Gi) { G[Gi.value - g0] = spx[Gi.value * O1]; });
      static_for<first_term_of(L, g0), first_term_of(L, g1)>([&](auto I) {
        constexpr int i = I.value;
        cfma<L.s[i], ADJ>(acc[L.u[i]], in[L.v[i]], G[L.g[i] - g0]);
      });
    });
  }
  const int kflip = ISK ? (__popc(mirror) & 1) : 0;
  static_for<0, QL>([&](auto U) {
    constexpr int u = U.value;
    C t = acc[u];
    if ((__popc(mirror & channel_axes<PWL>(u)) & 1) ^ kflip) { t.x = -t.x; t.y = -t.y; }
    S[u * LCC + X] = t;
  });
}

// x[k] *= w^k (INV: conj) for k = 1 .. A-1 with w = w_P^j.  FP64 reads every
// w^k from the table.  FP32 reads only w^(2^b) and forms w^k as the product
// over the set bits of k (at most 4 roundings for A <= 24, about 2.4e-7):
// A - 1 table values in flight next to the A column values do not fit the
// register budget of 4 waves per SIMD.
template <int A, bool INV, typename C>
__device__ __forceinline__ void x_twiddle(C* x, const C* __restrict__ tw, int j) {
  using T = typename scalar_of<C>::type;
  if constexpr (sizeof(T) == 8) {
#pragma unroll
    for (int k = 1; k < A; ++k)
      if (j) x[k] = INV ? cmulc(x[k], tw[j * k]) : cmul(x[k], tw[j * k]);
  } else {
    if (!j) return;
    constexpr int NB = A <= 8 ? 3 : A <= 16 ? 4 : 5;
    C wp[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) wp[b] = tw[j << b];
    static_for<1, A>([&](auto K) {
      constexpr int k = K.value;
      constexpr int lo = __builtin_ctz(k);
      C w = wp[lo];
      static_for<lo + 1, NB>([&](auto B) {
        if constexpr ((k >> B.value) & 1) w = cmul(w, wp[B.value]);
      });
      x[k] = INV ? cmulc(x[k], w) : cmul(x[k], w);
    });
  }
}

// ---------------------------------------------------------------------------
// axis pass with a two-step register FFT (P = A B), the x pass's scheme on a
// strided axis: thread (t, j) of a block of TX pencils x B threads takes the
// A elements j + B m of pencil t, does an A-point DFT and the twiddles in
// registers, writes them once to LDS (rows padded to B + 1), and after the one
// barrier does the B-point DFTs over j for k1 = j + B s < A, storing
// X[k1 + A k2] straight from registers.  Against the Stockham kernel: one LDS
// round trip and one barrier instead of one per radix stage, and a third of
// its LDS, so about twice as many pencils are in flight per CU.  It is the
// default schedule for FP32, where it was measured faster (C2 support box,
// lengths 96 and 112: FP32 operator 117 -> 111-114 us; C5 0.928 -> 0.906 ms);
// FP64 defaults to the Stockham kernel (the register kernel: 177 -> 182 us,
// its z-inverse pass 19 -> 23 us).  The tuner may pick another schedule (this
// one, or the Stockham stages in either radix order) for any pass
// (autotune_tiles).
// ---------------------------------------------------------------------------
// B = 16 from P = 160 (A <= 20): measured on the C4 batch (y at 160), 20 x 8
// against 10 x 16: mixed 8-rhs solve 1.334 -> 1.271 s
template <int P> struct AxisRfSplit { static constexpr int B = P >= 160 ? 16 : 8; };
template <int P, int TX> struct AxisRf {
  static constexpr int B = AxisRfSplit<P>::B;
  static constexpr int A = P / B;
  static constexpr int NT = TX * B < 64 ? 64 : TX * B;   // threads per block (whole wavefronts)
  static constexpr int LT = (A * (B + 1)) | 1;            // LDS values per pencil (odd stride)
};

template <typename T, int P, int TX>
constexpr size_t axis_rf_lds_bytes() {
  return (size_t)TX * AxisRf<P, TX>::LT * sizeof(V2<T>);
}

template <typename T, int IN, int OUT, bool INV, int P, int TX>
__global__ void __launch_bounds__((AxisRf<P, TX>::NT)) axis_pass_rf(AxisArgs a, const V2<T>* __restrict__ tw) {
  using C = V2<T>;
  using S = AxisRf<P, TX>;
  constexpr int A = S::A, B = S::B, LT = S::LT;
  static_assert(A * B == P && S::NT % 64 == 0 && S::NT >= TX * B, "axis split");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int64_t q = axis_item<line_group<T, TX>()>(blockIdx.x);
  if (q >= a.nitem) return;
  const int tile = (int)(q % a.ntile);
  const int64_t line = q / a.ntile;
  const int cb = blockIdx.y;
  const int slot = cb / a.ql, ch = cb - slot * a.ql;
  if (skip_slot(a.rmap, a.kdone, a.step, slot)) return;
  const int64_t rhs = a.rmap ? a.rmap[slot] : slot;
  const int x0 = tile * TX;
  if (zline_empty(a, x0, TX, line) || xtile_outside(a, x0, TX)) return;
  const int t = (int)threadIdx.x % TX, j = (int)threadIdx.x / TX;   // pencil, lane within the pencil
  const bool active = TX * B == S::NT || j < B;
  const int x = x0 + t;
  C* Lp = reinterpret_cast<C*>(smem_raw) + t * LT;
  if (active) {
    C v[A];
#pragma unroll
    for (int m = 0; m < A; ++m)   // forward: the input has Lin <= P/2 values, so m >= A/2 reads zeros
      v[m] = (INV || m < A / 2) ? axis_load<T, IN>(a, cb, rhs, ch, x, line, j + B * m) : mk<T>(0, 0);
    reg_fft<A, INV>(v);
    x_twiddle<A, INV>(v, tw, j);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) Lp[k1 * (B + 1) + j] = v[k1];
  }
  __syncthreads();
  if (!active || x >= a.n1) return;
  constexpr int NS = (A + B - 1) / B;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int k1 = j + B * s;
    if (A % B == 0 || k1 < A) {
      C y[B];
#pragma unroll
      for (int jj = 0; jj < B; ++jj) y[jj] = Lp[k1 * (B + 1) + jj];
      reg_fft<B, INV>(y);
#pragma unroll
      for (int k2 = 0; k2 < (INV ? B / 2 : B); ++k2)   // inverse: Lout <= P/2 values are kept
        if (k1 + A * k2 < a.Lout) axis_store<T, OUT>(a, cb, rhs, ch, x, line, k1 + A * k2, y[k2]);
    }
  }
}

// Staged spectrum: the block's two columns are consecutive work items, which
// the column order makes mirrors (or rhs) of one octant column; its spectrum
// column (NG x (P/2 + 1) values, contiguous in the octant layout) is loaded
// once per block into LDS, the loads issued with the column's own and
// overlapping the forward FFT, so the block product reads LDS instead of
// waiting on one global round trip per spectrum chunk.  Possible where the
// column's LDS plus the block's data fit two blocks per CU (FP64 PWL N at
// P = 64, 80; the C2 support box runs 80); the context picks the staged or
// the plain kernel (x_stage_for).
template <typename T, bool PWL, bool ISK, int P> struct XStage {
  using Sh = OpShape<PWL, ISK>;
  static constexpr int CPB = XBlock<PWL, ISK, P>::CPB;
  static constexpr size_t data = (size_t)CPB * Sh::QL * x_lcc(XSplit<P>::A, XSplit<P>::B) * sizeof(V2<T>);
  static constexpr int nsp = Sh::NG * (P / 2 + 1);   // complex values of a spectrum column
  static constexpr bool possible = PWL && !ISK && CPB == 2 && data + nsp * sizeof(V2<T>) + 64 <= 81920;
};

// per-column record of a block, written to LDS by the column's first thread:
// offset of its octant spectrum column (-1: no work) and its mirror bits
struct XColInfo {
  int64_t spoff;
  int bits, pad;
};

template <typename T, bool PWL, bool ISK, int P, bool STG>
constexpr size_t x_lds_bytes() {
  using X = XStage<T, PWL, ISK, P>;
  return X::data + (STG ? X::nsp * sizeof(V2<T>) : 0) + X::CPB * sizeof(XColInfo);
}

template <typename T, int NG>
__device__ __forceinline__ const V2<T>* x_column_spec(const XArgs<T>& a, int64_t wc, int& colbits) {
  colbits = 0;
  if (wc >= a.work) return nullptr;
  // work < 2^31 (checked at launch): 32-bit quotient and remainder
  const unsigned wu = (unsigned)wc, ns = (unsigned)a.nslot;
  if (skip_slot(a.rmap, a.kdone, a.step, (int)(wu % ns))) return nullptr;
  const int cm = a.colmap[wu / ns];
  const int Yc = cm & 0xffff, Zc = cm >> 16;
  const int yo = (2 * Yc <= a.P2) ? Yc : a.P2 - Yc;
  const int zo = (2 * Zc <= a.P3) ? Zc : a.P3 - Zc;
  colbits = ((2 * Yc > a.P2) ? 2 : 0) | ((2 * Zc > a.P3) ? 4 : 0);
  return a.spec + ((int64_t)zo * a.O2 + yo) * NG * a.O1;
}

// register budget: the product's live set is 2 ql complex values plus the
// spectrum chunk in flight; unbounded load hoisting would spill.  FP32 fits
// 128 VGPRs (4 waves per SIMD) with 6-value chunks and composed twiddles.
// FP64 is held at 2 by LDS (58 KB per block at P = 128) as well; at P = 80
// (40 KB) 3 waves fit with a few dwords of spill and were measured faster
// (C2 support box: 136 -> 113 us); at 96 and 112 the spill grows, and at 192
// (41 KB, one column per block) 3 waves spill 60 bytes and were slower (C5
// operator 1.51 ms at 2 waves, 1.55 ms at 3).  PWC (3 inputs and outputs per
// frequency) needs far fewer: FP64 fits 3 waves per SIMD, FP32 5.
// The B lanes of a pencil lie in one wavefront (B divides 64), and the FFT
// steps of the x pass read and write only their own pencil's LDS region: the
// steps need the LDS accesses of their own wavefront ordered, not a block
// barrier (LDS instructions of a wavefront execute in issue order; the fences
// keep the compiler from moving LDS accesses across).  Only the block product,
// which reads every channel of a column, needs the barriers around it.
__device__ __forceinline__ void pencil_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// With the staged spectrum LDS holds two blocks (six waves) per CU, so the
// twiddles of both FFT directions are loaded once into registers.
template <typename T, int P, bool PWL, bool STG> struct XWaves {
  static constexpr int value = STG ? (sizeof(T) == 8 ? 2 : 3)
                               : !PWL ? (sizeof(T) == 8 ? 3 : 5) : sizeof(T) == 8 ? (P == 80 ? 3 : 2) : 4;
};

// the twiddles x_twiddle multiplies by, w[k] for k = 1 .. A-1 (FP64 from the
// table, FP32 composed from w^(2^b) the same way), for a kernel that applies
// them in both directions
template <int A, typename C>
__device__ __forceinline__ void x_twiddle_load(C* w, const C* __restrict__ tw, int j) {
  using T = typename scalar_of<C>::type;
  if constexpr (sizeof(T) == 8) {
#pragma unroll
    for (int k = 1; k < A; ++k) w[k] = tw[j * k];
  } else {
    constexpr int NB = A <= 8 ? 3 : A <= 16 ? 4 : 5;
    C wp[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) wp[b] = tw[j << b];
    static_for<1, A>([&](auto K) {
      constexpr int k = K.value;
      constexpr int lo = __builtin_ctz(k);
      C v = wp[lo];
      static_for<lo + 1, NB>([&](auto Bb) {
        if constexpr ((k >> Bb.value) & 1) v = cmul(v, wp[Bb.value]);
      });
      w[k] = v;
    });
  }
}

// x[k] *= w[k] (INV: conj) for k = 1 .. A-1, twiddles already in registers
template <int A, bool INV, typename C>
__device__ __forceinline__ void x_twiddle_reg(C* x, const C* w, int j) {
  if (!j) return;
#pragma unroll
  for (int k = 1; k < A; ++k) x[k] = INV ? cmulc(x[k], w[k]) : cmul(x[k], w[k]);
}
template <typename T, bool PWL, bool ISK, bool ADJ, int P, bool STG>
__global__ void __launch_bounds__((XBlock<PWL, ISK, P>::threads))
    __attribute__((amdgpu_waves_per_eu(XWaves<T, P, PWL, STG>::value)))
x_fused_rf(XArgs<T> a, const V2<T>* __restrict__ tw) {
  using C = V2<T>;
  using Sh = OpShape<PWL, ISK>;
  constexpr int QL = Sh::QL, A = XSplit<P>::A, B = XSplit<P>::B;
  constexpr int TPC = XBlock<PWL, ISK, P>::TPC;       // threads per column
  constexpr int CPB = XBlock<PWL, ISK, P>::CPB;       // columns per block
  constexpr int NT = XBlock<PWL, ISK, P>::threads;    // threads per block
  constexpr int LCC = x_lcc(A, B);
  constexpr int O1 = P / 2 + 1;
  static_assert(CPB >= 1 && A * B == P && B % 2 == 0, "x split");
  static_assert(64 % B == 0, "a pencil's B lanes must lie in one wavefront (pencil_sync)");
  static_assert(!STG || XStage<T, PWL, ISK, P>::possible, "staged x pass");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* const Ls = reinterpret_cast<C*>(smem_raw) + (size_t)CPB * QL * LCC;   // staged spectrum [g][xo]
  constexpr int NSP = STG ? XStage<T, PWL, ISK, P>::nsp : 0;
  XColInfo* const xinfo = reinterpret_cast<XColInfo*>(Ls + NSP);
  const int tid = threadIdx.x;
  const int cl = tid / TPC;                  // column within the block
  const int r = tid - cl * TPC;
  const int c = r / B, j = r - c * B;        // channel, lane within the pencil
  C* Lc = reinterpret_cast<C*>(smem_raw) + (size_t)cl * QL * LCC;   // this column's region
  // work item of this column (XCD-chunked order, see XArgs); work < 2^31
  // (checked at launch): 32-bit quotient and remainder
  const int64_t q = (int64_t)(blockIdx.x % 8) * a.per_xcd + blockIdx.x / 8;
  const int64_t w = q * CPB + cl;
  bool valid = cl < CPB && w < a.work;
  int slot = 0, Y = 0, Z = 0;
  if (valid) {
    slot = (int)((unsigned)w % (unsigned)a.nslot);
    const int cm = a.colmap[(unsigned)w / (unsigned)a.nslot];
    Y = cm & 0xffff;
    Z = cm >> 16;
    valid = !skip_slot(a.rmap, a.kdone, a.step, slot);
  }
  const int n1 = a.n1;
  const int64_t gbase = (int64_t)slot * QL * a.cstride + (int64_t)n1 * (Y + (int64_t)a.P2 * Z) + c * a.cstride;
  if (cl < CPB && r == 0) {
    // the product items of this column read where its octant spectrum column
    // lies and which axes are mirrored from this record
    const int yo = (2 * Y <= a.P2) ? Y : a.P2 - Y;
    const int zo = (2 * Z <= a.P3) ? Z : a.P3 - Z;
    XColInfo ci;
    ci.spoff = valid ? ((int64_t)zo * a.O2 + yo) * Sh::NG * O1 : -1;
    ci.bits = ((2 * Y > a.P2) ? 2 : 0) | ((2 * Z > a.P3) ? 4 : 0);
    ci.pad = 0;
    xinfo[cl] = ci;
  }

  // forward step 1: A-point DFT over m of x[j + B m] (only m < A/2 are inside
  // the n1 <= P/2 inputs).  Staged kernel: every load unconditional (clamped
  // to an element of this pencil and masked after), so that with no branch
  // between them the FFT waits for the column's loads only and the twiddle
  // products for the twiddles, while the spectrum loads stay in flight; the
  // plain kernel loads only the elements inside [xlo, xhi] (the clamped loads
  // cost it registers: spills at 3 waves per SIMD, C2 x pass 94 -> 120 us).
  C x[A];
#pragma unroll
  for (int m = 0; m < A; ++m) {
    const int e = j + B * m;
    if constexpr (STG) {
      if (m < A / 2) {
        const C v = a.data[gbase + min(max(e, a.xlo), a.xhi)];
        x[m] = (valid && e >= a.xlo && e <= a.xhi) ? v : mk<T>(0, 0);
      } else {
        x[m] = mk<T>(0, 0);
      }
    } else {
      x[m] = (m < A / 2 && valid && e >= a.xlo && e <= a.xhi) ? a.data[gbase + e] : mk<T>(0, 0);
    }
  }
  constexpr int SPT = (NSP + NT - 1) / NT;
  C wt[STG ? A : 1];
  // the staged values travel as native vectors (a single load and LDS store
  // each; a copy of the vector struct would go through private memory)
  typedef T NV2 __attribute__((ext_vector_type(2)));
  NV2 sv[SPT > 0 ? SPT : 1];
  int64_t stoff = -1;   // offset of the staged octant column
  if constexpr (STG) {
    // twiddles of both directions, then the spectrum column of the block's
    // first column with work (block-uniform)
    x_twiddle_load<A>(wt, tw, j);
    int b0 = 0, b1 = 0;
    const C* s0 = x_column_spec<T, Sh::NG>(a, q * CPB, b0);
    const C* sst = s0 ? s0 : x_column_spec<T, Sh::NG>(a, q * CPB + 1, b1);
    stoff = sst ? (int64_t)(sst - a.spec) : -1;
    const C* sload = sst ? sst : a.spec;   // no column with work: any valid address, never used
    static_for<0, SPT>([&](auto S) {
      sv[S.value] = reinterpret_cast<const NV2*>(sload)[min(tid + S.value * NT, NSP - 1)];
    });
    reg_fft<A, false>(x);
    x_twiddle_reg<A, false>(x, wt, j);
  } else {
    reg_fft<A, false>(x);
    x_twiddle<A, false>(x, tw, j);
  }
  if (cl < CPB) {
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) Lc[c * LCC + k1 * (B + 1) + j] = x[k1];
  }
  pencil_sync();
  // forward step 2: B-point DFTs over j for k1 = j + B s < A (when B does not
  // divide A, the last round has A mod B of the B lanes of a pencil busy)
  constexpr int NS = (A + B - 1) / B;
  C y[NS][B];
  if (cl < CPB) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int k1 = j + B * s;
      if (A % B == 0 || k1 < A) {
#pragma unroll
        for (int jj = 0; jj < B; ++jj) y[s][jj] = Lc[c * LCC + k1 * (B + 1) + jj];
        reg_fft<B, false>(y[s]);
      }
    }
  }
  pencil_sync();
  if (cl < CPB) {
#pragma unroll
    for (int s = 0; s < NS; ++s)
      if (A % B == 0 || j + B * s < A) {
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) Lc[c * LCC + (j + B * s) + A * k2] = y[s][k2];
      }
  }
  if constexpr (STG) {
    static_for<0, SPT>([&](auto S) {
      const int i = tid + S.value * NT;
      if (NSP % NT == 0 || i < NSP) reinterpret_cast<NV2*>(Ls)[i] = sv[S.value];
    });
  }
  __syncthreads();
  // block product over the block's CPB * P frequencies, one frequency per
  // item with all ql outputs, in place; in the staged kernel a column of
  // another octant column than the staged one (a group boundary) reads its
  // spectrum from memory
  for (int i = tid; i < CPB * P; i += NT) {
    const int cc = i / P, X = i - cc * P;
    const XColInfo ci = xinfo[cc];
    if (ci.spoff < 0) continue;
    C* Sc = reinterpret_cast<C*>(smem_raw) + (size_t)cc * QL * LCC;
    if (STG && ci.spoff == stoff)
      product_freq<T, PWL, ISK, ADJ, P, true>(Sc, LCC, Ls, O1, X, ci.bits);
    else
      product_freq<T, PWL, ISK, ADJ, P>(Sc, LCC, a.spec + ci.spoff, O1, X, ci.bits);
  }
  __syncthreads();
  // inverse step 1
  if (cl < CPB) {
#pragma unroll
    for (int m = 0; m < A; ++m) x[m] = Lc[c * LCC + j + B * m];
  }
  reg_fft<A, true>(x);
  if constexpr (STG)
    x_twiddle_reg<A, true>(x, wt, j);
  else
    x_twiddle<A, true>(x, tw, j);
  pencil_sync();
  if (cl < CPB) {
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) Lc[c * LCC + k1 * (B + 1) + j] = x[k1];
  }
  pencil_sync();
  // inverse step 2; keep the outputs x = k1 + A k2 in [xlo, xhi] (xhi < n1;
  // k2 < B/2 covers n1 = P/2, a padded length P > 2 n1 keeps fewer)
  if (valid) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int k1 = j + B * s;
      if (A % B == 0 || k1 < A) {
#pragma unroll
        for (int jj = 0; jj < B; ++jj) y[s][jj] = Lc[c * LCC + k1 * (B + 1) + jj];
        reg_fft<B, true>(y[s]);
#pragma unroll
        for (int k2 = 0; k2 < B / 2; ++k2)
          if (k1 + A * k2 >= a.xlo && k1 + A * k2 <= a.xhi) a.data[gbase + k1 + A * k2] = y[s][k2];
      }
    }
  }
}
// ---------------------------------------------------------------------------
// dispatch over the specialised lengths
// ---------------------------------------------------------------------------
template <int... Ps> struct PList {};

template <typename F, int P0, int... Ps>
static bool dispatch_len(int P, F&& f, PList<P0, Ps...>) {
  if (P == P0) {
    f(std::integral_constant<int, P0>{});
    return true;
  }
  if constexpr (sizeof...(Ps) > 0) return dispatch_len(P, f, PList<Ps...>{});
  return false;
}

using AxisLens = PList<64, 80, 96, 112, 128, 160, 192, 224, 256, 320>;
using XLens = PList<64, 80, 96, 112, 128, 192, 256>;

template <typename K>
static void set_lds(K kernel, size_t lds) {
  if (lds > 65536)
    VIE_HIP_CHECK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
}

// tile widths (pencils per block) compiled for the axis passes; the context
// picks one per pass and precision (see autotune_tiles)
template <typename T> struct AxisTiles;
template <> struct AxisTiles<double> { using L = PList<4, 8>; };
template <> struct AxisTiles<float> { using L = PList<4, 8, 16>; };

template <typename T, int IN, int OUT, bool INV>
static bool axis_variant(int P, int TX, int sched, const AxisArgs& a, dim3 grid, size_t lds, hipStream_t s,
                         const void* tw) {
  const bool rf = sched == kSchedRF;
  const auto* twc = static_cast<const V2<T>*>(tw);
  return dispatch_len(P, [&](auto pc) {
    constexpr int PP = decltype(pc)::value;
    const bool ok = dispatch_len(TX, [&](auto tc) {
      constexpr int TXC = decltype(tc)::value;
      (void)lds;
      // grid.x = ntile * lines items, padded to whole rounds of 8 line groups
      constexpr unsigned R = 8u * line_group<T, TXC>();
      // the register kernel reads only the first half of a forward pencil and
      // writes only the first half of an inverse one
      if (rf && (INV ? a.Lout > PP / 2 : a.Lin > PP / 2))
        throw Error{VIE_ERR_ARG, "axis pass: pencil longer than half the FFT length"};
      AxisArgs b = a;
      b.nitem = grid.x;
      const dim3 g(((unsigned)grid.x + R - 1) / R * R, grid.y, grid.z);
      if (rf) {
        auto k = axis_pass_rf<T, IN, OUT, INV, PP, TXC>;
        const size_t bytes = axis_rf_lds_bytes<T, PP, TXC>();
        set_lds(k, bytes);
        hipLaunchKernelGGL(k, g, dim3(AxisRf<PP, TXC>::NT), bytes, s, b, twc);
      } else {
        auto k = sched == kSchedCTO ? axis_pass_ct<T, IN, OUT, INV, PP, TXC, 1> : axis_pass_ct<T, IN, OUT, INV, PP, TXC, 0>;
        const size_t bytes = axis_lds_bytes<T, PP, TXC>();
        set_lds(k, bytes);
        hipLaunchKernelGGL(k, g, dim3(kThreads), bytes, s, b, twc);
      }
    }, typename AxisTiles<T>::L{});
    if (!ok) throw Error{VIE_ERR_ARG, "axis pass: tile width without a compiled variant"};
  }, AxisLens{});
}

template <typename T>
static bool axis_typed(int in_kind, int out_kind, bool inv, int P, int TX, int sched, const AxisArgs& a, dim3 grid,
                       size_t lds, hipStream_t s, const void* tw) {
  if (in_packed32(in_kind) || out_packed32(out_kind))
    throw Error{VIE_ERR_ARG, "axis pass: packed FP32 kinds are launched by launch_axis_fast_fp32_p32"};
  if (!inv) {
    if (in_kind == IN_PACKED) return axis_variant<T, IN_PACKED, OUT_INTERNAL, false>(P, TX, sched, a, grid, lds, s, tw);
    if (in_kind == IN_DENSE) return axis_variant<T, IN_DENSE, OUT_INTERNAL, false>(P, TX, sched, a, grid, lds, s, tw);
    return axis_variant<T, IN_INTERNAL, OUT_INTERNAL, false>(P, TX, sched, a, grid, lds, s, tw);
  }
  if (out_kind == OUT_PACKED_VIE)
    return axis_variant<T, IN_INTERNAL, OUT_PACKED_VIE, true>(P, TX, sched, a, grid, lds, s, tw);
  if (out_kind == OUT_PACKED_VIEH)
    return axis_variant<T, IN_INTERNAL, OUT_PACKED_VIEH, true>(P, TX, sched, a, grid, lds, s, tw);
  if (out_kind == OUT_DENSE) return axis_variant<T, IN_INTERNAL, OUT_DENSE, true>(P, TX, sched, a, grid, lds, s, tw);
  return axis_variant<T, IN_INTERNAL, OUT_INTERNAL, true>(P, TX, sched, a, grid, lds, s, tw);
}

// the FP32 packed-I/O kinds of the first (z forward) and last (z inverse) pass
template <typename T>
static bool axis_typed_p32(int in_kind, int out_kind, bool inv, int P, int TX, int sched, const AxisArgs& a, dim3 grid,
                           size_t lds, hipStream_t s, const void* tw) {
  if (!inv && in_kind == IN_PACKED32 && out_kind == OUT_INTERNAL)
    return axis_variant<T, IN_PACKED32, OUT_INTERNAL, false>(P, TX, sched, a, grid, lds, s, tw);
  if (inv && in_kind == IN_INTERNAL && out_kind == OUT_PACKED_VIE32)
    return axis_variant<T, IN_INTERNAL, OUT_PACKED_VIE32, true>(P, TX, sched, a, grid, lds, s, tw);
  if (inv && in_kind == IN_INTERNAL && out_kind == OUT_PACKED_VIEH32)
    return axis_variant<T, IN_INTERNAL, OUT_PACKED_VIEH32, true>(P, TX, sched, a, grid, lds, s, tw);
  throw Error{VIE_ERR_ARG, "axis pass: unsupported packed FP32 pass"};
}

template <typename T, bool PWL, bool ISK, bool ADJ>
static bool x_variant(int P, const void* args, int64_t blocks, size_t lds, hipStream_t s, const void* tw) {
  XArgs<T> a = *static_cast<const XArgs<T>*>(args);
  const auto* twc = static_cast<const V2<T>*>(tw);
  (void)blocks;
  (void)lds;
  return dispatch_len(P, [&](auto pc) {
    constexpr int PP = decltype(pc)::value;
    constexpr int CPB = XBlock<PWL, ISK, PP>::CPB;
    const int64_t items = (a.work + CPB - 1) / CPB;
    a.per_xcd = (items + 7) / 8;
    if (a.O1 != PP / 2 + 1 || 2 * a.n1 > PP) throw Error{VIE_ERR_ARG, "x pass: length mismatch"};
    if (a.work >= ((int64_t)1 << 31)) throw Error{VIE_ERR_ARG, "x pass: more than 2^31 column items"};
    // an empty x range has no pencil element to read or write (the staged
    // kernel clamps its loads into [xlo, xhi])
    if (a.xlo > a.xhi || a.xlo < 0 || a.xhi >= a.n1) throw Error{VIE_ERR_ARG, "x pass: x range outside the pencil"};
    const dim3 grid((unsigned)(8 * a.per_xcd)), block(XBlock<PWL, ISK, PP>::threads);
    if constexpr (XStage<T, PWL, ISK, PP>::possible) {
      if (a.stage) {
        const size_t bytes = x_lds_bytes<T, PWL, ISK, PP, true>();
        auto k = x_fused_rf<T, PWL, ISK, ADJ, PP, true>;
        set_lds(k, bytes);
        hipLaunchKernelGGL(k, grid, block, bytes, s, a, twc);
        return;
      }
    }
    const size_t bytes = x_lds_bytes<T, PWL, ISK, PP, false>();
    auto k = x_fused_rf<T, PWL, ISK, ADJ, PP, false>;
    set_lds(k, bytes);
    hipLaunchKernelGGL(k, grid, block, bytes, s, a, twc);
  }, XLens{});
}

template <typename T>
static bool x_typed(bool pwl, bool isk, bool adj, int P, const void* args, int64_t blocks, size_t lds, hipStream_t s,
                    const void* tw) {
  if (pwl) {
    if (isk) return adj ? x_variant<T, true, true, true>(P, args, blocks, lds, s, tw)
                        : x_variant<T, true, true, false>(P, args, blocks, lds, s, tw);
    return adj ? x_variant<T, true, false, true>(P, args, blocks, lds, s, tw)
               : x_variant<T, true, false, false>(P, args, blocks, lds, s, tw);
  }
  if (isk) return adj ? x_variant<T, false, true, true>(P, args, blocks, lds, s, tw)
                      : x_variant<T, false, true, false>(P, args, blocks, lds, s, tw);
  return adj ? x_variant<T, false, false, true>(P, args, blocks, lds, s, tw)
             : x_variant<T, false, false, false>(P, args, blocks, lds, s, tw);
}

}  // namespace vie
