// Compile-time-length pass kernels for the FFT lengths of the benchmark and
// configuration grids (2 n_d in {64, 128, 160, 192, 224, 256, 320}); other
// lengths run the generic kernels of vie_kernels.hip.
//
// Same data flow as the generic kernels; differences:
//  * the Stockham plan, pencil count and tile width are constants, so all
//    index arithmetic folds and every stage loop unrolls;
//  * the fused x pass enumerates its block-product terms as template
//    constants (TermCT): a work item owns one frequency pair (X, P - X) and
//    the outputs of one vector component p, holds the ql inputs of that pair
//    in registers and issues all of its spectrum loads up front;
//  * the fused x pass walks its columns grouped by octant column (the up to
//    four mirror columns (+-Y, +-Z) that read the same octant spectrum are
//    adjacent), rhs fastest, in contiguous chunks per XCD, so the spectrum
//    column stays in that XCD's L2 while every mirror and rhs consumes it.

#include <utility>

#include "vie_fast.hpp"
#include "vie_passes.hpp"

namespace vie {

template <typename T, int IN, int OUT, bool INV, int P, int TX>
__global__ void __launch_bounds__(kThreads) axis_pass_ct(AxisArgs a, const V2<T>* __restrict__ tw) {
  using C = V2<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* bufA = reinterpret_cast<C*>(smem_raw);
  C* bufB = bufA + TX * P;
  const int tile = blockIdx.x % a.ntile;
  const int64_t line = blockIdx.x / a.ntile;
  const int cb = blockIdx.y;
  const int slot = cb / a.ql, ch = cb - slot * a.ql;
  if (skip_slot(a.rmap, a.kdone, a.step, slot)) return;
  const int64_t rhs = a.rmap ? a.rmap[slot] : slot;
  const int x0 = tile * TX;
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
  C* res = fft_ct<P, 1, TX, INV, true>(bufA, bufB, tw, tid);
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
template <int P> struct XSplit;
template <> struct XSplit<64> { static constexpr int A = 8, B = 8; };
template <> struct XSplit<128> { static constexpr int A = 16, B = 8; };
template <> struct XSplit<192> { static constexpr int A = 24, B = 8; };
template <> struct XSplit<256> { static constexpr int A = 16, B = 16; };

constexpr int kXThreads = 192;

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

// block product for the frequency pair (t, P - t) (t = 0: (0, P/2)) of one
// column, in place in the column's natural-order spectrum S[c][k] (stride LCC)
template <typename T, bool PWL, bool ISK, bool ADJ, int P>
__device__ __forceinline__ void product_pair(V2<T>* S, int LCC, const V2<T>* __restrict__ sp, int O1, int t,
                                             int colbits) {
  using C = V2<T>;
  using Sh = OpShape<PWL, ISK>;
  constexpr int QL = Sh::QL, H = P / 2;
  const int X1 = t, X2 = t ? P - t : H;
  C in1[QL], in2[QL], acc1[QL], acc2[QL];
#pragma unroll
  for (int v = 0; v < QL; ++v) {
    in1[v] = S[v * LCC + X1];
    in2[v] = S[v * LCC + X2];
    acc1[v] = mk<T>(0, 0);
    acc2[v] = mk<T>(0, 0);
  }
  constexpr TermList L = Terms<PWL, ISK, ADJ>::L;
  C G1 = mk<T>(0, 0), G2 = mk<T>(0, 0);
  static_for<0, L.n>([&](auto I) {
    constexpr int i = I.value;
    constexpr int g = L.g[i];
    if constexpr (i == 0 || L.g[i == 0 ? 0 : i - 1] != g) {
      constexpr int bits = gbits<ISK, PWL>(g);
      G1 = sp[g * O1 + t];
      G2 = t ? G1 : sp[g * O1 + H];
      // parity of the mirrored (Y, Z) column, and of the mirrored x frequency
      const bool ncol = __popc(bits & colbits) & 1;
      const bool nx2 = ncol != ((bits & 1) && t != 0);
      if (ncol) { G1.x = -G1.x; G1.y = -G1.y; }
      if (nx2) { G2.x = -G2.x; G2.y = -G2.y; }
    }
    constexpr int u = L.u[i], v = L.v[i], s = L.s[i];
    const C p1 = ADJ ? cmulc(in1[v], G1) : cmul(in1[v], G1);
    const C p2 = ADJ ? cmulc(in2[v], G2) : cmul(in2[v], G2);
    if constexpr (s > 0) {
      acc1[u].x += p1.x; acc1[u].y += p1.y; acc2[u].x += p2.x; acc2[u].y += p2.y;
    } else {
      acc1[u].x -= p1.x; acc1[u].y -= p1.y; acc2[u].x -= p2.x; acc2[u].y -= p2.y;
    }
  });
#pragma unroll
  for (int u = 0; u < QL; ++u) {
    S[u * LCC + X1] = acc1[u];
    S[u * LCC + X2] = acc2[u];
  }
}

template <typename T, bool PWL, bool ISK, bool ADJ, int P>
__global__ void __launch_bounds__(kXThreads) x_fused_rf(XArgs<T> a, const V2<T>* __restrict__ tw) {
  using C = V2<T>;
  using Sh = OpShape<PWL, ISK>;
  constexpr int QL = Sh::QL, A = XSplit<P>::A, B = XSplit<P>::B, H = P / 2;
  constexpr int TPC = QL * B;                // threads per column
  constexpr int CPB = kXThreads / TPC;       // columns per block
  constexpr int LCC = A * (B + 1);           // LDS complex values per channel (padded transpose)
  static_assert(CPB >= 1 && A % B == 0, "x split");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int tid = threadIdx.x;
  const int cl = tid / TPC;                  // column within the block
  const int r = tid - cl * TPC;
  const int c = r / B, j = r - c * B;        // channel, lane within the pencil
  C* Lc = reinterpret_cast<C*>(smem_raw) + (size_t)cl * QL * LCC;   // this column's region
  // work item of this column (XCD-chunked order, see XArgs)
  const int64_t q = (int64_t)(blockIdx.x % 8) * a.per_xcd + blockIdx.x / 8;
  const int64_t w = q * CPB + cl;
  bool valid = cl < CPB && w < a.work;
  int slot = 0, Y = 0, Z = 0;
  if (valid) {
    slot = (int)(w % a.nslot);
    const int cm = a.colmap[w / a.nslot];
    Y = cm & 0xffff;
    Z = cm >> 16;
    valid = !skip_slot(a.rmap, a.kdone, a.step, slot);
  }
  const int n1 = a.n1;
  const int64_t gbase = (int64_t)slot * QL * a.cstride + (int64_t)n1 * (Y + (int64_t)a.P2 * Z) + c * a.cstride;

  // forward step 1: A-point DFT over m of x[j + B m] (only m < A/2 are inside the n1 = P/2 inputs)
  C x[A];
#pragma unroll
  for (int m = 0; m < A; ++m) {
    const int e = j + B * m;
    x[m] = (valid && e < n1) ? a.data[gbase + e] : mk<T>(0, 0);
  }
  reg_fft<A, false>(x);
#pragma unroll
  for (int k1 = 1; k1 < A; ++k1)
    if (j) x[k1] = cmul(x[k1], tw[j * k1]);
  if (cl < CPB) {
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) Lc[c * LCC + k1 * (B + 1) + j] = x[k1];
  }
  __syncthreads();
  // forward step 2: B-point DFTs over j for k1 = j + B s
  constexpr int NS = A / B;
  C y[NS][B];
  if (cl < CPB) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int k1 = j + B * s;
#pragma unroll
      for (int jj = 0; jj < B; ++jj) y[s][jj] = Lc[c * LCC + k1 * (B + 1) + jj];
      reg_fft<B, false>(y[s]);
    }
  }
  __syncthreads();
  if (cl < CPB) {
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) Lc[c * LCC + (j + B * s) + A * k2] = y[s][k2];
  }
  __syncthreads();
  // block product: one frequency pair per item, all ql outputs, in place
  if (valid) {
    const int yo = (2 * Y <= a.P2) ? Y : a.P2 - Y;
    const int zo = (2 * Z <= a.P3) ? Z : a.P3 - Z;
    const int colbits = ((2 * Y > a.P2) ? 2 : 0) | ((2 * Z > a.P3) ? 4 : 0);
    const C* sp = a.spec + ((int64_t)zo * a.O2 + yo) * Sh::NG * a.O1;
    for (int t = r; t < H; t += TPC) product_pair<T, PWL, ISK, ADJ, P>(Lc, LCC, sp, a.O1, t, colbits);
  }
  __syncthreads();
  // inverse step 1
  if (cl < CPB) {
#pragma unroll
    for (int m = 0; m < A; ++m) x[m] = Lc[c * LCC + j + B * m];
  }
  reg_fft<A, true>(x);
#pragma unroll
  for (int k1 = 1; k1 < A; ++k1)
    if (j) x[k1] = cmulc(x[k1], tw[j * k1]);
  __syncthreads();
  if (cl < CPB) {
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) Lc[c * LCC + k1 * (B + 1) + j] = x[k1];
  }
  __syncthreads();
  // inverse step 2; keep the outputs k1 + A k2 < n1 (k2 < B/2)
  if (valid) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int k1 = j + B * s;
#pragma unroll
      for (int jj = 0; jj < B; ++jj) y[s][jj] = Lc[c * LCC + k1 * (B + 1) + jj];
      reg_fft<B, true>(y[s]);
#pragma unroll
      for (int k2 = 0; k2 < B / 2; ++k2) a.data[gbase + k1 + A * k2] = y[s][k2];
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

using AxisLens = PList<64, 128, 160, 192, 224, 256, 320>;
using XLens = PList<64, 128, 192, 256>;

template <typename K>
static void set_lds(K kernel, size_t lds) {
  if (lds > 65536)
    VIE_HIP_CHECK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
}

template <typename T, int IN, int OUT, bool INV>
static bool axis_variant(int P, int TX, const AxisArgs& a, dim3 grid, size_t lds, hipStream_t s, const void* tw) {
  const auto* twc = static_cast<const V2<T>*>(tw);
  return dispatch_len(P, [&](auto pc) {
    constexpr int PP = decltype(pc)::value;
    constexpr int TXC = sizeof(T) == 8 ? 8 : 16;
    if (TX != TXC) throw Error{VIE_ERR_ARG, "tile width mismatch"};
    auto k = axis_pass_ct<T, IN, OUT, INV, PP, TXC>;
    set_lds(k, lds);
    hipLaunchKernelGGL(k, grid, dim3(kThreads), lds, s, a, twc);
  }, AxisLens{});
}

template <typename T>
static bool axis_typed(int in_kind, int out_kind, bool inv, int P, int TX, const AxisArgs& a, dim3 grid, size_t lds,
                       hipStream_t s, const void* tw) {
  if (!inv) {
    if (in_kind == IN_PACKED) return axis_variant<T, IN_PACKED, OUT_INTERNAL, false>(P, TX, a, grid, lds, s, tw);
    if (in_kind == IN_DENSE) return axis_variant<T, IN_DENSE, OUT_INTERNAL, false>(P, TX, a, grid, lds, s, tw);
    return axis_variant<T, IN_INTERNAL, OUT_INTERNAL, false>(P, TX, a, grid, lds, s, tw);
  }
  if (out_kind == OUT_PACKED_VIE) return axis_variant<T, IN_INTERNAL, OUT_PACKED_VIE, true>(P, TX, a, grid, lds, s, tw);
  if (out_kind == OUT_PACKED_VIEH) return axis_variant<T, IN_INTERNAL, OUT_PACKED_VIEH, true>(P, TX, a, grid, lds, s, tw);
  if (out_kind == OUT_DENSE) return axis_variant<T, IN_INTERNAL, OUT_DENSE, true>(P, TX, a, grid, lds, s, tw);
  return axis_variant<T, IN_INTERNAL, OUT_INTERNAL, true>(P, TX, a, grid, lds, s, tw);
}

bool launch_axis_fast(int in_kind, int out_kind, bool inv, bool fp64, int P, int TX, const AxisArgs& a, dim3 grid,
                      size_t lds, hipStream_t s, const void* tw) {
  return fp64 ? axis_typed<double>(in_kind, out_kind, inv, P, TX, a, grid, lds, s, tw)
              : axis_typed<float>(in_kind, out_kind, inv, P, TX, a, grid, lds, s, tw);
}

template <typename T, bool PWL, bool ISK, bool ADJ>
static bool x_variant(int P, const void* args, int64_t blocks, size_t lds, hipStream_t s, const void* tw) {
  XArgs<T> a = *static_cast<const XArgs<T>*>(args);
  const auto* twc = static_cast<const V2<T>*>(tw);
  (void)blocks;
  (void)lds;
  return dispatch_len(P, [&](auto pc) {
    constexpr int PP = decltype(pc)::value;
    constexpr int QL = OpShape<PWL, ISK>::QL, A = XSplit<PP>::A, B = XSplit<PP>::B;
    constexpr int CPB = kXThreads / (QL * B);
    const int64_t items = (a.work + CPB - 1) / CPB;
    a.per_xcd = (items + 7) / 8;
    const size_t bytes = (size_t)CPB * QL * A * (B + 1) * sizeof(V2<T>);
    auto k = x_fused_rf<T, PWL, ISK, ADJ, PP>;
    set_lds(k, bytes);
    hipLaunchKernelGGL(k, dim3((unsigned)(8 * a.per_xcd)), dim3(kXThreads), bytes, s, a, twc);
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

bool launch_x_fast(bool fp64, bool pwl, bool isk, bool adj, int P, const void* args, int64_t blocks, size_t lds,
                   hipStream_t s, const void* tw) {
  return fp64 ? x_typed<double>(pwl, isk, adj, P, args, blocks, lds, s, tw)
              : x_typed<float>(pwl, isk, adj, P, args, blocks, lds, s, tw);
}

}  // namespace vie
