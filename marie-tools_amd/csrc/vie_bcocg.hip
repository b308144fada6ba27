// Device-resident block COCG on the VIE operator: the rhs of a batch share
// one block Krylov space.  No counterpart in the reference, whose solver is
// GMRES one rhs at a time (is_gmres_vie.m, built in vie_gmres.hip).
//
// The packed VIE operator A is symmetric in the weighted bilinear form
// [U, V] = U^T W V, W = diag(w) with the weights of vie_cocr.hip (the
// derivation is there).  Block CG (O'Leary 1980) carries over to that form:
// with residual block R and direction block P (n x s, one column per rhs),
//   alpha = [P, A P]^-1 [R, R],  X += P alpha,  R -= A P alpha,
//   beta = [R_old, R_old]^-1 [R, R],  P = R + P beta.
// The residual columns of different rhs become nearly dependent as the
// solve proceeds, so R is kept as R = Q C with Q^H Q = I (Euclidean) and
// C (s x s, upper triangular) on the host, and P = S C (Dubrulle 2001,
// "BCGrQ").  Substituting, with M = [S, A S] and N = [Q, Q]:
//   Z = A S,  X += S (M^-1 N C),  Q' = Q - Z M^-1 N,
//   Q' = Qn zeta (Cholesky QR),  C = zeta C,
//   S = Qn + S N^-1 zeta^T Nn  with Nn = [Qn, Qn].
// ||r_j|| = ||C e_j|| because Q is orthonormal.  One iteration applies the
// operator to all s columns at once (one batched application) and makes
// three passes over the n x s blocks: M; X and Q' with Q'^H Q' and
// [Q', Q']; Qn = Q' zeta^-1 with S.  The s x s algebra (LU solves, the
// Cholesky factor, triangular inverses) runs on the host in FP64 between
// the passes.
//
// Precision.  FP64 form: every block FP64.  When every column meets the
// tolerance, R is replaced by B - A X with the FP64 operator: Q, C = QR(R)
// and S = S C_old C^-1, which keeps the direction block P = S C; a solve ends
// only when every FP64 residual satisfies ||b_j - A x_j|| < tol ||b_j||.
// Mixed form: two phases.  The FP32 operator reads an FP32 copy of S
// (written by the S update) and writes Z in FP32, X, Q, S, C and every
// reduction FP64, until ||C||_F has fallen by kBlkSwitch; then the
// recurrence restarts from the FP64 residual and runs the FP64 form.  (With
// the FP32 operator to the end the block recurrence stalls near 1e-3 at the
// 96x96x112 torso, replacements or not.)
//
// Breakdown.  A singular M or N, or residual columns that are numerically
// dependent (a Cholesky pivot below 1e-14 of its column's squared norm: the
// block Krylov space is exhausted, or rhs are dependent) restarts the
// recurrence from the FP64 residual B - A X; after kBlkRestarts restarts, or
// when the residual itself is dependent (e.g. two equal rhs), the rhs of the
// group finish with vie_cocr warm-started from the block iterate.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <vector>

#include "vie_internal.hpp"
#include "vie_reduce.hpp"

namespace vie {

constexpr int kBlkMax = 32;      // rhs per block; larger batches run in groups
constexpr int kBlkRows = 64;     // rows of a tile
constexpr int kBlkLd = kBlkRows + 1;   // LDS row stride of a tile (bank spread)
constexpr int kBlkGrid = 512;    // workgroups of a pass (two per CU)
// mixed form: the FP32 operator until ||C||_F has fallen by this factor
constexpr double kBlkSwitch = 1e-1;
// restarts of the recurrence from the FP64 residual after breakdowns, before
// the group falls back to COCR
constexpr int kBlkRestarts = 3;

typedef std::complex<double> cd;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// acc += a b as four fused multiply-adds
__device__ __forceinline__ double2 cmac(double2 acc, double2 a, double2 b) {
  acc.x = __builtin_fma(a.x, b.x, acc.x);
  acc.x = __builtin_fma(-a.y, b.y, acc.x);
  acc.y = __builtin_fma(a.x, b.y, acc.y);
  acc.y = __builtin_fma(a.y, b.x, acc.y);
  return acc;
}
// acc += conj(a) b
__device__ __forceinline__ double2 cmacc(double2 acc, double2 a, double2 b) {
  acc.x = __builtin_fma(a.x, b.x, acc.x);
  acc.x = __builtin_fma(a.y, b.y, acc.x);
  acc.y = __builtin_fma(a.x, b.y, acc.y);
  acc.y = __builtin_fma(-a.y, b.x, acc.y);
  return acc;
}

// weights of the bilinear form per unknown i = c nS + s: wv[s] gw[c]
struct BlkW {
  const double2* wv;
  const double* gw;
  int64_t nS;
};

__global__ void k_blk_weights(const double2* __restrict__ mcr, int64_t nS, int adj, double2* __restrict__ wv) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= nS) return;
  const double2 m = mcr[s];
  if (adj) {
    wv[s] = make_double2(m.x, -m.y);
  } else {
    const double den = m.x * m.x + m.y * m.y;
    wv[s] = den > 0 ? make_double2(m.x / den, -m.y / den) : make_double2(0.0, 0.0);
  }
}

// A tile holds rows i0 .. i0+63 of the SB columns of an n x s block stored
// column-major (column k at base + k n): t[k * kBlkLd + r].  Columns k >= s
// and rows past n read as 0.
template <int SB, typename TV>
__device__ __forceinline__ void blk_load(double2* t, const TV* __restrict__ base, int64_t n, int s, int64_t i0) {
  for (int e = threadIdx.x; e < SB * kBlkRows; e += kGmBlock) {
    const int k = e / kBlkRows, r = e % kBlkRows;
    const int64_t i = i0 + r;
    t[k * kBlkLd + r] = (k < s && i < n) ? ld_v(base + (int64_t)k * n + i) : make_double2(0.0, 0.0);
  }
}

// A tile held in registers between its global loads and its LDS stores, so
// that a pass loads the next tile while it computes the current one.  Thread
// t holds elements e = t + 256 j, the same (column, row) pairs as blk_load.
template <int SB>
struct BlkRegs {
  double2 v[SB * kBlkRows / kGmBlock];
};
template <int SB, typename TV>
__device__ __forceinline__ void blk_fetch(BlkRegs<SB>& p, const TV* __restrict__ base, int64_t n, int s, int64_t i0) {
#pragma unroll
  for (int j = 0; j < SB * kBlkRows / kGmBlock; ++j) {
    const int e = threadIdx.x + j * kGmBlock;
    const int k = e / kBlkRows, r = e % kBlkRows;
    const int64_t i = i0 + r;
    p.v[j] = (k < s && i < n) ? ld_v(base + (int64_t)k * n + i) : make_double2(0.0, 0.0);
  }
}
template <int SB>
__device__ __forceinline__ void blk_put(double2* t, const BlkRegs<SB>& p, const double2* wr) {
#pragma unroll
  for (int j = 0; j < SB * kBlkRows / kGmBlock; ++j) {
    const int e = threadIdx.x + j * kGmBlock;
    const int k = e / kBlkRows, r = e % kBlkRows;
    t[k * kBlkLd + r] = wr ? cmul(p.v[j], wr[r]) : p.v[j];
  }
}

// weight of each row of the tile (rows past n: 0)
__device__ __forceinline__ void blk_row_weights(double2* wr, BlkW w, int64_t n, int64_t i0) {
  if (threadIdx.x < kBlkRows) {
    const int64_t i = i0 + threadIdx.x;
    double2 v = make_double2(0.0, 0.0);
    if (i < n) {
      const int64_t c = i / w.nS, s = i - c * w.nS;
      const double g = w.gw[c];
      v = w.wv[s];
      v.x *= g;
      v.y *= g;
    }
    wr[threadIdx.x] = v;
  }
}

// out = tile x (s x s matrix T, row-major with leading dimension SB, padded
// with zeros), out[r][b] = sum_k t[k][r] T[k][b], on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64:
// D[16x16] += A[16x4] B[4x16]; lane l supplies A[l%16][l/16] and
// B[l/16][l%16] and holds D[l/16 + 4v][l%16] in component v).  Wave w
// computes rows 16w .. 16w+15 of the tile, all SB columns: a complex product
// is four real ones, Re += Ar Br - Ai Bi, Im += Ar Bi + Ai Br.  The s x s
// operand comes per lane from device memory (one 16-byte load per 4 x 16
// fragment, cached).  Output column block cb, component v of lane l is tile
// row 16w + l/16 + 4v, column 16cb + l%16.
typedef double blk_d4 __attribute__((ext_vector_type(4)));
template <int SB>
__device__ __forceinline__ void blk_matmul_mfma(const double2* t, const double2* __restrict__ T, int s,
                                                blk_d4 (&re)[SB / 16], blk_d4 (&im)[SB / 16]) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int cb = 0; cb < SB / 16; ++cb) {
    re[cb] = blk_d4{0.0, 0.0, 0.0, 0.0};
    im[cb] = blk_d4{0.0, 0.0, 0.0, 0.0};
  }
  const int nkc = (s + 3) >> 2;
  for (int kc = 0; kc < nkc; ++kc) {
    const int k = 4 * kc + lk;   // rows k >= s of T are zero (padded by the host)
    const double2 a = t[k * kBlkLd + 16 * wave + li];
    const double nai = -a.y;
#pragma unroll
    for (int cb = 0; cb < SB / 16; ++cb) {
      const double2 b = T[k * SB + 16 * cb + li];
      re[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, re[cb], 0, 0, 0);
      re[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai, b.y, re[cb], 0, 0, 0);
      im[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, im[cb], 0, 0, 0);
      im[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, im[cb], 0, 0, 0);
    }
  }
}

// Gram accumulation over one tile: G[a][b] += sum_r op(u[a][r]) v[b][r]
// (op = conj when CONJ), optionally with the row weight wr[r] on v.  Wave w
// takes rows 16w .. 16w+15; lane (la, lb) of an 8 x 8 grid the PA x PA
// entries a = la PA + p, b = lb PA + q.
template <int SB, bool CONJ>
__device__ __forceinline__ void blk_gram(const double2* u, const double2* v, const double2* wr,
                                         double2 (&acc)[SB / 8][SB / 8]) {
  constexpr int PA = SB / 8;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int la = lane >> 3, lb = lane & 7;
#pragma unroll 2
  for (int rr = 0; rr < kBlkRows / 4; ++rr) {
    const int r = wave * (kBlkRows / 4) + rr;
    double2 ua[PA], vb[PA];
#pragma unroll
    for (int p = 0; p < PA; ++p) ua[p] = u[(la * PA + p) * kBlkLd + r];
#pragma unroll
    for (int q = 0; q < PA; ++q) {
      vb[q] = v[(lb * PA + q) * kBlkLd + r];
      if (wr) vb[q] = cmul(vb[q], wr[r]);
    }
#pragma unroll
    for (int p = 0; p < PA; ++p)
#pragma unroll
      for (int q = 0; q < PA; ++q) acc[p][q] = CONJ ? cmacc(acc[p][q], ua[p], vb[q]) : cmac(acc[p][q], ua[p], vb[q]);
  }
}

template <int SB>
__device__ __forceinline__ void blk_zero(double2 (&acc)[SB / 8][SB / 8]) {
#pragma unroll
  for (int p = 0; p < SB / 8; ++p)
#pragma unroll
    for (int q = 0; q < SB / 8; ++q) acc[p][q] = make_double2(0.0, 0.0);
}

// sum of the four waves' accumulators (through LDS scratch red[SB*SB]),
// stored to part[blockIdx.x][SB*SB]
template <int SB>
__device__ __forceinline__ void blk_gram_store(double2 (&acc)[SB / 8][SB / 8], double2* red,
                                               double2* __restrict__ part) {
  constexpr int PA = SB / 8;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int la = lane >> 3, lb = lane & 7;
  __syncthreads();
  for (int w = 0; w < 4; ++w) {
    if (wave == w) {
#pragma unroll
      for (int p = 0; p < PA; ++p)
#pragma unroll
        for (int q = 0; q < PA; ++q) {
          const int e = (la * PA + p) * SB + lb * PA + q;
          double2 t = acc[p][q];
          if (w > 0) {
            t.x += red[e].x;
            t.y += red[e].y;
          }
          red[e] = t;
        }
    }
    __syncthreads();
  }
  double2* dst = part + (int64_t)blockIdx.x * SB * SB;
  for (int e = threadIdx.x; e < SB * SB; e += kGmBlock) dst[e] = red[e];
}

// G = sum over blocks of part: ng Gram matrices of sb2 entries each, stored
// part[g][blk][sb2]; grid ceil(ng sb2 / 64), a fixed summation order
__global__ void __launch_bounds__(kGmBlock) k_blk_reduce(const double2* __restrict__ part, int nblk, int len, int sb2,
                                                         double2* __restrict__ G) {
  __shared__ double2 red[4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int e = blockIdx.x * 64 + lane;
  double2 s = make_double2(0.0, 0.0);
  if (e < len) {
    const int g = e / sb2;
    const double2* src = part + (int64_t)g * nblk * sb2 + (e - g * sb2);
    for (int b = wave; b < nblk; b += 4) {
      const double2 v = src[(int64_t)b * sb2];
      s.x += v.x;
      s.y += v.y;
    }
  }
  red[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && e < len) {
    double2 t = red[0][lane];
    for (int w = 1; w < 4; ++w) {
      t.x += red[w][lane].x;
      t.y += red[w][lane].y;
    }
    G[e] = t;
  }
}

// dynamic LDS of the passes: two tiles, the row weights
template <int SB>
constexpr size_t blk_lds() {
  return 2 * (size_t)SB * kBlkLd * sizeof(double2) + kBlkRows * sizeof(double2);
}

// M = S^T W Z (the bilinear Gram of the direction block and its image); the
// next tile is loaded while the current one is summed
template <int SB, typename TZ>
__global__ void __launch_bounds__(kGmBlock) k_blk_gram_sz(const double2* __restrict__ S, const TZ* __restrict__ Z,
                                                          int64_t n, int s, BlkW w, double2* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double2* t0 = reinterpret_cast<double2*>(smem_raw);
  double2* t1 = t0 + SB * kBlkLd;
  double2* wr = t1 + SB * kBlkLd;
  double2 acc[SB / 8][SB / 8];
  blk_zero<SB>(acc);
  const int64_t ntile = (n + kBlkRows - 1) / kBlkRows;
  BlkRegs<SB> ps, pz;
  int64_t tl = blockIdx.x;
  if (tl < ntile) {
    blk_fetch<SB>(ps, S, n, s, tl * kBlkRows);
    blk_fetch<SB>(pz, Z, n, s, tl * kBlkRows);
  }
  for (; tl < ntile; tl += gridDim.x) {
    const int64_t i0 = tl * kBlkRows;
    blk_row_weights(wr, w, n, i0);
    __syncthreads();
    blk_put<SB>(t0, ps, nullptr);
    blk_put<SB>(t1, pz, wr);
    __syncthreads();
    const int64_t tn = tl + gridDim.x;
    if (tn < ntile) {
      blk_fetch<SB>(ps, S, n, s, tn * kBlkRows);
      blk_fetch<SB>(pz, Z, n, s, tn * kBlkRows);
    }
    blk_gram<SB, false>(t0, t1, nullptr, acc);
    __syncthreads();
  }
  blk_gram_store<SB>(acc, t0, part);
}

// X += S F,  Q = Q - Z G,  and the partials of Q^H Q (first SB*SB of each
// block's part) and Q^T W Q (second)
template <int SB, typename TZ>
__global__ void __launch_bounds__(kGmBlock, 2) k_blk_step_a(const double2* __restrict__ S, const TZ* __restrict__ Z,
                                                         double2* __restrict__ X, double2* __restrict__ Q, int64_t n,
                                                         int s, const double2* __restrict__ F,
                                                         const double2* __restrict__ G, BlkW w,
                                                         double2* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double2* t0 = reinterpret_cast<double2*>(smem_raw);
  double2* t1 = t0 + SB * kBlkLd;
  double2* wr = t1 + SB * kBlkLd;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  double2 acc[SB / 8][SB / 8], acw[SB / 8][SB / 8];
  blk_zero<SB>(acc);
  blk_zero<SB>(acw);
  const int64_t ntile = (n + kBlkRows - 1) / kBlkRows;
  for (int64_t tl = blockIdx.x; tl < ntile; tl += gridDim.x) {
    const int64_t i0 = tl * kBlkRows;
    blk_row_weights(wr, w, n, i0);
    blk_load<SB>(t0, S, n, s, i0);
    blk_load<SB>(t1, Z, n, s, i0);
    __syncthreads();
    blk_d4 ore[SB / 16], oim[SB / 16];
    blk_matmul_mfma<SB>(t0, F, s, ore, oim);
#pragma unroll
    for (int cb = 0; cb < SB / 16; ++cb) {
      const int b = 16 * cb + li;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t i = i0 + 16 * wave + lk + 4 * v;
        if (b < s && i < n) {
          double2* px = X + (int64_t)b * n + i;
          const double2 xo = *px;
          *px = make_double2(xo.x + ore[cb][v], xo.y + oim[cb][v]);
        }
      }
    }
    blk_matmul_mfma<SB>(t1, G, s, ore, oim);
    __syncthreads();
#pragma unroll
    for (int cb = 0; cb < SB / 16; ++cb) {
      const int b = 16 * cb + li;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = 16 * wave + lk + 4 * v;
        const int64_t i = i0 + r;
        double2 q = make_double2(0.0, 0.0);
        if (b < s && i < n) {
          double2* pq = Q + (int64_t)b * n + i;
          const double2 qo = *pq;
          q = make_double2(qo.x - ore[cb][v], qo.y - oim[cb][v]);
          *pq = q;
        }
        t0[b * kBlkLd + r] = q;
      }
    }
    __syncthreads();
    blk_gram<SB, true>(t0, t0, nullptr, acc);
    blk_gram<SB, false>(t0, t0, wr, acw);
    __syncthreads();
  }
  blk_gram_store<SB>(acc, t0, part);
  blk_gram_store<SB>(acw, t0, part + (int64_t)gridDim.x * SB * SB);
}

// Q = B - T (T = A X; null: Q = B) and the partials of Q^H Q and Q^T W Q
template <int SB>
__global__ void __launch_bounds__(kGmBlock) k_blk_resid(const double2* __restrict__ Bv, const double2* __restrict__ T,
                                                        double2* __restrict__ Q, int64_t n, int s, BlkW w,
                                                        double2* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double2* t0 = reinterpret_cast<double2*>(smem_raw);
  double2* wr = t0 + 2 * SB * kBlkLd;
  constexpr int NC = SB / 4;
  const int wave = threadIdx.x >> 6, r = threadIdx.x & 63;
  double2 acc[SB / 8][SB / 8], acw[SB / 8][SB / 8];
  blk_zero<SB>(acc);
  blk_zero<SB>(acw);
  const int64_t ntile = (n + kBlkRows - 1) / kBlkRows;
  for (int64_t tl = blockIdx.x; tl < ntile; tl += gridDim.x) {
    const int64_t i0 = tl * kBlkRows, i = i0 + r;
    blk_row_weights(wr, w, n, i0);
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      const int b = wave * NC + j;
      double2 q = make_double2(0.0, 0.0);
      if (b < s && i < n) {
        const int64_t o = (int64_t)b * n + i;
        q = Bv[o];
        if (T) {
          const double2 t = T[o];
          q = make_double2(q.x - t.x, q.y - t.y);
        }
        Q[o] = q;
      }
      t0[(wave * NC + j) * kBlkLd + r] = q;
    }
    __syncthreads();
    blk_gram<SB, true>(t0, t0, nullptr, acc);
    blk_gram<SB, false>(t0, t0, wr, acw);
    __syncthreads();
  }
  blk_gram_store<SB>(acc, t0, part);
  blk_gram_store<SB>(acw, t0, part + (int64_t)gridDim.x * SB * SB);
}

// Q = Q T1;  S = (addq ? Q : 0) + (usep ? S P : 0);  S32 = S rounded (if given);
// the next tile is loaded while the current one is multiplied
template <int SB>
__global__ void __launch_bounds__(kGmBlock) k_blk_step_c(double2* __restrict__ Q, double2* __restrict__ S,
                                                         float2* __restrict__ S32, int64_t n, int s,
                                                         const double2* __restrict__ T1,
                                                         const double2* __restrict__ P, int addq, int usep) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double2* t0 = reinterpret_cast<double2*>(smem_raw);
  double2* t1 = t0 + SB * kBlkLd;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t ntile = (n + kBlkRows - 1) / kBlkRows;
  BlkRegs<SB> pq, ps;
  int64_t tl = blockIdx.x;
  if (tl < ntile) {
    blk_fetch<SB>(pq, Q, n, s, tl * kBlkRows);
    if (usep) blk_fetch<SB>(ps, S, n, s, tl * kBlkRows);
  }
  for (; tl < ntile; tl += gridDim.x) {
    const int64_t i0 = tl * kBlkRows;
    blk_put<SB>(t0, pq, nullptr);
    if (usep) blk_put<SB>(t1, ps, nullptr);
    __syncthreads();
    // the next tile's rows are not written by this one
    const int64_t tn = tl + gridDim.x;
    if (tn < ntile) {
      blk_fetch<SB>(pq, Q, n, s, tn * kBlkRows);
      if (usep) blk_fetch<SB>(ps, S, n, s, tn * kBlkRows);
    }
    blk_d4 qre[SB / 16], qim[SB / 16], pre[SB / 16], pim[SB / 16];
    blk_matmul_mfma<SB>(t0, T1, s, qre, qim);
    if (usep) blk_matmul_mfma<SB>(t1, P, s, pre, pim);
#pragma unroll
    for (int cb = 0; cb < SB / 16; ++cb) {
      const int b = 16 * cb + li;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t i = i0 + 16 * wave + lk + 4 * v;
        if (b < s && i < n) {
          const int64_t e = (int64_t)b * n + i;
          Q[e] = make_double2(qre[cb][v], qim[cb][v]);
          double2 x = addq ? make_double2(qre[cb][v], qim[cb][v]) : make_double2(0.0, 0.0);
          if (usep) x = make_double2(x.x + pre[cb][v], x.y + pim[cb][v]);
          S[e] = x;
          if (S32) S32[e] = make_float2((float)x.x, (float)x.y);
        }
      }
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------------------
// host side: s x s algebra (row-major, leading dimension s)
// --------------------------------------------------------------------------
typedef std::vector<cd> Mat;

static Mat mat_mul(const Mat& A, const Mat& B, int s) {
  Mat C((size_t)s * s, 0.0);
  for (int i = 0; i < s; ++i)
    for (int k = 0; k < s; ++k) {
      const cd a = A[(size_t)i * s + k];
      if (a == 0.0) continue;
      for (int j = 0; j < s; ++j) C[(size_t)i * s + j] += a * B[(size_t)k * s + j];
    }
  return C;
}

static Mat mat_t(const Mat& A, int s) {   // plain transpose (no conjugation)
  Mat T((size_t)s * s);
  for (int i = 0; i < s; ++i)
    for (int j = 0; j < s; ++j) T[(size_t)j * s + i] = A[(size_t)i * s + j];
  return T;
}

// A^-1 B by Gaussian elimination with partial pivoting; false when a pivot
// is zero or not finite
static bool mat_solve(Mat A, Mat B, int s, Mat& X) {
  for (int k = 0; k < s; ++k) {
    int p = k;
    double best = std::abs(A[(size_t)k * s + k]);
    for (int i = k + 1; i < s; ++i)
      if (std::abs(A[(size_t)i * s + k]) > best) {
        best = std::abs(A[(size_t)i * s + k]);
        p = i;
      }
    if (!(best > 0) || !std::isfinite(best)) return false;
    if (p != k)
      for (int j = 0; j < s; ++j) {
        std::swap(A[(size_t)k * s + j], A[(size_t)p * s + j]);
        std::swap(B[(size_t)k * s + j], B[(size_t)p * s + j]);
      }
    const cd inv = 1.0 / A[(size_t)k * s + k];
    for (int i = k + 1; i < s; ++i) {
      const cd f = A[(size_t)i * s + k] * inv;
      if (f == 0.0) continue;
      for (int j = k; j < s; ++j) A[(size_t)i * s + j] -= f * A[(size_t)k * s + j];
      for (int j = 0; j < s; ++j) B[(size_t)i * s + j] -= f * B[(size_t)k * s + j];
    }
  }
  X.assign((size_t)s * s, 0.0);
  for (int j = 0; j < s; ++j)
    for (int i = s - 1; i >= 0; --i) {
      cd t = B[(size_t)i * s + j];
      for (int k = i + 1; k < s; ++k) t -= A[(size_t)i * s + k] * X[(size_t)k * s + j];
      X[(size_t)i * s + j] = t / A[(size_t)i * s + i];
    }
  for (const cd& v : X)
    if (!std::isfinite(v.real()) || !std::isfinite(v.imag())) return false;
  return true;
}

// G = R^H R with R upper triangular (G Hermitian positive definite); false
// when a column is numerically dependent on the ones before it (its pivot
// <= 1e-14 of its squared norm G_jj: it lies within ~1e-7 rad of their span)
// or G is not finite
static bool mat_chol(const Mat& G, int s, Mat& R) {
  R.assign((size_t)s * s, 0.0);
  for (int j = 0; j < s; ++j) {
    cd d = G[(size_t)j * s + j];
    for (int k = 0; k < j; ++k) d -= std::conj(R[(size_t)k * s + j]) * R[(size_t)k * s + j];
    const double dr = d.real();
    if (!(dr > 1e-14 * G[(size_t)j * s + j].real()) || !std::isfinite(dr)) return false;
    const double rjj = std::sqrt(dr);
    R[(size_t)j * s + j] = rjj;
    for (int i = j + 1; i < s; ++i) {
      cd t = G[(size_t)j * s + i];
      for (int k = 0; k < j; ++k) t -= std::conj(R[(size_t)k * s + j]) * R[(size_t)k * s + i];
      R[(size_t)j * s + i] = t / rjj;
    }
  }
  return true;
}

static Mat mat_inv_upper(const Mat& R, int s) {
  Mat X((size_t)s * s, 0.0);
  for (int j = 0; j < s; ++j) {
    X[(size_t)j * s + j] = 1.0 / R[(size_t)j * s + j];
    for (int i = j - 1; i >= 0; --i) {
      cd t = 0.0;
      for (int k = i + 1; k <= j; ++k) t += R[(size_t)i * s + k] * X[(size_t)k * s + j];
      X[(size_t)i * s + j] = -t / R[(size_t)i * s + i];
    }
  }
  return X;
}

// --------------------------------------------------------------------------
// host side: one group of s <= kBlkMax rhs
// --------------------------------------------------------------------------
struct BlkWork {
  int64_t n;
  int s, sb, nblk;
  bool mixed;
  double2 *X, *Bv, *Q, *S, *Z64;   // [s][n] FP64 (Z64: A S in the FP64 form, A X at a replacement)
  float2 *S32, *Z32;               // [s][n] FP32 (mixed form)
  double2* part;                   // [2][nblk][sb sb]: Gram partials of the passes
  double2* G;                      // [2 sb sb]: reduced Gram matrices
  double2* mats;                   // [3][sb sb]: matrices the passes read
  BlkW w;
  cd* hpin;                        // pinned: [2 sb sb] Gram matrices, [3 sb sb] matrices
};

template <int SB>
struct BlkOps {
  static size_t lds() { return blk_lds<SB>(); }
  static void attrs() {
    static bool done = false;
    if (done) return;
    const int l = (int)blk_lds<SB>();
    VIE_HIP_CHECK(hipFuncSetAttribute((const void*)k_blk_gram_sz<SB, double2>, hipFuncAttributeMaxDynamicSharedMemorySize, l));
    VIE_HIP_CHECK(hipFuncSetAttribute((const void*)k_blk_gram_sz<SB, float2>, hipFuncAttributeMaxDynamicSharedMemorySize, l));
    VIE_HIP_CHECK(hipFuncSetAttribute((const void*)k_blk_step_a<SB, double2>, hipFuncAttributeMaxDynamicSharedMemorySize, l));
    VIE_HIP_CHECK(hipFuncSetAttribute((const void*)k_blk_step_a<SB, float2>, hipFuncAttributeMaxDynamicSharedMemorySize, l));
    VIE_HIP_CHECK(hipFuncSetAttribute((const void*)k_blk_resid<SB>, hipFuncAttributeMaxDynamicSharedMemorySize, l));
    VIE_HIP_CHECK(hipFuncSetAttribute((const void*)k_blk_step_c<SB>, hipFuncAttributeMaxDynamicSharedMemorySize, l));
    done = true;
  }
  static void gram_sz(Ctx& c, const BlkWork& w) {
    if (w.mixed)
      hipLaunchKernelGGL((k_blk_gram_sz<SB, float2>), dim3(w.nblk), dim3(kGmBlock), lds(), c.stream, w.S, w.Z32, w.n,
                         w.s, w.w, w.part);
    else
      hipLaunchKernelGGL((k_blk_gram_sz<SB, double2>), dim3(w.nblk), dim3(kGmBlock), lds(), c.stream, w.S, w.Z64,
                         w.n, w.s, w.w, w.part);
  }
  static void step_a(Ctx& c, const BlkWork& w) {
    const double2* F = w.mats;
    const double2* G = w.mats + SB * SB;
    if (w.mixed)
      hipLaunchKernelGGL((k_blk_step_a<SB, float2>), dim3(w.nblk), dim3(kGmBlock), lds(), c.stream, w.S, w.Z32, w.X,
                         w.Q, w.n, w.s, F, G, w.w, w.part);
    else
      hipLaunchKernelGGL((k_blk_step_a<SB, double2>), dim3(w.nblk), dim3(kGmBlock), lds(), c.stream, w.S, w.Z64,
                         w.X, w.Q, w.n, w.s, F, G, w.w, w.part);
  }
  static void resid(Ctx& c, const BlkWork& w, bool with_t) {
    hipLaunchKernelGGL(k_blk_resid<SB>, dim3(w.nblk), dim3(kGmBlock), lds(), c.stream, w.Bv,
                       with_t ? w.Z64 : nullptr, w.Q, w.n, w.s, w.w, w.part);
  }
  static void step_c(Ctx& c, const BlkWork& w, bool addq, bool usep) {
    hipLaunchKernelGGL(k_blk_step_c<SB>, dim3(w.nblk), dim3(kGmBlock), lds(), c.stream, w.Q, w.S,
                       w.mixed ? w.S32 : nullptr, w.n, w.s, w.mats, w.mats + SB * SB, addq ? 1 : 0, usep ? 1 : 0);
  }
};

// block widths: 16 (s <= 16) and 32 columns, the matrix-core column blocks
static void blk_attrs(int sb) {
  if (sb == 16) BlkOps<16>::attrs();
  else BlkOps<32>::attrs();
}

#define VIE_BLK_DISPATCH(sb, call) \
  do {                             \
    if ((sb) == 16) BlkOps<16>::call; \
    else BlkOps<32>::call;            \
  } while (0)

// reduce ng Gram matrices (part[g][blk][sb sb]) and read them back (synchronises); the s x s
// parts land in out[0..ng)
static void blk_read_grams(Ctx& c, const BlkWork& w, int ng, std::vector<Mat>& out) {
  const int len = ng * w.sb * w.sb;
  hipLaunchKernelGGL(k_blk_reduce, dim3((len + 63) / 64), dim3(kGmBlock), 0, c.stream, w.part, w.nblk, len,
                     w.sb * w.sb, w.G);
  VIE_HIP_CHECK(hipGetLastError());
  VIE_HIP_CHECK(hipMemcpyAsync(w.hpin, w.G, len * sizeof(double2), hipMemcpyDeviceToHost, c.stream));
  VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
  out.resize(ng);
  for (int g = 0; g < ng; ++g) {
    out[g].assign((size_t)w.s * w.s, 0.0);
    for (int a = 0; a < w.s; ++a)
      for (int b = 0; b < w.s; ++b) out[g][(size_t)a * w.s + b] = w.hpin[(size_t)g * w.sb * w.sb + a * w.sb + b];
  }
}

// pad the s x s matrices to sb x sb and send them to w.mats[0..k)
static void blk_send(Ctx& c, const BlkWork& w, std::initializer_list<const Mat*> ms) {
  cd* h = w.hpin + 2 * (size_t)w.sb * w.sb;
  int k = 0;
  for (const Mat* m : ms) {
    cd* d = h + (size_t)k * w.sb * w.sb;
    std::fill(d, d + (size_t)w.sb * w.sb, cd(0.0));
    if (m)
      for (int a = 0; a < w.s; ++a)
        for (int b = 0; b < w.s; ++b) d[a * w.sb + b] = (*m)[(size_t)a * w.s + b];
    ++k;
  }
  VIE_HIP_CHECK(hipMemcpyAsync(w.mats, h, (size_t)k * w.sb * w.sb * sizeof(double2), hipMemcpyHostToDevice, c.stream));
}

static void blk_apply(Ctx& c, const BlkWork& w, int op, int prec, const void* in, void* out, bool io32) {
  ApplyIO io;
  io.in = static_cast<const double2*>(in);
  io.in_rs = w.n;
  io.out = static_cast<double2*>(out);
  io.out_rs = w.n;
  io.io32 = io32;
  apply_operator(c, op, prec, io, w.s);
}

// Orthonormalisation of the block in Q from its Gram matrices G[0] = Q^H Q
// and G[1] = Q^T W Q (Cholesky QR): Q = Qn zeta, Qn = Q T1 with T1 = zeta^-1
// (applied by the pass that follows), Nn = [Qn, Qn] = T1^T G[1] T1.  One
// Cholesky pass leaves Qn orthonormal to O(cond(Q)^2 eps); the recurrence
// holds in any basis of the residual block (it carries R = Q C and P = S C
// exactly), so the orthonormalisation serves conditioning only.  False if
// a column of Q is numerically dependent on the ones before it.
struct QrOut {
  Mat zeta, Nn, T1;
};
static bool blk_qr(const std::vector<Mat>& G, int s, QrOut& o) {
  if (!mat_chol(G[0], s, o.zeta)) return false;
  o.T1 = mat_inv_upper(o.zeta, s);
  o.Nn = mat_mul(mat_t(o.T1, s), mat_mul(G[1], o.T1, s), s);
  return true;
}

// Solve one group; returns false at a breakdown that restarts did not cure
// (X holds the last iterate).  its/nmv/nrep/rv per rhs of the group are
// appended to.
static bool bcocg_group(Ctx& c, BlkWork& w, int op, int prec, double tol, int maxit, const std::vector<double>& bn,
                        bool x0nz, std::vector<int>& its, std::vector<int>& nmv, std::vector<int>& nrep,
                        std::vector<std::vector<double>>& rv, bool& converged) {
  const int s = w.s;
  converged = false;
  // mixed form: the FP32 operator until ||C||_F has fallen by `sw`, then a
  // restart from the FP64 residual and the FP64 operator to the end
  double sw = kBlkSwitch;
  if (const char* e = std::getenv("VIE_BCOCG_SWITCH"))   // VIE_BCOCG_SWITCH=f: another switch factor
    if (std::atof(e) > 0 && std::atof(e) < 1) sw = std::atof(e);
  std::vector<Mat> g;
  std::vector<double> rn(s);
  auto all_below = [&](const std::vector<double>& r) {
    for (int j = 0; j < s; ++j)
      if (!(r[j] < tol * bn[j])) return false;
    return true;
  };
  auto fro = [&](const Mat& C) {
    double t = 0;
    for (const cd& v : C) t += std::norm(v);
    return std::sqrt(t);
  };
  // R = B - A64 X (ax) or B, its norms appended; true when all meet tol
  auto true_residual = [&](bool ax) {
    if (ax) {
      blk_apply(c, w, op, VIE_PREC_FP64, w.X, w.Z64, false);
      for (int j = 0; j < s; ++j) nmv[j] += 1;
    }
    VIE_BLK_DISPATCH(w.sb, resid(c, w, ax));
    blk_read_grams(c, w, 2, g);
    for (int j = 0; j < s; ++j) {
      rn[j] = std::sqrt(std::max(0.0, g[0][(size_t)j * s + j].real()));
      rv[j].push_back(rn[j] / bn[j]);
    }
    return all_below(rn);
  };
  Mat C, N;
  double fn_rep = 0;
  QrOut qr;
  // a fresh recurrence from the residual in Q (its Gram in g[0]): Q, C = QR(R), S = Q
  auto fresh = [&]() {
    if (!blk_qr(g, s, qr)) return false;
    C = qr.zeta;
    N = qr.Nn;
    blk_send(c, w, {&qr.T1, nullptr});
    VIE_BLK_DISPATCH(w.sb, step_c(c, w, true, false));
    fn_rep = fro(C);
    return true;
  };
  if (true_residual(x0nz)) {
    converged = true;
    return true;
  }
  if (!fresh()) return false;
  const double fn0 = fn_rep;
  int it = 0, restarts = 0;
  while (it < maxit) {
    // Z = A S
    if (w.mixed) blk_apply(c, w, op, VIE_PREC_FP32, w.S32, w.Z32, true);
    else blk_apply(c, w, op, VIE_PREC_FP64, w.S, w.Z64, false);
    VIE_BLK_DISPATCH(w.sb, gram_sz(c, w));
    blk_read_grams(c, w, 1, g);
    Mat MN, Pm;
    bool ok = mat_solve(g[0], N, s, MN);
    if (ok) {
      const Mat F = mat_mul(MN, C, s);
      blk_send(c, w, {&F, &MN});
      VIE_BLK_DISPATCH(w.sb, step_a(c, w));
      blk_read_grams(c, w, 2, g);
      ++it;
      for (int j = 0; j < s; ++j) {
        its[j] += 1;
        nmv[j] += 1;
      }
      ok = blk_qr(g, s, qr) && mat_solve(N, mat_mul(mat_t(qr.zeta, s), qr.Nn, s), s, Pm);
    } else {
      for (int j = 0; j < s; ++j) nmv[j] += 1;   // the operator application was spent
    }
    if (!ok) {
      // breakdown (a singular M or N, or dependent residual columns, as when
      // the block Krylov space is exhausted): restart from the FP64 residual
      if (++restarts > kBlkRestarts) return false;
      for (int j = 0; j < s; ++j) nrep[j] += 1;
      if (true_residual(true)) {
        converged = true;
        return true;
      }
      if (!fresh()) return false;
      continue;
    }
    blk_send(c, w, {&qr.T1, &Pm});
    VIE_BLK_DISPATCH(w.sb, step_c(c, w, true, true));
    C = mat_mul(qr.zeta, C, s);
    N = qr.Nn;
    for (int j = 0; j < s; ++j) {
      double t = 0;
      for (int i = 0; i < s; ++i) t += std::norm(C[(size_t)i * s + j]);
      rn[j] = std::sqrt(t);
      rv[j].push_back(rn[j] / bn[j]);
    }
    if (w.mixed && !all_below(rn) && fro(C) < sw * fn0 && it < maxit) {
      // end of the FP32 phase: restart from the FP64 residual in FP64
      w.mixed = false;
      for (int j = 0; j < s; ++j) nrep[j] += 1;
      if (true_residual(true)) {
        converged = true;
        return true;
      }
      if (!fresh()) return false;
      continue;
    }
    if (!all_below(rn) && it < maxit) continue;
    // replacement: R = B - A64 X, Q, C = QR(R), S = S C_old C^-1 (P = S C kept)
    for (int j = 0; j < s; ++j) nrep[j] += 1;
    if (true_residual(true)) {
      converged = true;
      return true;
    }
    if (it >= maxit) return true;
    if (!blk_qr(g, s, qr)) return false;
    const Mat Pr = mat_mul(C, mat_inv_upper(qr.zeta, s), s);
    blk_send(c, w, {&qr.T1, &Pr});
    VIE_BLK_DISPATCH(w.sb, step_c(c, w, false, true));
    C = qr.zeta;
    N = qr.Nn;
    fn_rep = fro(C);
  }
  return true;
}

}  // namespace vie

using namespace vie;

extern "C" int vie_bcocg(vie_ctx* h, int32_t op, int32_t prec, const double* b, double* x, int32_t nrhs, double tol,
                         int32_t maxit, int32_t mem, double* resvec, int32_t* nres, int32_t* nmatvec,
                         int32_t* nreplace) {
  if (!h) return VIE_ERR_ARG;
  Ctx& c = h->c;
  try {
    c.err.clear();
    if (op != VIE_OP_VIE && op != VIE_OP_VIEH) throw Error{VIE_ERR_ARG, "vie_bcocg: op must be VIE_OP_VIE or VIE_OP_VIEH"};
    if (prec != VIE_PREC_FP64 && prec != VIE_PREC_FP32) throw Error{VIE_ERR_ARG, "unknown precision"};
    if (nrhs < 1 || maxit < 1) throw Error{VIE_ERR_ARG, "vie_bcocg: bad nrhs/maxit"};
    if (!(tol > 0)) throw Error{VIE_ERR_ARG, "vie_bcocg: tol must be > 0"};
    if (!b || !x) throw Error{VIE_ERR_ARG, "NULL vector"};
    if (!c.have_material) throw Error{VIE_ERR_STATE, "material not set"};
    if (c.box) {   // packed operators on the scatterer's support box
      const int rc = vie_bcocg(c.box, op, prec, b, x, nrhs, tol, maxit, mem, resvec, nres, nmatvec, nreplace);
      if (rc != VIE_OK) c.err = "support box: " + c.box->c.err;
      return rc;
    }
    VIE_HIP_CHECK(hipSetDevice(c.device));
    ensure_ready(c, 0, VIE_PREC_FP64);
    if (prec == VIE_PREC_FP32) ensure_ready(c, 0, VIE_PREC_FP32);
    if (!c.gm_stream) {
      VIE_HIP_CHECK(hipStreamCreateWithFlags(&c.gm_stream, hipStreamNonBlocking));
      VIE_HIP_CHECK(hipEventCreateWithFlags(&c.gm_order, hipEventDisableTiming));
      VIE_HIP_CHECK(hipEventCreateWithFlags(&c.gm_cycle_ev, hipEventDisableTiming));
    }
    VIE_HIP_CHECK(hipEventRecord(c.gm_order, c.stream));
    VIE_HIP_CHECK(hipStreamWaitEvent(c.gm_stream, c.gm_order, 0));
    const int64_t n = (int64_t)c.ql * c.nS;
    if (n == 0) throw Error{VIE_ERR_ARG, "no scatterer voxels"};
    const int sfull = std::min<int>(nrhs, kBlkMax);
    const int sbmax = sfull <= 16 ? 16 : 32;
    const size_t blk = (size_t)sfull * n;
    // workspace: X, B, Q, S, Z64 FP64 and S32, Z32 FP32 of one group
    c.bk_v.ensure(5 * blk * sizeof(double2) + (prec == VIE_PREC_FP32 ? 2 * blk * sizeof(float2) : 0));
    c.bk_part.ensure((size_t)kBlkGrid * 2 * sbmax * sbmax * sizeof(double2) +
                     5 * (size_t)sbmax * sbmax * sizeof(double2));
    const size_t pin = 5 * (size_t)sbmax * sbmax * sizeof(double2);
    if (c.bk_pin_bytes < pin) {
      if (c.bk_pin) VIE_HIP_CHECK(hipHostFree(c.bk_pin));
      c.bk_pin = nullptr;
      c.bk_pin_bytes = 0;
      VIE_HIP_CHECK(hipHostMalloc(&c.bk_pin, pin, hipHostMallocDefault));
      c.bk_pin_bytes = pin;
    }
    // the solve runs on the solver stream; the caller's stream is restored
    struct StreamSwap {
      Ctx& c;
      hipStream_t user;
      ~StreamSwap() {
        (void)hipStreamSynchronize(c.stream);
        c.stream = user;
      }
    } swap{c, c.stream};
    c.stream = c.gm_stream;
    double2* v = c.bk_v.as<double2>();
    c.bk_w.ensure((((size_t)c.nS * sizeof(double2) + 255) & ~(size_t)255) + c.ql * sizeof(double));
    double2* wv = c.bk_w.as<double2>();
    double* gw = reinterpret_cast<double*>(reinterpret_cast<char*>(wv) + (((size_t)c.nS * sizeof(double2) + 255) & ~(size_t)255));
    const bool adj = op == VIE_OP_VIEH;
    hipLaunchKernelGGL(k_blk_weights, dim3((unsigned)((c.nS + 255) / 256)), dim3(256), 0, c.stream,
                       c.mcr.as<double2>(), c.nS, adj ? 1 : 0, wv);
    std::vector<double> gwh(c.ql);
    for (int q = 0; q < c.ql; ++q) gwh[q] = adj ? c.invg_h[q] : 1.0 / c.invg_h[q];
    VIE_HIP_CHECK(hipMemcpyAsync(gw, gwh.data(), c.ql * sizeof(double), hipMemcpyHostToDevice, c.stream));
    prepare_operator(c, op, VIE_PREC_FP64, sfull);
    if (prec == VIE_PREC_FP32) prepare_operator(c, op, VIE_PREC_FP32, sfull);
    const hipMemcpyKind h2d = mem == VIE_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind d2h = mem == VIE_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    const int cap = 2 * maxit + 2;
    for (int g0 = 0; g0 < nrhs; g0 += kBlkMax) {
      const int s = std::min(kBlkMax, nrhs - g0);
      BlkWork w{};
      w.n = n;
      w.s = s;
      w.sb = s <= 16 ? 16 : 32;
      w.nblk = (int)std::min<int64_t>(kBlkGrid, (n + kBlkRows - 1) / kBlkRows);
      w.mixed = prec == VIE_PREC_FP32;
      w.X = v;
      w.Bv = v + blk;
      w.Q = v + 2 * blk;
      w.S = v + 3 * blk;
      w.Z64 = v + 4 * blk;
      if (w.mixed) {
        w.S32 = reinterpret_cast<float2*>(v + 5 * blk);
        w.Z32 = w.S32 + blk;
      }
      w.part = c.bk_part.as<double2>();
      w.G = w.part + (size_t)kBlkGrid * 2 * sbmax * sbmax;
      w.mats = w.G + 2 * (size_t)sbmax * sbmax;
      w.w = BlkW{wv, gw, c.nS};
      w.hpin = static_cast<cd*>(c.bk_pin);
      blk_attrs(w.sb);
      const size_t bytes = (size_t)s * n * sizeof(double2);
      const double* bg = b + (size_t)g0 * n * 2;
      double* xg = x + (size_t)g0 * n * 2;
      VIE_HIP_CHECK(hipMemcpyAsync(w.Bv, bg, bytes, h2d, c.stream));
      VIE_HIP_CHECK(hipMemcpyAsync(w.X, xg, bytes, h2d, c.stream));
      // ||b_j|| and whether x0 is nonzero: the Hermitian Gram of B and of X
      std::vector<Mat> gg;
      VIE_BLK_DISPATCH(w.sb, resid(c, w, false));   // Q = B
      blk_read_grams(c, w, 1, gg);
      std::vector<double> bn(s);
      for (int j = 0; j < s; ++j) {
        bn[j] = std::sqrt(std::max(0.0, gg[0][(size_t)j * s + j].real()));
        if (!(bn[j] > 0)) throw Error{VIE_ERR_ARG, "right-hand side is zero"};
      }
      BlkWork wx = w;
      wx.Bv = w.X;
      VIE_BLK_DISPATCH(w.sb, resid(c, wx, false));   // Q = X
      blk_read_grams(c, w, 1, gg);
      bool x0nz = false;
      for (int j = 0; j < s; ++j) x0nz = x0nz || gg[0][(size_t)j * s + j].real() > 0;
      std::vector<int> its(s, 0), nmv(s, 0), nrep(s, 0);
      std::vector<std::vector<double>> rv(s);
      bool converged = false;
      const bool ok = bcocg_group(c, w, op, prec, tol, maxit, bn, x0nz, its, nmv, nrep, rv, converged);
      if (!ok) {
        // breakdown of the block recurrence: finish each rhs with COCR from the block iterate
        int left = maxit;
        for (int j = 0; j < s; ++j) left = std::min(left, maxit - its[j]);
        if (left > 0) {
          std::vector<int32_t> nmv2(s), nrep2(s), nres2(s);
          const int cap2 = 2 * left + 2;
          std::vector<double> rv2((size_t)s * cap2);
          c.stream = swap.user;   // vie_cocr orders itself after the caller's stream
          VIE_HIP_CHECK(hipStreamSynchronize(c.gm_stream));
          const int rc = vie_cocr(h, op, prec, reinterpret_cast<const double*>(w.Bv), reinterpret_cast<double*>(w.X),
                                  s, tol, left, VIE_MEM_DEVICE, rv2.data(), nres2.data(), nmv2.data(), nrep2.data());
          c.stream = c.gm_stream;
          if (rc != VIE_OK) throw Error{rc, "vie_bcocg fallback: " + c.err};
          for (int j = 0; j < s; ++j) {
            nmv[j] += nmv2[j];
            nrep[j] += nrep2[j];
            for (int k = 1; k < nres2[j]; ++k) rv[j].push_back(rv2[(size_t)j * cap2 + k]);
          }
        }
      }
      VIE_HIP_CHECK(hipMemcpyAsync(xg, w.X, bytes, d2h, c.stream));
      VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
      for (int j = 0; j < s; ++j) {
        const int i = g0 + j;
        const int len = (int)std::min<size_t>(rv[j].size(), (size_t)cap);
        if (resvec)
          for (int k = 0; k < len; ++k) resvec[(size_t)i * cap + k] = rv[j][k];
        if (nres) nres[i] = len;
        if (nmatvec) nmatvec[i] = nmv[j];
        if (nreplace) nreplace[i] = nrep[j];
      }
    }
    return VIE_OK;
  } catch (const Error& e) {
    c.err = e.msg;
    return e.code;
  } catch (...) {
    c.err = "unknown error in vie_bcocg";
    return VIE_ERR_HIP;
  }
}
