// Device-resident COCR (conjugate orthogonal conjugate residual, Sogabe &
// Zhang 2007) on the VIE operator, batched over rhs: the short-recurrence
// solver of the mixed-precision batched form.  No counterpart in the
// reference, whose solver is GMRES (is_gmres_vie.m, built in vie_gmres.hip).
//
// Why a three-term recurrence applies.  The packed VIE operator is
// A = I - D N (mvp_vie.m:12-18) with D = diag(d), d_i = (Mc/Mr)_s invG_c for
// unknown i = c nS + s, and N the Galerkin matrix of a symmetric dyadic
// kernel between real basis functions, so N^T = N.  Hence
//   [A u, v] = u^T (D^-1 - N) v = [u, A v]   for [u, v] = sum_i u_i v_i / d_i,
// and the adjoint A^H = I - conj(N) conj(D) satisfies the same identity in
// [u, v] = sum_i u_i v_i conj(d_i).  A conjugate-residual recurrence in that
// bilinear form needs one operator application and a fixed number of vector
// operations per iteration, where GMRES's Gram-Schmidt grows with the step.
// A voxel with d = 0 (no contrast) has identity rows in A and b = 0 there
// (b = ce (Mc/Mr) E_inc, ie_solver_vie.m:18), so its unknowns stay 0 in every
// iterate and the recurrence runs on the other unknowns; its weight is taken
// as 0.
//
// Iteration, arranged so that one reduction per iteration suffices
// (Chronopoulos & Gear): after Ar = A r,
//   rho' = [r, Ar],  beta = rho'/rho,
//   sigma' = [Ap', Ap'] = [Ar, Ar] + 2 beta [Ar, Ap] + beta^2 sigma,  alpha = rho'/sigma',
//   p = r + beta p,  Ap = Ar + beta Ap,  xs += alpha p,  r -= alpha Ap
// (beta = 0 at the start of a recurrence: p = r, Ap = Ar).
//
// Precision.  FP64 form: FP64 operator, every vector FP64.  Mixed form: the
// FP32 operator (FP32 FFT and product) reading and writing its packed
// vectors in FP32; x, b and every recurrence scalar FP64; r, Ar, p, Ap and xs
// (the solution increment since the last replacement) stored FP32, every
// arithmetic operation on them FP64.  Reliable updates
// (van der Vorst & Ye 2000; the mixed-precision CG of Clark et al. 2010):
// whenever the recurrence residual has fallen by kRelDelta since the last
// replacement, and when it meets the tolerance, x += xs in FP64 and r is
// replaced by b - A x with the FP64 operator; the recurrence continues with
// the new r (p, Ap, rho and sigma kept).  A solve therefore ends only on a
// true FP64 residual ||b - A x|| < tol ||b||.
//
// Batching and graphs.  All rhs advance in lockstep; kSymIters iterations
// (the ||r|| test, the operator, the dot partials, the scalars, the vector
// update) are one captured graph.  An rhs that meets the tolerance or breaks
// down marks kdone, and the later kernels of the graph skip it (the operator
// included); between graphs the host reads the flags and the residual
// history, runs the FP64 replacements for the rhs that need one, and retires
// converged rhs (kdone = -1: they keep their slot, so one graph serves the
// whole solve).

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "vie_internal.hpp"
#include "vie_reduce.hpp"

namespace vie {

constexpr int kSymIters = 16;        // iterations per captured graph
// mixed form: replace r when ||r|| fell by this factor since the last
// replacement.  Measured at the 96x96x112 torso (32 dipoles, three rhs sets):
// 0.1 gives 403-409 iterations per rhs with the slowest at 438-456, 0.01
// gives 407-421 with the slowest at 459-593 (the batch waits for it); 0.2
// and 0.3 lose again (slowest up to 690)
constexpr double kRelDelta = 1e-1;
// Stall guard: an rhs whose best recurrence residual has not halved within
// kStallIters iterations (a normal solve halves it every 20-30) is replaced
// and its recurrence restarted from the FP64 residual.  In the indefinite
// bilinear form a near-breakdown (rho ~ 0, a large beta) can leave the
// recurrence making no progress for many iterations; with the FP32 operator
// this was measured on one rhs of a 32-rhs batch (1600 iterations against
// 385-520 for the others).
constexpr int kStallIters = 100;
// restarts of one rhs's recurrence: after kMaxStallRestarts a stalled rhs
// keeps its recurrence (the replacement still happens); a breakdown, which
// cannot continue without a restart, ends the rhs after kMaxBreakRestarts
constexpr int kMaxStallRestarts = 8;
constexpr int kMaxBreakRestarts = 32;
constexpr int kSymRows = 6;          // dot partial rows: [r,Ar], [Ar,Ar], [Ar,Ap] (re, im)

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// weight of unknown i = c nS + s in the bilinear form: wv[s] gw[c]
struct SymWeights {
  const double2* wv;   // [nS]: 1/(Mc/Mr) (0 where Mc/Mr = 0) for VIE, conj(Mc/Mr) for VIE^H
  const double* gw;    // [ql]: 1/invG (VIE) or invG (VIE^H)
  int64_t nS;
  int ql;
};

__global__ void k_sym_weights(const double2* __restrict__ mcr, int64_t nS, int adj, double2* __restrict__ wv) {
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

// v = b - ax (ax null: v = b) for the rhs of the slot map, stored to R when
// given (rounded to its precision), and the |v|^2 partials of the FP64 value
// into part[rhs][nblk]; grid (nblk, nslot)
template <typename TV>
__global__ void __launch_bounds__(kGmBlock) k_sym_resid(const double2* __restrict__ b, const double2* __restrict__ ax,
                                                        TV* __restrict__ R, int64_t n, const int* __restrict__ rmap,
                                                        double* __restrict__ part) {
  __shared__ double lds[64];
  const int rhs = rmap[blockIdx.y], nblk = gridDim.x;
  const double2* br = b + rhs * n;
  const double2* axr = ax ? ax + rhs * n : nullptr;
  TV* rr = R ? R + rhs * n : nullptr;
  double s[1] = {0};
  for (int64_t i = blockIdx.x * (int64_t)kGmBlock + threadIdx.x; i < n; i += (int64_t)nblk * kGmBlock) {
    double2 v = br[i];
    if (axr) {
      const double2 t = axr[i];
      v = make_double2(v.x - t.x, v.y - t.y);
    }
    if (rr) st_v(rr + i, v);
    s[0] += v.x * v.x + v.y * v.y;
  }
  const double r = block_sum<1>(s, lds, threadIdx.x);
  if (threadIdx.x == 0) part[(int64_t)rhs * nblk + blockIdx.x] = r;
}

// out[rhs] = sqrt(sum of the |r|^2 partials), one wavefront per slot
__global__ void __launch_bounds__(64) k_sym_norm_out(const double* __restrict__ part, int nblk,
                                                     const int* __restrict__ rmap, double* __restrict__ out) {
  const int rhs = rmap[blockIdx.x];
  double s = 0;
  for (int i = threadIdx.x; i < nblk; i += 64) s += part[(int64_t)rhs * nblk + i];
  s = wave_sum(s);
  if (threadIdx.x == 0) out[rhs] = sqrt(s);
}

// ||r|| of iteration `step` (from the partials of the last vector update or
// replacement) into the history column `col`; below the tolerance the rhs is
// done for the rest of the graph (kdone = step + 1)
__global__ void __launch_bounds__(64) k_sym_rnorm(const double* __restrict__ partr, int nblk,
                                                  const int* __restrict__ rmap, int* __restrict__ kdone, int step,
                                                  double* __restrict__ rv, int ldr, int col,
                                                  const double* __restrict__ tolabs) {
  const int rhs = rmap[blockIdx.x];
  if (gm_skip(kdone, rhs, step)) return;
  double s = 0;
  for (int i = threadIdx.x; i < nblk; i += 64) s += partr[(int64_t)rhs * nblk + i];
  s = wave_sum(s);
  if (threadIdx.x == 0) {
    const double rn = sqrt(s);
    rv[(int64_t)rhs * ldr + col] = rn;
    if (rn < tolabs[rhs]) kdone[rhs] = step + 1;
  }
}

// partials of [r, Ar], [Ar, Ar] and (not at the start of a recurrence)
// [Ar, Ap] in the weighted bilinear form, into part[rhs][kSymRows][nblk];
// grid (nblk, nslot), the channels walked in turn so the voxel weight is
// read with unit stride.  A recurrence starts (p = r) at the iteration whose
// step equals fresh[rhs].  Block 0 also moves the previous iteration's rho
// and sigma (written by k_sym_update into the `next` slots) to the slots the
// coming k_sym_update reads: no block of k_sym_update writes what another
// block of it reads.
template <typename TV>
__global__ void __launch_bounds__(kGmBlock)
    k_sym_dots(const TV* __restrict__ R, const TV* __restrict__ AR, const TV* __restrict__ AP, int64_t n,
               SymWeights w, const int* __restrict__ rmap, const int* __restrict__ kdone, int step,
               const int* __restrict__ fresh, double* __restrict__ part, double2* __restrict__ rho,
               double2* __restrict__ sig, const double2* __restrict__ rho_next, const double2* __restrict__ sig_next) {
  __shared__ double lds[8 * (kGmBlock / 64)];
  const int rhs = rmap[blockIdx.y], nblk = gridDim.x;
  if (gm_skip(kdone, rhs, step)) return;
  const bool fr = fresh[rhs] == step;
  if (!fr && blockIdx.x == 0 && threadIdx.x == 0) {
    rho[rhs] = rho_next[rhs];
    sig[rhs] = sig_next[rhs];
  }
  const TV* r = R + rhs * n;
  const TV* a = AR + rhs * n;
  const TV* ap = AP + rhs * n;
  double acc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) acc[q] = 0;
  for (int c = 0; c < w.ql; ++c) {
    const double g = w.gw[c];
    const int64_t base = (int64_t)c * w.nS;
    for (int64_t s = blockIdx.x * (int64_t)kGmBlock + threadIdx.x; s < w.nS; s += (int64_t)nblk * kGmBlock) {
      double2 wi = w.wv[s];
      wi.x *= g;
      wi.y *= g;
      const double2 ri = ld_v(r + base + s), ai = ld_v(a + base + s);
      const double2 aw = zmul(ai, wi);
      acc[0] += ri.x * aw.x - ri.y * aw.y;
      acc[1] += ri.x * aw.y + ri.y * aw.x;
      acc[2] += ai.x * aw.x - ai.y * aw.y;
      acc[3] += ai.x * aw.y + ai.y * aw.x;
      if (!fr) {
        const double2 pi = ld_v(ap + base + s);
        acc[4] += pi.x * aw.x - pi.y * aw.y;
        acc[5] += pi.x * aw.y + pi.y * aw.x;
      }
    }
  }
  const double t = block_sum_halving<8>(acc, lds, threadIdx.x);
  if (threadIdx.x < kSymRows) part[((int64_t)rhs * kSymRows + threadIdx.x) * nblk + blockIdx.x] = t;
}

// The recurrence scalars and the vector update of one iteration.  Every
// block sums the dot partials of k_sym_dots in the same order (one wavefront
// per row), so all blocks hold the same beta = rho'/rho, sigma' = [Ar, Ar] +
// 2 beta [Ar, Ap] + beta^2 sigma and alpha = rho'/sigma' (beta = 0 at the
// start of a recurrence); block 0 stores rho' and sigma' for the next
// iteration.  A zero or non-finite rho' or sigma' is a breakdown: no block
// updates, block 0 sets kdone = step and brk = 1 (the host restarts the
// recurrence after a replacement).  Then p = r + beta p, Ap = Ar + beta Ap
// (beta = 0: p = r, Ap = Ar, the old values unread), xs += alpha p,
// r -= alpha Ap with the stored (rounded) p and Ap, and the |r|^2 partials.
// grid (nblk, nslot)
template <typename TV>
__global__ void __launch_bounds__(kGmBlock)
    k_sym_update(TV* __restrict__ R, const TV* __restrict__ AR, TV* __restrict__ P, TV* __restrict__ AP,
                 TV* __restrict__ XS, int64_t n, const int* __restrict__ rmap, int* __restrict__ kdone, int step,
                 const int* __restrict__ fresh, const double* __restrict__ part, int nblk_d,
                 const double2* __restrict__ rho, const double2* __restrict__ sig, double2* __restrict__ rho_next,
                 double2* __restrict__ sig_next, int* __restrict__ brk, double* __restrict__ partr) {
  __shared__ double lds[64];
  __shared__ double sums[kSymRows];
  const int rhs = rmap[blockIdx.y], nblk = gridDim.x;
  if (gm_skip(kdone, rhs, step)) return;
  {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int row = wave; row < kSymRows; row += kGmBlock / 64) {
      const double* src = part + ((int64_t)rhs * kSymRows + row) * nblk_d;
      double t = 0;
      for (int i = lane; i < nblk_d; i += 64) t += src[i];
      t = wave_sum(t);
      if (lane == 0) sums[row] = t;
    }
    __syncthreads();
  }
  const double2 rr = make_double2(sums[0], sums[1]), aa = make_double2(sums[2], sums[3]);
  double2 be = make_double2(0.0, 0.0), sg = aa;
  if (fresh[rhs] != step) {
    const double2 r0 = rho[rhs];
    const double den = r0.x * r0.x + r0.y * r0.y;
    be = make_double2((rr.x * r0.x + rr.y * r0.y) / den, (rr.y * r0.x - rr.x * r0.y) / den);
    const double2 apr = make_double2(sums[4], sums[5]);
    const double2 b2 = zmul(be, be), s0 = sig[rhs];
    const double2 t1 = zmul(be, apr), t2 = zmul(b2, s0);
    sg = make_double2(aa.x + 2.0 * t1.x + t2.x, aa.y + 2.0 * t1.y + t2.y);
  }
  const double sm = sg.x * sg.x + sg.y * sg.y;
  const double rm = rr.x * rr.x + rr.y * rr.y;
  if (!(sm > 0) || !(rm > 0) || !__builtin_isfinite(sm) || !__builtin_isfinite(rm) ||
      !__builtin_isfinite(be.x) || !__builtin_isfinite(be.y)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      brk[rhs] = 1;
      kdone[rhs] = step;
    }
    return;
  }
  const double2 al = make_double2((rr.x * sg.x + rr.y * sg.y) / sm, (rr.y * sg.x - rr.x * sg.y) / sm);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    rho_next[rhs] = rr;
    sig_next[rhs] = sg;
  }
  const bool b0 = be.x == 0.0 && be.y == 0.0;
  TV* r = R + rhs * n;
  const TV* a = AR + rhs * n;
  TV* p = P + rhs * n;
  TV* ap = AP + rhs * n;
  TV* xs = XS + rhs * n;
  double s[1] = {0};
  for (int64_t i = blockIdx.x * (int64_t)kGmBlock + threadIdx.x; i < n; i += (int64_t)nblk * kGmBlock) {
    double2 ri = ld_v(r + i);
    const double2 ai = ld_v(a + i);
    double2 pn = ri, qn = ai;
    if (!b0) {
      const double2 po = zmul(be, ld_v(p + i)), qo = zmul(be, ld_v(ap + i));
      pn = make_double2(ri.x + po.x, ri.y + po.y);
      qn = make_double2(ai.x + qo.x, ai.y + qo.y);
    }
    pn = st_v(p + i, pn);
    qn = st_v(ap + i, qn);
    const double2 dx = zmul(al, pn), dr = zmul(al, qn);
    const double2 xo = ld_v(xs + i);
    st_v(xs + i, make_double2(xo.x + dx.x, xo.y + dx.y));
    ri = st_v(r + i, make_double2(ri.x - dr.x, ri.y - dr.y));
    s[0] += ri.x * ri.x + ri.y * ri.y;
  }
  const double t = block_sum<1>(s, lds, threadIdx.x);
  if (threadIdx.x == 0) partr[(int64_t)rhs * nblk + blockIdx.x] = t;
}

// x += xs, xs = 0 for the rhs of the slot map; grid (nblk, nslot)
template <typename TV>
__global__ void k_sym_accum(double2* __restrict__ X, TV* __restrict__ XS, int64_t n, const int* __restrict__ rmap) {
  const int rhs = rmap[blockIdx.y];
  double2* x = X + rhs * n;
  TV* xs = XS + rhs * n;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 a = x[i], d = ld_v(xs + i);
    x[i] = make_double2(a.x + d.x, a.y + d.y);
    st_v(xs + i, make_double2(0.0, 0.0));
  }
}

// xs = 0 for the rhs of the slot map; grid (nblk, nslot)
template <typename TV>
__global__ void k_sym_zero(TV* __restrict__ XS, int64_t n, const int* __restrict__ rmap) {
  TV* xs = XS + rmap[blockIdx.y] * n;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    st_v(xs + i, make_double2(0.0, 0.0));
}

// --------------------------------------------------------------------------
// host side
// --------------------------------------------------------------------------
struct SymWork {
  int64_t n;
  int B, nblk;
  bool b32;                          // r, Ar, p, Ap, xs stored FP32 (the FP32 operator's packed I/O)
  double2 *X, *Bv, *T;               // [B][n] FP64: x, b, A x of a replacement
  void *R, *AR, *P, *AP, *XS;        // [B][n], double2 or (b32) float2
  double2* wv;                       // [nS]
  double* gw;                        // [ql]
  double *part, *partr;              // [B][kSymRows][nblk], [B][nblk]
  double2 *rho, *sig, *rho_next, *sig_next;
  double *rv, *tolabs, *rn;          // [B][kSymIters + 1], [B], [B]
  int *kdone, *rmap, *brk;           // [B]
  int* fresh;                        // [B]: step (1..K) of the next graph at which the recurrence starts, else 0
};

static void sym_alloc(Ctx& c, SymWork& w, int64_t n, int B, bool b32) {
  w.n = n;
  w.B = B;
  // blocks per rhs of the vector kernels: about 2048 blocks in all (enough to
  // stream at full HBM rate) and few partial sums per rhs for the small
  // reduction kernels
  w.nblk = std::max(16, std::min(nblocks_for(n), 2048 / B));
  w.b32 = b32;
  const size_t vec = (size_t)n * sizeof(double2), tv = (size_t)n * (b32 ? sizeof(float2) : sizeof(double2));
  c.gm_r.ensure((size_t)B * vec);
  c.gm_x.ensure((size_t)B * vec);
  c.gm_b.ensure((size_t)B * vec);
  c.gm_V.ensure(5 * (size_t)B * tv);
  w.T = c.gm_r.as<double2>();
  w.X = c.gm_x.as<double2>();
  w.Bv = c.gm_b.as<double2>();
  char* v = c.gm_V.as<char>();
  w.R = v;
  w.AR = v + (size_t)B * tv;
  w.P = v + 2 * (size_t)B * tv;
  w.AP = v + 3 * (size_t)B * tv;
  w.XS = v + 4 * (size_t)B * tv;
  c.gm_red.ensure((size_t)B * w.nblk * (kSymRows + 1) * sizeof(double));
  w.part = c.gm_red.as<double>();
  w.partr = w.part + (size_t)B * kSymRows * w.nblk;
  const size_t small = (size_t)c.nS * sizeof(double2) + c.ql * sizeof(double) + (size_t)B * 4 * sizeof(double2) +
                       (size_t)B * (kSymIters + 3) * sizeof(double) + (size_t)B * 4 * sizeof(int) +
                       16 * 256;   // each of the 14 pieces below is rounded up to 256 bytes
  c.gm_small.ensure(small);
  char* q = c.gm_small.as<char>();
  auto take = [&](size_t sz) { char* t = q; q += (sz + 255) & ~(size_t)255; return t; };
  w.wv = (double2*)take((size_t)c.nS * sizeof(double2));
  w.gw = (double*)take(c.ql * sizeof(double));
  w.rho = (double2*)take(B * sizeof(double2));
  w.sig = (double2*)take(B * sizeof(double2));
  w.rho_next = (double2*)take(B * sizeof(double2));
  w.sig_next = (double2*)take(B * sizeof(double2));
  w.rv = (double*)take((size_t)B * (kSymIters + 1) * sizeof(double));
  w.tolabs = (double*)take(B * sizeof(double));
  w.rn = (double*)take(B * sizeof(double));
  w.kdone = (int*)take(B * sizeof(int));
  w.rmap = (int*)take(B * sizeof(int));
  w.fresh = (int*)take(B * sizeof(int));
  w.brk = (int*)take(B * sizeof(int));
}

// pinned host block: per rhs kdone, brk, the residual history of a graph, ||r||
struct SymHost {
  int *kd, *brk;
  double *rv, *rn;
};

static SymHost sym_host(Ctx& c, int B) {
  const size_t need = 2 * (size_t)B * sizeof(int) + (size_t)B * (kSymIters + 2) * sizeof(double) + 64;
  if (c.gm_pin_bytes < need) {
    if (c.gm_pin) VIE_HIP_CHECK(hipHostFree(c.gm_pin));
    c.gm_pin = nullptr;
    c.gm_pin_bytes = 0;
    VIE_HIP_CHECK(hipHostMalloc(&c.gm_pin, need, hipHostMallocDefault));
    c.gm_pin_bytes = need;
  }
  SymHost h;
  h.rv = static_cast<double*>(c.gm_pin);
  h.rn = h.rv + (size_t)B * (kSymIters + 1);
  h.kd = reinterpret_cast<int*>(h.rn + B);
  h.brk = h.kd + B;
  return h;
}

static std::vector<intptr_t> sym_signature(const Ctx& c, const SymWork& w) {
  std::vector<intptr_t> s = {(intptr_t)w.R, (intptr_t)w.AR, (intptr_t)w.P, (intptr_t)w.AP, (intptr_t)w.XS,
                             (intptr_t)w.b32, (intptr_t)w.part, (intptr_t)w.partr, (intptr_t)w.wv, (intptr_t)w.gw,
                             (intptr_t)w.rho, (intptr_t)w.rv, (intptr_t)w.kdone, (intptr_t)w.rmap, (intptr_t)w.n,
                             (intptr_t)w.B, (intptr_t)c.gm_pin, (intptr_t)c.wsA1.p, (intptr_t)c.wsA2.p,
                             (intptr_t)c.vox2pk.p, (intptr_t)c.mcr.p, (intptr_t)c.invg.p, (intptr_t)c.nS,
                             (intptr_t)c.zline.p, (intptr_t)c.prune, (intptr_t)c.xlo, (intptr_t)c.xhi};
  for (int k = 0; k < 2; ++k) {
    s.push_back((intptr_t)c.spec[k].d64.p);
    s.push_back((intptr_t)c.spec[k].d32.p);
  }
  for (int p = 0; p < 2; ++p)
    for (int b = 0; b < 2; ++b)
      for (int k = 0; k < 4; ++k) s.push_back(c.tile[p][b][k]);
  for (int p = 0; p < 2; ++p)
    for (int k = 0; k < 4; ++k) s.push_back(c.fam[p][k]);
  s.push_back(c.force_fam);
  return s;
}

// K iterations of the recurrence for the ns slots of w.rmap, then the last
// ||r|| and the copies of the flags and the history to pinned memory
template <typename TV>
static void enqueue_iters(Ctx& c, const SymWork& w, int op, int prec, int ns, int K, const SymHost& hb) {
  const int64_t n = w.n;
  const SymWeights sw{w.wv, w.gw, c.nS, c.ql};
  TV* R = static_cast<TV*>(w.R);
  TV* AR = static_cast<TV*>(w.AR);
  TV* P = static_cast<TV*>(w.P);
  TV* AP = static_cast<TV*>(w.AP);
  TV* XS = static_cast<TV*>(w.XS);
  const dim3 g(w.nblk, ns);
  for (int j = 0; j < K; ++j) {
    hipLaunchKernelGGL(k_sym_rnorm, dim3(ns), dim3(64), 0, c.stream, w.partr, w.nblk, w.rmap, w.kdone, j, w.rv,
                       kSymIters + 1, j, w.tolabs);
    ApplyIO io;
    io.in = reinterpret_cast<const double2*>(R);
    io.in_rs = n;
    io.out = reinterpret_cast<double2*>(AR);
    io.out_rs = n;
    io.io32 = w.b32;
    io.rmap = w.rmap;
    io.kdone = w.kdone;
    io.step = j + 1;
    apply_operator(c, op, prec, io, ns);
    hipLaunchKernelGGL(k_sym_dots<TV>, g, dim3(kGmBlock), 0, c.stream, R, AR, AP, n, sw, w.rmap, w.kdone, j + 1,
                       w.fresh, w.part, w.rho, w.sig, w.rho_next, w.sig_next);
    hipLaunchKernelGGL(k_sym_update<TV>, g, dim3(kGmBlock), 0, c.stream, R, AR, P, AP, XS, n, w.rmap, w.kdone, j + 1,
                       w.fresh, w.part, w.nblk, w.rho, w.sig, w.rho_next, w.sig_next, w.brk, w.partr);
  }
  hipLaunchKernelGGL(k_sym_rnorm, dim3(ns), dim3(64), 0, c.stream, w.partr, w.nblk, w.rmap, w.kdone, K, w.rv,
                     kSymIters + 1, K, w.tolabs);
  VIE_HIP_CHECK(hipGetLastError());
  VIE_HIP_CHECK(hipMemcpyAsync(hb.kd, w.kdone, w.B * sizeof(int), hipMemcpyDeviceToHost, c.stream));
  VIE_HIP_CHECK(hipMemcpyAsync(hb.brk, w.brk, w.B * sizeof(int), hipMemcpyDeviceToHost, c.stream));
  VIE_HIP_CHECK(hipMemcpyAsync(hb.rv, w.rv, (size_t)w.B * (kSymIters + 1) * sizeof(double), hipMemcpyDeviceToHost,
                               c.stream));
}

// ||r|| of the rhs in `set` from their |r|^2 partials, to hb.rn (synchronises)
static void sym_norms_host(Ctx& c, SymWork& w, const std::vector<int>& set, const SymHost& hb) {
  VIE_HIP_CHECK(hipMemcpyAsync(w.rmap, set.data(), set.size() * sizeof(int), hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(k_sym_norm_out, dim3((unsigned)set.size()), dim3(64), 0, c.stream, w.partr, w.nblk, w.rmap, w.rn);
  VIE_HIP_CHECK(hipMemcpyAsync(hb.rn, w.rn, w.B * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
}

// r = b - A64 x for the rhs in `set`, stored in the recurrence precision
// (|r|^2 partials of the FP64 residual refreshed)
template <typename TV>
static void sym_true_residual(Ctx& c, SymWork& w, int op, const std::vector<int>& set) {
  const int ns = (int)set.size();
  VIE_HIP_CHECK(hipMemcpyAsync(w.rmap, set.data(), ns * sizeof(int), hipMemcpyHostToDevice, c.stream));
  ApplyIO io;
  io.in = w.X;
  io.in_rs = w.n;
  io.out = w.T;
  io.out_rs = w.n;
  io.rmap = w.rmap;
  apply_operator(c, op, VIE_PREC_FP64, io, ns);
  hipLaunchKernelGGL(k_sym_resid<TV>, dim3(w.nblk, ns), dim3(kGmBlock), 0, c.stream, w.Bv, w.T, static_cast<TV*>(w.R),
                     w.n, w.rmap, w.partr);
  VIE_HIP_CHECK(hipGetLastError());
}

template <typename TV>
static void cocr_solve(Ctx& c, SymWork& w, int op, int prec, double tol, int maxit, const std::vector<double>& bn,
                       const std::vector<double>& xn, std::vector<std::vector<double>>& rv, std::vector<int>& nmv,
                       std::vector<int>& nrep) {
  const int B = w.B;
  const int64_t n = w.n;
  double delta = prec == VIE_PREC_FP32 ? kRelDelta : 0.0;
  if (const char* dl = std::getenv("VIE_COCR_DELTA"))   // VIE_COCR_DELTA=f: another replacement factor
    if (prec == VIE_PREC_FP32 && std::atof(dl) > 0 && std::atof(dl) < 1) delta = std::atof(dl);
  int stall_iters = kStallIters;   // VIE_COCR_STALL=n: another window (tests force restarts with it)
  if (const char* st = std::getenv("VIE_COCR_STALL"))
    if (std::atoi(st) > 0) stall_iters = std::atoi(st);
  SymHost hb = sym_host(c, B);
  std::vector<int> all(B);
  for (int i = 0; i < B; ++i) all[i] = i;
  // r = b - A x0 where x0 != 0, else r = b; xs = 0
  std::vector<int> nz, zr;
  for (int i = 0; i < B; ++i) (xn[i] > 0 ? nz : zr).push_back(i);
  if (!zr.empty()) {
    VIE_HIP_CHECK(hipMemcpyAsync(w.rmap, zr.data(), zr.size() * sizeof(int), hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_sym_resid<TV>, dim3(w.nblk, (int)zr.size()), dim3(kGmBlock), 0, c.stream, w.Bv, nullptr,
                       static_cast<TV*>(w.R), n, w.rmap, w.partr);
    VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
  }
  if (!nz.empty()) {
    sym_true_residual<TV>(c, w, op, nz);
    for (int i : nz) nmv[i] += 1;
  }
  VIE_HIP_CHECK(hipMemcpyAsync(w.rmap, all.data(), B * sizeof(int), hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(k_sym_zero<TV>, dim3(w.nblk, B), dim3(kGmBlock), 0, c.stream, static_cast<TV*>(w.XS), n, w.rmap);
  sym_norms_host(c, w, all, hb);
  std::vector<double> rmax(B), tolh(B), rbest(B);
  std::vector<int> its(B, 0), fresh(B, 1), zero(B, 0), live, itbest(B, 0);
  for (int i = 0; i < B; ++i) {
    rv[i].push_back(hb.rn[i] / bn[i]);
    rmax[i] = hb.rn[i];
    rbest[i] = hb.rn[i];
    tolh[i] = tol * bn[i];
    if (hb.rn[i] >= tolh[i]) live.push_back(i);
  }
  VIE_HIP_CHECK(hipMemcpyAsync(w.tolabs, tolh.data(), B * sizeof(double), hipMemcpyHostToDevice, c.stream));
  VIE_HIP_CHECK(hipMemcpyAsync(w.fresh, fresh.data(), B * sizeof(int), hipMemcpyHostToDevice, c.stream));
  VIE_HIP_CHECK(hipMemcpyAsync(w.brk, zero.data(), B * sizeof(int), hipMemcpyHostToDevice, c.stream));
  if (live.empty()) return;
  // the slot map of the graphs holds every rhs that starts live; rhs that end
  // are retired in place (kdone = -1)
  const std::vector<int> slots = live;
  const int ns = (int)slots.size();
  // workspaces and tile widths of both operator precisions fixed before capture
  prepare_operator(c, op, VIE_PREC_FP64, ns);
  if (prec != VIE_PREC_FP64) prepare_operator(c, op, prec, ns);
  std::vector<char> on(B, 0);
  for (int i : live) on[i] = 1;
  std::vector<int> nrestart(B, 0);
  while (!live.empty()) {
    int K = kSymIters;
    for (int i : live) K = std::min(K, maxit - its[i]);
    std::vector<int> kd0(B);
    for (int i = 0; i < B; ++i) kd0[i] = on[i] ? 0 : -1;
    VIE_HIP_CHECK(hipMemcpyAsync(w.rmap, slots.data(), ns * sizeof(int), hipMemcpyHostToDevice, c.stream));
    VIE_HIP_CHECK(hipMemcpyAsync(w.kdone, kd0.data(), B * sizeof(int), hipMemcpyHostToDevice, c.stream));
    // recurrences (re)starting at the graph's first iteration; none after it
    VIE_HIP_CHECK(hipMemcpyAsync(w.fresh, fresh.data(), B * sizeof(int), hipMemcpyHostToDevice, c.stream));
    std::fill(fresh.begin(), fresh.end(), 0);
    auto enqueue = [&]() { enqueue_iters<TV>(c, w, op, prec, ns, K, hb); };
    if (c.use_graphs) {
      auto sig = sym_signature(c, w);
      if (sig != c.sym_sig) {
        for (auto& kv : c.sym_graphs) (void)hipGraphExecDestroy(kv.second);
        c.sym_graphs.clear();
        c.sym_sig.swap(sig);
      }
      const auto key = std::make_tuple(op, prec, ns, K, w.b32 ? 1 : 0);
      auto it = c.sym_graphs.find(key);
      if (it == c.sym_graphs.end()) {
        VIE_HIP_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        try {
          enqueue();
        } catch (...) {
          hipGraph_t gr = nullptr;
          (void)hipStreamEndCapture(c.stream, &gr);
          if (gr) (void)hipGraphDestroy(gr);
          throw;
        }
        hipGraph_t gr = nullptr;
        VIE_HIP_CHECK(hipStreamEndCapture(c.stream, &gr));
        hipGraphExec_t ex = nullptr;
        const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        VIE_HIP_CHECK(e);
        it = c.sym_graphs.emplace(key, ex).first;
      }
      VIE_HIP_CHECK(hipGraphLaunch(it->second, c.stream));
    } else {
      enqueue();
    }
    VIE_HIP_CHECK(hipEventRecord(c.gm_cycle_ev, c.stream));
    VIE_HIP_CHECK(hipEventSynchronize(c.gm_cycle_ev));
    // read the graph's outcome per live rhs
    std::vector<int> upd, brk_set;
    for (int i : live) {
      const int its0 = its[i];
      const int kd = hb.kd[i];
      const bool broke = hb.brk[i] != 0;
      // kd = j + 1: the ||r|| test (converged) or the scalars (breakdown) of
      // iteration j stopped it; iterations 0..j-1 are complete, and a
      // breakdown also spent iteration j's operator application
      const int done = kd > 0 ? kd - 1 : K;
      const int spent = kd > 0 && broke ? done + 1 : done;
      its[i] += spent;
      nmv[i] += spent;
      const double* h = hb.rv + (size_t)i * (kSymIters + 1);
      for (int col = 1; col <= done; ++col) {
        rv[i].push_back(h[col] / bn[i]);
        if (h[col] < 0.5 * rbest[i]) {
          rbest[i] = h[col];
          itbest[i] = its0 + col;
        }
      }
      const double rcur = h[done];
      const bool conv = kd > 0 && !broke;
      const bool stall = its[i] - itbest[i] > stall_iters;
      if (conv || broke || stall || its[i] >= maxit || rcur < delta * rmax[i]) upd.push_back(i);
      if (broke || stall) brk_set.push_back(i);
    }
    if (upd.empty()) continue;
    // replacement: x += xs, r = b - A64 x
    const int nu = (int)upd.size();
    VIE_HIP_CHECK(hipMemcpyAsync(w.rmap, upd.data(), nu * sizeof(int), hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_sym_accum<TV>, dim3(w.nblk, nu), dim3(kGmBlock), 0, c.stream, w.X, static_cast<TV*>(w.XS), n,
                       w.rmap);
    sym_true_residual<TV>(c, w, op, upd);
    sym_norms_host(c, w, upd, hb);
    bool fresh_changed = false;
    std::vector<int> live2;
    std::vector<char> ended(B, 0);
    for (int i : upd) {
      nmv[i] += 1;
      nrep[i] += 1;
      const double rn = hb.rn[i];
      rv[i].push_back(rn / bn[i]);
      rmax[i] = rn;
      if (rn < tolh[i] || its[i] >= maxit) ended[i] = 1;
    }
    for (int i : brk_set) {
      if (ended[i]) continue;
      // restart the recurrence from the replaced residual (breakdown or stall)
      rbest[i] = hb.rn[i];
      itbest[i] = its[i];
      const bool broke = hb.brk[i] != 0;
      if (nrestart[i] >= (broke ? kMaxBreakRestarts : kMaxStallRestarts)) {
        if (broke) ended[i] = 1;
        continue;
      }
      ++nrestart[i];
      fresh[i] = 1;
      fresh_changed = true;
    }
    if (fresh_changed)
      VIE_HIP_CHECK(hipMemcpyAsync(w.brk, zero.data(), B * sizeof(int), hipMemcpyHostToDevice, c.stream));
    for (int i : live) {
      if (ended[i]) on[i] = 0;
      else live2.push_back(i);
    }
    live.swap(live2);
  }
  VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
}

}  // namespace vie

using namespace vie;

extern "C" int vie_cocr(vie_ctx* h, int32_t op, int32_t prec, const double* b, double* x, int32_t nrhs, double tol,
                        int32_t maxit, int32_t mem, double* resvec, int32_t* nres, int32_t* nmatvec,
                        int32_t* nreplace) {
  if (!h) return VIE_ERR_ARG;
  Ctx& c = h->c;
  try {
    c.err.clear();
    if (op != VIE_OP_VIE && op != VIE_OP_VIEH) throw Error{VIE_ERR_ARG, "vie_cocr: op must be VIE_OP_VIE or VIE_OP_VIEH"};
    if (prec != VIE_PREC_FP64 && prec != VIE_PREC_FP32) throw Error{VIE_ERR_ARG, "unknown precision"};
    if (nrhs < 1 || maxit < 1) throw Error{VIE_ERR_ARG, "vie_cocr: bad nrhs/maxit"};
    if (!(tol > 0)) throw Error{VIE_ERR_ARG, "vie_cocr: tol must be > 0"};
    if (!b || !x) throw Error{VIE_ERR_ARG, "NULL vector"};
    if (!c.have_material) throw Error{VIE_ERR_STATE, "material not set"};
    if (nrhs > 65535 / c.ql) throw Error{VIE_ERR_ARG, "nrhs too large"};
    if (c.box) {   // packed operators on the scatterer's support box
      const int rc = vie_cocr(c.box, op, prec, b, x, nrhs, tol, maxit, mem, resvec, nres, nmatvec, nreplace);
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
    struct StreamSwap {
      Ctx& c;
      hipStream_t user;
      ~StreamSwap() {
        (void)hipStreamSynchronize(c.stream);
        c.stream = user;
      }
    } swap{c, c.stream};
    c.stream = c.gm_stream;
    const int64_t n = (int64_t)c.ql * c.nS;
    if (n == 0) throw Error{VIE_ERR_ARG, "no scatterer voxels"};
    SymWork w{};
    sym_alloc(c, w, n, nrhs, prec == VIE_PREC_FP32);
    const size_t bytes = (size_t)nrhs * n * sizeof(double2);
    const hipMemcpyKind h2d = mem == VIE_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    VIE_HIP_CHECK(hipMemcpyAsync(w.Bv, b, bytes, h2d, c.stream));
    VIE_HIP_CHECK(hipMemcpyAsync(w.X, x, bytes, h2d, c.stream));
    // bilinear-form weights
    const bool adj = op == VIE_OP_VIEH;
    hipLaunchKernelGGL(k_sym_weights, dim3((unsigned)((c.nS + 255) / 256)), dim3(256), 0, c.stream,
                       c.mcr.as<double2>(), c.nS, adj ? 1 : 0, w.wv);
    std::vector<double> gw(c.ql);
    for (int q = 0; q < c.ql; ++q) gw[q] = adj ? c.invg_h[q] : 1.0 / c.invg_h[q];
    VIE_HIP_CHECK(hipMemcpyAsync(w.gw, gw.data(), c.ql * sizeof(double), hipMemcpyHostToDevice, c.stream));
    // ||b|| and ||x0||
    SymHost hb = sym_host(c, nrhs);
    std::vector<int> all(nrhs);
    for (int i = 0; i < nrhs; ++i) all[i] = i;
    std::vector<double> bn(nrhs), xn(nrhs);
    VIE_HIP_CHECK(hipMemcpyAsync(w.rmap, all.data(), nrhs * sizeof(int), hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_sym_resid<double2>, dim3(w.nblk, nrhs), dim3(kGmBlock), 0, c.stream, w.Bv, nullptr,
                       (double2*)nullptr, n, w.rmap, w.partr);
    sym_norms_host(c, w, all, hb);
    for (int i = 0; i < nrhs; ++i) bn[i] = hb.rn[i];
    hipLaunchKernelGGL(k_sym_resid<double2>, dim3(w.nblk, nrhs), dim3(kGmBlock), 0, c.stream, w.X, nullptr,
                       (double2*)nullptr, n, w.rmap, w.partr);
    sym_norms_host(c, w, all, hb);
    for (int i = 0; i < nrhs; ++i) xn[i] = hb.rn[i];
    for (int i = 0; i < nrhs; ++i)
      if (!(bn[i] > 0)) throw Error{VIE_ERR_ARG, "right-hand side is zero"};
    std::vector<std::vector<double>> rv(nrhs);
    std::vector<int> nmv(nrhs, 0), nrep(nrhs, 0);
    if (w.b32) cocr_solve<float2>(c, w, op, prec, tol, maxit, bn, xn, rv, nmv, nrep);
    else cocr_solve<double2>(c, w, op, prec, tol, maxit, bn, xn, rv, nmv, nrep);
    const hipMemcpyKind d2h = mem == VIE_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    // on the solver stream (non-blocking: a hipMemcpy on the null stream would
    // not be ordered after the solve), complete before the call returns
    VIE_HIP_CHECK(hipMemcpyAsync(x, w.X, bytes, d2h, c.stream));
    VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
    const int cap = 2 * maxit + 2;
    for (int i = 0; i < nrhs; ++i) {
      const int len = (int)std::min<size_t>(rv[i].size(), (size_t)cap);
      if (resvec)
        for (int k = 0; k < len; ++k) resvec[(size_t)i * cap + k] = rv[i][k];
      if (nres) nres[i] = len;
      if (nmatvec) nmatvec[i] = nmv[i];
      if (nreplace) nreplace[i] = nrep[i];
    }
    return VIE_OK;
  } catch (const Error& e) {
    c.err = e.msg;
    return e.code;
  } catch (...) {
    c.err = "unknown error in vie_cocr";
    return VIE_ERR_HIP;
  }
}
