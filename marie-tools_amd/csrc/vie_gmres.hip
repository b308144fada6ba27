// Device-resident restarted GMRES on the VIE operator (batched over rhs).
//
// The plain form follows is_gmres_vie.m:1-33 / is_iter_gmres_vie.m:1-40:
// GMRES(m) with one pass of classical Gram-Schmidt (H(1:k,k) = V' w;
// w = w - V H), H(k+1,k) = ||w||, a least-squares solve of the small
// Hessenberg system every step, break when its residual drops below
// tol*||b||, x += V y, and restart from the recurrence residual r = V res.
// The least-squares residual is tracked with Givens rotations on the device
// (it equals norm(rhs - Hc*y) of the reference's backslash solve); at the end
// of a cycle a device kernel solves R y = g by back substitution and forms
// res = rhs - Hc y, so a whole restart cycle (start vector, m Arnoldi steps,
// the small solve, x += V y, r = V res, and the copies of the cycle's state to
// pinned host memory) is one captured graph, launched once per cycle.
//
// All rhs of a batch advance in lockstep through one cycle; an rhs that
// meets its tolerance inside the cycle marks kdone and every later kernel of
// the cycle (the batched operator included) skips it.  An rhs whose solve has
// ended is retired (kdone = -1) between cycles: it keeps its slot, so the
// same captured cycle serves every cycle of the solve, and skips all work.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <vector>

#include "vie_internal.hpp"
#include "vie_reduce.hpp"

namespace vie {

using zc = std::complex<double>;

constexpr int kInnerRestart = 20;   // restart length of the mixed-precision correction solves
constexpr int kMaxRestart = 120;    // largest GMRES restart length (LDS of k_finalize_scale)
constexpr int kDotSweep = 16;       // basis vectors per sweep of w in k_dots

// Basis-vector read marked non-temporal, for the last reader of V in a step
// (k_update; k_combine once per cycle): k_dots has just read the same
// vectors, and the hint keeps them from displacing the operator's spectrum
// and intermediates in the Infinity Cache before the next matvec (measured,
// C2 GMRES to 1e-6: FP64 38.6 -> 37.4 ms, mixed 28.5 -> 27.4 ms).
__device__ __forceinline__ double2 ld_basis_last(const double2* p) {
  typedef double nv2 __attribute__((ext_vector_type(2)));
  const nv2 t = __builtin_nontemporal_load(reinterpret_cast<const nv2*>(p));
  return make_double2(t.x, t.y);
}
__device__ __forceinline__ double2 ld_v_last(const double2* p) { return ld_basis_last(p); }
__device__ __forceinline__ double2 ld_v_last(const float2* p) {
  typedef float nv2 __attribute__((ext_vector_type(2)));
  const nv2 t = __builtin_nontemporal_load(reinterpret_cast<const nv2*>(p));
  return make_double2(t.x, t.y);
}

// partial sums of |x|^2, grid (nblk, nslot)
__global__ void __launch_bounds__(kGmBlock) k_norm2_part(const double2* __restrict__ x, int64_t n, int64_t rs,
                                                         const int* __restrict__ rmap, double* __restrict__ part) {
  __shared__ double lds[64];
  const int slot = blockIdx.y, nblk = gridDim.x;
  const int rhs = rmap[slot];
  const double2* v = x + rhs * rs;
  double s[1] = {0};
  for (int64_t i = blockIdx.x * (int64_t)kGmBlock + threadIdx.x; i < n; i += (int64_t)nblk * kGmBlock) {
    const double2 a = v[i];
    s[0] += a.x * a.x + a.y * a.y;
  }
  const double r = block_sum<1>(s, lds, threadIdx.x);
  if (threadIdx.x == 0) part[slot * nblk + blockIdx.x] = r;
}

// out[rhs] = sqrt(sum of partials) and, when inv is given, inv[rhs] = 1/out
// (0 for a zero vector), one block per slot
__global__ void k_norm_final(const double* __restrict__ part, int nblk, const int* __restrict__ rmap,
                             double* __restrict__ out, double* __restrict__ inv) {
  __shared__ double lds[64];
  const int slot = blockIdx.x;
  double s[1] = {0};
  for (int i = threadIdx.x; i < nblk; i += kGmBlock) s[0] += part[slot * nblk + i];
  const double r = block_sum<1>(s, lds, threadIdx.x);
  if (threadIdx.x == 0) {
    const double nr = sqrt(r);
    out[rmap[slot]] = nr;
    if (inv) inv[rmap[slot]] = nr > 0 ? 1.0 / nr : 0.0;
  }
}

// dst[rhs] = src[rhs] * scal[rhs] (scal is a reciprocal) in the basis
// precision TV.  grid (nblk, nslot)
template <typename TV>
__global__ void k_scale(const double2* __restrict__ src, int64_t src_rs, TV* __restrict__ dst, int64_t dst_rs,
                        int64_t n, const int* __restrict__ rmap, const double* __restrict__ scal,
                        const int* __restrict__ kdone, int step) {
  const int rhs = rmap[blockIdx.y];
  if (kdone && gm_skip(kdone, rhs, step)) return;
  const double f = scal[rhs];
  const double2* s = src + rhs * src_rs;
  TV* d = dst + rhs * dst_rs;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 v = s[i];
    v.x *= f;
    v.y *= f;
    st_v(d + i, v);
  }
}

// partial dots conj(V_j) . w for j = 0..step, grid (nblk, nslot), grid-stride
// over the vector (coalesced across blocks), JB basis vectors per sweep of w;
// part [slot][2 (step + 1)][nblk] (real, imaginary per dot).  w is stored in
// the basis precision TV (the mixed form's FP32 operator writes it in FP32)
template <int JB, typename TV>
__global__ void __launch_bounds__(kGmBlock) k_dots(const TV* __restrict__ V, int64_t vrs, const TV* __restrict__ W,
                                                   int64_t n, int step, const int* __restrict__ rmap,
                                                   const int* __restrict__ kdone, double* __restrict__ part, int ldp) {
  __shared__ double lds[2 * JB * (kGmBlock / 64)];
  const int slot = blockIdx.y, nblk = gridDim.x;
  const int rhs = rmap[slot];
  if (gm_skip(kdone, rhs, step)) return;
  const TV* Vb = V + rhs * vrs;
  const TV* w = W + rhs * n;
  double* outp = part + (int64_t)slot * ldp * nblk + blockIdx.x;
  for (int j0 = 0; j0 <= step; j0 += JB) {
    double acc[2 * JB];
#pragma unroll
    for (int q = 0; q < 2 * JB; ++q) acc[q] = 0;
    for (int64_t i = blockIdx.x * (int64_t)kGmBlock + threadIdx.x; i < n; i += (int64_t)nblk * kGmBlock) {
      const double2 wv = ld_v(w + i);
#pragma unroll
      for (int q = 0; q < JB; ++q) {
        if (j0 + q <= step) {
          const double2 v = ld_v(Vb + (int64_t)(j0 + q) * n + i);
          acc[2 * q] += v.x * wv.x + v.y * wv.y;       // conj(v) * w
          acc[2 * q + 1] += v.x * wv.y - v.y * wv.x;
        }
      }
    }
    const double r = block_sum_halving<2 * JB>(acc, lds, threadIdx.x);
    if (threadIdx.x < 2 * JB) {
      const int q = threadIdx.x >> 1;
      if (j0 + q <= step) outp[(int64_t)(2 * (j0 + q) + (threadIdx.x & 1)) * nblk] = r;
    }
  }
}

// H(0..step, step) = sums of the dot partials: one wavefront per real or
// imaginary part, grid (nslot, ceil(2 (step + 1) / 4)), 256 threads
__global__ void __launch_bounds__(kGmBlock) k_dots_final(const double* __restrict__ part, int nblk, int ldp, int step,
                                                         int m, const int* __restrict__ rmap,
                                                         const int* __restrict__ kdone, double2* __restrict__ H) {
  const int slot = blockIdx.x;
  const int rhs = rmap[slot];
  if (gm_skip(kdone, rhs, step)) return;
  const int q = blockIdx.y * (kGmBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (q >= 2 * (step + 1)) return;
  const double* src = part + ((int64_t)slot * ldp + q) * nblk;
  double s = 0;
  for (int b = lane; b < nblk; b += 64) s += src[b];
  s = wave_sum(s);
  if (lane == 0) {
    double* h = reinterpret_cast<double*>(H + (int64_t)rhs * (m + 1) * m + (int64_t)(m + 1) * step);
    h[q] = s;
  }
}

// w -= V(:, 0..step) h (w stored back in the basis precision) ; partial |w|^2
// of the stored w.  grid (nblk, nslot)
template <typename TV>
__global__ void __launch_bounds__(kGmBlock) k_update(const TV* __restrict__ V, int64_t vrs, TV* __restrict__ W,
                                                     int64_t n, int step, int m, const int* __restrict__ rmap,
                                                     const int* __restrict__ kdone, const double2* __restrict__ H,
                                                     double* __restrict__ part) {
  __shared__ double2 hs[128];
  __shared__ double lds[64];
  const int slot = blockIdx.y, nblk = gridDim.x;
  const int rhs = rmap[slot];
  if (gm_skip(kdone, rhs, step)) return;
  const double2* hcol = H + (int64_t)rhs * (m + 1) * m + (int64_t)(m + 1) * step;
  for (int q = threadIdx.x; q <= step; q += kGmBlock) hs[q] = hcol[q];
  __syncthreads();
  const TV* Vb = V + rhs * vrs;
  TV* w = W + rhs * n;
  double s[1] = {0};
  for (int64_t i = blockIdx.x * (int64_t)kGmBlock + threadIdx.x; i < n; i += (int64_t)nblk * kGmBlock) {
    double2 a = ld_v(w + i);
    for (int j = 0; j <= step; ++j) {
      const double2 v = ld_v_last(Vb + (int64_t)j * n + i);
      const double2 h = hs[j];
      a.x -= v.x * h.x - v.y * h.y;
      a.y -= v.x * h.y + v.y * h.x;
    }
    const double2 t = st_v(w + i, a);
    s[0] += t.x * t.x + t.y * t.y;
  }
  const double r = block_sum<1>(s, lds, threadIdx.x);
  if (threadIdx.x == 0) part[slot * nblk + blockIdx.x] = r;
}

// Givens step of one slot (block 0 of k_finalize_scale): H(step+1, step) =
// sqrt(r) = ||w||, the earlier rotations applied to column `step`, the new
// rotation, g and the residual history, the convergence flag
__device__ __forceinline__ void givens_step(double r, int step, int m, int rhs, int* __restrict__ kdone,
                                            double2* __restrict__ H, double2* __restrict__ Rm, double* __restrict__ cs,
                                            double2* __restrict__ sn, double2* __restrict__ g, double* __restrict__ rv,
                                            const double* __restrict__ tolabs, double* __restrict__ hinv,
                                            double2* col, double* cst, double2* snt) {
  const int64_t ld = m + 1;
  double2* Hr = H + (int64_t)rhs * ld * m;
  double2* Rr = Rm + (int64_t)rhs * ld * m;
  double* c = cs + (int64_t)rhs * m;
  double2* sv = sn + (int64_t)rhs * m;
  double2* gv = g + (int64_t)rhs * ld;
  // the column and the earlier rotations are staged in LDS by the whole block;
  // one thread then applies the rotations in order
  for (int i = threadIdx.x; i <= step; i += kGmBlock) col[i] = Hr[ld * step + i];
  for (int i = threadIdx.x; i < step; i += kGmBlock) {
    cst[i] = c[i];
    snt[i] = sv[i];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double hn = sqrt(r);
  Hr[ld * step + step + 1] = make_double2(hn, 0.0);
  for (int i = 0; i < step; ++i) {
    const double2 a = col[i], b = col[i + 1];
    const double ci = cst[i];
    const double2 si = snt[i];
    // [a; b] <- [c s; -conj(s) c] [a; b]
    double2 na, nb;
    na.x = ci * a.x + (si.x * b.x - si.y * b.y);
    na.y = ci * a.y + (si.x * b.y + si.y * b.x);
    nb.x = -(si.x * a.x + si.y * a.y) + ci * b.x;
    nb.y = -(si.x * a.y - si.y * a.x) + ci * b.y;
    col[i] = na;
    col[i + 1] = nb;
  }
  // new rotation zeroing the real subdiagonal hn
  const double2 a = col[step];
  const double am = hypot(a.x, a.y);
  double cn;
  double2 snew, rr;
  if (hn == 0.0) {
    cn = 1.0; snew = make_double2(0, 0); rr = a;
  } else if (am == 0.0) {
    cn = 0.0; snew = make_double2(1.0, 0.0); rr = make_double2(hn, 0.0);
  } else {
    const double t = hypot(am, hn);
    const double2 al = make_double2(a.x / am, a.y / am);
    cn = am / t;
    snew = make_double2(al.x * hn / t, al.y * hn / t);
    rr = make_double2(al.x * t, al.y * t);
  }
  for (int i = 0; i < step; ++i) Rr[ld * step + i] = col[i];
  c[step] = cn;
  sv[step] = snew;
  Rr[ld * step + step] = rr;
  Rr[ld * step + step + 1] = make_double2(0, 0);
  const double2 gj = gv[step];
  gv[step] = make_double2(cn * gj.x, cn * gj.y);
  // g(step+1) = -conj(s) g(step)
  gv[step + 1] = make_double2(-(snew.x * gj.x + snew.y * gj.y), -(snew.x * gj.y - snew.y * gj.x));
  const double res = hypot(gv[step + 1].x, gv[step + 1].y);
  rv[(int64_t)rhs * m + step] = res;
  hinv[rhs] = hn > 0 ? 1.0 / hn : 0.0;
  if (res < tolabs[rhs]) kdone[rhs] = step + 1;
}

// H(step+1, step) = ||w||, the Givens update of the least-squares residual
// and the convergence flag (block 0 of each slot), and v_{step+1} = w / ||w||
// (every block).  grid (nblk, nslot).  Every block sums the ||w||^2 partials
// in the same order, so all scale by the reciprocal of the ||w|| that block 0
// stores in H; block 0 sets kdone to step + 1, which no block of this step
// takes for done (gm_skip needs kdone <= step).
template <typename TV>
__global__ void __launch_bounds__(kGmBlock)
    k_finalize_scale(const double* __restrict__ part, int nblk, int step, int m, const int* __restrict__ rmap,
                     int* __restrict__ kdone, double2* __restrict__ H, double2* __restrict__ Rm,
                     double* __restrict__ cs, double2* __restrict__ sn, double2* __restrict__ g,
                     double* __restrict__ rv, const double* __restrict__ tolabs, double* __restrict__ hinv,
                     const TV* __restrict__ W, int64_t n, TV* __restrict__ Vn, int64_t vrs) {
  __shared__ double lds[64];
  __shared__ double2 col[kMaxRestart + 1];   // column `step` of H, rotated in place
  __shared__ double cst[kMaxRestart];
  __shared__ double2 snt[kMaxRestart];
  __shared__ double rsum;
  const int slot = blockIdx.y;
  const int rhs = rmap[slot];
  if (gm_skip(kdone, rhs, step)) return;
  double s[1] = {0};
  for (int i = threadIdx.x; i < nblk; i += kGmBlock) s[0] += part[slot * nblk + i];
  // block_sum returns the total in thread 0 only: every thread scales by it
  const double r0 = block_sum<1>(s, lds, threadIdx.x);
  if (threadIdx.x == 0) rsum = r0;
  __syncthreads();
  const double r = rsum;
  const double hn = sqrt(r);
  // the reference divides by H(k+1,k) unguarded; an exact breakdown gives a zero vector here
  const double f = hn > 0 ? 1.0 / hn : 0.0;
  if (blockIdx.x == 0) givens_step(r, step, m, rhs, kdone, H, Rm, cs, sn, g, rv, tolabs, hinv, col, cst, snt);
  const TV* src = W + rhs * n;
  TV* dst = Vn + rhs * vrs;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 v = ld_v(src + i);
    v.x *= f;
    v.y *= f;
    st_v(dst + i, v);
  }
}

// per-cycle reset: kdone = 0, g = [||r||, 0, ...]
__global__ void k_cycle_init(const int* __restrict__ rmap, int nslot, int m, int* __restrict__ kdone,
                             double2* __restrict__ g, const double* __restrict__ rn) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslot) return;
  const int rhs = rmap[s];
  if (kdone[rhs] < 0) return;   // retired
  kdone[rhs] = 0;
  g[(int64_t)rhs * (m + 1)] = make_double2(rn[rhs], 0.0);
}

// End of a restart cycle, one wavefront per slot: y = R \ g (R the rotated
// Hessenberg matrix, back substitution over the k = kdone (or m) steps of the
// cycle), res = rhs - Hc y with rhs = [||r||; 0; ...] (the reference forms both
// with Hc \ rhs, is_iter_gmres_vie.m:25-27), written as the coefficients of
// x += V y (coef, kk) and r = V res (coef2, kk2).  Retired rhs get k = 0.
__global__ void __launch_bounds__(64)
    k_cycle_solve(int m, const int* __restrict__ rmap, const int* __restrict__ kdone, const double2* __restrict__ H,
                  const double2* __restrict__ Rm, const double2* __restrict__ g, const double* __restrict__ nrm,
                  double2* __restrict__ coef, double2* __restrict__ coef2, int* __restrict__ kk,
                  int* __restrict__ kk2) {
  __shared__ double2 y[kMaxRestart];
  const int slot = blockIdx.x, lane = threadIdx.x;
  const int rhs = rmap[slot];
  const int kd = kdone[rhs];
  if (kd < 0) {
    if (lane == 0) { kk[slot] = 0; kk2[slot] = 0; }
    return;
  }
  const int k = kd > 0 ? kd : m;
  const int64_t ld = m + 1;
  const double2* Rr = Rm + (int64_t)rhs * ld * m;
  const double2* Hr = H + (int64_t)rhs * ld * m;
  const double2* gv = g + (int64_t)rhs * ld;
  for (int i = k - 1; i >= 0; --i) {
    double sr = 0.0, si = 0.0;
    for (int jj = i + 1 + lane; jj < k; jj += 64) {
      const double2 r = Rr[ld * jj + i], v = y[jj];
      sr += r.x * v.x - r.y * v.y;
      si += r.x * v.y + r.y * v.x;
    }
    sr = wave_sum(sr);
    si = wave_sum(si);
    if (lane == 0) {
      const double2 a = make_double2(gv[i].x - sr, gv[i].y - si), b = Rr[ld * i + i];
      const double den = b.x * b.x + b.y * b.y;
      y[i] = make_double2((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
    }
    __syncthreads();
  }
  double2* cy = coef + (int64_t)slot * ld;
  double2* cr = coef2 + (int64_t)slot * ld;
  for (int i = lane; i < k; i += 64) cy[i] = y[i];
  for (int i = lane; i <= k; i += 64) {
    double rr = i == 0 ? nrm[rhs] : 0.0, ri = 0.0;
    for (int jj = i > 0 ? i - 1 : 0; jj < k; ++jj) {   // Hc(i, jj) = 0 for i > jj + 1
      const double2 h = Hr[ld * jj + i], v = y[jj];
      rr -= h.x * v.x - h.y * v.y;
      ri -= h.x * v.y + h.y * v.x;
    }
    cr[i] = make_double2(rr, ri);
  }
  if (lane == 0) { kk[slot] = k; kk2[slot] = k + 1; }
}

// out[rhs] (+)= sum_{j < kk[slot]} V_j coef[slot][j], grid (nblk, nslot)
template <typename TV>
__global__ void k_combine(const TV* __restrict__ V, int64_t vrs, int64_t n, const int* __restrict__ rmap,
                          const int* __restrict__ kk, const double2* __restrict__ coef, int ldc, double2* __restrict__ out,
                          int64_t out_rs, int accumulate, const int* __restrict__ kdone) {
  __shared__ double2 cf[128];
  const int slot = blockIdx.y;
  const int rhs = rmap[slot];
  if (kdone[rhs] < 0) return;   // retired: x and r stay as they are
  const int k = kk[slot];
  for (int q = threadIdx.x; q < k; q += blockDim.x) cf[q] = coef[(int64_t)slot * ldc + q];
  __syncthreads();
  const TV* Vb = V + rhs * vrs;
  double2* o = out + rhs * out_rs;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 a = accumulate ? o[i] : make_double2(0, 0);
    for (int j = 0; j < k; ++j) {
      const double2 v = ld_v_last(Vb + (int64_t)j * n + i);
      const double2 c = cf[j];
      a.x += v.x * c.x - v.y * c.y;
      a.y += v.x * c.y + v.y * c.x;
    }
    o[i] = a;
  }
}

// y[rhs] = a*y[rhs] + b*x[rhs] (real a, b), grid (nblk, nslot)
__global__ void k_axpby(const double2* __restrict__ x, double2* __restrict__ y, int64_t n, int64_t rs,
                        const int* __restrict__ rmap, double a, double b) {
  const int rhs = rmap[blockIdx.y];
  const double2* xv = x + rhs * rs;
  double2* yv = y + rhs * rs;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    // a zero coefficient never reads its operand (it may be uninitialised)
    double2 yi = make_double2(0.0, 0.0);
    if (a != 0.0) {
      const double2 t = yv[i];
      yi.x = a * t.x;
      yi.y = a * t.y;
    }
    if (b != 0.0) {
      const double2 xi = xv[i];
      yi.x += b * xi.x;
      yi.y += b * xi.y;
    }
    yv[i] = yi;
  }
}

// --------------------------------------------------------------------------
// host side
// --------------------------------------------------------------------------
struct GmWork {
  int64_t n;   // packed length ql*nS
  int B, m;
  int nblk;
  void* V;                          // Krylov basis [B][m+1][n], double2 or (b32) float2
  bool b32;                         // FP32 basis (correction solves of the mixed form): the FP32
                                    // operator reads V and writes W in FP32
  void* W;                          // operator output w [B][n], in the basis precision
  double2 *R, *X, *Bv, *T;          // R, X, Bv, T [B][n]
  double *part;                     // reduction partials
  int ldp;
  // small state
  double2 *H, *Rm, *sn, *g, *coef, *coef2;
  double *cs, *rv, *tolabs, *hinv, *nrm, *bnorm;
  int *kdone, *rmap, *kk, *kk2;
};

// Workspaces for B rhs, restart m of the small state, and a basis of mv + 1
// vectors per rhs in FP64 or (b32) FP32.
static void gm_alloc(Ctx& c, GmWork& w, int64_t n, int B, int m, int mv, bool b32) {
  w.n = n; w.B = B; w.m = m;
  w.nblk = nblocks_for(n);
  w.ldp = 2 * (m + 1);
  const size_t vec = (size_t)n * sizeof(double2);
  c.gm_V.ensure((size_t)B * (mv + 1) * (size_t)n * (b32 ? sizeof(float2) : sizeof(double2)));
  c.gm_w.ensure((size_t)B * vec);
  c.gm_r.ensure((size_t)B * vec);
  c.gm_x.ensure((size_t)B * vec);
  c.gm_b.ensure((size_t)B * vec);
  c.gm_t.ensure((size_t)B * vec);
  w.V = c.gm_V.p;
  w.b32 = b32;
  w.W = c.gm_w.p;
  w.R = c.gm_r.as<double2>();
  w.X = c.gm_x.as<double2>();
  w.Bv = c.gm_b.as<double2>();
  w.T = c.gm_t.as<double2>();
  c.gm_red.ensure((size_t)B * w.nblk * w.ldp * sizeof(double));
  w.part = c.gm_red.as<double>();
  const size_t hm = (size_t)(m + 1) * m;
  const size_t bytes = (size_t)B * (2 * hm + m + 3 * (m + 1)) * sizeof(double2) +
                       (size_t)B * (m + m + 5) * sizeof(double) + (size_t)B * 4 * sizeof(int) + 512;
  c.gm_small.ensure(bytes);
  char* p = c.gm_small.as<char>();
  auto take = [&](size_t sz) { char* q = p; p += (sz + 15) & ~(size_t)15; return q; };
  w.H = (double2*)take((size_t)B * hm * sizeof(double2));
  w.Rm = (double2*)take((size_t)B * hm * sizeof(double2));
  w.sn = (double2*)take((size_t)B * m * sizeof(double2));
  w.g = (double2*)take((size_t)B * (m + 1) * sizeof(double2));
  w.coef = (double2*)take((size_t)B * (m + 1) * sizeof(double2));
  w.coef2 = (double2*)take((size_t)B * (m + 1) * sizeof(double2));
  w.cs = (double*)take((size_t)B * m * sizeof(double));
  w.rv = (double*)take((size_t)B * m * sizeof(double));
  w.tolabs = (double*)take((size_t)B * sizeof(double));
  w.hinv = (double*)take((size_t)B * sizeof(double));
  w.nrm = (double*)take((size_t)B * sizeof(double));
  w.bnorm = (double*)take((size_t)B * sizeof(double));
  w.kdone = (int*)take((size_t)B * sizeof(int));
  w.rmap = (int*)take((size_t)B * sizeof(int));
  w.kk = (int*)take((size_t)B * sizeof(int));
  w.kk2 = (int*)take((size_t)B * sizeof(int));
}

// norms of vectors x[rmap[s]] (stride rs) into w.nrm[rhs]
// (and their reciprocals into inv when given)
static void gm_norms(Ctx& c, const GmWork& w, const double2* x, int64_t rs, int nslot, double* inv = nullptr) {
  hipLaunchKernelGGL(k_norm2_part, dim3(w.nblk, nslot), dim3(kGmBlock), 0, c.stream, x, w.n, rs, w.rmap, w.part);
  hipLaunchKernelGGL(k_norm_final, dim3(nslot), dim3(kGmBlock), 0, c.stream, w.part, w.nblk, w.rmap, w.nrm, inv);
}

static void upload_rmap(Ctx& c, GmWork& w, const std::vector<int>& act) {
  VIE_HIP_CHECK(hipMemcpyAsync(w.rmap, act.data(), act.size() * sizeof(int), hipMemcpyHostToDevice, c.stream));
}

// r[rhs] = b[rhs] - A x[rhs] for the listed rhs (operator at precision `prec`)
static void gm_residual(Ctx& c, GmWork& w, int op, int prec, const double2* b, const double2* x, double2* r,
                        int nslot) {
  ApplyIO io;
  io.in = x; io.out = r; io.in_rs = w.n; io.out_rs = w.n; io.rmap = w.rmap;
  apply_operator(c, op, prec, io, nslot);
  const int g = w.nblk;
  // r = b - A x  :  r = -1 * r + 1 * b
  hipLaunchKernelGGL(k_axpby, dim3(g, nslot), dim3(kGmBlock), 0, c.stream, b, r, w.n, w.n, w.rmap, -1.0, 1.0);
}

// --------------------------------------------------------------------------
// graph capture of blocks of Arnoldi steps
// --------------------------------------------------------------------------
void drop_graphs(Ctx& c) {
  for (auto& kv : c.gm_graphs) (void)hipGraphExecDestroy(kv.second);
  c.gm_graphs.clear();
  for (auto& kv : c.sym_graphs) (void)hipGraphExecDestroy(kv.second);
  c.sym_graphs.clear();
}

// device pointers and values the captured launches bake in
static std::vector<intptr_t> graph_signature(const Ctx& c, const GmWork& w) {
  std::vector<intptr_t> s = {(intptr_t)w.V, (intptr_t)w.b32, (intptr_t)w.W, (intptr_t)w.R, (intptr_t)w.X, (intptr_t)w.part,
                             (intptr_t)w.H, (intptr_t)w.Rm, (intptr_t)w.cs, (intptr_t)w.sn, (intptr_t)w.g,
                             (intptr_t)w.rv, (intptr_t)w.tolabs, (intptr_t)w.hinv, (intptr_t)w.nrm, (intptr_t)w.kdone,
                             (intptr_t)w.rmap, (intptr_t)w.coef, (intptr_t)w.coef2, (intptr_t)w.kk, (intptr_t)w.kk2,
                             (intptr_t)w.n, (intptr_t)w.B, (intptr_t)c.gm_pin, (intptr_t)c.wsA1.p, (intptr_t)c.wsA2.p,
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

// The cycle's results in pinned host memory, written by the copies at the end
// of every cycle: kdone and the residual history of the cycle of every rhs;
// `retire` holds the kdone values the host sends back between cycles.
struct CycleHost {
  int* kd;
  double* rv;
  int* retire;
};

static CycleHost cycle_host(Ctx& c, int B, int m) {
  const size_t need = 2 * (size_t)B * sizeof(int) + (size_t)B * m * sizeof(double) + 64;
  if (c.gm_pin_bytes < need) {
    if (c.gm_pin) VIE_HIP_CHECK(hipHostFree(c.gm_pin));
    c.gm_pin = nullptr;
    c.gm_pin_bytes = 0;
    VIE_HIP_CHECK(hipHostMalloc(&c.gm_pin, need, hipHostMallocDefault));
    c.gm_pin_bytes = need;
  }
  char* p = static_cast<char*>(c.gm_pin);
  CycleHost h;
  h.rv = reinterpret_cast<double*>(p);
  h.kd = reinterpret_cast<int*>(h.rv + (size_t)B * m);
  h.retire = h.kd + B;
  return h;
}

// runs one restart cycle as a graph launch, capturing it on first use
template <typename F>
static void run_cycle_graph(Ctx& c, const GmWork& w, int op, int prec, int ns, F&& enqueue) {
  const auto key = std::make_tuple(op, prec, ns, w.m, w.b32 ? 1 : 0);
  auto it = c.gm_graphs.find(key);
  if (it == c.gm_graphs.end()) {
    VIE_HIP_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    try {
      enqueue();
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(c.stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    hipGraph_t g = nullptr;
    VIE_HIP_CHECK(hipStreamEndCapture(c.stream, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    VIE_HIP_CHECK(e);
    it = c.gm_graphs.emplace(key, ex).first;
  }
  VIE_HIP_CHECK(hipGraphLaunch(it->second, c.stream));
}

// One Arnoldi step: w = A V(:, j), CGS, Givens, v_{j+1} (basis element TV).
template <typename TV>
static void enqueue_step_t(Ctx& c, const GmWork& w, int op, int prec, int ns, int j) {
  const int64_t n = w.n, vrs = (int64_t)(w.m + 1) * n;
  TV* V = static_cast<TV*>(w.V);
  TV* W = static_cast<TV*>(w.W);
  ApplyIO io;
  // an FP32 basis goes through the operator's FP32 packed input and output
  io.in = reinterpret_cast<const double2*>(V + (int64_t)j * n); io.in_rs = vrs;
  io.out = reinterpret_cast<double2*>(W); io.out_rs = n;
  io.io32 = w.b32;
  io.rmap = w.rmap; io.kdone = w.kdone; io.step = j;
  apply_operator(c, op, prec, io, ns);
  hipLaunchKernelGGL((k_dots<kDotSweep, TV>), dim3(w.nblk, ns), dim3(kGmBlock), 0, c.stream, V, vrs, W, n, j, w.rmap,
                     w.kdone, w.part, w.ldp);
  hipLaunchKernelGGL(k_dots_final, dim3(ns, (2 * (j + 1) + 3) / 4), dim3(kGmBlock), 0, c.stream, w.part, w.nblk,
                     w.ldp, j, w.m, w.rmap, w.kdone, w.H);
  hipLaunchKernelGGL(k_update<TV>, dim3(w.nblk, ns), dim3(kGmBlock), 0, c.stream, V, vrs, W, n, j, w.m, w.rmap,
                     w.kdone, w.H, w.part);
  hipLaunchKernelGGL(k_finalize_scale<TV>, dim3(w.nblk, ns), dim3(kGmBlock), 0, c.stream, w.part, w.nblk, j, w.m,
                     w.rmap, w.kdone, w.H, w.Rm, w.cs, w.sn, w.g, w.rv, w.tolabs, w.hinv, W, n,
                     V + (int64_t)(j + 1) * n, vrs);
  VIE_HIP_CHECK(hipGetLastError());
}

// One restart cycle: V(:,1) = r / ||r||, rhs(1) = ||r||, m steps (a rhs that
// converges inside the cycle skips the rest), y = R \ g and res = rhs - H y
// on the device, x += V y, r = V res, and the cycle's state to pinned memory.
template <typename TV>
static void enqueue_cycle_t(Ctx& c, const GmWork& w, int op, int prec, int ns, const CycleHost& hb) {
  const int m = w.m;
  const int64_t n = w.n, vrs = (int64_t)(m + 1) * n;
  TV* V = static_cast<TV*>(w.V);
  gm_norms(c, w, w.R, n, ns, w.hinv);
  hipLaunchKernelGGL(k_scale<TV>, dim3(w.nblk, ns), dim3(kGmBlock), 0, c.stream, w.R, n, V, vrs, n, w.rmap, w.hinv,
                     (const int*)w.kdone, 0);
  hipLaunchKernelGGL(k_cycle_init, dim3((ns + 63) / 64), dim3(64), 0, c.stream, w.rmap, ns, m, w.kdone, w.g, w.nrm);
  for (int j = 0; j < m; ++j) enqueue_step_t<TV>(c, w, op, prec, ns, j);
  hipLaunchKernelGGL(k_cycle_solve, dim3(ns), dim3(64), 0, c.stream, m, w.rmap, w.kdone, w.H, w.Rm, w.g, w.nrm,
                     w.coef, w.coef2, w.kk, w.kk2);
  hipLaunchKernelGGL(k_combine<TV>, dim3(w.nblk, ns), dim3(kGmBlock), 0, c.stream, V, vrs, n, w.rmap, w.kk, w.coef,
                     m + 1, w.X, n, 1, (const int*)w.kdone);
  hipLaunchKernelGGL(k_combine<TV>, dim3(w.nblk, ns), dim3(kGmBlock), 0, c.stream, V, vrs, n, w.rmap, w.kk2, w.coef2,
                     m + 1, w.R, n, 0, (const int*)w.kdone);
  VIE_HIP_CHECK(hipGetLastError());
  VIE_HIP_CHECK(hipMemcpyAsync(hb.kd, w.kdone, w.B * sizeof(int), hipMemcpyDeviceToHost, c.stream));
  VIE_HIP_CHECK(hipMemcpyAsync(hb.rv, w.rv, (size_t)w.B * m * sizeof(double), hipMemcpyDeviceToHost, c.stream));
}

// Restarted GMRES on A x = b for the rhs in `act` (indices into the work
// arrays).  x (w.X) holds the start; r (w.R) holds b - A x on entry.  Appends
// to resvec (normalised by bnorm) and fills iters.  Returns matvecs done.
static int64_t gmres_core(Ctx& c, GmWork& w, int op, int prec, std::vector<int> act, double tol, int maxit,
                          const std::vector<double>& bnorm, std::vector<std::vector<double>>& resvec,
                          std::vector<int>& p_out, std::vector<int>& outit) {
  const int m = w.m;
  int64_t nmv = 0;
  auto still = [&](int b) { return resvec[b].back() > tol && outit[b] < maxit; };
  {
    std::vector<int> a2;
    for (int b : act) if (still(b)) a2.push_back(b);
    act.swap(a2);
  }
  if (act.empty()) return 0;
  const int ns = (int)act.size();
  std::vector<double> tolh(w.B);
  for (int b : act) tolh[b] = tol * bnorm[b];
  VIE_HIP_CHECK(hipMemcpyAsync(w.tolabs, tolh.data(), w.B * sizeof(double), hipMemcpyHostToDevice, c.stream));
  upload_rmap(c, w, act);
  CycleHost hb = cycle_host(c, w.B, m);
  // every slot starts live (kdone 0); the rest of the batch is not in the slot map
  VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
  for (int b = 0; b < w.B; ++b) hb.retire[b] = 0;
  VIE_HIP_CHECK(hipMemcpyAsync(w.kdone, hb.retire, w.B * sizeof(int), hipMemcpyHostToDevice, c.stream));

  auto enqueue_cycle = [&]() {
    if (w.b32) enqueue_cycle_t<float2>(c, w, op, prec, ns, hb);
    else enqueue_cycle_t<double2>(c, w, op, prec, ns, hb);
  };
  // the operator's workspaces and tile widths are fixed before any capture
  prepare_operator(c, op, prec, ns);
  if (c.use_graphs) {
    auto sig = graph_signature(c, w);
    if (sig != c.gm_sig) {
      drop_graphs(c);
      c.gm_sig.swap(sig);
    }
  }
  std::vector<char> live(w.B, 0);
  for (int b : act) live[b] = 1;
  int nlive = ns;
  while (nlive > 0) {
    for (int b : act)
      if (live[b]) outit[b] += 1;
    if (c.use_graphs) run_cycle_graph(c, w, op, prec, ns, enqueue_cycle);
    else enqueue_cycle();
    VIE_HIP_CHECK(hipEventRecord(c.gm_cycle_ev, c.stream));
    VIE_HIP_CHECK(hipEventSynchronize(c.gm_cycle_ev));
    bool retired = false;
    for (int b : act) {
      if (!live[b]) continue;
      const int k = hb.kd[b] > 0 ? hb.kd[b] : m;
      nmv += k;
      for (int i = 0; i < k; ++i) resvec[b].push_back(hb.rv[(size_t)b * m + i] / bnorm[b]);
      p_out[b] = k;
      if (!still(b)) {
        live[b] = 0;
        --nlive;
        retired = true;
      }
    }
    if (retired && nlive > 0) {
      for (int b = 0; b < w.B; ++b) hb.retire[b] = live[b] ? 0 : -1;
      VIE_HIP_CHECK(hipMemcpyAsync(w.kdone, hb.retire, w.B * sizeof(int), hipMemcpyHostToDevice, c.stream));
    }
  }
  VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
  return nmv;
}

}  // namespace vie

using namespace vie;

extern "C" int vie_gmres(vie_ctx* h, int32_t op, int32_t prec, const double* b, double* x, int32_t nrhs, double tol,
                         int32_t restart, int32_t maxit, int32_t mem, double* resvec, int32_t* nres, int32_t* iters,
                         int32_t* nmatvec) {
  if (!h) return VIE_ERR_ARG;
  Ctx& c = h->c;
  try {
    c.err.clear();
    if (op != VIE_OP_VIE && op != VIE_OP_VIEH) throw Error{VIE_ERR_ARG, "vie_gmres: op must be VIE_OP_VIE or VIE_OP_VIEH"};
    if (prec != VIE_PREC_FP64 && prec != VIE_PREC_FP32) throw Error{VIE_ERR_ARG, "unknown precision"};
    if (nrhs < 1 || restart < 1 || restart > kMaxRestart || maxit < 1)
      throw Error{VIE_ERR_ARG, "vie_gmres: bad nrhs/restart/maxit"};
    if (!(tol > 0)) throw Error{VIE_ERR_ARG, "vie_gmres: tol must be > 0"};
    if (!b || !x) throw Error{VIE_ERR_ARG, "NULL vector"};
    if (!c.have_material) throw Error{VIE_ERR_STATE, "material not set"};
    if (nrhs > 65535 / c.ql) throw Error{VIE_ERR_ARG, "nrhs too large"};
    if (c.box) {   // packed operators on the scatterer's support box
      const int rc = vie_gmres(c.box, op, prec, b, x, nrhs, tol, restart, maxit, mem, resvec, nres, iters, nmatvec);
      if (rc != VIE_OK) c.err = "support box: " + c.box->c.err;
      return rc;
    }
    VIE_HIP_CHECK(hipSetDevice(c.device));
    ensure_ready(c, 0, VIE_PREC_FP64);
    if (prec == VIE_PREC_FP32) ensure_ready(c, 0, VIE_PREC_FP32);
    // run on the context's GMRES stream, ordered after the caller's stream;
    // the call returns after that stream has drained
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
    const int m = restart;
    // the plain form keeps an FP64 basis of restart + 1 vectors per rhs; the
    // mixed form's correction solves an FP32 basis of min(restart, kInnerRestart) + 1
    int m_in = std::min(m, kInnerRestart);
    if (const char* ir = std::getenv("VIE_INNER_RESTART"))   // restart length of the correction solves
      if (std::atoi(ir) > 0) m_in = std::min(m, std::atoi(ir));
    GmWork w{};
    if (prec == VIE_PREC_FP64) gm_alloc(c, w, n, nrhs, m, m, false);
    else gm_alloc(c, w, n, nrhs, m, m_in, true);
    const size_t bytes = (size_t)nrhs * n * sizeof(double2);
    const hipMemcpyKind h2d = mem == VIE_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    VIE_HIP_CHECK(hipMemcpyAsync(w.Bv, b, bytes, h2d, c.stream));
    VIE_HIP_CHECK(hipMemcpyAsync(w.X, x, bytes, h2d, c.stream));
    std::vector<int> all(nrhs);
    for (int i = 0; i < nrhs; ++i) all[i] = i;
    upload_rmap(c, w, all);
    // ||b|| and ||x0||
    std::vector<double> bn(nrhs), xn(nrhs);
    gm_norms(c, w, w.Bv, n, nrhs);
    VIE_HIP_CHECK(hipMemcpyAsync(bn.data(), w.nrm, nrhs * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
    gm_norms(c, w, w.X, n, nrhs);
    VIE_HIP_CHECK(hipMemcpyAsync(xn.data(), w.nrm, nrhs * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
    for (int i = 0; i < nrhs; ++i)
      if (!(bn[i] > 0)) throw Error{VIE_ERR_ARG, "right-hand side is zero"};
    std::vector<std::vector<double>> rv(nrhs);
    std::vector<int> p(nrhs, 1), outit(nrhs, 0);
    std::vector<int64_t> nmv(nrhs, 0);
    // r = b - A x0 where x0 != 0 (is_gmres_vie.m:5-9), else r = b
    std::vector<int> nz, zr;
    for (int i = 0; i < nrhs; ++i) (xn[i] > 0 ? nz : zr).push_back(i);
    if (!zr.empty()) {
      upload_rmap(c, w, zr);
      hipLaunchKernelGGL(k_axpby, dim3(w.nblk, (int)zr.size()), dim3(kGmBlock), 0, c.stream, w.Bv, w.R, n, n, w.rmap,
                         0.0, 1.0);
      VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
    }
    if (!nz.empty()) {
      upload_rmap(c, w, nz);
      gm_residual(c, w, op, VIE_PREC_FP64, w.Bv, w.X, w.R, (int)nz.size());
      for (int i : nz) nmv[i] += 1;
    }
    VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
    upload_rmap(c, w, all);
    gm_norms(c, w, w.R, n, nrhs);
    std::vector<double> rn(nrhs);
    VIE_HIP_CHECK(hipMemcpyAsync(rn.data(), w.nrm, nrhs * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
    for (int i = 0; i < nrhs; ++i) rv[i].push_back(rn[i] / bn[i]);

    if (prec == VIE_PREC_FP64) {
      const int64_t mv = gmres_core(c, w, op, VIE_PREC_FP64, all, tol, maxit, bn, rv, p, outit);
      (void)mv;
      // per-rhs matvec count: sum of inner steps taken
      for (int i = 0; i < nrhs; ++i) nmv[i] += (int64_t)rv[i].size() - 1;
    } else {
      // FP64 iterative refinement around the FP32-operator GMRES.  The
      // correction solve runs with FP64 Arnoldi on r; its tolerance asks for
      // the outer target but not below what the FP32 operator resolves.  It
      // restarts every min(m, 20) steps: the Gram-Schmidt traffic grows with
      // the step, and a correction solve needs only a few digits, so the
      // shorter cycle is cheaper per digit; the cycle budget keeps the matvec
      // cap m * maxit.
      const int max_refine = 20;
      GmWork w_in = w;
      w_in.m = m_in;
      const int maxit_in = (int)std::min<int64_t>((int64_t)maxit * ((m + w_in.m - 1) / w_in.m), 1 << 30);
      std::vector<int> act;
      for (int i = 0; i < nrhs; ++i) if (rv[i].back() > tol) act.push_back(i);
      int it = 0;
      std::vector<std::vector<double>> inner(nrhs);
      std::vector<int> pin(nrhs, 1), oin(nrhs, 0);
      while (!act.empty() && it < max_refine) {
        ++it;
        // d = 0 (X is reused for d: keep x in T)
        const int ns = (int)act.size();
        upload_rmap(c, w, act);
        hipLaunchKernelGGL(k_axpby, dim3(w.nblk, ns), dim3(kGmBlock), 0, c.stream, w.X, w.T, n, n, w.rmap, 0.0, 1.0);
        hipLaunchKernelGGL(k_axpby, dim3(w.nblk, ns), dim3(kGmBlock), 0, c.stream, w.X, w.X, n, n, w.rmap, 0.0, 0.0);
        std::vector<double> rnorm(nrhs, 1.0);
        std::vector<double> tin(nrhs, tol);
        double tol_in = 1.0;
        for (int i : act) {
          const double rel = rv[i].back();
          // absolute target 0.5 tol ||b||, relative to ||r||; the FP32 operator floor is ~1e-6
          double t = 0.5 * tol / rel;
          if (t < 2e-6) t = 2e-6;
          if (t > 0.5) t = 0.5;
          tol_in = std::min(tol_in, t);
          rnorm[i] = rel * bn[i];
          inner[i].assign(1, 1.0);
          oin[i] = 0;
        }
        // inner GMRES on A d = r with ||r|| as the reference norm
        std::vector<double> rn_abs(nrhs);
        for (int i = 0; i < nrhs; ++i) rn_abs[i] = rnorm[i];
        gmres_core(c, w_in, op, VIE_PREC_FP32, act, tol_in, maxit_in, rn_abs, inner, pin, oin);
        for (int i : act) nmv[i] += (int64_t)inner[i].size() - 1;
        // gmres_core leaves the slot map of its last cycle (a subset of act)
        upload_rmap(c, w, act);
        // x = x_prev + d ; r = b - A64 x
        hipLaunchKernelGGL(k_axpby, dim3(w.nblk, ns), dim3(kGmBlock), 0, c.stream, w.T, w.X, n, n, w.rmap, 1.0, 1.0);
        gm_residual(c, w, op, VIE_PREC_FP64, w.Bv, w.X, w.R, ns);
        for (int i : act) nmv[i] += 1;
        gm_norms(c, w, w.R, n, ns);
        VIE_HIP_CHECK(hipMemcpyAsync(rn.data(), w.nrm, nrhs * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
        std::vector<int> a2;
        for (int i : act) {
          rv[i].push_back(rn[i] / bn[i]);
          p[i] = pin[i];
          outit[i] = it;
          if (rv[i].back() > tol) a2.push_back(i);
        }
        act.swap(a2);
      }
    }
    VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
    const hipMemcpyKind d2h = mem == VIE_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    VIE_HIP_CHECK(hipMemcpy(x, w.X, bytes, d2h));
    const int cap = restart * maxit + 1;
    for (int i = 0; i < nrhs; ++i) {
      const int len = (int)std::min<size_t>(rv[i].size(), (size_t)cap);
      if (resvec)
        for (int k = 0; k < len; ++k) resvec[(size_t)i * cap + k] = rv[i][k];
      if (nres) nres[i] = len;
      if (iters) { iters[2 * i] = p[i]; iters[2 * i + 1] = outit[i]; }
      if (nmatvec) nmatvec[i] = (int32_t)nmv[i];
    }
    return VIE_OK;
  } catch (const Error& e) {
    c.err = e.msg;
    return e.code;
  } catch (...) {
    c.err = "unknown error in vie_gmres";
    return VIE_ERR_HIP;
  }
}
