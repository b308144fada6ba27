// Block-wide reductions shared by the device Krylov solvers (GMRES in
// vie_gmres.hip, COCR in vie_cocr.hip): every reduction kernel runs
// kGmBlock = 256 threads, four wavefronts of 64.
#pragma once

#include <hip/hip_runtime.h>

namespace vie {

constexpr int kGmBlock = 256;

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// block-wide sums of NV doubles per thread: thread q < NV receives the sum of
// vals[q] over the block (the value returned in other threads is 0)
template <int NV>
__device__ __forceinline__ double block_sum(const double* vals, double* lds, int tid) {
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double s = wave_sum(vals[q]);
    if (lane == 0) lds[wave * NV + q] = s;
  }
  __syncthreads();
  double s = 0;
  if (tid < NV)
    for (int w = 0; w < kGmBlock / 64; ++w) s += lds[w * NV + tid];
  __syncthreads();
  return s;
}

// Sum over the 64 lanes of a wavefront of each of C values per lane (C a power
// of two): at the exchange with lane ^ M a lane keeps one half of its values
// and sends the other, so C - 1 + log2(64 / C) shuffles replace 6 C.  Call
// with M = 32; afterwards every lane holds the total of value index
// (lane >> (6 - log2 C)) & (C - 1) in v[0].
template <int C, int M>
__device__ __forceinline__ void wave_sum_halving(double* v, int lane) {
  if constexpr (C > 1) {
    const bool up = (lane & M) != 0;
#pragma unroll
    for (int k = 0; k < C / 2; ++k) {
      const double keep = up ? v[k + C / 2] : v[k];
      const double send = up ? v[k] : v[k + C / 2];
      v[k] = keep + __shfl_xor(send, M, 64);
    }
    wave_sum_halving<C / 2, M / 2>(v, lane);
  } else if constexpr (M >= 1) {
    v[0] += __shfl_xor(v[0], M, 64);
    wave_sum_halving<1, M / 2>(v, lane);
  }
}

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }

// block-wide sums of NV doubles per thread (NV a power of two <= 64) by
// halving: thread q < NV receives the sum of vals[q] over the block (vals is
// overwritten; the value returned in other threads is 0)
template <int NV>
__device__ __forceinline__ double block_sum_halving(double* vals, double* lds, int tid) {
  static_assert(NV >= 1 && NV <= 64 && (NV & (NV - 1)) == 0, "block_sum_halving: NV");
  constexpr int SH = 6 - ilog2(NV);
  const int lane = tid & 63, wave = tid >> 6;
  wave_sum_halving<NV, 32>(vals, lane);
  if ((lane & ((1 << SH) - 1)) == 0) lds[wave * NV + (lane >> SH)] = vals[0];
  __syncthreads();
  double s = 0;
  if (tid < NV)
    for (int w = 0; w < kGmBlock / 64; ++w) s += lds[w * NV + tid];
  __syncthreads();
  return s;
}

// grid of the vector kernels: about two elements per thread, at most 2048
// blocks (8 per CU)
inline int nblocks_for(int64_t n) {
  int64_t b = (n + 2 * kGmBlock - 1) / (2 * kGmBlock);
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (int)b;
}

// a rhs whose solve step has already finished (kdone in (0, step]) or that is
// retired (kdone < 0) does no work
__device__ __forceinline__ bool gm_skip(const int* kdone, int rhs, int step) {
  const int kd = kdone[rhs];
  return kd != 0 && kd <= step;
}

// Krylov vector element access: vectors are stored as double2 or float2
// (the mixed form's FP32-stored vectors); arithmetic is FP64 either way
__device__ __forceinline__ double2 ld_v(const double2* p) { return *p; }
__device__ __forceinline__ double2 ld_v(const float2* p) {
  const float2 t = *p;
  return make_double2(t.x, t.y);
}
// stores v rounded to the basis precision and returns the stored value, widened
__device__ __forceinline__ double2 st_v(double2* p, double2 v) {
  *p = v;
  return v;
}
__device__ __forceinline__ double2 st_v(float2* p, double2 v) {
  const float2 t = make_float2((float)v.x, (float)v.y);
  *p = t;
  return make_double2(t.x, t.y);
}

}  // namespace vie
