// Device building blocks for the VIE circulant operator on CDNA4 (gfx950):
// complex helpers, small DFT butterflies and a mixed-radix Stockham FFT that
// runs on pencils staged in LDS.
//
// Twiddles come from a per-length table exp(-2*pi*i*k/P), k = 0..P-1,
// generated on the host in FP64 from the exact angle (see vie_api.hip); the
// butterfly constants below are cos/sin of 2*pi*k/R written out as literals.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace vie {

template <typename T> struct vec2;
template <> struct vec2<double> { using type = double2; };
template <> struct vec2<float> { using type = float2; };
template <typename T> using V2 = typename vec2<T>::type;

template <typename T> __device__ __forceinline__ V2<T> mk(T x, T y) { V2<T> r; r.x = x; r.y = y; return r; }
template <typename C> struct scalar_of;
template <> struct scalar_of<double2> { using type = double; };
template <> struct scalar_of<float2> { using type = float; };

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { C r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { C r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  C r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
// a * conj(b)
template <typename C> __device__ __forceinline__ C cmulc(C a, C b) {
  C r; r.x = a.x * b.x + a.y * b.y; r.y = a.y * b.x - a.x * b.y; return r;
}
template <typename C, typename S> __device__ __forceinline__ C cscale(C a, S s) {
  C r; r.x = a.x * s; r.y = a.y * s; return r;
}
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename C> __device__ __forceinline__ C rot90(C a) {
  C r;
  if (INV) { r.x = -a.y; r.y = a.x; } else { r.x = a.y; r.y = -a.x; }
  return r;
}

constexpr int kMaxStages = 10;

// One FFT length: P = radix[0] * radix[1] * ... ; twiddle table in global memory.
struct FftPlan {
  int P;
  int nst;
  int radix[kMaxStages];
};

// ---------------------------------------------------------------------------
// small DFTs: X[q] = sum_r x[r] exp(sgn * 2 pi i r q / R), sgn = -1 forward
// ---------------------------------------------------------------------------
template <bool INV, typename C> __device__ __forceinline__ void dft2(C* x) {
  C a = x[0], b = x[1];
  x[0] = cadd(a, b); x[1] = csub(a, b);
}

template <bool INV, typename C> __device__ __forceinline__ void dft4(C* x) {
  C t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]);
  C t2 = cadd(x[1], x[3]), t3 = rot90<INV>(csub(x[1], x[3]));
  x[0] = cadd(t0, t2); x[2] = csub(t0, t2);
  x[1] = cadd(t1, t3); x[3] = csub(t1, t3);
}

template <bool INV, typename C> __device__ __forceinline__ void dft8(C* x) {
  using T = typename scalar_of<C>::type;
  const T h = (T)0.70710678118654752440;   // 1/sqrt(2)
  C e[4] = {x[0], x[2], x[4], x[6]};
  C o[4] = {x[1], x[3], x[5], x[7]};
  dft4<INV>(e);
  dft4<INV>(o);
  // o[q] *= w8^q, w8 = exp(-+ i pi / 4)
  C o1 = o[1], o3 = o[3];
  if (INV) {
    o[1].x = h * (o1.x - o1.y); o[1].y = h * (o1.x + o1.y);
    o[3].x = -h * (o3.x + o3.y); o[3].y = h * (o3.x - o3.y);
  } else {
    o[1].x = h * (o1.x + o1.y); o[1].y = h * (o1.y - o1.x);
    o[3].x = h * (o3.y - o3.x); o[3].y = -h * (o3.x + o3.y);
  }
  o[2] = rot90<INV>(o[2]);
#pragma unroll
  for (int q = 0; q < 4; ++q) { x[q] = cadd(e[q], o[q]); x[q + 4] = csub(e[q], o[q]); }
}

// odd prime radix via conjugate-symmetric pairs:
// X[q] = x0 + sum_{r=1}^{(R-1)/2} (x_r + x_{R-r}) cos(2 pi r q / R) + sgn i (x_r - x_{R-r}) sin(2 pi r q / R)
template <int R> struct OddTab;
template <> struct OddTab<3> {
  static __device__ __forceinline__ double c(int k) { return -0.5; }
  static __device__ __forceinline__ double s(int k) { return 0.86602540378443864676; }
};
template <> struct OddTab<5> {
  static __device__ __forceinline__ double c(int k) {
    return k == 1 ? 0.30901699437494742410 : -0.80901699437494742410;
  }
  static __device__ __forceinline__ double s(int k) {
    return k == 1 ? 0.95105651629515357212 : 0.58778525229247312917;
  }
};
template <> struct OddTab<7> {
  static __device__ __forceinline__ double c(int k) {
    return k == 1 ? 0.62348980185873353053 : (k == 2 ? -0.22252093395631440429 : -0.90096886790241912624);
  }
  static __device__ __forceinline__ double s(int k) {
    return k == 1 ? 0.78183148246802980871 : (k == 2 ? 0.97492791218182360702 : 0.43388373911755812048);
  }
};

template <int R, bool INV, typename C> __device__ __forceinline__ void dft_odd(C* x) {
  using T = typename scalar_of<C>::type;
  constexpr int H = (R - 1) / 2;
  C sp[H], sm[H];
#pragma unroll
  for (int r = 1; r <= H; ++r) { sp[r - 1] = cadd(x[r], x[R - r]); sm[r - 1] = csub(x[r], x[R - r]); }
  C y[R];
  C s0 = x[0];
#pragma unroll
  for (int r = 0; r < H; ++r) s0 = cadd(s0, sp[r]);
  y[0] = s0;
#pragma unroll
  for (int q = 1; q <= H; ++q) {
    C a = x[0];
    C b; b.x = 0; b.y = 0;
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      const int k = (r * q) % R;               // angle index in 0..R-1
      const int kk = k <= H ? k : R - k;       // cos(2 pi k / R) = cos(2 pi (R-k) / R)
      const T cr = (T)OddTab<R>::c(kk);
      const T sr = (T)(k <= H ? OddTab<R>::s(kk) : -OddTab<R>::s(kk));
      a.x += cr * sp[r - 1].x; a.y += cr * sp[r - 1].y;
      b.x += sr * sm[r - 1].x; b.y += sr * sm[r - 1].y;
    }
    // forward: a - i b ; inverse: a + i b.  X[R-q] gets the opposite sign.
    C ib = rot90<INV>(b);  // forward: -i b
    y[q] = cadd(a, ib);
    y[R - q] = csub(a, ib);
  }
#pragma unroll
  for (int r = 0; r < R; ++r) x[r] = y[r];
}

template <int R, bool INV, typename C> __device__ __forceinline__ void dft(C* x) {
  if constexpr (R == 2) dft2<INV>(x);
  else if constexpr (R == 4) dft4<INV>(x);
  else if constexpr (R == 8) dft8<INV>(x);
  else dft_odd<R, INV>(x);
}

// ---------------------------------------------------------------------------
// Stockham stage over `npen` pencils of length P held in LDS.
//   LAYOUT_PF (pencil-fastest): element (pen, a) at a * npen + pen
//   otherwise (contiguous pencils): element (pen, a) at pen * P + a
// Work item i -> (pen, j) with the fast index matching the layout so that a
// wavefront touches consecutive LDS words.
// ---------------------------------------------------------------------------
template <int R, bool INV, bool PF, typename C>
__device__ __forceinline__ void stockham_stage(const C* __restrict__ src, C* __restrict__ dst,
                                               const C* __restrict__ tw, int P, int Ns, int npen,
                                               int tid, int nthr) {
  const int PR = P / R;
  const int twstep = P / (Ns * R);
  const int nitems = npen * PR;
  for (int i = tid; i < nitems; i += nthr) {
    int pen, j;
    if (PF) { pen = i % npen; j = i / npen; } else { pen = i / PR; j = i - pen * PR; }
    const int k = j % Ns;
    C x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int a = j + r * PR;
      x[r] = PF ? src[a * npen + pen] : src[pen * P + a];
    }
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const C w = tw[r * k * twstep];
        x[r] = INV ? cmulc(x[r], w) : cmul(x[r], w);
      }
    }
    dft<R, INV>(x);
    const int o = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int a = o + r * Ns;
      if (PF) dst[a * npen + pen] = x[r]; else dst[pen * P + a] = x[r];
    }
  }
}

// Full FFT over all stages; ping-pongs between bufA and bufB, returns the
// buffer holding the result.  Caller syncs before (data in bufA) and the
// routine syncs after every stage.
template <bool INV, bool PF, typename C>
__device__ C* fft_lds(C* bufA, C* bufB, const FftPlan& pl, const C* __restrict__ tw, int npen,
                      int tid, int nthr) {
  C* src = bufA;
  C* dst = bufB;
  int Ns = 1;
  for (int s = 0; s < pl.nst; ++s) {
    const int R = pl.radix[s];
    switch (R) {
      case 8: stockham_stage<8, INV, PF>(src, dst, tw, pl.P, Ns, npen, tid, nthr); break;
      case 4: stockham_stage<4, INV, PF>(src, dst, tw, pl.P, Ns, npen, tid, nthr); break;
      case 2: stockham_stage<2, INV, PF>(src, dst, tw, pl.P, Ns, npen, tid, nthr); break;
      case 3: stockham_stage<3, INV, PF>(src, dst, tw, pl.P, Ns, npen, tid, nthr); break;
      case 5: stockham_stage<5, INV, PF>(src, dst, tw, pl.P, Ns, npen, tid, nthr); break;
      case 7: stockham_stage<7, INV, PF>(src, dst, tw, pl.P, Ns, npen, tid, nthr); break;
      default: break;
    }
    __syncthreads();
    Ns *= R;
    C* t = src; src = dst; dst = t;
  }
  return src;
}

}  // namespace vie
