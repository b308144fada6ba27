// HIP kernels of the VIE circulant operator (gfx950) and their launchers.
//
// One operator application on a voxel grid n1 x n2 x n3 with ql channels
// runs five passes over per-channel arrays laid out x-fastest:
//   1. z forward   : input (n1, n2, n3) [user dense / packed gather]  -> A1 (n1, n2, P3)
//   2. y forward   : A1 (n1, n2, P3)                                 -> A2 (n1, P2, P3)
//   3. x fused     : A2 columns: pad n1 -> P1, FFT, block product with the
//                    octant Green spectrum over all ql channels, inverse FFT,
//                    keep n1                                        -> A2 (in place)
//   4. y inverse   : A2 (n1, P2, P3) keep n2                         -> A1 (n1, n2, P3)
//   5. z inverse   : A1 keep n3, scale 1/(P1 P2 P3), epilogue         -> user dense / packed
// The zero padding of the circulant embedding is never stored: passes 1-2
// read only the n_d nonzero inputs, passes 3-5 write only the n_d outputs
// the crop keeps (mvp_N_pwl_tucker.m:16,36-37 pad with fftn(.,[m1 m2 m3])
// and crop with (1:n1,1:n2,1:n3)).
//
// Pencils along y and z are strided, so passes 1, 2, 4 and 5 work on tiles
// of TX consecutive x and keep the tile pencil-fastest in LDS (every global
// access is TX contiguous complex values).  Pass 3 runs along the contiguous
// x axis, which is why the spectral product lives there: its LDS holds all
// ql channel pencils of one (Y, Z) column and nothing is strided.

#include <algorithm>

#include "vie_passes.hpp"

namespace vie {

template <typename T, int IN, int OUT, bool INV>
__global__ void __launch_bounds__(256) axis_pass(AxisArgs a, FftPlan pl, const V2<T>* __restrict__ tw,
                                                 int TX) {
  using C = V2<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* bufA = reinterpret_cast<C*>(smem_raw);
  C* bufB = bufA + TX * pl.P;
  const int tile = blockIdx.x % a.ntile;
  const int64_t line = blockIdx.x / a.ntile;
  const int cb = blockIdx.y;  // batch slot * ql + channel
  const int slot = cb / a.ql, ch = cb - slot * a.ql;
  if (skip_slot(a.rmap, a.kdone, a.step, slot)) return;
  const int64_t rhs = a.rmap ? a.rmap[slot] : slot;
  const int x0 = tile * TX;
  if (zline_empty(a, x0, TX, line) || xtile_outside(a, x0, TX)) return;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int P = pl.P;

  for (int i = tid; i < TX * P; i += nthr) {
    const int t = i % TX, aa = i / TX;
    bufA[aa * TX + t] = axis_load<T, IN>(a, cb, rhs, ch, x0 + t, line, aa);
  }
  __syncthreads();
  C* res = fft_lds<INV, true>(bufA, bufB, pl, tw, TX, tid, nthr);

  for (int i = tid; i < TX * a.Lout; i += nthr) {
    const int t = i % TX, aa = i / TX;
    const int x = x0 + t;
    if (x >= a.n1) continue;
    axis_store<T, OUT>(a, cb, rhs, ch, x, line, aa, res[aa * TX + t]);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) x_fused_pass(XArgs<T> a, FftPlan pl, const V2<T>* __restrict__ tw) {
  using C = V2<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int P = pl.P;
  C* bufA = reinterpret_cast<C*>(smem_raw);
  C* bufB = bufA + a.ql * P;
  const int col = blockIdx.x;
  const int b = blockIdx.y;
  if (skip_slot(a.rmap, a.kdone, a.step, b)) return;
  const int Y = col % a.P2, Z = col / a.P2;
  const int yo = (2 * Y <= a.P2) ? Y : a.P2 - Y;
  const int zo = (2 * Z <= a.P3) ? Z : a.P3 - Z;
  const int colbits = ((2 * Y > a.P2) ? 2 : 0) | ((2 * Z > a.P3) ? 4 : 0);
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int64_t colbase = (int64_t)b * a.ql * a.cstride + (int64_t)a.n1 * col;

  for (int i = tid; i < a.ql * P; i += nthr) {
    const int c = i / P, x = i - c * P;
    C v = mk<T>(0, 0);
    if (x >= a.xlo && x <= a.xhi) v = a.data[colbase + c * a.cstride + x];
    bufA[i] = v;
  }
  __syncthreads();
  C* X = fft_lds<false, false>(bufA, bufB, pl, tw, a.ql, tid, nthr);
  C* Yb = (X == bufA) ? bufB : bufA;

  const int H = P / 2;
  const C* sp = a.spec + ((int64_t)zo * a.O2 + yo) * a.nG * a.O1;
  for (int i = tid; i < a.ql * H; i += nthr) {
    const int u = i / H, t = i - u * H;
    const int X1 = t, X2 = t ? P - t : H;
    const int xo2 = t ? t : H;
    const int mir2 = t ? 1 : 0;
    T a1x = 0, a1y = 0, a2x = 0, a2y = 0;
    for (int k = 0; k < a.nterm; ++k) {
      const int e = a.terms[u * kMaxTerms + k];
      const int v = e & 0xff, g = (e >> 8) & 0xff;
      const int bits = a.gpar[g];
      int neg = (e >> 16) & 1;
      neg ^= __popc(bits & colbits) & 1;
      const int neg2 = neg ^ (mir2 & bits & 1);
      const C G1 = sp[(int64_t)g * a.O1 + t];
      const C G2 = t ? G1 : sp[(int64_t)g * a.O1 + xo2];
      const C x1 = X[v * P + X1];
      const C x2 = X[v * P + X2];
      C p1 = a.conj ? cmulc(x1, G1) : cmul(x1, G1);
      C p2 = a.conj ? cmulc(x2, G2) : cmul(x2, G2);
      if (neg) { a1x -= p1.x; a1y -= p1.y; } else { a1x += p1.x; a1y += p1.y; }
      if (neg2) { a2x -= p2.x; a2y -= p2.y; } else { a2x += p2.x; a2y += p2.y; }
    }
    Yb[u * P + X1] = mk<T>(a1x, a1y);
    Yb[u * P + X2] = mk<T>(a2x, a2y);
  }
  __syncthreads();
  C* Yo = (Yb == bufA) ? bufB : bufA;
  C* R = fft_lds<true, false>(Yb, Yo, pl, tw, a.ql, tid, nthr);

  for (int i = tid; i < a.ql * a.n1; i += nthr) {
    const int c = i / a.n1, x = i - c * a.n1;
    if (x < a.xlo || x > a.xhi) continue;
    a.data[colbase + c * a.cstride + x] = R[c * P + x];
  }
}

// ---------------------------------------------------------------------------
// spectrum set-up: out(xo, yo, zo) = sum_abc core(a,b,c) F1(a,xo) F2(b,yo) F3(c,zo)
// F_d are r_d x O_d column-major.  Written into the octant layout [zo][yo][g][xo].
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  double2 r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}

__global__ void mode1_kernel(const double2* __restrict__ G, const double2* __restrict__ F1, double2* __restrict__ T1,
                             int r1, int r2, int r3, int O1) {
  const int64_t n = (int64_t)O1 * r2 * r3;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int xo = i % O1;
    const int64_t bc = i / O1;
    double2 acc; acc.x = 0; acc.y = 0;
    for (int aa = 0; aa < r1; ++aa) {
      const double2 p = zmul(G[aa + r1 * bc], F1[aa + (int64_t)r1 * xo]);
      acc.x += p.x; acc.y += p.y;
    }
    T1[i] = acc;
  }
}

__global__ void mode2_kernel(const double2* __restrict__ T1, const double2* __restrict__ F2, double2* __restrict__ T2,
                             int O1, int r2, int r3, int O2) {
  const int64_t n = (int64_t)O1 * O2 * r3;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int xo = i % O1;
    const int yo = (i / O1) % O2;
    const int64_t c = i / ((int64_t)O1 * O2);
    double2 acc; acc.x = 0; acc.y = 0;
    for (int bb = 0; bb < r2; ++bb) {
      const double2 p = zmul(T1[xo + (int64_t)O1 * (bb + r2 * c)], F2[bb + (int64_t)r2 * yo]);
      acc.x += p.x; acc.y += p.y;
    }
    T2[i] = acc;
  }
}

__global__ void mode3_kernel(const double2* __restrict__ T2, const double2* __restrict__ F3, double2* __restrict__ out,
                             int O1, int O2, int r3, int O3, int nG, int g) {
  const int64_t n = (int64_t)O1 * O2 * O3;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int xo = i % O1;
    const int yo = (i / O1) % O2;
    const int zo = i / ((int64_t)O1 * O2);
    double2 acc; acc.x = 0; acc.y = 0;
    for (int cc = 0; cc < r3; ++cc) {
      const double2 p = zmul(T2[xo + (int64_t)O1 * (yo + (int64_t)O2 * cc)], F3[cc + (int64_t)r3 * zo]);
      acc.x += p.x; acc.y += p.y;
    }
    out[(((int64_t)zo * O2 + yo) * nG + g) * O1 + xo] = acc;
  }
}

__global__ void to_fp32_kernel(const double2* __restrict__ in, float2* __restrict__ out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = in[i];
    out[i] = make_float2((float)v.x, (float)v.y);
  }
}

static int grid_for(int64_t n, int block = 256) {
  int64_t g = (n + block - 1) / block;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

void build_spectrum_term(Ctx& c, int opk, int g, const double2* d_core, int r1, int r2, int r3,
                         const double2* d_F1, const double2* d_F2, const double2* d_F3, double2* d_t1,
                         double2* d_t2) {
  const int O1 = c.O[0], O2 = c.O[1], O3 = c.O[2];
  Spectrum& s = c.spec[opk];
  hipLaunchKernelGGL(mode1_kernel, dim3(grid_for((int64_t)O1 * r2 * r3)), dim3(256), 0, c.stream, d_core, d_F1,
                     d_t1, r1, r2, r3, O1);
  hipLaunchKernelGGL(mode2_kernel, dim3(grid_for((int64_t)O1 * O2 * r3)), dim3(256), 0, c.stream, d_t1, d_F2,
                     d_t2, O1, r2, r3, O2);
  hipLaunchKernelGGL(mode3_kernel, dim3(grid_for((int64_t)O1 * O2 * O3)), dim3(256), 0, c.stream, d_t2, d_F3,
                     s.d64.as<double2>(), O1, O2, r3, O3, s.nG, g);
  VIE_HIP_CHECK(hipGetLastError());
}

__global__ void crop_kernel(const double2* __restrict__ src, double2* __restrict__ dst, int n1, int n2, int64_t nsrc,
                            int w1, int w2, int w3, int64_t nout) {
  const int64_t nw = (int64_t)w1 * w2 * w3;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nout; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = i / nw, r = i - k * nw;
    const int m1 = (int)(r % w1);
    const int m2 = (int)((r / w1) % w2);
    const int m3 = (int)(r / ((int64_t)w1 * w2));
    dst[i] = src[k * nsrc + m1 + (int64_t)n1 * (m2 + (int64_t)n2 * m3)];
  }
}

void crop_offsets(Ctx& c, const double2* src, double2* dst, const int n[3], const int w[3], int nk) {
  for (int d = 0; d < 3; ++d)
    if (w[d] < 1 || w[d] > n[d]) throw Error{VIE_ERR_ARG, "internal: box extent outside the grid"};
  const int64_t nsrc = (int64_t)n[0] * n[1] * n[2];
  const int64_t nout = (int64_t)w[0] * w[1] * w[2] * nk;
  hipLaunchKernelGGL(crop_kernel, dim3(grid_for(nout)), dim3(256), 0, c.stream, src, dst, n[0], n[1], nsrc, w[0],
                     w[1], w[2], nout);
  VIE_HIP_CHECK(hipGetLastError());
}

void tucker_real_space(Ctx& c, const double2* d_core, int r1, int r2, int r3, const double2* dV1,
                       const double2* dV2, const double2* dV3, double2* dt1, double2* dt2, double2* out) {
  const int n1 = c.n[0], n2 = c.n[1], n3 = c.n[2];
  hipLaunchKernelGGL(mode1_kernel, dim3(grid_for((int64_t)n1 * r2 * r3)), dim3(256), 0, c.stream, d_core, dV1, dt1,
                     r1, r2, r3, n1);
  hipLaunchKernelGGL(mode2_kernel, dim3(grid_for((int64_t)n1 * n2 * r3)), dim3(256), 0, c.stream, dt1, dV2, dt2, n1,
                     r2, r3, n2);
  // with one term (nG = 1, g = 0) the octant layout [zo][yo][g][xo] is [n3][n2][n1]
  hipLaunchKernelGGL(mode3_kernel, dim3(grid_for((int64_t)n1 * n2 * n3)), dim3(256), 0, c.stream, dt2, dV3, out, n1,
                     n2, r3, n3, 1, 0);
  VIE_HIP_CHECK(hipGetLastError());
}

void convert_spectrum_fp32(Ctx& c, int opk) {
  Spectrum& s = c.spec[opk];
  const int64_t n = (int64_t)s.nG * c.O[0] * c.O[1] * c.O[2];
  s.d32.ensure(n * sizeof(float2));
  hipLaunchKernelGGL(to_fp32_kernel, dim3(grid_for(n)), dim3(256), 0, c.stream, s.d64.as<double2>(),
                     s.d32.as<float2>(), n);
  VIE_HIP_CHECK(hipGetLastError());
  s.have32 = true;
}

// tile widths compiled for the axis passes (the compile-time kernels hold the
// same lists); the generic kernel takes any width
static const int kTiles64[] = {4, 8};
static const int kTiles32[] = {4, 8, 16};
constexpr int kTileDefault = 8;
constexpr size_t kMaxLds = 160 * 1024;

// width of axis pass k (0 z-fwd, 1 y-fwd, 2 y-inv, 3 z-inv) along an axis of
// FFT length P: the chosen (or default) width, lowered to the largest compiled
// width whose two LDS pencil buffers and twiddle table fit
// FFT schedule of axis pass k (z forward, y forward, y inverse, z inverse):
// kSchedCT / kSchedCTO Stockham stages (largest radix / odd radices first),
// kSchedRF two-step register split.  The default is the register split for
// FP32 and the Stockham stages, largest radix first, for FP64.
static int sched_for(const Ctx& c, bool fp64, int k) {
  if (c.force_fam) return c.force_fam;
  const int f = c.fam[fp64 ? 0 : 1][k];
  return f ? f : (fp64 ? kSchedCT : kSchedRF);
}

static int tile_for(const Ctx& c, bool fp64, int nrhs, int k, int P) {
  const int want = c.tile[fp64 ? 0 : 1][nrhs > 1 ? 1 : 0][k];
  const int w = want ? want : kTileDefault;
  const int* list = fp64 ? kTiles64 : kTiles32;
  const int nl = fp64 ? 2 : 3;
  const size_t es = fp64 ? sizeof(double2) : sizeof(float2);
  int best = 0;
  for (int i = 0; i < nl; ++i)
    if (list[i] <= w && (2 * (size_t)list[i] + 1) * P * es <= kMaxLds) best = list[i];
  if (!best) throw Error{VIE_ERR_ARG, "FFT length too long for the axis-pass LDS"};
  return best;
}

static void record(const ApplyIO& io, int i, hipStream_t s) {
  if (io.ev) VIE_HIP_CHECK(hipEventRecord(io.ev[i], s));
}

// Kernels whose LDS exceeds the 64 KiB default opt in to the larger CDNA4 LDS.
template <typename... KArgs, typename... Args>
static void launch(void (*k)(KArgs...), dim3 g, dim3 b, size_t lds, hipStream_t s, Args... args) {
  if (lds > 65536)
    VIE_HIP_CHECK(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k, g, b, lds, s, args...);
}

void workspace_bytes(const Ctx& c, int prec, int nrhs, size_t* a1, size_t* a2) {
  const size_t es = prec == VIE_PREC_FP64 ? sizeof(double2) : sizeof(float2);
  const size_t chans = (size_t)nrhs * c.ql;
  *a1 = chans * (size_t)c.n[0] * c.n[1] * c.P[2] * es;
  *a2 = chans * (size_t)c.n[0] * c.P[1] * c.P[2] * es;
}

template <typename T>
static void run_passes(Ctx& c, int op, const ApplyIO& io, int nrhs) {
  const double2* d_in = io.in;
  double2* d_out = io.out;
  using C = V2<T>;
  const int n1 = c.n[0], n2 = c.n[1], n3 = c.n[2];
  const int P1 = c.P[0], P2 = c.P[1], P3 = c.P[2];
  const bool fp64 = sizeof(T) == 8;
  const int chans = nrhs * c.ql;
  const bool packed = (op == VIE_OP_VIE || op == VIE_OP_VIEH);
  const int opk = (op == VIE_OP_K || op == VIE_OP_KH) ? 1 : 0;
  const int tabl = packed ? (op == VIE_OP_VIE ? VIE_OP_N : VIE_OP_NH) : op;
  const bool adj = (op == VIE_OP_NH || op == VIE_OP_KH || op == VIE_OP_VIEH);
  if (io.io32 && (fp64 || !packed)) throw Error{VIE_ERR_ARG, "internal: FP32 packed vectors need a packed FP32 operator"};
  Spectrum& s = c.spec[opk];
  const C* spec = sizeof(T) == 8 ? (const C*)s.d64.p : (const C*)s.d32.p;
  const C* tw[3];
  for (int d = 0; d < 3; ++d) tw[d] = sizeof(T) == 8 ? (const C*)c.tw64[d].p : (const C*)c.tw32[d].p;
  C* A1 = io.ws1 ? static_cast<C*>(io.ws1) : c.wsA1.as<C>();
  C* A2 = io.ws2 ? static_cast<C*>(io.ws2) : c.wsA2.as<C>();
  const int64_t N = c.nvox();
  const int64_t a1c = (int64_t)n1 * n2 * P3;  // per-channel sizes
  const int64_t a2c = (int64_t)n1 * P2 * P3;

  AxisArgs base{};
  base.n1 = n1;
  base.ql = c.ql;
  base.vox2pk = c.vox2pk.as<int>();
  base.nS = c.nS;
  base.mcr = c.mcr.as<double2>();
  base.invg = c.invg.as<double>();
  base.scale = 1.0 / ((double)P1 * P2 * P3);
  base.in_rs = io.in_rs ? io.in_rs : (int64_t)c.ql * c.nS;
  base.out_rs = io.out_rs ? io.out_rs : (int64_t)c.ql * c.nS;
  base.rmap = io.rmap;
  base.kdone = io.kdone;
  base.step = io.step;
  const unsigned char* zl = (packed && c.prune) ? c.zline.as<unsigned char>() : nullptr;
  // x range of the pencils the y and x passes touch (all of n1 without pruning)
  base.xlo = zl ? c.xlo : 0;
  base.xhi = zl ? c.xhi : n1 - 1;

  record(io, 0, c.stream);
  // 1. z forward
  {
    const int TX = tile_for(c, fp64, nrhs, 0, P3);
    AxisArgs a = base;
    a.ntile = (n1 + TX - 1) / TX;
    a.in = d_in; a.out = A1;
    a.Lin = n3; a.Lout = P3;
    a.sin_ = (int64_t)n1 * n2; a.sout = (int64_t)n1 * n2;
    a.lsi = n1; a.lso = n1;
    a.csi = N; a.cso = a1c;
    a.prescale = (op == VIE_OP_VIEH);
    a.zline = zl;
    const size_t lds = 2 * (size_t)TX * P3 * sizeof(C);
    dim3 grid(a.ntile * n2, chans);
    const int ik = packed ? (io.io32 ? IN_PACKED32 : IN_PACKED) : IN_DENSE;
    if (c.use_fast && launch_axis_fast(ik, OUT_INTERNAL, false, fp64, P3, TX, sched_for(c, fp64, 0), a, grid, lds,
                                       c.stream, tw[2])) {
    } else if (packed && io.io32) {
      if constexpr (sizeof(T) == 4)
        launch((axis_pass<T, IN_PACKED32, OUT_INTERNAL, false>), grid, dim3(256), lds, c.stream, a, c.plan[2], tw[2],
               TX);
    } else if (packed) {
      launch((axis_pass<T, IN_PACKED, OUT_INTERNAL, false>), grid, dim3(256), lds, c.stream, a, c.plan[2], tw[2], TX);
    } else {
      launch((axis_pass<T, IN_DENSE, OUT_INTERNAL, false>), grid, dim3(256), lds, c.stream, a, c.plan[2], tw[2], TX);
    }
  }
  record(io, 1, c.stream);
  // 2. y forward
  {
    const int TX = tile_for(c, fp64, nrhs, 1, P2);
    AxisArgs a = base;
    a.ntile = (n1 + TX - 1) / TX;
    a.in = A1; a.out = A2;
    a.ymask_in = zl;
    a.Lin = n2; a.Lout = P2;
    a.sin_ = n1; a.sout = n1;
    a.lsi = (int64_t)n1 * n2; a.lso = (int64_t)n1 * P2;
    a.csi = a1c; a.cso = a2c;
    const size_t lds = 2 * (size_t)TX * P2 * sizeof(C);
    dim3 grid(a.ntile * P3, chans);
    if (!(c.use_fast && launch_axis_fast(IN_INTERNAL, OUT_INTERNAL, false, fp64, P2, TX, sched_for(c, fp64, 1), a, grid,
                                         lds, c.stream, tw[1])))
      launch((axis_pass<T, IN_INTERNAL, OUT_INTERNAL, false>), grid, dim3(256), lds, c.stream, a, c.plan[1], tw[1], TX);
  }
  record(io, 2, c.stream);
  // 3. x fused: FFT, block product, inverse FFT
  {
    XArgs<T> a{};
    a.data = A2;
    a.cstride = a2c;
    a.n1 = n1; a.P2 = P2; a.P3 = P3; a.O1 = c.O[0]; a.O2 = c.O[1];
    a.ql = c.ql; a.nG = s.nG;
    a.spec = spec;
    a.terms = c.terms.as<int>() + tabl * c.ql * kMaxTerms;
    a.nterm = c.nterm[tabl];
    a.gpar = s.gpar.as<int>();
    a.conj = adj ? 1 : 0;
    a.rmap = io.rmap;
    a.kdone = io.kdone;
    a.step = io.step;
    a.colmap = c.xcolmap.as<int>();
    a.nslot = nrhs;
    a.xlo = base.xlo;
    a.xhi = base.xhi;
    a.stage = c.xstage ? 1 : 0;
    a.work = (int64_t)P2 * P3 * nrhs;
    a.per_xcd = (a.work + 7) / 8;
    const size_t lds = 2 * (size_t)c.ql * P1 * sizeof(C);
    const bool isk = opk == 1, pwl = c.basis == VIE_BASIS_PWL;
    if (!(c.use_fast && launch_x_fast(fp64, pwl, isk, adj, P1, &a, 8 * a.per_xcd, lds, c.stream, tw[0]))) {
      dim3 grid(P2 * P3, nrhs);
      launch((x_fused_pass<T>), grid, dim3(256), lds, c.stream, a, c.plan[0], tw[0]);
    }
  }
  record(io, 3, c.stream);
  // 4. y inverse (keep n2)
  {
    const int TX = tile_for(c, fp64, nrhs, 2, P2);
    AxisArgs a = base;
    a.ntile = (n1 + TX - 1) / TX;
    a.in = A2; a.out = A1;
    a.ymask_out = zl;
    a.Lin = P2; a.Lout = n2;
    a.sin_ = n1; a.sout = n1;
    a.lsi = (int64_t)n1 * P2; a.lso = (int64_t)n1 * n2;
    a.csi = a2c; a.cso = a1c;
    const size_t lds = 2 * (size_t)TX * P2 * sizeof(C);
    dim3 grid(a.ntile * P3, chans);
    if (!(c.use_fast && launch_axis_fast(IN_INTERNAL, OUT_INTERNAL, true, fp64, P2, TX, sched_for(c, fp64, 2), a, grid,
                                         lds, c.stream, tw[1])))
      launch((axis_pass<T, IN_INTERNAL, OUT_INTERNAL, true>), grid, dim3(256), lds, c.stream, a, c.plan[1], tw[1], TX);
  }
  record(io, 4, c.stream);
  // 5. z inverse (keep n3) + epilogue
  {
    const int TX = tile_for(c, fp64, nrhs, 3, P3);
    AxisArgs a = base;
    a.ntile = (n1 + TX - 1) / TX;
    a.in = A1; a.out = d_out;
    a.zline = zl;
    a.Lin = P3; a.Lout = n3;
    a.sin_ = (int64_t)n1 * n2; a.sout = (int64_t)n1 * n2;
    a.lsi = n1; a.lso = n1;
    a.csi = a1c; a.cso = N;
    a.jin = d_in;
    const size_t lds = 2 * (size_t)TX * P3 * sizeof(C);
    dim3 grid(a.ntile * n2, chans);
    const int ok = op == VIE_OP_VIE ? (io.io32 ? OUT_PACKED_VIE32 : OUT_PACKED_VIE)
                   : op == VIE_OP_VIEH ? (io.io32 ? OUT_PACKED_VIEH32 : OUT_PACKED_VIEH) : OUT_DENSE;
    if (c.use_fast && launch_axis_fast(IN_INTERNAL, ok, true, fp64, P3, TX, sched_for(c, fp64, 3), a, grid, lds, c.stream,
                                       tw[2])) {
    } else if (io.io32) {
      if constexpr (sizeof(T) == 4) {
        if (op == VIE_OP_VIE)
          launch((axis_pass<T, IN_INTERNAL, OUT_PACKED_VIE32, true>), grid, dim3(256), lds, c.stream, a, c.plan[2],
                 tw[2], TX);
        else
          launch((axis_pass<T, IN_INTERNAL, OUT_PACKED_VIEH32, true>), grid, dim3(256), lds, c.stream, a, c.plan[2],
                 tw[2], TX);
      }
    } else if (op == VIE_OP_VIE) {
      launch((axis_pass<T, IN_INTERNAL, OUT_PACKED_VIE, true>), grid, dim3(256), lds, c.stream, a, c.plan[2], tw[2], TX);
    } else if (op == VIE_OP_VIEH) {
      launch((axis_pass<T, IN_INTERNAL, OUT_PACKED_VIEH, true>), grid, dim3(256), lds, c.stream, a, c.plan[2], tw[2],
             TX);
    } else {
      launch((axis_pass<T, IN_INTERNAL, OUT_DENSE, true>), grid, dim3(256), lds, c.stream, a, c.plan[2], tw[2], TX);
    }
  }
  record(io, 5, c.stream);
  VIE_HIP_CHECK(hipGetLastError());
}

// Schedule and tile-width autotuner.  Every candidate (FFT schedule of the
// axis passes: Stockham stages in either radix order or the register split,
// x compiled tile width) runs the whole operator three times on
// zero scratch vectors of the caller's shape with events around each pass;
// each axis pass keeps the fastest candidate.  Runs once per (precision, one
// rhs / several rhs) and context, on the first application.  The schedule of a
// pass is fixed by the first tuning of its precision (or by VIE_AXIS_FFT); a
// later tuning for the other batch shape picks only the width.
void autotune_tiles(Ctx& c, int op, int prec, int nrhs) {
  const bool fp64 = prec == VIE_PREC_FP64;
  const int pi = fp64 ? 0 : 1, b = nrhs > 1 ? 1 : 0;
  const int* list = fp64 ? kTiles64 : kTiles32;
  const int nl = fp64 ? 2 : 3;
  const int defam = fp64 ? kSchedCT : kSchedRF;   // Stockham stages for FP64, register split for FP32
  const bool packed = (op == VIE_OP_VIE || op == VIE_OP_VIEH);
  const size_t len = (size_t)nrhs * c.ql * (packed ? (size_t)c.nS : (size_t)c.nvox()) * sizeof(double2);
  int fam0[4];
  for (int k = 0; k < 4; ++k) fam0[k] = c.force_fam ? c.force_fam : c.fam[pi][k];
  struct Cand { int fam, tx; };
  std::vector<Cand> cand;
  for (int f = kSchedCT; f <= kSchedCTO; ++f) {
    bool used = false;   // some pass may still take schedule f
    for (int k = 0; k < 4; ++k) used = used || fam0[k] == 0 || fam0[k] == f;
    if (used)
      for (int i = 0; i < nl; ++i) cand.push_back({f, list[i]});
  }
  const int nc = (int)cand.size();
  // Every run of every candidate is queued back to back with its own six
  // events and the host waits once at the end: each timed pass starts right
  // behind the previous kernel, so no span holds host launch latency (the
  // first pass of a run timed after an idle GPU would).
  constexpr int kReps = 3;
  struct Scratch {
    DBuf in, out;
    std::vector<hipEvent_t> ev;
    ~Scratch() {
      in.release();
      out.release();
      for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    }
  } sc;
  sc.ev.assign((size_t)nc * kReps * 6, nullptr);
  sc.in.ensure(len);
  sc.out.ensure(len);
  VIE_HIP_CHECK(hipMemsetAsync(sc.in.p, 0, len, c.stream));
  for (auto& e : sc.ev) VIE_HIP_CHECK(hipEventCreate(&e));
  size_t a1, a2;
  workspace_bytes(c, prec, nrhs, &a1, &a2);
  c.wsA1.ensure(a1);
  c.wsA2.ensure(a2);
  const int span[4][2] = {{0, 1}, {1, 2}, {3, 4}, {4, 5}};   // events around the axis passes
  std::vector<float> tm((size_t)nc * 4);   // [candidate][pass] fastest of the runs
  ApplyIO io;
  io.in = sc.in.as<double2>();
  io.out = sc.out.as<double2>();
  const int force = c.force_fam;
  c.force_fam = 0;   // candidate runs set the schedule of every pass themselves
  for (int i = 0; i < nc; ++i) {
    for (int k = 0; k < 4; ++k) {
      c.fam[pi][k] = cand[i].fam;
      c.tile[pi][b][k] = cand[i].tx;
    }
    for (int rep = 0; rep < kReps; ++rep) {
      io.ev = sc.ev.data() + ((size_t)i * kReps + rep) * 6;
      if (fp64) run_passes<double>(c, op, io, nrhs);
      else run_passes<float>(c, op, io, nrhs);
    }
  }
  c.force_fam = force;
  VIE_HIP_CHECK(hipEventSynchronize(sc.ev.back()));
  for (int i = 0; i < nc; ++i)
    for (int k = 0; k < 4; ++k) {
      float tmin = 3.0e38f;
      for (int rep = 0; rep < kReps; ++rep) {
        const hipEvent_t* e = sc.ev.data() + ((size_t)i * kReps + rep) * 6;
        float ms = 0;
        VIE_HIP_CHECK(hipEventElapsedTime(&ms, e[span[k][0]], e[span[k][1]]));
        tmin = std::min(tmin, ms);
      }
      tm[(size_t)i * 4 + k] = tmin;
    }
  // per pass: the default (the pass's schedule, or the precision's default
  // one, at the default width) unless another allowed candidate is at least
  // 5 % faster (differences inside the timing noise keep the default; narrower
  // tiles also fetch whole cache lines for half-used segments)
  for (int k = 0; k < 4; ++k) {
    const int f0 = fam0[k] ? fam0[k] : defam;
    int pick = -1;
    float best = 3.0e38f;
    for (int i = 0; i < nc; ++i)
      if (cand[i].fam == f0 && cand[i].tx == kTileDefault) { pick = i; best = tm[(size_t)i * 4 + k]; }
    for (int i = 0; i < nc; ++i) {
      const bool allowed = fam0[k] == 0 || cand[i].fam == fam0[k];
      if (allowed && i != pick && tm[(size_t)i * 4 + k] < 0.95f * best) {
        best = tm[(size_t)i * 4 + k];
        pick = i;
      }
    }
    if (pick < 0) throw Error{VIE_ERR_HIP, "internal: no default tuning candidate"};
    c.tile[pi][b][k] = cand[pick].tx;
    c.fam[pi][k] = cand[pick].fam;
  }
}

void prepare_operator(Ctx& c, int op, int prec, int nrhs) {
  size_t a1, a2;
  workspace_bytes(c, prec, nrhs, &a1, &a2);
  c.wsA1.ensure(a1);
  c.wsA2.ensure(a2);
  const int pi = prec == VIE_PREC_FP64 ? 0 : 1, b = nrhs > 1 ? 1 : 0;
  if (c.tile[pi][b][0] == 0) {
    if (c.force_tile || !c.autotune) {
      for (int k = 0; k < 4; ++k) c.tile[pi][b][k] = c.force_tile ? c.force_tile : kTileDefault;
    } else {
      autotune_tiles(c, op, prec, nrhs);
    }
  }
}

void apply_operator(Ctx& c, int op, int prec, const ApplyIO& io, int nrhs) {
  prepare_operator(c, op, prec, nrhs);
  if (prec == VIE_PREC_FP64)
    run_passes<double>(c, op, io, nrhs);
  else
    run_passes<float>(c, op, io, nrhs);
}

void apply_operator_ws(Ctx& c, int op, int prec, const ApplyIO& io, int nrhs) {
  if (!io.ws1 || !io.ws2) throw Error{VIE_ERR_ARG, "internal: no workspaces"};
  if (prec == VIE_PREC_FP64)
    run_passes<double>(c, op, io, nrhs);
  else
    run_passes<float>(c, op, io, nrhs);
}

}  // namespace vie
