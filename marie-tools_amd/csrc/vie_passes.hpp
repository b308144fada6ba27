// Arguments and element access shared by the pass kernels (generic-length
// kernels in vie_kernels.hip, compile-time-length kernels in vie_fast_*.hip).
#pragma once

#include "vie_internal.hpp"

namespace vie {

// Packed user vectors are complex FP64; the *32 kinds (FP32 operator only)
// read and write them as complex FP32 (the mixed-precision solver's vectors).
enum InKind { IN_INTERNAL = 0, IN_DENSE = 1, IN_PACKED = 2, IN_PACKED32 = 3 };
enum OutKind { OUT_INTERNAL = 0, OUT_DENSE = 1, OUT_PACKED = 2, OUT_PACKED_VIE = 3, OUT_PACKED_VIEH = 4,
               OUT_PACKED_VIE32 = 5, OUT_PACKED_VIEH32 = 6 };

constexpr bool in_packed32(int k) { return k == IN_PACKED32; }
constexpr bool out_packed32(int k) { return k == OUT_PACKED_VIE32 || k == OUT_PACKED_VIEH32; }

// One axis pass over (rhs, channel) blocks of pencils along a strided axis.
struct AxisArgs {
  const void* in;
  void* out;
  int n1;          // tile axis extent (x)
  int Lin, Lout;   // pencil values read / written
  int64_t sin_, sout;  // pencil-axis stride (in / out)
  int64_t lsi, lso;    // line stride (in / out)
  int64_t csi, cso;    // stride between (rhs, channel) blocks (in / out)
  int ntile;           // ceil(n1 / TX)
  int ql;
  // packed addressing (first / last pass of the VIE operator)
  const int* vox2pk;
  int64_t nS;
  const double2* mcr;
  const double* invg;
  const void* jin;     // packed input for the identity term (double2, or float2 for the *32 kinds)
  double scale;
  int prescale;        // IN_PACKED: multiply by conj(Mc/Mr) invG (adjoint VIE)
  int64_t in_rs, out_rs;  // rhs stride of packed user vectors (complex elements)
  const int* rmap;        // batch slot -> rhs index of the packed user vectors (or null)
  const int* kdone;       // GMRES: skip rhs whose cycle finished before `step` (or null)
  int step;
  // scatterer-support pruning of the packed operators: zline[x + n1 y] = 1
  // when the z line (x, y) holds a scatterer voxel.  z passes: a block whose
  // TX x's have no such line at its y does nothing (its input is zero, its
  // output is never gathered).  y passes: element (x, y) of those lines reads
  // as zero (ymask_in) or is not stored (ymask_out).  Null: no pruning.
  const unsigned char* zline;
  const unsigned char* ymask_in;
  const unsigned char* ymask_out;
  int64_t nitem;  // compile-time-length kernels: (tile, line) items before the grid is padded
  // y passes of the packed operators: tiles with no x in [xlo, xhi] (the x
  // range of the scatterer) do nothing — y-forward's output there is zero and
  // never read (the x pass reads only that range), y-inverse's is never stored
  int xlo, xhi;
};

__device__ __forceinline__ bool xtile_outside(const AxisArgs& a, int x0, int TX) {
  return x0 > a.xhi || x0 + TX <= a.xlo;
}

// z-pass block of tile x0 on line (y) with no scatterer line: nothing to do
__device__ __forceinline__ bool zline_empty(const AxisArgs& a, int x0, int TX, int64_t line) {
  if (!a.zline) return false;
  for (int t = 0; t < TX; ++t) {
    const int x = x0 + t;
    if (x < a.n1 && a.zline[x + (int64_t)a.n1 * line]) return false;
  }
  return true;
}

// The fused x pass over (Y, Z) columns.
template <typename T>
struct XArgs {
  V2<T>* data;            // A2, (n1, P2, P3) per (rhs, channel)
  int64_t cstride;        // n1 * P2 * P3
  int n1, P2, P3, O1, O2, ql, nG;
  const V2<T>* spec;      // octant spectra [zo][yo][g][xo]
  const int* terms;       // [ql][kMaxTerms] (generic kernel)
  int nterm;
  const int* gpar;        // [nG] parity bits (generic kernel)
  int conj;               // adjoint: conj(spectrum) (generic kernel)
  const int* rmap;
  const int* kdone;
  int step;
  // work order of the specialised kernel: column list grouped by octant
  // column (mirror columns adjacent), rhs fastest, contiguous chunks per XCD
  const int* colmap;      // Y | Z << 16
  int nslot;
  int64_t work;           // columns * nslot
  int64_t per_xcd;        // ceil(work / 8)
  // x values read and written: the rest of each pencil is zero on input (no
  // scatterer has such an x) and unused on output
  int xlo, xhi;
  // specialised kernel: stage each block's octant spectrum column in LDS
  // where that variant exists (XStage)
  int stage;
};

// A GMRES rhs whose restart cycle has already converged does no work.
__device__ __forceinline__ bool skip_slot(const int* rmap, const int* kdone, int step, int slot) {
  if (!kdone) return false;
  const int rhs = rmap ? rmap[slot] : slot;
  const int kd = kdone[rhs];
  return kd != 0 && kd <= step;
}

// element (x, line, aa) of block cb read by an axis pass; zero outside the input
template <typename T, int IN>
__device__ __forceinline__ V2<T> axis_load(const AxisArgs& a, int cb, int64_t rhs, int ch, int x, int64_t line,
                                           int aa) {
  using C = V2<T>;
  if (aa >= a.Lin || x >= a.n1) return mk<T>(0, 0);
  const int64_t off = x + line * a.lsi + aa * a.sin_;
  if constexpr (IN == IN_INTERNAL) {
    if (a.ymask_in && !a.ymask_in[x + (int64_t)a.n1 * aa]) return mk<T>(0, 0);
    return static_cast<const C*>(a.in)[cb * a.csi + off];
  } else if constexpr (IN == IN_DENSE) {
    const double2 d = static_cast<const double2*>(a.in)[cb * a.csi + off];
    return mk<T>((T)d.x, (T)d.y);
  } else {
    const int pk = a.vox2pk[off];
    if (pk < 0) return mk<T>(0, 0);
    const int64_t idx = rhs * a.in_rs + ch * a.nS + pk;
    double2 d;
    if constexpr (IN == IN_PACKED32) {
      const float2 f = static_cast<const float2*>(a.in)[idx];
      d = make_double2(f.x, f.y);
    } else {
      d = static_cast<const double2*>(a.in)[idx];
    }
    if (a.prescale) {
      const double2 m = a.mcr[pk];
      const double g = a.invg[ch];
      double2 r;
      r.x = g * (m.x * d.x + m.y * d.y);  // conj(m) * d
      r.y = g * (m.x * d.y - m.y * d.x);
      d = r;
    }
    return mk<T>((T)d.x, (T)d.y);
  }
}

// element (x, line, aa) of block cb written by an axis pass, with the epilogue
// of the last pass (1/M scale, identity term, contrast, invG, gather)
template <typename T, int OUT>
__device__ __forceinline__ void axis_store(const AxisArgs& a, int cb, int64_t rhs, int ch, int x, int64_t line, int aa,
                                           V2<T> v) {
  using C = V2<T>;
  const int64_t off = x + line * a.lso + aa * a.sout;
  if constexpr (OUT == OUT_INTERNAL) {
    if (a.ymask_out && !a.ymask_out[x + (int64_t)a.n1 * aa]) return;
    static_cast<C*>(a.out)[cb * a.cso + off] = v;
  } else if constexpr (OUT == OUT_DENSE) {
    double2 r;
    r.x = (double)v.x * a.scale;
    r.y = (double)v.y * a.scale;
    static_cast<double2*>(a.out)[cb * a.cso + off] = r;
  } else {
    const int pk = a.vox2pk[off];
    if (pk < 0) return;
    const int64_t idx = rhs * a.out_rs + ch * a.nS + pk;
    const int64_t jdx = rhs * a.in_rs + ch * a.nS + pk;
    double2 r;
    r.x = (double)v.x * a.scale;
    r.y = (double)v.y * a.scale;
    constexpr bool P32 = out_packed32(OUT);
    double2 j = make_double2(0.0, 0.0);
    if constexpr (OUT != OUT_PACKED) {
      if constexpr (P32) {
        const float2 f = static_cast<const float2*>(a.jin)[jdx];
        j = make_double2(f.x, f.y);
      } else {
        j = static_cast<const double2*>(a.jin)[jdx];
      }
    }
    if constexpr (OUT == OUT_PACKED_VIE || OUT == OUT_PACKED_VIE32) {
      const double2 m = a.mcr[pk];
      const double g = a.invg[ch];
      const double px = g * (m.x * r.x - m.y * r.y);
      const double py = g * (m.x * r.y + m.y * r.x);
      r.x = j.x - px;
      r.y = j.y - py;
    } else if constexpr (OUT == OUT_PACKED_VIEH || OUT == OUT_PACKED_VIEH32) {
      r.x = j.x - r.x;
      r.y = j.y - r.y;
    }
    if constexpr (P32) static_cast<float2*>(a.out)[idx] = make_float2((float)r.x, (float)r.y);
    else static_cast<double2*>(a.out)[idx] = r;
  }
}

// FFT lengths with compile-time-length kernels (axis passes / fused x pass)
bool fast_axis_length(int P);
bool fast_x_length(int P);

// dst[k][m1 + w1 (m2 + w2 m3)] = src[k][m1 + n1 (m2 + n2 m3)] for m_d < w_d,
// k < nk: the kernel offsets of a box of extent w inside a grid of extent n
void crop_offsets(Ctx& c, const double2* src, double2* dst, const int n[3], const int w[3], int nk);

// launchers of the compile-time-length kernels (vie_fast_*.hip); return false
// when the length has no specialisation and the generic kernel must run
constexpr int kSchedCT = 1, kSchedRF = 2, kSchedCTO = 3;   // axis-pass FFT schedules
// sched: kSchedCT Stockham stages (axis_pass_ct, largest radix first),
// kSchedRF the two-step register FFT (axis_pass_rf), kSchedCTO Stockham stages
// with the odd radices first
bool launch_axis_fast(int in_kind, int out_kind, bool inv, bool fp64, int P, int TX, int sched, const AxisArgs& a,
                      dim3 grid, size_t lds, hipStream_t s, const void* tw);
bool launch_x_fast(bool fp64, bool pwl, bool isk, bool adj, int P, const void* args, int64_t blocks, size_t lds,
                   hipStream_t s, const void* tw);
bool launch_axis_fast_fp64(int in_kind, int out_kind, bool inv, int P, int TX, int sched, const AxisArgs& a, dim3 grid,
                           size_t lds, hipStream_t s, const void* tw);
bool launch_axis_fast_fp32(int in_kind, int out_kind, bool inv, int P, int TX, int sched, const AxisArgs& a, dim3 grid,
                           size_t lds, hipStream_t s, const void* tw);
// the FP32 packed-I/O kinds (IN_PACKED32, OUT_PACKED_VIE32, OUT_PACKED_VIEH32), FP32 operator only
bool launch_axis_fast_fp32_p32(int in_kind, int out_kind, bool inv, int P, int TX, int sched, const AxisArgs& a,
                               dim3 grid, size_t lds, hipStream_t s, const void* tw);
bool launch_x_fast_fp64(bool pwl, bool isk, bool adj, int P, const void* args, int64_t blocks, size_t lds,
                        hipStream_t s, const void* tw);
bool launch_x_fast_fp32(bool pwl, bool isk, bool adj, int P, const void* args, int64_t blocks, size_t lds,
                        hipStream_t s, const void* tw);

}  // namespace vie
