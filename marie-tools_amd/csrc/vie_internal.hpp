// Internal declarations shared by the HIP translation units of libvie_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "vie_fft.hpp"
#include "../../include/vie_op.h"

namespace vie {

// Error carried from the launch helpers to the C ABI boundary.
struct Error {
  int code;
  std::string msg;
};

#define VIE_HIP_CHECK(expr)                                                        \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      throw ::vie::Error{VIE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
  } while (0)

// Device allocation that reports VIE_ERR_OOM (and honours the fault-injection
// hook VIE_FAULT_INJECT_OOM=1 used by the host-side failure tests).
void* dmalloc(size_t bytes);
void dfree(void* p);

// A growable device buffer.
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void ensure(size_t need) {
    if (need <= bytes) return;
    dfree(p);
    p = nullptr;
    bytes = 0;
    p = dmalloc(need);
    bytes = need;
  }
  void release() { dfree(p); p = nullptr; bytes = 0; }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

// One spectral term of an output channel: input channel v, unique spectrum g,
// sign.  Packed as v | g << 8 | neg << 16.
inline int pack_term(int v, int g, int neg) { return v | (g << 8) | (neg << 16); }
// row length of the coupling tables (an output channel couples to at most 12 inputs)
constexpr int kMaxTerms = 12;

// Operator data built once per (op, precision).
struct Spectrum {
  int nG = 0;          // unique spectra
  bool ready = false;  // all terms set
  std::vector<char> have;  // per term set?
  DBuf d64;            // octant spectra FP64 [zo][yo][g][xo]
  DBuf d32;            // FP32 copy (made on first FP32 use)
  bool have32 = false;
  DBuf gpar;           // per g parity bits: 1 = odd along x, 2 = y, 4 = z
};

struct Ctx {
  int n[3];
  int P[3];
  int O[3];   // octant extent P/2 + 1
  int basis;
  int ql;
  int npair;  // 1 (PWC) or 10 (PWL)
  double res;
  int device;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;

  Spectrum spec[2];     // [0] N, [1] K
  DBuf terms;           // coupling tables for ops N, K, NH, KH: [4][ql][nterm]
  int nterm[4];

  // compile-time-length kernels where a specialisation exists (VIE_GENERIC_KERNELS=1 turns them off)
  bool use_fast = true;
  bool xstage = true;   // x pass: stage the octant spectrum column in LDS where that kernel exists
  // work order of the fused x pass: (Y, Z) columns grouped by octant column
  DBuf xcolmap;

  // FFT plans and twiddles
  FftPlan plan[3];
  DBuf tw64[3], tw32[3];

  // material
  int64_t nS = 0;
  DBuf vox2pk;          // int32 [N]
  DBuf mcr;             // double2 [nS]
  DBuf invg;            // double [ql]
  std::vector<double> invg_h;  // host copy of invg
  DBuf zline;           // uint8 [n1 n2]: 1 where the z line (x, y) holds a scatterer voxel
  int xlo = 0, xhi = -1;  // x range of the scatterer voxels (empty when there are none)
  bool prune = true;    // VIE_PRUNE=0: no scatterer-support pruning in the packed operators
  bool have_material = false;

  // Support box.  The packed operators (VIE, VIE^H, their GMRES) need only
  // the kernel offsets inside the scatterer's bounding box: when a circulant
  // of that box (FFT lengths >= 2 w_d) is smaller than the grid's, they run on
  // a child context over the box, with the spectrum of the N kernel cropped to
  // offsets < w_d and the material cropped to the box; packed vectors keep
  // their order, so they pass through unchanged.  The real-space N kernel is
  // kept for this (as given to vie_set_green_dense, or rebuilt from each
  // Tucker term).  VIE_BOX=0 or VIE_PRUNE=0: full grid.
  vie_ctx* box = nullptr;
  bool use_box = true;
  int box_lo[3] = {0, 0, 0}, box_ext[3] = {0, 0, 0};  // scatterer bounding box (origin, extent)
  int box_w[3] = {0, 0, 0};                          // extent of the current child
  DBuf green_src;                      // real-space N kernel [n1 n2 n3][npair][6], all offsets < n_d
  std::vector<char> green_src_have;    // per term of N: present in green_src
  bool green_src_ok = false;           // every term present
  bool box_green_dirty = false;        // the child's spectrum is stale
  std::vector<double> box_Mr, box_Mc;  // material cropped to the box (interleaved complex)
  std::vector<int64_t> box_idx;        // idxS in box coordinates

  // axis-pass tile widths (pencils per block), per precision (0 FP64, 1 FP32),
  // batch shape (0: one rhs, 1: several) and axis pass (z forward, y forward,
  // y inverse, z inverse); 0 = not tuned yet.  The tiling does not change the
  // arithmetic of any pencil, so every width gives bitwise the same result.
  int tile[2][2][4] = {};
  // FFT schedule of each compile-time-length axis pass, per precision:
  // 0 = not fixed yet (the default), 1 = Stockham mixed-radix stages
  // (axis_pass_ct), 2 = two-step A x B register split (axis_pass_rf).  The
  // schedules round differently, so it is fixed per precision and pass, not
  // per batch shape: batched and single applications keep the same arithmetic.
  int fam[2][4] = {};
  bool autotune = true;   // VIE_AUTOTUNE=0: keep the default width and schedule
  int force_tile = 0;     // VIE_AXIS_TX=w: use width w wherever it is compiled
  int force_fam = 0;      // VIE_AXIS_FFT=ct|rf: use that schedule in every axis pass

  // workspaces (per precision): A1 (n1, n2, P3) and A2 (n1, P2, P3) per channel
  DBuf wsA1, wsA2;
  // host-pointer staging
  DBuf stage_in, stage_out;
  // GMRES workspace
  DBuf gm_V, gm_w, gm_x, gm_r, gm_b, gm_red, gm_small, gm_t;
  // GMRES runs on its own stream (graph capture needs a non-legacy stream);
  // one captured graph per restart cycle keyed by (op, prec, rhs slots,
  // restart, basis precision), valid while gm_sig (the device and pinned
  // pointers and tile widths it bakes in) is unchanged.  VIE_GRAPHS=0
  // launches the same kernels eagerly.
  hipStream_t gm_stream = nullptr;
  hipEvent_t gm_order = nullptr;
  // end of a cycle's copies to gm_pin (pinned: flags, residual history and
  // ||r|| of the cycle, and the retire flags the host sends back)
  hipEvent_t gm_cycle_ev = nullptr;
  void* gm_pin = nullptr;
  size_t gm_pin_bytes = 0;
  bool use_graphs = true;
  std::map<std::tuple<int, int, int, int, int>, hipGraphExec_t> gm_graphs;
  std::vector<intptr_t> gm_sig;
  // COCR (vie_cocr.hip): one captured graph per block of iterations keyed by
  // (op, prec, rhs slots, iterations, vector precision), valid while sym_sig
  // is unchanged; it uses the GMRES workspaces above
  std::map<std::tuple<int, int, int, int, int>, hipGraphExec_t> sym_graphs;
  std::vector<intptr_t> sym_sig;
  // block COCG (vie_bcocg.hip): the blocks of one rhs group, the Gram
  // partials and small matrices, the bilinear-form weights, and a pinned
  // block for the s x s matrices exchanged with the host every pass
  DBuf bk_v, bk_part, bk_w;
  void* bk_pin = nullptr;
  size_t bk_pin_bytes = 0;

  int64_t nvox() const { return (int64_t)n[0] * n[1] * n[2]; }
};

// One batched operator application.  Packed user vectors of batch slot s
// live at in + rmap[s] * in_rs (rmap null: s * in_rs); strides of 0 mean
// back-to-back vectors.  kdone/step let GMRES skip converged rhs.
struct ApplyIO {
  const double2* in = nullptr;
  double2* out = nullptr;
  int64_t in_rs = 0, out_rs = 0;
  const int* rmap = nullptr;
  const int* kdone = nullptr;
  int step = 0;
  // packed ops at FP32 precision only: in and out (and the identity term's
  // input) are complex FP32 vectors (float2) instead of complex FP64
  bool io32 = false;
  hipEvent_t* ev = nullptr;   // optional: 6 events recorded around the five passes
  // optional pass workspaces (A1, A2) in place of the context's buffers
  void* ws1 = nullptr;
  void* ws2 = nullptr;
};

// launchers (vie_kernels.hip)
void build_spectrum_term(Ctx& c, int opk, int g, const double2* d_core, int r1, int r2, int r3,
                         const double2* d_F1, const double2* d_F2, const double2* d_F3,
                         double2* d_t1, double2* d_t2);
void convert_spectrum_fp32(Ctx& c, int opk);
// out[m1 + n1 (m2 + n2 m3)] = sum_abc core(a,b,c) V1(a,m1) V2(b,m2) V3(c,m3), V_d
// r_d x n_d column-major: one Tucker term back in real space (t1, t2: scratch
// of n1 r2 r3 and n1 n2 r3)
void tucker_real_space(Ctx& c, const double2* d_core, int r1, int r2, int r3, const double2* dV1,
                       const double2* dV2, const double2* dV3, double2* dt1, double2* dt2, double2* out);
// sizes the workspaces and fixes the tile widths for (prec, nrhs); after it
// apply_operator with the same shape allocates and synchronises nothing (so
// it may be captured into a graph)
void prepare_operator(Ctx& c, int op, int prec, int nrhs);
void apply_operator(Ctx& c, int op, int prec, const ApplyIO& io, int nrhs);
// the five passes with the workspaces io.ws1 / io.ws2 and the tile widths
// already fixed (the defaults where none were tuned): sizes, tunes and
// allocates nothing
void apply_operator_ws(Ctx& c, int op, int prec, const ApplyIO& io, int nrhs);
// picks the tile width of every axis pass for (prec, batch shape of nrhs) by
// timing each compiled width on scratch vectors; writes them to c.tile
void autotune_tiles(Ctx& c, int op, int prec, int nrhs);
void workspace_bytes(const Ctx& c, int prec, int nrhs, size_t* a1, size_t* a2);

// (vie_api.hip) throws VIE_ERR_STATE unless every term of operator opk
// (0 = N, 1 = K) is set; makes the FP32 copy of the spectrum when asked.
void ensure_ready(Ctx& c, int opk, int prec);
// (vie_gmres.hip) destroys the captured GMRES and COCR graphs
void drop_graphs(Ctx& c);


}  // namespace vie

// the opaque handle of the C ABI
struct vie_ctx {
  vie::Ctx c;
};
