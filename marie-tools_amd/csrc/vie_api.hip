// C ABI of libvie_hip.so (declared in include/vie_op.h): context life cycle,
// FFT plans, Green-spectrum set-up, material, and operator application.

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>

#include "vie_passes.hpp"

namespace vie {

// --------------------------------------------------------------------------
// device memory
// --------------------------------------------------------------------------
void* dmalloc(size_t bytes) {
  const char* inj = std::getenv("VIE_FAULT_INJECT_OOM");
  if (inj && inj[0] == '1') throw Error{VIE_ERR_OOM, "device allocation refused (VIE_FAULT_INJECT_OOM=1)"};
  if (bytes == 0) bytes = 16;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
    (void)hipGetLastError();
    throw Error{VIE_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes: out of device memory"};
  }
  if (e != hipSuccess) throw Error{VIE_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e)};
  return p;
}

void dfree(void* p) {
  if (p) (void)hipFree(p);
}

// --------------------------------------------------------------------------
// scratch of vie_apply_noalloc
// --------------------------------------------------------------------------
// The pass workspaces of the allocation-free entry point live in this array;
// every device the library runs on holds its own copy until the process ends.
// In host code `scratch` names a host-side stand-in, not device memory: the
// device address comes from hipGetSymbolAddress, once per device at context
// creation (scratch_init), and only that address is used.
constexpr size_t kScratchBytes = (size_t)VIE_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device `dev` (looked up once)
static void scratch_init(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw Error{VIE_ERR_ARG, "device ordinal out of range"};
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return;
  void* p = nullptr;
  VIE_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw Error{VIE_ERR_HIP, "hipGetSymbolAddress(scratch) gave no address"};
  g_scratch_ptr[dev] = p;
}

static void* scratch_of(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (dev >= 0 && dev < kMaxDevices) ? g_scratch_ptr[dev] : nullptr;
}

// --------------------------------------------------------------------------
// tables
// --------------------------------------------------------------------------
// symmetric 3x3 -> N component (mvp_N_pwc_tucker.m:12), 0-based
static const int kPQN[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
// curl coupling of K, signed 1-based (mvp_K_pwc_tucker.m:12)
static const int kPQK[3][3] = {{0, -3, 2}, {3, 0, -1}, {-2, 1, 0}};
// (l, lp) -> signed 1-based pair (mvp_N_pwl_tucker.m:13)
static const int kLLP[4][4] = {{1, 5, 6, 7}, {-5, 2, 8, 9}, {-6, 8, 3, 10}, {-7, 9, 10, 4}};
// adjoint pair signs (mvp_N_herm_adj_pwl_tucker.m:14)
static const int kSLP[4][4] = {{1, -1, -1, -1}, {-1, 1, 1, 1}, {-1, 1, 1, 1}, {-1, 1, 1, 1}};
// per-axis parity of N and K components (assembly_fft_circ_tucker_pwc.m:9-15)
static const int kGN[3][6] = {{+1, -1, -1, +1, +1, +1}, {+1, -1, +1, +1, -1, +1}, {+1, +1, -1, +1, -1, +1}};
static const int kGK[3][3] = {{-1, +1, +1}, {+1, -1, +1}, {+1, +1, -1}};
// per-axis parity of the ten (l, lp) pairs (assembly_fft_circ_tucker_pwl.m:45-47)
static const int kBT[3][10] = {{+1, +1, +1, +1, -1, +1, +1, -1, -1, +1},
                               {+1, +1, +1, +1, +1, -1, +1, -1, +1, -1},
                               {+1, +1, +1, +1, +1, +1, -1, +1, -1, -1}};

static int sgn(int v) { return v < 0 ? -1 : 1; }

// parity of unique spectrum g of operator opk (0 = N, 1 = K) along axis d
static int parity_of(const Ctx& c, int opk, int g, int d) {
  const int ncomp = opk == 0 ? 6 : 3;
  const int j = g % ncomp, i = g / ncomp;
  int s = opk == 0 ? kGN[d][j] : kGK[d][j];
  if (c.basis == VIE_BASIS_PWL) s *= kBT[d][i];
  return s;
}

// coupling tables: [op N,K,NH,KH][ql][kMaxTerms]
static void build_terms(Ctx& c) {
  std::vector<int> tab(4 * c.ql * kMaxTerms, 0);
  for (int op = 0; op < 4; ++op) {
    const bool isK = (op == VIE_OP_K || op == VIE_OP_KH);
    const bool adj = (op == VIE_OP_NH || op == VIE_OP_KH);
    int nt = 0;
    for (int u = 0; u < c.ql; ++u) {
      int k = 0;
      const int p = c.basis == VIE_BASIS_PWL ? u / 4 : u;
      const int l = c.basis == VIE_BASIS_PWL ? u % 4 : 0;
      for (int q = 0; q < 3; ++q) {
        if (isK && q == p) continue;
        const int nlp = c.basis == VIE_BASIS_PWL ? 4 : 1;
        for (int lp = 0; lp < nlp; ++lp) {
          const int v = c.basis == VIE_BASIS_PWL ? 4 * q + lp : q;
          int s = 1, g;
          const int comp = isK ? std::abs(kPQK[p][q]) - 1 : kPQN[p][q];
          const int ncomp = isK ? 3 : 6;
          if (isK) s *= sgn(kPQK[p][q]);
          if (c.basis == VIE_BASIS_PWL) {
            const int i = std::abs(kLLP[l][lp]) - 1;
            s *= sgn(kLLP[l][lp]);
            if (adj) s *= kSLP[l][lp];
            g = i * ncomp + comp;
          } else {
            g = comp;
          }
          // K^H multiplies by conj(-K) (mvp_K_herm_adj_pwc_tucker.m:26)
          if (isK && adj) s = -s;
          tab[(op * c.ql + u) * kMaxTerms + k] = pack_term(v, g, s < 0 ? 1 : 0);
          ++k;
        }
      }
      nt = k;
    }
    c.nterm[op] = nt;
  }
  c.terms.ensure(tab.size() * sizeof(int));
  VIE_HIP_CHECK(hipMemcpy(c.terms.p, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice));
}

static bool smooth7(int P) {
  for (int f : {2, 3, 5, 7})
    while (P % f == 0) P /= f;
  return P == 1;
}

static FftPlan make_plan(int P) {
  FftPlan pl{};
  pl.P = P;
  int rem = P, k = 0;
  for (int r : {8, 4, 2, 3, 5, 7}) {
    while (rem % r == 0 && rem > 1) {
      if (k >= kMaxStages) throw Error{VIE_ERR_ARG, "FFT length needs too many stages"};
      pl.radix[k++] = r;
      rem /= r;
    }
  }
  if (rem != 1) throw Error{VIE_ERR_ARG, "FFT length is not 2,3,5,7-smooth"};
  pl.nst = k;
  return pl;
}

static void make_twiddles(Ctx& c, int d) {
  const int P = c.P[d];
  std::vector<double> w64(2 * P);
  std::vector<float> w32(2 * P);
  for (int k = 0; k < P; ++k) {
    // exp(-2 pi i k / P), angle reduced exactly before the FP64 sin/cos
    const long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)P;
    const double cr = (double)std::cos(ang), si = (double)-std::sin(ang);
    w64[2 * k] = cr;
    w64[2 * k + 1] = si;
    w32[2 * k] = (float)cr;
    w32[2 * k + 1] = (float)si;
  }
  c.tw64[d].ensure(P * sizeof(double2));
  c.tw32[d].ensure(P * sizeof(float2));
  VIE_HIP_CHECK(hipMemcpy(c.tw64[d].p, w64.data(), P * sizeof(double2), hipMemcpyHostToDevice));
  VIE_HIP_CHECK(hipMemcpy(c.tw32[d].p, w32.data(), P * sizeof(float2), hipMemcpyHostToDevice));
}

static void alloc_spectrum(Ctx& c, int opk) {
  Spectrum& s = c.spec[opk];
  if (s.d64.p) return;
  const int ncomp = opk == 0 ? 6 : 3;
  s.nG = ncomp * c.npair;
  s.have.assign(s.nG, 0);
  const int64_t n = (int64_t)s.nG * c.O[0] * c.O[1] * c.O[2];
  s.d64.ensure(n * sizeof(double2));
  std::vector<int> bits(s.nG);
  for (int g = 0; g < s.nG; ++g) {
    int b = 0;
    for (int d = 0; d < 3; ++d)
      if (parity_of(c, opk, g, d) < 0) b |= (1 << d);
    bits[g] = b;
  }
  s.gpar.ensure(s.nG * sizeof(int));
  VIE_HIP_CHECK(hipMemcpy(s.gpar.p, bits.data(), s.nG * sizeof(int), hipMemcpyHostToDevice));
}

void ensure_ready(Ctx& c, int opk, int prec) {
  Spectrum& s = c.spec[opk];
  if (!s.d64.p) throw Error{VIE_ERR_STATE, opk == 0 ? "N operator spectrum not set" : "K operator spectrum not set"};
  if (!s.ready) {
    for (int g = 0; g < s.nG; ++g)
      if (!s.have[g]) throw Error{VIE_ERR_STATE, "operator spectrum term " + std::to_string(g) + " not set"};
    s.ready = true;
  }
  if (prec == VIE_PREC_FP32 && !s.have32) convert_spectrum_fp32(c, opk);
}

}  // namespace vie

using namespace vie;

static thread_local std::string g_create_err;

template <typename F>
static int guarded(vie_ctx* h, F&& f) {
  try {
    if (h) h->c.err.clear();
    f();
    return VIE_OK;
  } catch (const Error& e) {
    if (h) h->c.err = e.msg; else g_create_err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    if (h) h->c.err = "host allocation failed"; else g_create_err = "host allocation failed";
    return VIE_ERR_OOM;
  } catch (...) {
    if (h) h->c.err = "unknown error"; else g_create_err = "unknown error";
    return VIE_ERR_HIP;
  }
}

// invG per channel for the voxel edge c.res (mvp_invG_pwc.m:3, mvp_invG_pwl.m:3-7)
static void set_gram(Ctx& c) {
  std::vector<double> ig(c.ql);
  const double r3 = c.res * c.res * c.res;
  for (int u = 0; u < c.ql; ++u) {
    if (c.basis == VIE_BASIS_PWC) ig[u] = 1.0 / r3;
    else ig[u] = (u % 4 == 0) ? 1.0 / r3 : 1.0 / (r3 / 12.0);
  }
  c.invg.ensure(c.ql * sizeof(double));
  VIE_HIP_CHECK(hipMemcpy(c.invg.p, ig.data(), c.ql * sizeof(double), hipMemcpyHostToDevice));
  c.invg_h = ig;
}

// Sets up a context for the grid of d (FFT lengths 2 n_d rounded up to
// 2,3,5,7-smooth, or Pforce when given: a support-box child).
static void init_ctx(Ctx& c, const vie_desc& d, const int* Pforce) {
  c.n[0] = d.n1; c.n[1] = d.n2; c.n[2] = d.n3;
  c.basis = d.basis;
  c.ql = d.basis == VIE_BASIS_PWL ? 12 : 3;
  c.npair = d.basis == VIE_BASIS_PWL ? 10 : 1;
  c.res = d.res;
  c.device = d.device;
  for (int k = 0; k < 3; ++k) {
    int P = 2 * c.n[k];
    while (!smooth7(P)) P += 2;
    if (Pforce) P = Pforce[k];
    c.P[k] = P;
    c.O[k] = P / 2 + 1;
    c.plan[k] = make_plan(P);
  }
  if (c.P[0] > 512) throw Error{VIE_ERR_ARG, "n1 too large for the LDS-resident x pass (P1 <= 512)"};
  // the fused x pass keeps 2 * ql * P1 complex values in LDS
  if (2ull * c.ql * c.P[0] * sizeof(double2) > 160 * 1024)
    throw Error{VIE_ERR_ARG, "n1 too large for the fused x pass at this basis (LDS)"};
  // probe allocation first so that the fault-injection hook fires before any HIP call
  dfree(dmalloc(16));
  VIE_HIP_CHECK(hipSetDevice(c.device));
  scratch_init(c.device);
  VIE_HIP_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  c.own_stream = true;
  for (int k = 0; k < 3; ++k) make_twiddles(c, k);
  build_terms(c);
  const char* gen = std::getenv("VIE_GENERIC_KERNELS");
  c.use_fast = !(gen && gen[0] == '1');
  // axis-pass tile widths: VIE_AXIS_TX=w fixes the width (lowered to a
  // compiled one), VIE_AUTOTUNE=0 keeps the default; else they are tuned on
  // the first application of each precision and batch shape
  const char* tx = std::getenv("VIE_AXIS_TX");
  if (tx && *tx) c.force_tile = std::max(4, std::atoi(tx));
  const char* at = std::getenv("VIE_AUTOTUNE");
  c.autotune = !(at && at[0] == '0');
  // VIE_AXIS_FFT=ct (Stockham stages, largest radix first), rf (two-step
  // register split) or cto (Stockham stages, odd radices first) fixes the FFT
  // schedule of every compile-time-length axis pass
  const char* fs = std::getenv("VIE_AXIS_FFT");
  if (fs && std::strcmp(fs, "ct") == 0) c.force_fam = kSchedCT;
  if (fs && std::strcmp(fs, "rf") == 0) c.force_fam = kSchedRF;
  if (fs && std::strcmp(fs, "cto") == 0) c.force_fam = kSchedCTO;
  const char* gr = std::getenv("VIE_GRAPHS");
  c.use_graphs = !(gr && gr[0] == '0');
  const char* pr = std::getenv("VIE_PRUNE");
  c.prune = !(pr && pr[0] == '0');
  const char* xs = std::getenv("VIE_XSTAGE");
  c.xstage = !(xs && xs[0] == '0');
  const char* bx = std::getenv("VIE_BOX");
  c.use_box = !(bx && bx[0] == '0');
  // fused-x-pass column order: the mirror columns (Y, P2-Y) x (Z, P3-Z) of
  // each octant column (yo, zo) are consecutive, the groups of four first,
  // then those of two, then the single ones, so that every aligned pair of
  // columns belongs to one octant column (the x pass stages that spectrum
  // column once per block)
  {
    const int P2 = c.P[1], P3 = c.P[2];
    std::vector<int> cm;
    cm.reserve((size_t)P2 * P3);
    for (int size : {4, 2, 1})
      for (int zo = 0; zo <= P3 / 2; ++zo)
        for (int yo = 0; yo <= P2 / 2; ++yo) {
          const int ny = (yo == 0 || 2 * yo == P2) ? 1 : 2;
          const int nz = (zo == 0 || 2 * zo == P3) ? 1 : 2;
          if (ny * nz != size) continue;
          for (int iz = 0; iz < nz; ++iz)
            for (int iy = 0; iy < ny; ++iy) {
              const int Y = iy ? P2 - yo : yo, Z = iz ? P3 - zo : zo;
              cm.push_back(Y | (Z << 16));
            }
        }
    if (cm.size() != (size_t)P2 * P3) throw Error{VIE_ERR_HIP, "internal: column map size"};
    c.xcolmap.ensure(cm.size() * sizeof(int));
    VIE_HIP_CHECK(hipMemcpy(c.xcolmap.p, cm.data(), cm.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  set_gram(c);
}

// FFT length of the support box along axis d for extent w: 2 w rounded up to
// 2,3,5,7-smooth, or the smallest length >= 2 w with compile-time kernels (the
// fused x pass for d = 0, the axis passes otherwise) when that is at most 1.6
// times longer (those kernels are the faster per point)
static int box_length(int w, int d) {
  int P = 2 * w;
  while (!smooth7(P)) P += 2;
  for (int L = 2 * w; L <= 1024 && 5 * L <= 8 * P; L += 2)
    if (d == 0 ? fast_x_length(L) : fast_axis_length(L)) return L;
  return P;
}

static void box_check(vie_ctx* b, int rc) {
  if (rc != VIE_OK) throw Error{rc, "support box: " + b->c.err};
}

static void drop_box(Ctx& c) {
  if (c.box) {
    vie_destroy(c.box);
    c.box = nullptr;
  }
  for (int d = 0; d < 3; ++d) c.box_w[d] = 0;
}

// (Re)builds the support-box child after the material or a dense kernel
// changed: kept only when its circulant is at most 0.9 of the grid's.
static void update_box(Ctx& c) {
  if (!c.use_box || !c.prune || !c.have_material || c.nS == 0 || !c.green_src_ok) { drop_box(c); return; }
  int w[3], P[3];
  double full = 1.0, part = 1.0;
  for (int d = 0; d < 3; ++d) {
    w[d] = c.box_ext[d];
    P[d] = box_length(w[d], d);
    full *= c.P[d];
    part *= P[d];
  }
  const vie_desc bd{w[0], w[1], w[2], c.basis, c.res, c.device};
  if (part > 0.9 * full || 2ull * c.ql * P[0] * sizeof(double2) > 160 * 1024) { drop_box(c); return; }
  bool same = c.box != nullptr;
  for (int d = 0; d < 3 && same; ++d) same = c.box_w[d] == w[d] && c.box->c.P[d] == P[d];
  if (!same) {
    drop_box(c);
    vie_ctx* b = new vie_ctx();
    try {
      init_ctx(b->c, bd, P);
      b->c.use_box = false;
      box_check(b, vie_set_stream(b, c.stream));
    } catch (...) {
      vie_destroy(b);
      throw;
    }
    c.box = b;
    for (int d = 0; d < 3; ++d) c.box_w[d] = w[d];
    c.box_green_dirty = true;
  }
  if (c.box_green_dirty) {   // the packed operators use N only
    const int nk = c.npair * 6;
    DBuf tmp;
    tmp.ensure((size_t)w[0] * w[1] * w[2] * nk * sizeof(double2));
    crop_offsets(c, c.green_src.as<double2>(), tmp.as<double2>(), c.n, w, nk);
    box_check(c.box, vie_set_green_dense(c.box, VIE_OP_N, static_cast<const double*>(tmp.p), VIE_MEM_DEVICE));
    VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
    c.box_green_dirty = false;
  }
  box_check(c.box, vie_set_material(c.box, c.box_Mr.data(), c.box_Mc.data(), c.box_idx.data(), c.nS));
}

// The factors U_d of a T-struct term are the length-2n_d DFTs of the embedded
// real-space factors e = [u; 0; s u(n:-1:2)] (assembly_fft_circ_tucker_*.m:22-27).
// E (r x 2n, column-major, interleaved) receives e for every rank component:
// the inverse DFT of the rows of U.
static void tucker_factor_real(const double* U, int r, int n, std::vector<double>& E) {
  const int P = 2 * n;
  std::vector<double> cs(P), sn(P);
  for (int k = 0; k < P; ++k) {
    const long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)P;
    cs[k] = (double)std::cos(ang);
    sn[k] = (double)std::sin(ang);
  }
  E.assign(2 * (size_t)r * P, 0.0);
  for (int m = 0; m < P; ++m)
    for (int a = 0; a < r; ++a) {
      double re = 0.0, im = 0.0;
      for (int f = 0; f < P; ++f) {   // U(a, f) e^{+2 pi i f m / P}
        const size_t e = 2 * ((size_t)a + (size_t)r * f);
        const int k = (int)(((long long)f * m) % P);
        re += U[e] * cs[k] - U[e + 1] * sn[k];
        im += U[e] * sn[k] + U[e + 1] * cs[k];
      }
      E[2 * ((size_t)a + (size_t)r * m)] = re / P;
      E[2 * ((size_t)a + (size_t)r * m) + 1] = im / P;
    }
}

// The same kernel factor embedded in a circulant of length P >= 2n - 1: offsets
// 0..n-1 at the front, -1..-(n-1) at the back, zeros between; F (r x O,
// O = P/2 + 1) receives the first O columns of its length-P DFT.  The circulant
// of any such length applies the same linear convolution to an n-point input.
static void tucker_factor_reembed(const std::vector<double>& E, int r, int n, int P, std::vector<double>& F) {
  const int O = P / 2 + 1;
  std::vector<double> cs(P), sn(P);
  for (int k = 0; k < P; ++k) {
    const long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)P;
    cs[k] = (double)std::cos(ang);
    sn[k] = (double)std::sin(ang);
  }
  F.assign(2 * (size_t)r * O, 0.0);
  for (int f = 0; f < O; ++f)
    for (int a = 0; a < r; ++a) {
      double re = 0.0, im = 0.0;
      for (int m = 0; m < n; ++m) {   // e(m) w^{m f} + e(2n - m) w^{-m f}, w = e^{-2 pi i / P}
        const int k = (int)(((long long)f * m) % P);
        const size_t ep = 2 * ((size_t)a + (size_t)r * m);
        re += E[ep] * cs[k] + E[ep + 1] * sn[k];
        im += E[ep + 1] * cs[k] - E[ep] * sn[k];
        if (m > 0) {
          const size_t en = 2 * ((size_t)a + (size_t)r * (2 * n - m));
          re += E[en] * cs[k] - E[en + 1] * sn[k];
          im += E[en + 1] * cs[k] + E[en] * sn[k];
        }
      }
      F[2 * ((size_t)a + (size_t)r * f)] = re;
      F[2 * ((size_t)a + (size_t)r * f) + 1] = im;
    }
}

extern "C" {

int vie_create(const vie_desc* d, vie_ctx** out) {
  if (!out) { g_create_err = "out is NULL"; return VIE_ERR_ARG; }
  *out = nullptr;
  vie_ctx* h = nullptr;
  int rc = guarded(nullptr, [&] {
    if (!d) throw Error{VIE_ERR_ARG, "desc is NULL"};
    if (d->n1 < 1 || d->n2 < 1 || d->n3 < 1) throw Error{VIE_ERR_ARG, "grid dimensions must be >= 1"};
    if (d->basis != VIE_BASIS_PWC && d->basis != VIE_BASIS_PWL) throw Error{VIE_ERR_ARG, "basis must be PWC or PWL"};
    if (!(d->res > 0)) throw Error{VIE_ERR_ARG, "res must be > 0"};
    if ((int64_t)d->n1 * d->n2 * d->n3 > (int64_t)1 << 30) throw Error{VIE_ERR_ARG, "grid too large"};
    h = new vie_ctx();
    init_ctx(h->c, *d, nullptr);
  });
  if (rc != VIE_OK) {
    if (h) vie_destroy(h);
    return rc;
  }
  *out = h;
  return VIE_OK;
}

int vie_destroy(vie_ctx* h) {
  if (!h) return VIE_OK;
  Ctx& c = h->c;
  if (c.box) {
    vie_destroy(c.box);
    c.box = nullptr;
  }
  (void)hipSetDevice(c.device);
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  for (auto* b : {&c.green_src, &c.terms, &c.xcolmap, &c.vox2pk, &c.mcr, &c.invg, &c.zline, &c.wsA1, &c.wsA2, &c.stage_in, &c.stage_out, &c.gm_V,
                  &c.gm_w, &c.gm_x, &c.gm_r, &c.gm_b, &c.gm_red, &c.gm_small, &c.gm_t, &c.bk_v, &c.bk_part, &c.bk_w})
    b->release();
  for (int k = 0; k < 3; ++k) { c.tw64[k].release(); c.tw32[k].release(); }
  for (int k = 0; k < 2; ++k) { c.spec[k].d64.release(); c.spec[k].d32.release(); c.spec[k].gpar.release(); }
  drop_graphs(c);
  if (c.gm_stream) {
    (void)hipStreamSynchronize(c.gm_stream);
    (void)hipStreamDestroy(c.gm_stream);
  }
  if (c.gm_order) (void)hipEventDestroy(c.gm_order);
  if (c.gm_cycle_ev) (void)hipEventDestroy(c.gm_cycle_ev);
  if (c.gm_pin) (void)hipHostFree(c.gm_pin);
  if (c.bk_pin) (void)hipHostFree(c.bk_pin);
  if (c.own_stream && c.stream) (void)hipStreamDestroy(c.stream);
  delete h;
  return VIE_OK;
}

const char* vie_last_error(const vie_ctx* h) { return h ? h->c.err.c_str() : g_create_err.c_str(); }

int vie_set_stream(vie_ctx* h, void* s) {
  if (!h) return VIE_ERR_ARG;
  return guarded(h, [&] {
    Ctx& c = h->c;
    if (c.own_stream && c.stream) {
      VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
      VIE_HIP_CHECK(hipStreamDestroy(c.stream));
    }
    c.stream = static_cast<hipStream_t>(s);
    c.own_stream = false;
    if (c.box) box_check(c.box, vie_set_stream(c.box, s));
  });
}

int vie_set_gram(vie_ctx* h, double res) {
  if (!h) return VIE_ERR_ARG;
  return guarded(h, [&] {
    Ctx& c = h->c;
    if (!(res > 0) || !std::isfinite(res)) throw Error{VIE_ERR_ARG, "res must be > 0"};
    VIE_HIP_CHECK(hipSetDevice(c.device));
    // the operators read invG on c.stream: nothing queued may still use the old table
    VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
    c.res = res;
    set_gram(c);
    if (c.box) box_check(c.box, vie_set_gram(c.box, res));
  });
}

int vie_fft_dims(const vie_ctx* h, int32_t* P1, int32_t* P2, int32_t* P3) {
  if (!h) return VIE_ERR_ARG;
  if (P1) *P1 = h->c.P[0];
  if (P2) *P2 = h->c.P[1];
  if (P3) *P3 = h->c.P[2];
  return VIE_OK;
}

int vie_set_green_tucker(vie_ctx* h, int32_t op, int32_t term, const double* core, int32_t r1, int32_t r2,
                         int32_t r3, const double* U1, const double* U2, const double* U3) {
  if (!h) return VIE_ERR_ARG;
  return guarded(h, [&] {
    Ctx& c = h->c;
    if (op != VIE_OP_N && op != VIE_OP_K) throw Error{VIE_ERR_ARG, "op must be VIE_OP_N or VIE_OP_K"};
    const int opk = op == VIE_OP_N ? 0 : 1;
    if (r1 < 0 || r2 < 0 || r3 < 0 || r1 > c.n[0] || r2 > c.n[1] || r3 > c.n[2])
      throw Error{VIE_ERR_ARG, "Tucker ranks out of range"};
    // a term of rank 0 along some mode is the zero tensor (a HOSVD without a
    // tolerance gives it for a component that parity forces to vanish, e.g.
    // one odd along an axis of a single voxel): held as a rank-1 term with a
    // zero core and zero factors
    std::vector<double> zcore, zU[3];
    if (r1 == 0 || r2 == 0 || r3 == 0) {
      r1 = r2 = r3 = 1;
      zcore.assign(2, 0.0);
      for (int d = 0; d < 3; ++d) zU[d].assign(4 * (size_t)c.n[d], 0.0);
      core = zcore.data();
      U1 = zU[0].data();
      U2 = zU[1].data();
      U3 = zU[2].data();
    }
    if (!core || !U1 || !U2 || !U3) throw Error{VIE_ERR_ARG, "NULL Tucker factor"};
    VIE_HIP_CHECK(hipSetDevice(c.device));
    alloc_spectrum(c, opk);
    Spectrum& s = c.spec[opk];
    if (term < 0 || term >= s.nG) throw Error{VIE_ERR_ARG, "term index out of range"};
    const double* U[3] = {U1, U2, U3};
    const int r[3] = {r1, r2, r3};
    // real-space factors: for the support box (N only) and for every axis whose
    // FFT length is not 2 n_d (2 n_d not 2,3,5,7-smooth, e.g. the pFFT grids of
    // pfft_extend_vie_domain.m), where the factor is re-embedded at P_d
    const bool want_box = opk == 0 && c.use_box && c.prune;
    std::vector<double> E[3], F[3];
    const double* Fd[3];
    for (int d = 0; d < 3; ++d) {
      const bool reembed = c.P[d] != 2 * c.n[d];
      if (reembed || want_box) tucker_factor_real(U[d], r[d], c.n[d], E[d]);
      if (reembed) {
        tucker_factor_reembed(E[d], r[d], c.n[d], c.P[d], F[d]);
        Fd[d] = F[d].data();
      } else {
        Fd[d] = U[d];   // the first O_d columns of the r_d x 2n_d column-major factor are contiguous
      }
    }
    const int O1 = c.O[0], O2 = c.O[1], O3 = c.O[2];
    const size_t ncore = (size_t)r1 * r2 * r3;
    const size_t nf1 = (size_t)r1 * O1, nf2 = (size_t)r2 * O2, nf3 = (size_t)r3 * O3;
    const size_t nt1 = (size_t)O1 * r2 * r3, nt2 = (size_t)O1 * O2 * r3;
    DBuf tmp;
    const size_t tot = ncore + nf1 + nf2 + nf3 + nt1 + nt2;
    tmp.ensure(tot * sizeof(double2));
    double2* dcore = tmp.as<double2>();
    double2* dF1 = dcore + ncore;
    double2* dF2 = dF1 + nf1;
    double2* dF3 = dF2 + nf2;
    double2* dt1 = dF3 + nf3;
    double2* dt2 = dt1 + nt1;
    VIE_HIP_CHECK(hipMemcpyAsync(dcore, core, ncore * sizeof(double2), hipMemcpyHostToDevice, c.stream));
    VIE_HIP_CHECK(hipMemcpyAsync(dF1, Fd[0], nf1 * sizeof(double2), hipMemcpyHostToDevice, c.stream));
    VIE_HIP_CHECK(hipMemcpyAsync(dF2, Fd[1], nf2 * sizeof(double2), hipMemcpyHostToDevice, c.stream));
    VIE_HIP_CHECK(hipMemcpyAsync(dF3, Fd[2], nf3 * sizeof(double2), hipMemcpyHostToDevice, c.stream));
    build_spectrum_term(c, opk, term, dcore, r1, r2, r3, dF1, dF2, dF3, dt1, dt2);
    VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
    tmp.release();
    s.have[term] = 1;
    if (want_box) {
      // the term in real space for the support box: the factor at the offsets m < n_d
      // (the first n_d columns of E)
      const int64_t N = c.nvox();
      if (c.green_src_have.size() != (size_t)s.nG) {
        c.green_src_have.assign(s.nG, 0);
        c.green_src_ok = false;
      }
      c.green_src.ensure((size_t)N * s.nG * sizeof(double2));
      const size_t nv = (size_t)r1 * c.n[0] + (size_t)r2 * c.n[1] + (size_t)r3 * c.n[2];
      DBuf rt;
      rt.ensure((ncore + nv + (size_t)c.n[0] * r2 * r3 + (size_t)c.n[0] * c.n[1] * r3) * sizeof(double2));
      double2* rcore = rt.as<double2>();
      double2* dV1 = rcore + ncore;
      double2* dV2 = dV1 + (size_t)r1 * c.n[0];
      double2* dV3 = dV2 + (size_t)r2 * c.n[1];
      double2* rt1 = dV3 + (size_t)r3 * c.n[2];
      double2* rt2 = rt1 + (size_t)c.n[0] * r2 * r3;
      VIE_HIP_CHECK(hipMemcpyAsync(rcore, core, ncore * sizeof(double2), hipMemcpyHostToDevice, c.stream));
      VIE_HIP_CHECK(hipMemcpyAsync(dV1, E[0].data(), (size_t)r1 * c.n[0] * sizeof(double2), hipMemcpyHostToDevice, c.stream));
      VIE_HIP_CHECK(hipMemcpyAsync(dV2, E[1].data(), (size_t)r2 * c.n[1] * sizeof(double2), hipMemcpyHostToDevice, c.stream));
      VIE_HIP_CHECK(hipMemcpyAsync(dV3, E[2].data(), (size_t)r3 * c.n[2] * sizeof(double2), hipMemcpyHostToDevice, c.stream));
      // kernel layout [n1 n2 n3][pair][component], term = pair * 6 + component
      const int pair = term / 6, comp = term % 6;
      tucker_real_space(c, rcore, r1, r2, r3, dV1, dV2, dV3, rt1, rt2,
                        c.green_src.as<double2>() + (int64_t)(pair + c.npair * comp) * N);
      VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
      c.green_src_have[term] = 1;
      bool all = true;
      for (char hv : c.green_src_have) all = all && hv;
      c.green_src_ok = all;
      c.box_green_dirty = true;
      if (all) update_box(c);
    }
    s.ready = false;
    s.have32 = false;
  });
}

int vie_set_green_dense(vie_ctx* h, int32_t op, const double* A, int32_t mem) {
  if (!h) return VIE_ERR_ARG;
  return guarded(h, [&] {
    Ctx& c = h->c;
    if (op != VIE_OP_N && op != VIE_OP_K) throw Error{VIE_ERR_ARG, "op must be VIE_OP_N or VIE_OP_K"};
    if (!A) throw Error{VIE_ERR_ARG, "A is NULL"};
    const int opk = op == VIE_OP_N ? 0 : 1;
    VIE_HIP_CHECK(hipSetDevice(c.device));
    alloc_spectrum(c, opk);
    Spectrum& s = c.spec[opk];
    const int n1 = c.n[0], n2 = c.n[1], n3 = c.n[2];
    const int O1 = c.O[0], O2 = c.O[1], O3 = c.O[2];
    const int64_t N = c.nvox();
    // circulant factors F_d(m, f) = w^{m f} + [m >= 1] s w^{-m f}, w = exp(-2 pi i / P_d),
    // for both parities s = +1 / -1 of every axis
    std::vector<double> F[3][2];
    for (int d = 0; d < 3; ++d) {
      const int n = c.n[d], O = c.O[d], P = c.P[d];
      for (int sp = 0; sp < 2; ++sp) {
        const double sg = sp == 0 ? 1.0 : -1.0;
        F[d][sp].assign(2 * (size_t)n * O, 0.0);
        for (int f = 0; f < O; ++f)
          for (int m = 0; m < n; ++m) {
            const long long mf = ((long long)m * f) % P;
            const long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)mf / P;
            const double cr = (double)std::cos(ang), si = (double)std::sin(ang);
            double re, im;
            if (m == 0) { re = 1.0; im = 0.0; }
            else { re = cr + sg * cr; im = -si + sg * si; }   // w^{mf} = cos - i sin, w^{-mf} = cos + i sin
            F[d][sp][2 * ((size_t)m + (size_t)n * f)] = re;
            F[d][sp][2 * ((size_t)m + (size_t)n * f) + 1] = im;
          }
      }
    }
    const size_t nf[3] = {(size_t)n1 * O1, (size_t)n2 * O2, (size_t)n3 * O3};
    const size_t nt1 = (size_t)O1 * n2 * n3, nt2 = (size_t)O1 * O2 * n3;
    const int64_t ntotA = N * c.npair * (opk == 0 ? 6 : 3);
    DBuf tmp;
    const size_t tot = 2 * (nf[0] + nf[1] + nf[2]) + nt1 + nt2 + (mem == VIE_MEM_DEVICE ? 0 : (size_t)ntotA);
    tmp.ensure(tot * sizeof(double2));
    double2* dF[3][2];
    double2* p = tmp.as<double2>();
    for (int d = 0; d < 3; ++d)
      for (int sp = 0; sp < 2; ++sp) {
        dF[d][sp] = p;
        VIE_HIP_CHECK(hipMemcpyAsync(p, F[d][sp].data(), nf[d] * sizeof(double2), hipMemcpyHostToDevice, c.stream));
        p += nf[d];
      }
    double2* dt1 = p; p += nt1;
    double2* dt2 = p; p += nt2;
    const double2* dA;
    if (mem == VIE_MEM_DEVICE) {
      dA = reinterpret_cast<const double2*>(A);
    } else {
      VIE_HIP_CHECK(hipMemcpyAsync(p, A, ntotA * sizeof(double2), hipMemcpyHostToDevice, c.stream));
      dA = p;
    }
    for (int g = 0; g < s.nG; ++g) {
      const int ncomp = opk == 0 ? 6 : 3;
      const int j = g % ncomp, i = g / ncomp;
      const double2* core = dA + (int64_t)(i + c.npair * j) * N;
      const int s1 = parity_of(c, opk, g, 0) < 0, s2 = parity_of(c, opk, g, 1) < 0, s3 = parity_of(c, opk, g, 2) < 0;
      build_spectrum_term(c, opk, g, core, n1, n2, n3, dF[0][s1], dF[1][s2], dF[2][s3], dt1, dt2);
      s.have[g] = 1;
    }
    // the N kernel is kept for the support-box spectrum (cropped to the box offsets)
    const bool keep = opk == 0 && c.use_box && c.prune;
    if (keep) {
      c.green_src.ensure((size_t)ntotA * sizeof(double2));
      VIE_HIP_CHECK(hipMemcpyAsync(c.green_src.p, dA, (size_t)ntotA * sizeof(double2), hipMemcpyDeviceToDevice,
                                   c.stream));
    }
    VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
    tmp.release();
    s.ready = false;
    s.have32 = false;
    if (opk == 0) {
      c.green_src_have.assign(s.nG, keep ? 1 : 0);
      c.green_src_ok = keep;
      c.box_green_dirty = true;
      update_box(c);
    }
  });
}

int vie_set_material(vie_ctx* h, const double* Mr, const double* Mc, const int64_t* idxS, int64_t nS) {
  if (!h) return VIE_ERR_ARG;
  return guarded(h, [&] {
    Ctx& c = h->c;
    const int64_t N = c.nvox();
    if (!Mr || !Mc || (!idxS && nS > 0)) throw Error{VIE_ERR_ARG, "NULL material argument"};
    if (nS < 0 || nS > N) throw Error{VIE_ERR_ARG, "nS out of range"};
    std::vector<int> v2p(N, -1);
    std::vector<double> mcr(2 * (size_t)(nS > 0 ? nS : 1));
    for (int64_t s = 0; s < nS; ++s) {
      const int64_t v = idxS[s];
      if (v < 0 || v >= N) throw Error{VIE_ERR_ARG, "idxS entry out of range"};
      if (s > 0 && idxS[s - 1] >= v) throw Error{VIE_ERR_ARG, "idxS must be strictly increasing"};
      v2p[v] = (int)s;
      // (1./Mr).*Mc (mvp_vie.m:17)
      const double ar = Mr[2 * v], ai = Mr[2 * v + 1];
      const double den = ar * ar + ai * ai;
      if (den == 0) throw Error{VIE_ERR_ARG, "Mr is zero at a scatterer voxel"};
      const double ir = ar / den, ii = -ai / den;
      const double br = Mc[2 * v], bi = Mc[2 * v + 1];
      mcr[2 * s] = ir * br - ii * bi;
      mcr[2 * s + 1] = ir * bi + ii * br;
    }
    // z lines (x, y) that hold a scatterer voxel
    const int64_t nxy = (int64_t)c.n[0] * c.n[1];
    std::vector<unsigned char> zl(nxy, 0);
    for (int64_t s = 0; s < nS; ++s) zl[idxS[s] % nxy] = 1;
    int xlo = c.n[0], xhi = -1;
    for (int64_t s = 0; s < nS; ++s) {
      const int x = (int)(idxS[s] % c.n[0]);
      xlo = std::min(xlo, x);
      xhi = std::max(xhi, x);
    }
    VIE_HIP_CHECK(hipSetDevice(c.device));
    VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
    c.vox2pk.ensure(N * sizeof(int));
    c.mcr.ensure((nS > 0 ? nS : 1) * sizeof(double2));
    c.zline.ensure(nxy);
    VIE_HIP_CHECK(hipMemcpy(c.zline.p, zl.data(), nxy, hipMemcpyHostToDevice));
    VIE_HIP_CHECK(hipMemcpy(c.vox2pk.p, v2p.data(), N * sizeof(int), hipMemcpyHostToDevice));
    if (nS > 0) VIE_HIP_CHECK(hipMemcpy(c.mcr.p, mcr.data(), nS * sizeof(double2), hipMemcpyHostToDevice));
    c.nS = nS;
    c.xlo = nS > 0 ? xlo : 0;
    c.xhi = xhi;
    c.have_material = true;
    // support box: bounding box of the scatterer, material and idxS in its coordinates
    c.box_Mr.clear();
    c.box_Mc.clear();
    c.box_idx.clear();
    if (c.use_box && c.prune && nS > 0) {
      int lo[3] = {c.n[0], c.n[1], c.n[2]}, hi[3] = {-1, -1, -1};
      for (int64_t s = 0; s < nS; ++s) {
        const int64_t v = idxS[s];
        const int m[3] = {(int)(v % c.n[0]), (int)((v / c.n[0]) % c.n[1]), (int)(v / ((int64_t)c.n[0] * c.n[1]))};
        for (int d = 0; d < 3; ++d) {
          lo[d] = std::min(lo[d], m[d]);
          hi[d] = std::max(hi[d], m[d]);
        }
      }
      for (int d = 0; d < 3; ++d) {
        c.box_lo[d] = lo[d];
        c.box_ext[d] = hi[d] - lo[d] + 1;
      }
      const int w0 = c.box_ext[0], w1 = c.box_ext[1], w2 = c.box_ext[2];
      const int64_t nb = (int64_t)w0 * w1 * w2;
      c.box_Mr.resize(2 * (size_t)nb);
      c.box_Mc.resize(2 * (size_t)nb);
      for (int64_t b = 0; b < nb; ++b) {
        const int m1 = (int)(b % w0), m2 = (int)((b / w0) % w1), m3 = (int)(b / ((int64_t)w0 * w1));
        const int64_t v = (lo[0] + m1) + (int64_t)c.n[0] * ((lo[1] + m2) + (int64_t)c.n[1] * (lo[2] + m3));
        for (int k = 0; k < 2; ++k) {
          c.box_Mr[2 * b + k] = Mr[2 * v + k];
          c.box_Mc[2 * b + k] = Mc[2 * v + k];
        }
      }
      c.box_idx.resize((size_t)nS);
      for (int64_t s = 0; s < nS; ++s) {
        const int64_t v = idxS[s];
        const int m1 = (int)(v % c.n[0]) - lo[0], m2 = (int)((v / c.n[0]) % c.n[1]) - lo[1];
        const int m3 = (int)(v / ((int64_t)c.n[0] * c.n[1])) - lo[2];
        c.box_idx[s] = m1 + (int64_t)w0 * (m2 + (int64_t)w1 * m3);
      }
    }
    update_box(c);
  });
}

int64_t vie_num_scatterers(const vie_ctx* h) { return h ? h->c.nS : -1; }

int vie_apply(vie_ctx* h, int32_t op, int32_t prec, const double* in, double* out, int32_t nrhs, int32_t mem) {
  if (!h) return VIE_ERR_ARG;
  return guarded(h, [&] {
    Ctx& c = h->c;
    if (op < VIE_OP_N || op > VIE_OP_VIEH) throw Error{VIE_ERR_ARG, "unknown op"};
    if (prec != VIE_PREC_FP64 && prec != VIE_PREC_FP32) throw Error{VIE_ERR_ARG, "unknown precision"};
    if (nrhs < 1 || nrhs > 65535 / c.ql) throw Error{VIE_ERR_ARG, "nrhs out of range"};
    if (!in || !out) throw Error{VIE_ERR_ARG, "NULL vector"};
    if (in == out) throw Error{VIE_ERR_ARG, "in and out must not alias"};
    if (mem != VIE_MEM_HOST && mem != VIE_MEM_DEVICE) throw Error{VIE_ERR_ARG, "unknown mem kind"};
    const bool packed = (op == VIE_OP_VIE || op == VIE_OP_VIEH);
    const int opk = (op == VIE_OP_K || op == VIE_OP_KH) ? 1 : 0;
    if (packed && !c.have_material) throw Error{VIE_ERR_STATE, "material not set"};
    if (packed && c.box) {
      box_check(c.box, vie_apply(c.box, op, prec, in, out, nrhs, mem));
      return;
    }
    VIE_HIP_CHECK(hipSetDevice(c.device));
    ensure_ready(c, opk, prec);
    const size_t len = (size_t)nrhs * c.ql * (packed ? (size_t)c.nS : (size_t)c.nvox());
    if (packed && c.nS == 0) return;
    ApplyIO io;
    if (mem == VIE_MEM_DEVICE) {
      io.in = reinterpret_cast<const double2*>(in);
      io.out = reinterpret_cast<double2*>(out);
      apply_operator(c, op, prec, io, nrhs);
    } else {
      c.stage_in.ensure(len * sizeof(double2));
      c.stage_out.ensure(len * sizeof(double2));
      VIE_HIP_CHECK(hipMemcpyAsync(c.stage_in.p, in, len * sizeof(double2), hipMemcpyHostToDevice, c.stream));
      io.in = c.stage_in.as<double2>();
      io.out = c.stage_out.as<double2>();
      apply_operator(c, op, prec, io, nrhs);
      VIE_HIP_CHECK(hipMemcpyAsync(out, c.stage_out.p, len * sizeof(double2), hipMemcpyDeviceToHost, c.stream));
      VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
    }
  });
}

int vie_apply_noalloc(vie_ctx* h, int32_t op, int32_t prec, const double* in, double* out, int32_t nrhs) {
  if (!h) return VIE_ERR_ARG;
  return guarded(h, [&] {
    Ctx& c = h->c;
    if (op < VIE_OP_N || op > VIE_OP_VIEH) throw Error{VIE_ERR_ARG, "unknown op"};
    if (prec != VIE_PREC_FP64 && prec != VIE_PREC_FP32) throw Error{VIE_ERR_ARG, "unknown precision"};
    if (nrhs < 1 || nrhs > 65535 / c.ql) throw Error{VIE_ERR_ARG, "nrhs out of range"};
    if (!in || !out) throw Error{VIE_ERR_ARG, "NULL vector"};
    if (in == out) throw Error{VIE_ERR_ARG, "in and out must not alias"};
    const bool packed = (op == VIE_OP_VIE || op == VIE_OP_VIEH);
    const int opk = (op == VIE_OP_K || op == VIE_OP_KH) ? 1 : 0;
    if (packed && !c.have_material) throw Error{VIE_ERR_STATE, "material not set"};
    if (packed && c.box) {
      box_check(c.box, vie_apply_noalloc(c.box, op, prec, in, out, nrhs));
      return;
    }
    // the spectrum must be complete, and for FP32 its FP32 copy made at set-up
    // (making it would allocate)
    Spectrum& s = c.spec[opk];
    if (!s.d64.p) throw Error{VIE_ERR_STATE, opk == 0 ? "N operator spectrum not set" : "K operator spectrum not set"};
    for (int g = 0; g < s.nG; ++g)
      if (!s.have[g]) throw Error{VIE_ERR_STATE, "operator spectrum term " + std::to_string(g) + " not set"};
    if (prec == VIE_PREC_FP32 && !s.have32)
      throw Error{VIE_ERR_STATE, "FP32 spectrum not made: call vie_autotune (or vie_apply) with VIE_PREC_FP32 first"};
    if (packed && c.nS == 0) return;
    size_t a1, a2;
    workspace_bytes(c, prec, nrhs, &a1, &a2);
    const size_t a1r = (a1 + 255) & ~(size_t)255;
    if (a1r + a2 > kScratchBytes)
      throw Error{VIE_ERR_ARG, "workspace of " + std::to_string(a1r + a2) + " bytes exceeds the scratch of " +
                                   std::to_string(kScratchBytes) + " bytes; use vie_apply"};
    VIE_HIP_CHECK(hipSetDevice(c.device));
    unsigned char* base = static_cast<unsigned char*>(scratch_of(c.device));
    if (!base) throw Error{VIE_ERR_STATE, "scratch not set up on this device"};
    ApplyIO io;
    io.in = reinterpret_cast<const double2*>(in);
    io.out = reinterpret_cast<double2*>(out);
    io.ws1 = base;
    io.ws2 = base + a1r;
    apply_operator_ws(c, op, prec, io, nrhs);
  });
}

int vie_autotune(vie_ctx* h, int32_t op, int32_t prec, int32_t nrhs, int32_t* tiles) {
  if (!h) return VIE_ERR_ARG;
  return guarded(h, [&] {
    Ctx& c = h->c;
    if (op < VIE_OP_N || op > VIE_OP_VIEH) throw Error{VIE_ERR_ARG, "unknown op"};
    if (prec != VIE_PREC_FP64 && prec != VIE_PREC_FP32) throw Error{VIE_ERR_ARG, "unknown precision"};
    if (nrhs < 1 || nrhs > 65535 / c.ql) throw Error{VIE_ERR_ARG, "nrhs out of range"};
    const bool packed = (op == VIE_OP_VIE || op == VIE_OP_VIEH);
    const int opk = (op == VIE_OP_K || op == VIE_OP_KH) ? 1 : 0;
    if (packed && (!c.have_material || c.nS == 0)) throw Error{VIE_ERR_STATE, "material not set"};
    if (packed && c.box) {
      box_check(c.box, vie_autotune(c.box, op, prec, nrhs, tiles));
      return;
    }
    VIE_HIP_CHECK(hipSetDevice(c.device));
    ensure_ready(c, opk, prec);
    const int pi = prec == VIE_PREC_FP64 ? 0 : 1, b = nrhs > 1 ? 1 : 0;
    prepare_operator(c, op, prec, nrhs);   // tunes unless already fixed for this shape
    VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
    if (tiles)
      for (int k = 0; k < 4; ++k) tiles[k] = c.tile[pi][b][k];
  });
}

int vie_axis_schedule(const vie_ctx* h, int32_t prec, int32_t* sched) {
  if (!h || !sched || (prec != VIE_PREC_FP64 && prec != VIE_PREC_FP32)) return VIE_ERR_ARG;
  // the packed operators and GMRES run on the support box: report its schedule
  const Ctx& c = h->c.box ? h->c.box->c : h->c;
  for (int k = 0; k < 4; ++k) sched[k] = c.force_fam ? c.force_fam : c.fam[prec == VIE_PREC_FP64 ? 0 : 1][k];
  return VIE_OK;
}

int vie_box_dims(const vie_ctx* h, int32_t* lo, int32_t* w, int32_t* P) {
  if (!h) return VIE_ERR_ARG;
  const Ctx& c = h->c;
  for (int d = 0; d < 3; ++d) {
    if (lo) lo[d] = c.box ? c.box_lo[d] : 0;
    if (w) w[d] = c.box ? c.box_w[d] : 0;
    if (P) P[d] = c.box ? c.box->c.P[d] : 0;
  }
  return VIE_OK;
}

}  // extern "C"
