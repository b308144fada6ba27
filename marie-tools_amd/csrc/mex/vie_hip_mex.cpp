// MATLAB gateway of libvie_hip.so (R2018a interleaved-complex MEX API).
//
// It keeps the call signatures MARIE's operator handles already have, so the
// binding in parse_inputs.m:6-40 changes and the call sites do not:
//   Jout = vie_hip_mex('mvp_N', Jin, T)           % mvp_N_{pwc,pwl}_tucker(Jin,T)
//   Jout = vie_hip_mex('mvp_K', Jin, T)           % mvp_K_{pwc,pwl}_tucker(Jin,T)
//   Jout = vie_hip_mex('mvp_herm_adj_N', Jin, T)  % mvp_N_herm_adj_*_tucker(Jin,T)
//   Jout = vie_hip_mex('mvp_herm_adj_K', Jin, T)  % mvp_K_herm_adj_*_tucker(Jin,T)
//   Jout = vie_hip_mex('mvp_vie', Jb, EP, dims, fN)          % mvp_vie.m
//   [x, iter, resvec] = vie_hip_mex('gmres_vie', b, EP, dims, fN, restart, tol, maxit, x0, prec)
//   [X, nmatvec, resvec] = vie_hip_mex('cocr_vie', B, EP, dims, fN, tol, maxit, X0, prec)
//                                                 % B: [ql*nS, nports], every port in one batched
//                                                 % solve (the loop of ie_solver_vie.m:16-21)
//   [X, nmatvec, resvec] = vie_hip_mex('bcocg_vie', B, EP, dims, fN, tol, maxit, X0, prec)
//                                                 % the same with block COCG (vie_bcocg): the
//                                                 % ports share one block Krylov space; prec
//                                                 % defaults to FP64
//   vie_hip_mex('release')                        % drop every cached context
// T is the Tucker T-struct of assembly_fft_circ_tucker_*.m; the gateway
// expands it on the device once.  Contexts are cached by content: the grid,
// the operator, the basis and a 64-bit fingerprint of every term's ranks and
// factor data (vie_mex_args.cpp), so a T-struct rebuilt at an address MATLAB
// reused is never mistaken for the old one, and mvp_N, mvp_vie and gmres_vie
// share the context of one T.  At most four contexts are kept (least recently
// used first out).  The material of a context (EP.Mr, EP.Mc, dims.idxS) is
// compared byte for byte with the last one given before it is reused.
// dims.idxS may be int32 (the shipped body models, geo_body_domain.m:13),
// uint32, int64, uint64, double or single; an unsorted idxS is supported by
// permuting packed vectors on the way in and out.
//
// Library status codes become MATLAB errors with identifiers vie_hip:arg,
// vie_hip:hip, vie_hip:oom, vie_hip:state, so solver_wsvie.m's try/catch
// GPU -> CPU fallback (solver_wsvie.m:10-23) keeps working.
//
// Build (where MATLAB is installed; libvie_hip.so built by the Makefile):
//   mex -R2018a vie_hip_mex.cpp vie_mex_args.cpp -I../../../include -L../../marie_vie -lvie_hip
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "mex.h"
#include "vie_mex_args.hpp"
#include "vie_op.h"

namespace {

using vie_mex::CtxCache;
using vie_mex::NumClass;

void destroy_ctx(void* p) { vie_destroy(static_cast<vie_ctx*>(p)); }

CtxCache g_cache(destroy_ctx, 4);
bool g_exit_hook = false;

void release_all() { g_cache.clear(); }

void fail(int rc, vie_ctx* ctx) {
  const char* id = rc == VIE_ERR_OOM ? "vie_hip:oom" : rc == VIE_ERR_ARG ? "vie_hip:arg"
                   : rc == VIE_ERR_STATE ? "vie_hip:state" : "vie_hip:hip";
  mexErrMsgIdAndTxt(id, "libvie_hip: %s", vie_last_error(ctx));
}

void check(int rc, vie_ctx* ctx) {
  if (rc != VIE_OK) fail(rc, ctx);
}

NumClass class_of(const mxArray* a) {
  switch (mxGetClassID(a)) {
    case mxDOUBLE_CLASS: return NumClass::Double;
    case mxSINGLE_CLASS: return NumClass::Single;
    case mxINT32_CLASS: return NumClass::Int32;
    case mxUINT32_CLASS: return NumClass::UInt32;
    case mxINT64_CLASS: return NumClass::Int64;
    case mxUINT64_CLASS: return NumClass::UInt64;
    default: return NumClass::Other;
  }
}

size_t data_bytes(const mxArray* a) { return mxGetNumberOfElements(a) * mxGetElementSize(a); }

// complex copy of a real or complex double array (interleaved)
std::vector<double> complex_copy(const mxArray* a, const char* what) {
  if (!mxIsDouble(a) || mxIsSparse(a)) mexErrMsgIdAndTxt("vie_hip:arg", "%s must be a full double array", what);
  const size_t n = mxGetNumberOfElements(a);
  std::vector<double> v(2 * n, 0.0);
  if (mxIsComplex(a)) {
    std::memcpy(v.data(), mxGetComplexDoubles(a), 2 * n * sizeof(double));
  } else {
    const double* r = mxGetDoubles(a);
    for (size_t i = 0; i < n; ++i) v[2 * i] = r[i];
  }
  return v;
}

const mxArray* field(const mxArray* s, size_t e, const char* name) {
  const mxArray* f = mxGetField(s, e, name);
  if (!f) mexErrMsgIdAndTxt("vie_hip:arg", "struct field '%s' missing", name);
  return f;
}

int basis_of(const mxArray* T, int op) {
  if (!mxIsStruct(T)) mexErrMsgIdAndTxt("vie_hip:arg", "T must be the Tucker T-struct");
  const size_t nt = mxGetNumberOfElements(T);
  const size_t ncomp = op == VIE_OP_N ? 6 : 3;
  if (nt == ncomp) return VIE_BASIS_PWC;
  if (nt == 10 * ncomp && mxGetM(T) == 10) return VIE_BASIS_PWL;
  mexErrMsgIdAndTxt("vie_hip:arg", "T has %d terms; expected %d (PWC) or 10x%d (PWL)", (int)nt, (int)ncomp,
                    (int)ncomp);
  return -1;
}

// ranks of term e, with its factors checked against the grid: U_d is r_d x 2 n_d
void term_ranks(const mxArray* T, size_t e, const int n[3], int r[3]) {
  const mxArray* G = field(T, e, "G");
  const mwSize* gd = mxGetDimensions(G);
  const int nd = (int)mxGetNumberOfDimensions(G);
  r[0] = (int)gd[0];
  r[1] = nd > 1 ? (int)gd[1] : 1;
  r[2] = nd > 2 ? (int)gd[2] : 1;
  const char* names[3] = {"U1", "U2", "U3"};
  for (int d = 0; d < 3; ++d) {
    const mxArray* U = field(T, e, names[d]);
    if (!mxIsDouble(U) || mxGetM(U) != (size_t)r[d] || mxGetN(U) != 2 * (size_t)n[d])
      mexErrMsgIdAndTxt("vie_hip:arg", "T(%d).%s must be a double %d x %d array (rank x 2 n%d)", (int)e + 1,
                        names[d], r[d], 2 * n[d], d + 1);
  }
  if (!mxIsDouble(G)) mexErrMsgIdAndTxt("vie_hip:arg", "T(%d).G must be double", (int)e + 1);
}

// content key of T for the grid n and operator op
vie_mex::TKey key_of(const mxArray* T, const int n[3], int op, int basis) {
  vie_mex::TKey k;
  k.n1 = n[0]; k.n2 = n[1]; k.n3 = n[2]; k.op = op; k.basis = basis;
  uint64_t h = vie_mex::kFingerprintSeed;
  const size_t nt = mxGetNumberOfElements(T);
  const char* names[4] = {"G", "U1", "U2", "U3"};
  for (size_t e = 0; e < nt; ++e) {
    int r[3];
    term_ranks(T, e, n, r);
    h = vie_mex::fingerprint(r, sizeof(r), h);
    for (const char* nm : names) {
      const mxArray* f = field(T, e, nm);
      const int cplx = mxIsComplex(f) ? 1 : 0;
      h = vie_mex::fingerprint(&cplx, sizeof(cplx), h);
      h = vie_mex::fingerprint(mxGetData(f), data_bytes(f), h);
    }
  }
  k.fp = h;
  return k;
}

// Cached context of T (op 'N' or 'K') on grid n, built on first use.  A new
// context's Gram diagonal is left at res 1 until mvp_vie sets it.
CtxCache::Entry* context_for(const mxArray* T, const int n[3], int op) {
  if (!g_exit_hook) {
    mexAtExit(release_all);
    g_exit_hook = true;
  }
  const int basis = basis_of(T, op);
  const vie_mex::TKey key = key_of(T, n, op, basis);
  if (CtxCache::Entry* hit = g_cache.find(key)) return hit;
  const vie_desc d{n[0], n[1], n[2], basis, 1.0, 0};
  vie_ctx* ctx = nullptr;
  check(vie_create(&d, &ctx), nullptr);
  // T(i,j) is column-major in MATLAB: element (i,j) at i + 10*j; the library
  // numbers PWL terms i*ncomp + j
  const int ncomp = op == VIE_OP_N ? 6 : 3;
  const int npair = basis == VIE_BASIS_PWL ? 10 : 1;
  for (int i = 0; i < npair; ++i)
    for (int j = 0; j < ncomp; ++j) {
      const size_t e = basis == VIE_BASIS_PWL ? (size_t)i + 10 * (size_t)j : (size_t)j;
      int r[3];
      term_ranks(T, e, n, r);
      std::vector<double> cg = complex_copy(field(T, e, "G"), "T.G"), c1 = complex_copy(field(T, e, "U1"), "T.U1"),
                          c2 = complex_copy(field(T, e, "U2"), "T.U2"), c3 = complex_copy(field(T, e, "U3"), "T.U3");
      const int term = basis == VIE_BASIS_PWL ? i * ncomp + j : j;
      const int rc = vie_set_green_tucker(ctx, op, term, cg.data(), r[0], r[1], r[2], c1.data(), c2.data(), c3.data());
      if (rc != VIE_OK) {
        const std::string msg = vie_last_error(ctx);
        vie_destroy(ctx);
        mexErrMsgIdAndTxt(rc == VIE_ERR_OOM ? "vie_hip:oom" : "vie_hip:arg", "libvie_hip: %s", msg.c_str());
      }
    }
  CtxCache::Entry* e = g_cache.insert(key, ctx);
  e->gram_res = 1.0;
  return e;
}

int op_of(const std::string& name) {
  if (name == "mvp_N") return VIE_OP_N;
  if (name == "mvp_K") return VIE_OP_K;
  if (name == "mvp_herm_adj_N") return VIE_OP_NH;
  if (name == "mvp_herm_adj_K") return VIE_OP_KH;
  return -1;
}

// Sets the material of e's context from EP and dims unless it holds exactly
// these arrays already; returns nS.
size_t set_material(CtxCache::Entry* e, const mxArray* EP, const mxArray* dims, int64_t nvox) {
  vie_ctx* ctx = static_cast<vie_ctx*>(e->ctx);
  const mxArray* Mr = field(EP, 0, "Mr");
  const mxArray* Mc = field(EP, 0, "Mc");
  const mxArray* idxS = field(dims, 0, "idxS");
  if ((int64_t)mxGetNumberOfElements(Mr) != nvox || (int64_t)mxGetNumberOfElements(Mc) != nvox)
    mexErrMsgIdAndTxt("vie_hip:arg", "EP.Mr and EP.Mc must have n1*n2*n3 = %lld elements", (long long)nvox);
  if (!mxIsDouble(Mr) || !mxIsDouble(Mc)) mexErrMsgIdAndTxt("vie_hip:arg", "EP.Mr and EP.Mc must be double");
  if (mxIsComplex(idxS)) mexErrMsgIdAndTxt("vie_hip:arg", "dims.idxS must be real");
  const int cr = (int)mxGetClassID(Mr), cc = (int)mxGetClassID(Mc), ci = (int)mxGetClassID(idxS);
  if (e->have_material && e->mr.same(mxGetData(Mr), data_bytes(Mr), cr, mxIsComplex(Mr)) &&
      e->mc.same(mxGetData(Mc), data_bytes(Mc), cc, mxIsComplex(Mc)) &&
      e->idx.same(mxGetData(idxS), data_bytes(idxS), ci, false))
    return mxGetNumberOfElements(idxS);
  std::vector<int64_t> sorted, perm;
  std::string err;
  if (!vie_mex::idx_to_zero_based(class_of(idxS), mxGetData(idxS), mxGetNumberOfElements(idxS), nvox, sorted, perm,
                                  err))
    mexErrMsgIdAndTxt("vie_hip:arg", "dims.idxS: %s", err.c_str());
  std::vector<double> mr = complex_copy(Mr, "EP.Mr"), mc = complex_copy(Mc, "EP.Mc");
  e->have_material = false;
  check(vie_set_material(ctx, mr.data(), mc.data(), sorted.data(), (int64_t)sorted.size()), ctx);
  e->mr.assign(mxGetData(Mr), data_bytes(Mr), cr, mxIsComplex(Mr));
  e->mc.assign(mxGetData(Mc), data_bytes(Mc), cc, mxIsComplex(Mc));
  e->idx.assign(mxGetData(idxS), data_bytes(idxS), ci, false);
  e->perm.swap(perm);
  e->have_material = true;
  return sorted.size();
}

// packed vector in the library's (sorted idxS) order
std::vector<double> packed_in(const mxArray* a, const CtxCache::Entry* e, int ql, const char* what, int nvec = 1) {
  std::vector<double> v = complex_copy(a, what);
  if (e->perm.empty()) return v;
  std::vector<double> s(v.size());
  vie_mex::packed_to_sorted(v.data(), s.data(), e->perm, ql, nvec);
  return s;
}

// copies nvec packed vectors in the library's order to a new n x nvec complex output
mxArray* packed_out(const std::vector<double>& v, const CtxCache::Entry* e, int ql, size_t n, int nvec = 1) {
  mxArray* out = mxCreateDoubleMatrix(n, nvec, mxCOMPLEX);
  double* o = reinterpret_cast<double*>(mxGetComplexDoubles(out));
  if (e->perm.empty()) std::memcpy(o, v.data(), 2 * n * nvec * sizeof(double));
  else vie_mex::packed_from_sorted(v.data(), o, e->perm, ql, nvec);
  return out;
}

int scalar_int(const mxArray* a, const char* what) {
  if (!mxIsNumeric(a) || mxGetNumberOfElements(a) != 1) mexErrMsgIdAndTxt("vie_hip:arg", "%s must be a scalar", what);
  return (int)mxGetScalar(a);
}

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (nrhs < 1 || !mxIsChar(prhs[0])) mexErrMsgIdAndTxt("vie_hip:arg", "first argument: command name");
  char buf[64];
  mxGetString(prhs[0], buf, sizeof(buf));
  const std::string cmd(buf);
  if (cmd == "release") { release_all(); return; }

  const int op = op_of(cmd);
  if (op >= 0) {
    if (nrhs != 3) mexErrMsgIdAndTxt("vie_hip:arg", "usage: Jout = vie_hip_mex('%s', Jin, T)", buf);
    const mxArray* Jin = prhs[1];
    const mwSize* jd = mxGetDimensions(Jin);
    const size_t nd = mxGetNumberOfDimensions(Jin);
    if (nd < 3 || nd > 4) mexErrMsgIdAndTxt("vie_hip:arg", "Jin must be [n1,n2,n3,ql]");
    const int n[3] = {(int)jd[0], (int)jd[1], (int)jd[2]};
    const int opk = (op == VIE_OP_N || op == VIE_OP_NH) ? VIE_OP_N : VIE_OP_K;
    CtxCache::Entry* e = context_for(prhs[2], n, opk);
    const size_t ql = e->key.basis == VIE_BASIS_PWL ? 12 : 3;
    if ((nd == 4 ? (size_t)jd[3] : 1) != ql)
      mexErrMsgIdAndTxt("vie_hip:arg", "Jin must have ql = %d components for this T", (int)ql);
    vie_ctx* ctx = static_cast<vie_ctx*>(e->ctx);
    std::vector<double> in = complex_copy(Jin, "Jin");
    plhs[0] = mxCreateNumericArray(nd, jd, mxDOUBLE_CLASS, mxCOMPLEX);
    check(vie_apply(ctx, op, VIE_PREC_FP64, in.data(), (double*)mxGetComplexDoubles(plhs[0]), 1, VIE_MEM_HOST), ctx);
    return;
  }

  if (cmd == "mvp_vie" || cmd == "gmres_vie" || cmd == "cocr_vie" || cmd == "bcocg_vie") {
    if (nrhs < 5) mexErrMsgIdAndTxt("vie_hip:arg", "usage: vie_hip_mex('%s', Jb, EP, dims, fN, ...)", buf);
    const mxArray* dims = prhs[3];
    if (!mxIsStruct(dims) || !mxIsStruct(prhs[2])) mexErrMsgIdAndTxt("vie_hip:arg", "EP and dims must be structs");
    const int n[3] = {scalar_int(field(dims, 0, "n1"), "dims.n1"), scalar_int(field(dims, 0, "n2"), "dims.n2"),
                      scalar_int(field(dims, 0, "n3"), "dims.n3")};
    const double res = mxGetScalar(field(dims, 0, "res"));
    CtxCache::Entry* e = context_for(prhs[4], n, VIE_OP_N);
    vie_ctx* ctx = static_cast<vie_ctx*>(e->ctx);
    if (e->gram_res != res) {   // invG of mvp_vie.m:17 for this voxel size
      check(vie_set_gram(ctx, res), ctx);
      e->gram_res = res;
    }
    const int ql = e->key.basis == VIE_BASIS_PWL ? 12 : 3;
    const size_t nS = set_material(e, prhs[2], dims, (int64_t)n[0] * n[1] * n[2]);
    const size_t nb = (size_t)ql * nS;
    if (cmd == "cocr_vie" || cmd == "bcocg_vie") {   // B: one column per port, all solved in one batch
      const bool block = cmd == "bcocg_vie";
      const int nv = (int)mxGetN(prhs[1]);
      if (mxGetM(prhs[1]) != nb || nv < 1)
        mexErrMsgIdAndTxt("vie_hip:arg", "B must be [ql*nS, nports] with ql*nS = %d", (int)nb);
      std::vector<double> B = packed_in(prhs[1], e, ql, "B", nv);
      const double tol = nrhs > 5 ? mxGetScalar(prhs[5]) : 1e-5;
      const int maxit = nrhs > 6 ? scalar_int(prhs[6], "maxit") : 1000;
      if (maxit < 1) mexErrMsgIdAndTxt("vie_hip:arg", "maxit must be >= 1");
      std::vector<double> X(2 * nb * nv, 0.0);
      if (nrhs > 7 && !mxIsEmpty(prhs[7])) {
        if (mxGetM(prhs[7]) != nb || (int)mxGetN(prhs[7]) != nv)
          mexErrMsgIdAndTxt("vie_hip:arg", "X0 must match B");
        X = packed_in(prhs[7], e, ql, "X0", nv);
      }
      const int prec = nrhs > 8 ? scalar_int(prhs[8], "prec") : (block ? VIE_PREC_FP64 : VIE_PREC_FP32);
      const size_t cap = 2 * (size_t)maxit + 2;
      std::vector<double> resvec(cap * nv);
      std::vector<int32_t> nres(nv), nmv(nv);
      check((block ? vie_bcocg : vie_cocr)(ctx, VIE_OP_VIE, prec, B.data(), X.data(), nv, tol, maxit, VIE_MEM_HOST,
                                           resvec.data(), nres.data(), nmv.data(), nullptr),
            ctx);
      plhs[0] = packed_out(X, e, ql, nb, nv);
      if (nlhs > 1) {
        plhs[1] = mxCreateDoubleMatrix(1, nv, mxREAL);
        for (int i = 0; i < nv; ++i) mxGetDoubles(plhs[1])[i] = nmv[i];
      }
      if (nlhs > 2) {   // one column per port, NaN past its last entry
        size_t len = 0;
        for (int i = 0; i < nv; ++i) len = std::max(len, (size_t)nres[i]);
        plhs[2] = mxCreateDoubleMatrix(len, nv, mxREAL);
        double* r = mxGetDoubles(plhs[2]);
        for (int i = 0; i < nv; ++i)
          for (size_t k = 0; k < len; ++k) r[i * len + k] = k < (size_t)nres[i] ? resvec[i * cap + k] : mxGetNaN();
      }
      return;
    }
    if (mxGetNumberOfElements(prhs[1]) != nb)
      mexErrMsgIdAndTxt("vie_hip:arg", "Jb has %d elements; ql*nS = %d", (int)mxGetNumberOfElements(prhs[1]), (int)nb);
    std::vector<double> b = packed_in(prhs[1], e, ql, "Jb");
    std::vector<double> out(2 * nb, 0.0);
    if (cmd == "mvp_vie") {
      check(vie_apply(ctx, VIE_OP_VIE, VIE_PREC_FP64, b.data(), out.data(), 1, VIE_MEM_HOST), ctx);
      plhs[0] = packed_out(out, e, ql, nb);
      return;
    }
    const int restart = nrhs > 5 ? scalar_int(prhs[5], "restart") : 50;
    const double tol = nrhs > 6 ? mxGetScalar(prhs[6]) : 1e-5;
    const int maxit = nrhs > 7 ? scalar_int(prhs[7], "maxit") : 200;
    if (restart < 1 || maxit < 1) mexErrMsgIdAndTxt("vie_hip:arg", "restart and maxit must be >= 1");
    if (nrhs > 8 && !mxIsEmpty(prhs[8])) {
      if (mxGetNumberOfElements(prhs[8]) != nb)
        mexErrMsgIdAndTxt("vie_hip:arg", "x0 has %d elements; ql*nS = %d", (int)mxGetNumberOfElements(prhs[8]),
                          (int)nb);
      out = packed_in(prhs[8], e, ql, "x0");
    }
    const int prec = nrhs > 9 ? scalar_int(prhs[9], "prec") : VIE_PREC_FP64;
    std::vector<double> resvec((size_t)restart * maxit + 1);
    int32_t nres = 0, iters[2] = {0, 0}, nmv = 0;
    check(vie_gmres(ctx, VIE_OP_VIE, prec, b.data(), out.data(), 1, tol, restart, maxit, VIE_MEM_HOST, resvec.data(),
                    &nres, iters, &nmv), ctx);
    plhs[0] = packed_out(out, e, ql, nb);
    if (nlhs > 1) {
      plhs[1] = mxCreateDoubleMatrix(1, 2, mxREAL);
      mxGetDoubles(plhs[1])[0] = iters[0];
      mxGetDoubles(plhs[1])[1] = iters[1];
    }
    if (nlhs > 2) {
      plhs[2] = mxCreateDoubleMatrix(nres, 1, mxREAL);
      std::memcpy(mxGetDoubles(plhs[2]), resvec.data(), nres * sizeof(double));
    }
    return;
  }
  mexErrMsgIdAndTxt("vie_hip:arg", "unknown command '%s'", buf);
}
