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
//   vie_hip_mex('release')                        % drop every cached context
// T is the Tucker T-struct of assembly_fft_circ_tucker_*.m; the gateway
// expands it on the device once and caches the context keyed by the data
// pointers of T(1).G and T(end).U3 and the grid size, so repeated calls with
// the same T (every GMRES iteration) reuse it.  The material of mvp_vie is
// cached by the data pointers of EP.Mr / EP.Mc.
//
// Library status codes become MATLAB errors with identifiers vie_hip:arg,
// vie_hip:hip, vie_hip:oom, vie_hip:state, so solver_wsvie.m's try/catch
// GPU -> CPU fallback (solver_wsvie.m:10-23) keeps working.
//
// Build (where MATLAB is installed; libvie_hip.so built by the Makefile):
//   mex -R2018a vie_hip_mex.cpp -I../../../include -L../../marie_vie -lvie_hip
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "mex.h"
#include "vie_op.h"

namespace {

struct Cached {
  vie_ctx* ctx = nullptr;
  const void* mr = nullptr;
  const void* mc = nullptr;
};

using Key = std::tuple<const void*, const void*, int, int, int, int, int, double>;
std::map<Key, Cached> g_cache;

void release_all() {
  for (auto& kv : g_cache) vie_destroy(kv.second.ctx);
  g_cache.clear();
}

void fail(int rc, vie_ctx* ctx) {
  const char* id = rc == VIE_ERR_OOM ? "vie_hip:oom" : rc == VIE_ERR_ARG ? "vie_hip:arg"
                   : rc == VIE_ERR_STATE ? "vie_hip:state" : "vie_hip:hip";
  mexErrMsgIdAndTxt(id, "libvie_hip: %s", vie_last_error(ctx));
}

void check(int rc, vie_ctx* ctx) {
  if (rc != VIE_OK) fail(rc, ctx);
}

// complex copy of a real or complex double array (interleaved)
std::vector<double> complex_copy(const mxArray* a) {
  const size_t n = mxGetNumberOfElements(a);
  std::vector<double> v(2 * n, 0.0);
  if (!mxIsDouble(a)) mexErrMsgIdAndTxt("vie_hip:arg", "expected a double array");
  if (mxIsComplex(a)) {
    std::memcpy(v.data(), mxGetComplexDoubles(a), 2 * n * sizeof(double));
  } else {
    const double* r = mxGetDoubles(a);
    for (size_t i = 0; i < n; ++i) v[2 * i] = r[i];
  }
  return v;
}

// Context for the T-struct T (n1, n2, n3 from the size of Jin); op 'N' or 'K'.
// res only enters invG of mvp_vie; the plain N/K handles pass 1.
vie_ctx* context_for(const mxArray* T, int n1, int n2, int n3, int op, double res) {
  if (!mxIsStruct(T)) mexErrMsgIdAndTxt("vie_hip:arg", "T must be the Tucker T-struct");
  const size_t nt = mxGetNumberOfElements(T);
  const int ncomp = op == VIE_OP_N ? 6 : 3;
  const int basis = nt == (size_t)ncomp ? VIE_BASIS_PWC : VIE_BASIS_PWL;
  const mxArray* g0 = mxGetField(T, 0, "G");
  const mxArray* ul = mxGetField(T, nt - 1, "U3");
  Key key{mxGetData(g0), mxGetData(ul), n1, n2, n3, op, basis, res};
  auto it = g_cache.find(key);
  if (it != g_cache.end()) return it->second.ctx;
  vie_desc d{n1, n2, n3, basis, res, 0};
  vie_ctx* ctx = nullptr;
  check(vie_create(&d, &ctx), nullptr);
  // T(i,j) is column-major in MATLAB: element (i,j) at i + 10*j; the library
  // numbers PWL terms i*ncomp + j
  const int npair = basis == VIE_BASIS_PWL ? 10 : 1;
  for (int i = 0; i < npair; ++i)
    for (int j = 0; j < ncomp; ++j) {
      const size_t e = basis == VIE_BASIS_PWL ? (size_t)i + 10 * (size_t)j : (size_t)j;
      const mxArray* G = mxGetField(T, e, "G");
      const mxArray* U1 = mxGetField(T, e, "U1");
      const mxArray* U2 = mxGetField(T, e, "U2");
      const mxArray* U3 = mxGetField(T, e, "U3");
      const mwSize* gd = mxGetDimensions(G);
      const int nd = (int)mxGetNumberOfDimensions(G);
      const int r1 = (int)gd[0], r2 = nd > 1 ? (int)gd[1] : 1, r3 = nd > 2 ? (int)gd[2] : 1;
      std::vector<double> cg = complex_copy(G), c1 = complex_copy(U1), c2 = complex_copy(U2), c3 = complex_copy(U3);
      const int term = basis == VIE_BASIS_PWL ? i * ncomp + j : j;
      int rc = vie_set_green_tucker(ctx, op, term, cg.data(), r1, r2, r3, c1.data(), c2.data(), c3.data());
      if (rc != VIE_OK) {
        std::string msg = vie_last_error(ctx);
        vie_destroy(ctx);
        mexErrMsgIdAndTxt("vie_hip:arg", "libvie_hip: %s", msg.c_str());
      }
    }
  g_cache[key] = Cached{ctx, nullptr, nullptr};
  return ctx;
}

int op_of(const std::string& name) {
  if (name == "mvp_N") return VIE_OP_N;
  if (name == "mvp_K") return VIE_OP_K;
  if (name == "mvp_herm_adj_N") return VIE_OP_NH;
  if (name == "mvp_herm_adj_K") return VIE_OP_KH;
  return -1;
}

// set (or reuse) the material of a context from EP and dims
void set_material(vie_ctx* ctx, const mxArray* EP, const mxArray* dims) {
  const mxArray* Mr = mxGetField(EP, 0, "Mr");
  const mxArray* Mc = mxGetField(EP, 0, "Mc");
  for (auto& kv : g_cache)
    if (kv.second.ctx == ctx && kv.second.mr == mxGetData(Mr) && kv.second.mc == mxGetData(Mc)) return;
  const mxArray* idxS = mxGetField(dims, 0, "idxS");
  const size_t nS = mxGetNumberOfElements(idxS);
  std::vector<int64_t> idx(nS);
  const double* p = mxGetDoubles(idxS);
  for (size_t s = 0; s < nS; ++s) idx[s] = (int64_t)p[s] - 1;  // MATLAB 1-based
  std::vector<double> mr = complex_copy(Mr), mc = complex_copy(Mc);
  check(vie_set_material(ctx, mr.data(), mc.data(), idx.data(), (int64_t)nS), ctx);
  for (auto& kv : g_cache)
    if (kv.second.ctx == ctx) { kv.second.mr = mxGetData(Mr); kv.second.mc = mxGetData(Mc); }
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
    if (mxGetNumberOfDimensions(Jin) < 3) mexErrMsgIdAndTxt("vie_hip:arg", "Jin must be [n1,n2,n3,ql]");
    const int opk = (op == VIE_OP_N || op == VIE_OP_NH) ? VIE_OP_N : VIE_OP_K;
    vie_ctx* ctx = context_for(prhs[2], (int)jd[0], (int)jd[1], (int)jd[2], opk, 1.0);
    std::vector<double> in = complex_copy(Jin);
    plhs[0] = mxCreateNumericArray(mxGetNumberOfDimensions(Jin), jd, mxDOUBLE_CLASS, mxCOMPLEX);
    check(vie_apply(ctx, op, VIE_PREC_FP64, in.data(), (double*)mxGetComplexDoubles(plhs[0]), 1, VIE_MEM_HOST), ctx);
    return;
  }

  if (cmd == "mvp_vie" || cmd == "gmres_vie") {
    if (nrhs < 5) mexErrMsgIdAndTxt("vie_hip:arg", "usage: vie_hip_mex('%s', Jb, EP, dims, fN, ...)", buf);
    const mxArray* dims = prhs[3];
    const int n1 = (int)mxGetScalar(mxGetField(dims, 0, "n1"));
    const int n2 = (int)mxGetScalar(mxGetField(dims, 0, "n2"));
    const int n3 = (int)mxGetScalar(mxGetField(dims, 0, "n3"));
    const double res = mxGetScalar(mxGetField(dims, 0, "res"));
    vie_ctx* ctx = context_for(prhs[4], n1, n2, n3, VIE_OP_N, res);
    set_material(ctx, prhs[2], dims);
    std::vector<double> b = complex_copy(prhs[1]);
    const mwSize n = mxGetNumberOfElements(prhs[1]);
    if (cmd == "mvp_vie") {
      plhs[0] = mxCreateDoubleMatrix(n, 1, mxCOMPLEX);
      check(vie_apply(ctx, VIE_OP_VIE, VIE_PREC_FP64, b.data(), (double*)mxGetComplexDoubles(plhs[0]), 1,
                      VIE_MEM_HOST), ctx);
      return;
    }
    const int restart = nrhs > 5 ? (int)mxGetScalar(prhs[5]) : 50;
    const double tol = nrhs > 6 ? mxGetScalar(prhs[6]) : 1e-5;
    const int maxit = nrhs > 7 ? (int)mxGetScalar(prhs[7]) : 200;
    std::vector<double> x(2 * n, 0.0);
    if (nrhs > 8) x = complex_copy(prhs[8]);
    const int prec = nrhs > 9 ? (int)mxGetScalar(prhs[9]) : VIE_PREC_FP64;
    std::vector<double> resvec((size_t)restart * maxit + 1);
    int32_t nres = 0, iters[2] = {0, 0}, nmv = 0;
    check(vie_gmres(ctx, VIE_OP_VIE, prec, b.data(), x.data(), 1, tol, restart, maxit, VIE_MEM_HOST, resvec.data(),
                    &nres, iters, &nmv), ctx);
    plhs[0] = mxCreateDoubleMatrix(n, 1, mxCOMPLEX);
    std::memcpy(mxGetComplexDoubles(plhs[0]), x.data(), 2 * n * sizeof(double));
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
