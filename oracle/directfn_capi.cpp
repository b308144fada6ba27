// C entry points over the reference's DIRECTFN singular surface-surface
// integrals (Cpp_Assembly/src/src_cpp/create_{ST,EA,VA}.cpp), for the oracle's
// ctypes binding (oracle/near_field.py).  Test infrastructure only: built by
// oracle/Makefile from the reference sources in place into oracle/_ref/, never
// shipped and never loaded by the product.
//
// Each entry point takes the ordered corner points of the square pair, the
// 1D quadrature order N (used for all four DIRECTFN directions, as
// singular_{ST,EA,VA}_lin.m pass Np four times), k0, the voxel edge dx, the
// square centres, the normals, the reduced kernel type (1..4), the test and
// basis monomial indices (0 = constant, 1..3 = x, y, z), and writes the
// complex result to out[0] (real) and out[1] (imaginary).  Argument order and
// meaning follow the MEX gateways solve_{st,ea,va}_mex.cpp, which pass
// (rq_c, rp_c, nq, np) straight to create_{ST,EA,VA}.
#include <complex>

using std::complex;

void create_ST(const double r1[], const double r2[], const double r3[], const double r4[], int N1, int N2, int N3,
               int N4, double k0, double dx, double rq_c[], double rp_c[], double nq[], double np[], int ker_type,
               int l, int lp, complex<double> I[]);
void create_EA(const double r1[], const double r2[], const double r3[], const double r4[], const double r5[],
               const double r6[], const int N1, const int N2, const int N3, const int N4, double k0, double dx,
               double rq_c[], double rp_c[], double nq[], double np[], int ker_type, int l, int lp, complex<double> I[]);
void create_VA(const double r1[], const double r2[], const double r3[], const double r4[], const double r5[],
               const double r6[], const double r7[], const int N1, const int N2, const int N3, const int N4, double k0,
               double dx, double rq_c[], double rp_c[], double nq[], double np[], int ker_type, int l, int lp,
               complex<double> I[]);

namespace {
void copy3(const double* s, double* d) {
  for (int i = 0; i < 3; ++i) d[i] = s[i];
}
}  // namespace

extern "C" {

// pts: 4 x 3 corners (r1..r4)
int dfn_st(const double* pts, int N, double k0, double dx, const double* rq_c, const double* rp_c, const double* nq,
           const double* np, int ker_type, int l, int lp, double* out) {
  double cq[3], cp[3], vq[3], vp[3];
  copy3(rq_c, cq); copy3(rp_c, cp); copy3(nq, vq); copy3(np, vp);
  complex<double> I[1];
  create_ST(pts, pts + 3, pts + 6, pts + 9, N, N, N, N, k0, dx, cq, cp, vq, vp, ker_type, l, lp, I);
  out[0] = I[0].real();
  out[1] = I[0].imag();
  return 0;
}

// pts: 6 x 3 corners (r1..r6)
int dfn_ea(const double* pts, int N, double k0, double dx, const double* rq_c, const double* rp_c, const double* nq,
           const double* np, int ker_type, int l, int lp, double* out) {
  double cq[3], cp[3], vq[3], vp[3];
  copy3(rq_c, cq); copy3(rp_c, cp); copy3(nq, vq); copy3(np, vp);
  complex<double> I[1];
  create_EA(pts, pts + 3, pts + 6, pts + 9, pts + 12, pts + 15, N, N, N, N, k0, dx, cq, cp, vq, vp, ker_type, l, lp,
            I);
  out[0] = I[0].real();
  out[1] = I[0].imag();
  return 0;
}

// pts: 7 x 3 corners (r1..r7)
int dfn_va(const double* pts, int N, double k0, double dx, const double* rq_c, const double* rp_c, const double* nq,
           const double* np, int ker_type, int l, int lp, double* out) {
  double cq[3], cp[3], vq[3], vp[3];
  copy3(rq_c, cq); copy3(rp_c, cp); copy3(nq, vq); copy3(np, vp);
  complex<double> I[1];
  create_VA(pts, pts + 3, pts + 6, pts + 9, pts + 12, pts + 15, pts + 18, N, N, N, N, k0, dx, cq, cp, vq, vp,
            ker_type, l, lp, I);
  out[0] = I[0].real();
  out[1] = I[0].imag();
  return 0;
}

}  // extern "C"
