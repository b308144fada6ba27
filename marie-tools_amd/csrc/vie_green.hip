// Set-up: far- and medium-field Galerkin integrals of the VIE kernels.
//
// For each voxel offset m in a box [0,e1) x [0,e2) x [0,e3) the kernel
// integrates, with an nq-point Gauss-Legendre rule in each of the six
// coordinates (the rule of VV_Nop.m:1-49 / VV_Kop.m:1-65 as assembly_N.m:20-52
// and assembly_K.m use it),
//   A(m, pair, comp) = (h/2)^6 sum_q W_q T_l(X) B_lp(Xp) f_comp(r_m + h/2 (X - Xp))
// where f are the six N components (d_i d_j + k0^2 delta_ij) g or the three K
// components d_j g of g = exp(-i k0 R) / (4 pi R) (G_N.m:3-19, G_K.m:3-12),
// T_l / B_lp the monomials 1, X/2, Y/2, Z/2 (test) and 1, Xp/2, Yp/2, Zp/2
// (basis), and the ten (l, lp) pairs of VV_Nop.m:27-28.
// One block per offset; the nq^6 points are split over the block's threads
// and reduced through wavefront shuffles and LDS.

#include "vie_internal.hpp"

namespace vie {

constexpr int kGreenBlock = 256;
__constant__ int c_ind_l[10] = {0, 1, 2, 3, 0, 0, 0, 1, 1, 2};
__constant__ int c_ind_lp[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};

template <int NP, int NC>
__global__ void __launch_bounds__(kGreenBlock) green_vv_kernel(double2* __restrict__ A, int n1, int n2, int n3, int e1,
                                                               int e2, int e3, int nq, const double* __restrict__ xq,
                                                               const double* __restrict__ wq, double h, double k0) {
  __shared__ double xs[16], ws[16];
  __shared__ double red[kGreenBlock / 64][2 * NP * NC];
  const int tid = threadIdx.x;
  if (tid < nq) { xs[tid] = xq[tid]; ws[tid] = wq[tid]; }
  __syncthreads();
  const int bid = blockIdx.x;
  const int mx = bid % e1, my = (bid / e1) % e2, mz = bid / (e1 * e2);
  double acc[2 * NP * NC];
#pragma unroll
  for (int q = 0; q < 2 * NP * NC; ++q) acc[q] = 0;
  const int nq3 = nq * nq * nq;
  const int64_t npts = (int64_t)nq3 * nq3;
  const double hh = 0.5 * h;
  const double inv4pi = 0.07957747154594766788;  // 1 / (4 pi)
  for (int64_t p = tid; p < npts; p += kGreenBlock) {
    const int a = (int)(p % nq3), b = (int)(p / nq3);
    const int ix = a % nq, iy = (a / nq) % nq, iz = a / (nq * nq);
    const int jx = b % nq, jy = (b / nq) % nq, jz = b / (nq * nq);
    const double X = xs[ix], Y = xs[iy], Z = xs[iz], Xp = xs[jx], Yp = xs[jy], Zp = xs[jz];
    const double dx = mx * h + hh * (X - Xp), dy = my * h + hh * (Y - Yp), dz = mz * h + hh * (Z - Zp);
    const double R2 = dx * dx + dy * dy + dz * dz;
    if (R2 == 0.0) continue;  // coincident points only occur on offsets the near-field pass overwrites
    const double R = sqrt(R2);
    const double W = ws[ix] * ws[iy] * ws[iz] * ws[jx] * ws[jy] * ws[jz];
    double sn, cs;
    sincos(k0 * R, &sn, &cs);
    // Gexp = exp(-i k0 R) / (4 pi), times the quadrature weight
    const double ger = cs * inv4pi * W, gei = -sn * inv4pi * W;
    const double iR = 1.0 / R, iR2 = iR * iR, iR3 = iR2 * iR;
    // c = -(1/R^3 + i k0 / R^2)
    const double cr = -iR3, ci = -k0 * iR2;
    double f[2 * NC];
    if (NC == 6) {
      // a = 3/R^5 + 3 i k0 / R^4 - k0^2 / R^3
      const double ar = 3.0 * iR3 * iR2 - k0 * k0 * iR3, ai = 3.0 * k0 * iR2 * iR2;
      const double xx = dx * dx, yy = dy * dy, zz = dz * dz;
      // components before the Gexp factor
      const double t1r = -(ar * (yy + zz) + 2 * cr), t1i = -(ai * (yy + zz) + 2 * ci);
      const double t4r = -(ar * (xx + zz) + 2 * cr), t4i = -(ai * (xx + zz) + 2 * ci);
      const double t6r = -(ar * (xx + yy) + 2 * cr), t6i = -(ai * (xx + yy) + 2 * ci);
      const double t2r = ar * dx * dy, t2i = ai * dx * dy;
      const double t3r = ar * dx * dz, t3i = ai * dx * dz;
      const double t5r = ar * dy * dz, t5i = ai * dy * dz;
      const double tr[6] = {t1r, t2r, t3r, t4r, t5r, t6r}, ti[6] = {t1i, t2i, t3i, t4i, t5i, t6i};
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        f[2 * c] = ger * tr[c] - gei * ti[c];
        f[2 * c + 1] = ger * ti[c] + gei * tr[c];
      }
    } else {
      const double d[3] = {dx, dy, dz};
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const double tr = cr * d[c], ti = ci * d[c];
        f[2 * c] = ger * tr - gei * ti;
        f[2 * c + 1] = ger * ti + gei * tr;
      }
    }
    if (NP == 1) {
#pragma unroll
      for (int q = 0; q < 2 * NC; ++q) acc[q] += f[q];
    } else {
      const double T[4] = {1.0, 0.5 * X, 0.5 * Y, 0.5 * Z};
      const double B[4] = {1.0, 0.5 * Xp, 0.5 * Yp, 0.5 * Zp};
#pragma unroll
      for (int pr = 0; pr < NP; ++pr) {
        const double tb = T[c_ind_l[pr]] * B[c_ind_lp[pr]];
#pragma unroll
        for (int q = 0; q < 2 * NC; ++q) acc[2 * NC * pr + q] += tb * f[q];
      }
    }
  }
  // block reduction
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int q = 0; q < 2 * NP * NC; ++q) {
    double v = acc[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) red[wave][q] = v;
  }
  __syncthreads();
  const double J = hh * hh * hh * hh * hh * hh;
  const int64_t N = (int64_t)n1 * n2 * n3;
  const int64_t vox = mx + (int64_t)n1 * (my + (int64_t)n2 * mz);
  for (int q = tid; q < NP * NC; q += kGreenBlock) {
    double sr = 0, si = 0;
    for (int w = 0; w < kGreenBlock / 64; ++w) { sr += red[w][2 * q]; si += red[w][2 * q + 1]; }
    const int pr = q / NC, c = q % NC;
    A[vox + N * (pr + (int64_t)NP * c)] = make_double2(J * sr, J * si);
  }
}

}  // namespace vie

using namespace vie;

extern "C" int vie_green_quadrature(vie_ctx* h, int32_t op, double k0, int32_t nq, const double* nodes,
                                    const double* weights, int32_t e1, int32_t e2, int32_t e3, double* A, int32_t mem) {
  if (!h) return VIE_ERR_ARG;
  Ctx& c = h->c;
  try {
    c.err.clear();
    if (op != VIE_OP_N && op != VIE_OP_K) throw Error{VIE_ERR_ARG, "op must be VIE_OP_N or VIE_OP_K"};
    if (nq < 1 || nq > 16) throw Error{VIE_ERR_ARG, "nq must be in 1..16"};
    if (!nodes || !weights || !A) throw Error{VIE_ERR_ARG, "NULL argument"};
    if (e1 < 1 || e2 < 1 || e3 < 1 || e1 > c.n[0] || e2 > c.n[1] || e3 > c.n[2])
      throw Error{VIE_ERR_ARG, "offset box outside the grid"};
    VIE_HIP_CHECK(hipSetDevice(c.device));
    const int NC = op == VIE_OP_N ? 6 : 3;
    const int64_t N = c.nvox();
    const size_t bytes = (size_t)N * c.npair * NC * sizeof(double2);
    DBuf tmp, q;
    q.ensure(32 * sizeof(double));
    VIE_HIP_CHECK(hipMemcpyAsync(q.p, nodes, nq * sizeof(double), hipMemcpyHostToDevice, c.stream));
    VIE_HIP_CHECK(hipMemcpyAsync(q.as<double>() + 16, weights, nq * sizeof(double), hipMemcpyHostToDevice, c.stream));
    double2* dA;
    if (mem == VIE_MEM_DEVICE) {
      dA = reinterpret_cast<double2*>(A);
    } else {
      tmp.ensure(bytes);
      dA = tmp.as<double2>();
      VIE_HIP_CHECK(hipMemcpyAsync(dA, A, bytes, hipMemcpyHostToDevice, c.stream));
    }
    const int nb = e1 * e2 * e3;
    const double* dx = q.as<double>();
    const double* dw = q.as<double>() + 16;
    const double hres = c.res;
    if (c.npair == 1 && NC == 6)
      hipLaunchKernelGGL((green_vv_kernel<1, 6>), dim3(nb), dim3(kGreenBlock), 0, c.stream, dA, c.n[0], c.n[1], c.n[2],
                         e1, e2, e3, nq, dx, dw, hres, k0);
    else if (c.npair == 1)
      hipLaunchKernelGGL((green_vv_kernel<1, 3>), dim3(nb), dim3(kGreenBlock), 0, c.stream, dA, c.n[0], c.n[1], c.n[2],
                         e1, e2, e3, nq, dx, dw, hres, k0);
    else if (NC == 6)
      hipLaunchKernelGGL((green_vv_kernel<10, 6>), dim3(nb), dim3(kGreenBlock), 0, c.stream, dA, c.n[0], c.n[1],
                         c.n[2], e1, e2, e3, nq, dx, dw, hres, k0);
    else
      hipLaunchKernelGGL((green_vv_kernel<10, 3>), dim3(nb), dim3(kGreenBlock), 0, c.stream, dA, c.n[0], c.n[1],
                         c.n[2], e1, e2, e3, nq, dx, dw, hres, k0);
    VIE_HIP_CHECK(hipGetLastError());
    if (mem != VIE_MEM_DEVICE)
      VIE_HIP_CHECK(hipMemcpyAsync(A, dA, bytes, hipMemcpyDeviceToHost, c.stream));
    VIE_HIP_CHECK(hipStreamSynchronize(c.stream));
    return VIE_OK;
  } catch (const Error& e) {
    c.err = e.msg;
    return e.code;
  } catch (...) {
    c.err = "unknown error in vie_green_quadrature";
    return VIE_ERR_HIP;
  }
}
