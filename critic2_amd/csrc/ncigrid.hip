// NCIPLOT's grid-interpolation scheme (nci@proc.f90:565-571): density,
// gradient and Hessian of a loaded grid from its tricubic interpolant
// (grid3%interp in the default tricubic mode, grid3mod@proc.f90:1695-1752 and
// :2649-2799) at each node of the NCI grid (= the loaded grid,
// nci@proc.f90:418-428).
//
// The node's containing cell is found with the reference's own arithmetic,
// which rounds: a node can land in the cell below it (then u is 1 - eps) and,
// where a rounded-below-zero coordinate wraps to 1.0, in the cell at the
// origin with u = n (the reference evaluates that cell's cubic there).  The
// node search must therefore round exactly like the reference, and this file
// is compiled without floating-point contraction.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "common.hpp"
#include "nci_dev.hpp"

namespace c2 {

// nci@proc.f90:565-571 through field%grd and grid3%interp in its default
// tricubic mode (grid3mod@proc.f90:2649-2799).  The tricubic of Lekien and
// Marsden with central-difference corner derivatives is, per axis, the cubic
// Hermite interpolant whose end slopes are (f(1)-f(-1))/2 and (f(2)-f(0))/2:
// a 4-point stencil with weights w_s(u), s = -1..2.  Value, gradient and
// Hessian are separable contractions of the 4x4x4 neighbourhood.
struct GridInterp {
  double xm[9];   // xmat(:,c) = x2c(:,c) / n(c), nci@proc.f90:418-428
  double c2x[9];  // matinv(x2c): crystal%m_c2x and the grid's c2xl
};

// Fractional coordinates of node (i,j,k) exactly as nciplot and field%grd
// form them: x = 0 + i*xmat(:,1) + j*xmat(:,2) + k*xmat(:,3), then
// matmul(m_c2x, x), the 1d-4 border fold and modulo(.,1d0).  This file is
// built without FP contraction, so every product and sum rounds as in the
// reference and floor() below picks the reference's cell.
__device__ __forceinline__ void grid_node_frac(const GridInterp &T, int i, int j, int k, double wx[3]) {
  double xc[3];
  for (int r = 0; r < 3; r++)
    xc[r] = ((0.0 + (double)i * T.xm[r]) + (double)j * T.xm[r + 3]) + (double)k * T.xm[r + 6];
  for (int r = 0; r < 3; r++) {
    double v = ((0.0 + T.c2x[r] * xc[0]) + T.c2x[r + 3] * xc[1]) + T.c2x[r + 6] * xc[2];
    if (v < -1e-4 || v > 1.0 + 1e-4) v = v - floor(v);
    wx[r] = v - floor(v);
  }
}

// periodic index in 0..n-1 for v in -n..2n-1
__device__ __forceinline__ int wrap0(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// weights of the value and first and second derivatives at offset u
__device__ __forceinline__ void hermite_cr(double u, double w[4], double d[4], double dd[4]) {
  const double u2 = u * u, u3 = u2 * u;
  const double h00 = 2.0 * u3 - 3.0 * u2 + 1.0, h01 = 3.0 * u2 - 2.0 * u3;
  const double h10 = u3 - 2.0 * u2 + u, h11 = u3 - u2;
  const double p00 = 6.0 * u2 - 6.0 * u, p01 = 6.0 * u - 6.0 * u2;
  const double p10 = 3.0 * u2 - 4.0 * u + 1.0, p11 = 3.0 * u2 - 2.0 * u;
  const double q00 = 12.0 * u - 6.0, q01 = 6.0 - 12.0 * u, q10 = 6.0 * u - 4.0, q11 = 6.0 * u - 2.0;
  w[0] = -0.5 * h10;
  w[1] = h00 - 0.5 * h11;
  w[2] = h01 + 0.5 * h10;
  w[3] = 0.5 * h11;
  d[0] = -0.5 * p10;
  d[1] = p00 - 0.5 * p11;
  d[2] = p01 + 0.5 * p10;
  d[3] = 0.5 * p11;
  dd[0] = -0.5 * q10;
  dd[1] = q00 - 0.5 * q11;
  dd[2] = q01 + 0.5 * q10;
  dd[3] = 0.5 * q11;
}

__global__ void __launch_bounds__(256) k_nci_grid(Geo G, GridInterp T, const double *__restrict__ rho,
                                                  NciOut O) {
  const int b = blockIdx.x;
  const int b1 = b % G.nb1, b2 = (b / G.nb1) % G.nb2, b3 = b / (G.nb1 * G.nb2);
  const int t = threadIdx.x;
  const int i = b1 * BX + (t % BX), j = b2 * BY + ((t / BX) % BY), k = b3 * BZ + t / (BX * BY);
  if (i >= G.n1 || j >= G.n2 || k >= G.n3) return;
  double wx[3];
  grid_node_frac(T, i, j, k, wx);
  const int nn[3] = {G.n1, G.n2, G.n3};
  int base[3];
  double w[3][4], d[3][4], dd[3][4];
  for (int c = 0; c < 3; c++) {
    const double tt = wx[c] * (double)nn[c];
    long fl = (long)floor(tt) % nn[c];
    if (fl < 0) fl += nn[c];
    hermite_cr(tt - (double)fl, w[c], d[c], dd[c]);  // offset from the cell's lower node
    base[c] = (int)fl;
  }
  int ia[4], jb[4];
  size_t kc[4];
  for (int s = 0; s < 4; s++) {
    ia[s] = wrap0(base[0] + s - 1, G.n1);
    jb[s] = wrap0(base[1] + s - 1, G.n2);
    kc[s] = (size_t)wrap0(base[2] + s - 1, G.n3) * G.n2;
  }
  double f = 0, gx = 0, gy = 0, gz = 0, hxx = 0, hyy = 0, hzz = 0, hxy = 0, hxz = 0, hyz = 0;
  for (int s = 0; s < 4; s++) {
    double y00 = 0, y10 = 0, y01 = 0, y20 = 0, y11 = 0, y02 = 0;
    for (int q = 0; q < 4; q++) {
      double z0 = 0, z1 = 0, z2 = 0;
      for (int v = 0; v < 4; v++) {
        const double g = rho[(size_t)ia[s] + (size_t)G.n1 * ((size_t)jb[q] + kc[v])];
        z0 += w[2][v] * g;
        z1 += d[2][v] * g;
        z2 += dd[2][v] * g;
      }
      y00 += w[1][q] * z0;
      y10 += d[1][q] * z0;
      y01 += w[1][q] * z1;
      y20 += dd[1][q] * z0;
      y11 += d[1][q] * z1;
      y02 += w[1][q] * z2;
    }
    f += w[0][s] * y00;
    gx += d[0][s] * y00;
    gy += w[0][s] * y10;
    gz += w[0][s] * y01;
    hxx += dd[0][s] * y00;
    hyy += w[0][s] * y20;
    hzz += w[0][s] * y02;
    hxy += d[0][s] * y10;
    hxz += d[0][s] * y01;
    hyz += w[0][s] * y11;
  }
  // fractional -> per-unit-cell derivatives, then Cartesian:
  // yp = c2xl^T yp, ypp = c2xl^T ypp c2xl (grid3mod@proc.f90:1746-1749)
  const double n1 = G.n1, n2 = G.n2, n3 = G.n3;
  const double yp[3] = {gx * n1, gy * n2, gz * n3};
  const double H[9] = {hxx * n1 * n1, hxy * n1 * n2, hxz * n1 * n3, hxy * n1 * n2, hyy * n2 * n2,
                       hyz * n2 * n3, hxz * n1 * n3, hyz * n2 * n3, hzz * n3 * n3};  // symmetric
  const double *C = T.c2x;  // column-major: C[q + 3r] = c2x(q,r)
  double gc[3], M[9], Hc[9];
  for (int r = 0; r < 3; r++) gc[r] = C[3 * r] * yp[0] + C[3 * r + 1] * yp[1] + C[3 * r + 2] * yp[2];
  for (int r = 0; r < 3; r++)  // M = C^T H
    for (int c = 0; c < 3; c++)
      M[r + 3 * c] = C[3 * r] * H[3 * c] + C[3 * r + 1] * H[1 + 3 * c] + C[3 * r + 2] * H[2 + 3 * c];
  for (int r = 0; r < 3; r++)  // Hc = M C
    for (int c = 0; c < 3; c++)
      Hc[r + 3 * c] = M[r] * C[3 * c] + M[r + 3] * C[1 + 3 * c] + M[r + 6] * C[2 + 3 * c];
  const double gmod = sqrt(gc[0] * gc[0] + gc[1] * gc[1] + gc[2] * gc[2]);
  const double cst = 2.0 * pow(3.0 * PI_D * PI_D, 1.0 / 3.0);
  const double den = f > 1e-80 ? f : 1e-80;
  const double sv = gmod / (cst * pow(den, 4.0 / 3.0));
  const double l2 = eig_mid_t<double>(Hc[0], Hc[4], Hc[8], Hc[3], Hc[6], Hc[7]);
  const double a = fabs(f) * 100.0;
  const double crho = (l2 < 0.0 || (l2 == 0.0 && signbit(l2))) ? -a : a;
  store_out<double>(O, out_index(G, O.zfast, i, j, k), crho, sv);
}

void nci_grid_launch(c2gpu_ctx *ctx, const Geo &G, const double x2c[9], const double c2x[9],
                     const double *d_rho, const NciOut &O) {
  GridInterp T;
  const int n[3] = {G.n1, G.n2, G.n3};
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) T.xm[r + 3 * c] = x2c[r + 3 * c] / (double)n[c];
  std::memcpy(T.c2x, c2x, sizeof(T.c2x));
  const int nbk = G.nb1 * G.nb2 * G.nb3;
  C2_LAUNCH(ctx, k_nci_grid, dim3(nbk), dim3(256), 0, G, T, d_rho, O);
}

}  // namespace c2
