// Device definitions shared by the NCI sources: the 8x8x4 node bricks, the
// output layout of nciplot's crho/cgrad arrays (nci@proc.f90:502-505) and the
// middle eigenvalue of the Hessian.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace c2 {

static constexpr int BX = 8, BY = 8, BZ = 4;  // brick of 256 nodes
static const double PI_D = 3.14159265358979323846264338327950288;

// The nodes: those of a periodic grid of the cell x2c (box == 0), at
// x2c (i/n1, j/n2, k/n3), or those of an nciplot cube (box == 1), at
// x0 + i xm(:,1) + j xm(:,2) + k xm(:,3) (nci@proc.f90:548).
struct Geo {
  int n1, n2, n3;
  int nb1, nb2, nb3;  // bricks per axis
  double x2c[9];
  int box;
  double x0[3], xm[9];
};

// Cartesian position at real-valued node coordinates (u, v, w).
__device__ __forceinline__ void geo_cart(const Geo &G, double u, double v, double w, double xc[3]) {
  if (G.box) {
#pragma unroll
    for (int r = 0; r < 3; r++) xc[r] = G.x0[r] + u * G.xm[r] + v * G.xm[r + 3] + w * G.xm[r + 6];
  } else {
    const double f0 = u / G.n1, f1 = v / G.n2, f2 = w / G.n3;
#pragma unroll
    for (int r = 0; r < 3; r++) xc[r] = G.x2c[r] * f0 + G.x2c[r + 3] * f1 + G.x2c[r + 6] * f2;
  }
}

__device__ __forceinline__ void node_cart(const Geo &G, int i, int j, int k, double xc[3]) {
  geo_cart(G, (double)i, (double)j, (double)k, xc);
}

// Middle eigenvalue of a symmetric 3x3 matrix (trigonometric closed form of
// the characteristic cubic).
template <class T>
__device__ inline T eig_mid_t(T h00, T h11, T h22, T h01, T h02, T h12) {
  const T p1 = h01 * h01 + h02 * h02 + h12 * h12;
  if (p1 == T(0)) {  // diagonal
    T a = h00, b = h11, c = h22, t;
    if (a > b) { t = a; a = b; b = t; }
    if (b > c) { t = b; b = c; c = t; }
    if (a > b) { t = a; a = b; b = t; }
    return b;
  }
  const T q = (h00 + h11 + h22) / T(3);
  const T d0 = h00 - q, d1 = h11 - q, d2 = h22 - q;
  const T p2 = d0 * d0 + d1 * d1 + d2 * d2 + T(2) * p1;
  const T p = sqrt(p2 / T(6));
  const T ip = T(1) / p;
  const T b00 = d0 * ip, b11 = d1 * ip, b22 = d2 * ip, b01 = h01 * ip, b02 = h02 * ip, b12 = h12 * ip;
  T detb = b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b12 * b02) + b02 * (b01 * b12 - b11 * b02);
  T rr = detb / T(2);
  rr = rr < T(-1) ? T(-1) : (rr > T(1) ? T(1) : rr);
  const T phi = acos(rr) / T(3);
  const T l1 = q + T(2) * p * cos(phi);
  const T l3 = q + T(2) * p * cos(phi + T(2.0943951023931954923));
  return T(3) * q - l1 - l3;
}

struct NciOut {
  void *crho, *cgrad;  // double* or float*
  int is_f32;
  int zfast;
  void *rhoat = nullptr;  // promolecular density (same type), when wanted
};

__device__ __forceinline__ size_t out_index(const Geo &G, int zfast, int i, int j, int k) {
  return zfast ? (size_t)k + (size_t)G.n3 * ((size_t)j + (size_t)G.n2 * i)
               : (size_t)i + (size_t)G.n1 * ((size_t)j + (size_t)G.n2 * k);
}

template <class T>
__device__ __forceinline__ void store_out(const NciOut &O, size_t idx, T crho, T cgrad) {
  if (O.is_f32) {
    ((float *)O.crho)[idx] = (float)crho;
    ((float *)O.cgrad)[idx] = (float)cgrad;
  } else {
    ((double *)O.crho)[idx] = (double)crho;
    ((double *)O.cgrad)[idx] = (double)cgrad;
  }
}

}  // namespace c2
