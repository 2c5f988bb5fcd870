// Basin multipoles (INTEGRABLE ... MULTIPOLE, integration@proc.f90:1302-1362,
// the Bader branch): per basin a, Q_lm(a) = sum over its points of
// R_lm(r) f(p), r the shortest lattice translate of p/n - xattr(a) in
// Cartesian coordinates (crystal%shortest) and R_lm the real regular solid
// harmonics of genrlm_real (tools_math@proc.f90:273-306), ordered
// (C00, C11, C10, S11, C22, C21, C20, S21, S22, ...): within each l,
// index l^2 + l - m holds C_lm and l^2 + l + m holds S_lm (m > 0).
//
// The harmonics are evaluated from Cartesian coordinates with the
// recurrences of the Racah-normalised regular solid harmonics
//   C11 = x, S11 = y,
//   C/S_{m+1,m+1} = sqrt((2m+1)/(2m+2)) (x C_mm - y S_mm, y C_mm + x S_mm),
//   R_{l+1,m} = ((2l+1) z R_lm - sqrt((l+m)(l-m)) r^2 R_{l-1,m})
//               / sqrt((l+m+1)(l-m+1)),
// the same functions the reference builds from spherical angles and
// spherical harmonics (tosphere, genylm); the sums agree to rounding.
//
// The YT branch (integration@proc.f90:1314-1336) weighs each point by
// yt_weights' w_a(p): 1 at the interior points of basin a (the labels of a
// c2gpu_yt call), the stored surface weight at its surface points (label 0).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "common.hpp"

namespace c2 {

int ws_facets(const double m[9], int vec[3 * 14], double area[14]);

namespace {

constexpr int MB = 256;
constexpr int LMAX_MAX = 10;
constexpr int NLM_MAX = (LMAX_MAX + 1) * (LMAX_MAX + 1);

struct MpGeo {
  int n1, n2, n3, lmax, nattr;
  double x2c[9];
  bool ortho;
};

// Real regular solid harmonics up to lmax at (x, y, z), genrlm_real order.
__device__ __forceinline__ void solid_harmonics(int lmax, double x, double y, double z, double *R) {
  const double r2 = x * x + y * y + z * z;
  R[0] = 1.0;
  if (lmax == 0) return;
  // m = 0 column
  double cm1 = 1.0, c0 = z;  // C_{l-1,0}, C_{l,0}
  R[2] = z;                  // l = 1: index 1 + 1
  for (int l = 1; l < lmax; l++) {
    const double c1 = ((2 * l + 1) * z * c0 - l * r2 * cm1) / (l + 1);
    R[(l + 1) * (l + 1) + (l + 1)] = c1;
    cm1 = c0;
    c0 = c1;
  }
  // m >= 1: diagonal, then up each column
  double cmm = x, smm = y;  // C_mm, S_mm at m = 1
  for (int m = 1; m <= lmax; m++) {
    if (m > 1) {
      const double f = sqrt((2.0 * m - 1.0) / (2.0 * m));
      const double cn = f * (x * cmm - y * smm), sn = f * (y * cmm + x * smm);
      cmm = cn;
      smm = sn;
    }
    R[m * m + m - m] = cmm;
    R[m * m + m + m] = smm;
    double cp = 0.0, sp = 0.0, cc = cmm, sc = smm;  // l-1 and l values
    for (int l = m; l < lmax; l++) {
      const double a = (2 * l + 1) * z;
      const double b = sqrt((double)(l + m) * (l - m)) * r2;
      const double d = sqrt((double)(l + m + 1) * (l - m + 1));
      const double cn = (a * cc - b * cp) / d, sn = (a * sc - b * sp) / d;
      const int base = (l + 1) * (l + 1) + (l + 1);
      R[base - m] = cn;
      R[base + m] = sn;
      cp = cc;
      sp = sc;
      cc = cn;
      sc = sn;
    }
  }
}

// The Cartesian shortest lattice translate of p/n - xa (crystal%shortest):
// the nint fold (half away from zero), then for non-orthogonal cells the
// first strictly shorter neighbouring translate.
__device__ __forceinline__ void shortest_offset(const MpGeo &G, int i, int j, int k, const double *xa,
                                                double v[3]) {
  double dv[3] = {(double)i / G.n1 - xa[0], (double)j / G.n2 - xa[1], (double)k / G.n3 - xa[2]};
  for (int c = 0; c < 3; c++) dv[c] = dv[c] - round(dv[c]);
  for (int c = 0; c < 3; c++) v[c] = G.x2c[c] * dv[0] + G.x2c[c + 3] * dv[1] + G.x2c[c + 6] * dv[2];
  if (!G.ortho) {
    double best = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    double bv[3] = {v[0], v[1], v[2]};
    for (int a = -1; a <= 1; a++)
      for (int b = -1; b <= 1; b++)
        for (int c = -1; c <= 1; c++) {
          if (!a && !b && !c) continue;
          double t[3];
          for (int e = 0; e < 3; e++) t[e] = v[e] + G.x2c[e] * a + G.x2c[e + 3] * b + G.x2c[e + 6] * c;
          const double d = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
          if (d < best) {
            best = d;
            bv[0] = t[0];
            bv[1] = t[1];
            bv[2] = t[2];
          }
        }
    v[0] = bv[0];
    v[1] = bv[1];
    v[2] = bv[2];
  }
}

// Bins of a workgroup: in LDS, or with gbins (more bins than the LDS
// holds) in the workgroup's own row of global memory.
__global__ void __launch_bounds__(MB) k_mpole(MpGeo G, uint64_t npts, const int32_t *__restrict__ lab,
                                               const double *__restrict__ f, const double *__restrict__ xattr,
                                               double *__restrict__ out, double *__restrict__ gbins) {
  extern __shared__ double lbins[];  // nattr * nlm
  const int nlm = (G.lmax + 1) * (G.lmax + 1);
  const int nb = G.nattr * nlm;
  double *bins = gbins ? gbins + (size_t)blockIdx.x * nb : lbins;
  for (int q = threadIdx.x; q < nb; q += MB) bins[q] = 0.0;
  __syncthreads();
  double R[NLM_MAX];
  for (uint64_t p = (uint64_t)blockIdx.x * MB + threadIdx.x; p < npts; p += (uint64_t)gridDim.x * MB) {
    const int32_t ix = lab[p];
    if (ix < 1 || ix > G.nattr) continue;
    const int i = (int)(p % (uint64_t)G.n1);
    const uint64_t r = p / (uint64_t)G.n1;
    const int j = (int)(r % (uint64_t)G.n2);
    const int k = (int)(r / (uint64_t)G.n2);
    double v[3];
    shortest_offset(G, i, j, k, &xattr[3 * (ix - 1)], v);
    solid_harmonics(G.lmax, v[0], v[1], v[2], R);
    const double fv = f[p];
    double *bb = bins + (size_t)(ix - 1) * nlm;
    for (int q = 0; q < nlm; q++) atomicAdd(&bb[q], R[q] * fv);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nb; q += MB)
    if (bins[q] != 0.0) atomicAdd(&out[q], bins[q]);
}

// The surface points of a YT partition (integration@proc.f90:1314-1336): each
// entry (basin b, weight w) of point p adds R_lm(shortest(p/n - xattr(b)))
// f(p) w to basin b; entries with |w| < 1d-15 are skipped as the reference
// skips them.  Binned per workgroup in LDS.
__global__ void __launch_bounds__(MB) k_mpole_yt_surface(MpGeo G, uint32_t nias, const uint32_t *__restrict__ ias,
                                                          const uint64_t *__restrict__ woff,
                                                          const uint16_t *__restrict__ wc,
                                                          const uint32_t *__restrict__ wb,
                                                          const double *__restrict__ ww, const double *__restrict__ f,
                                                          const double *__restrict__ xattr, double *__restrict__ out,
                                                          double *__restrict__ gbins) {
  extern __shared__ double lbins[];  // nattr * nlm
  const int nlm = (G.lmax + 1) * (G.lmax + 1);
  const int nb = G.nattr * nlm;
  double *bins = gbins ? gbins + (size_t)blockIdx.x * nb : lbins;
  for (int q = threadIdx.x; q < nb; q += MB) bins[q] = 0.0;
  __syncthreads();
  double R[NLM_MAX];
  for (uint32_t t = blockIdx.x * MB + threadIdx.x; t < nias; t += gridDim.x * MB) {
    const uint32_t p = ias[t];
    const int i = (int)(p % (uint32_t)G.n1);
    const uint32_t r = p / (uint32_t)G.n1;
    const int j = (int)(r % (uint32_t)G.n2);
    const int k = (int)(r / (uint32_t)G.n2);
    const double fv = f[p];
    const uint64_t so = woff[t];
    const int ce = wc[t];
    for (int e = 0; e < ce; e++) {
      const int a = (int)wb[so + e];
      const double w = ww[so + e];
      if (a < 1 || a > G.nattr || fabs(w) < 1e-15) continue;
      double v[3];
      shortest_offset(G, i, j, k, &xattr[3 * (a - 1)], v);
      solid_harmonics(G.lmax, v[0], v[1], v[2], R);
      double *bb = bins + (size_t)(a - 1) * nlm;
      const double fw = fv * w;
      for (int q = 0; q < nlm; q++) atomicAdd(&bb[q], R[q] * fw);
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nb; q += MB)
    if (bins[q] != 0.0) atomicAdd(&out[q], bins[q]);
}

}  // namespace

namespace {

// Multipoles of the labelled points (every point p with 1 <= label <= nattr
// adds to its basin), and with yt the surface points of the last c2gpu_yt
// call through their weights.
void mpole_impl(c2gpu_ctx *ctx, const char *who, bool yt, const int32_t *label, const double *f, const int n[3],
                const double x2c[9], int nattr, const double *xattr, int lmax, int flags, double *mpole_out) {
  const std::string w(who);
  if (!label || !f || !n || !x2c || !xattr || !mpole_out) fail(C2GPU_ERR_ARG, w + ": NULL argument");
  if (lmax < 0 || lmax > LMAX_MAX) fail(C2GPU_ERR_ARG, w + ": 0 <= lmax <= 10");
  if (nattr < 1) fail(C2GPU_ERR_ARG, w + ": nattr >= 1");
  for (int a = 0; a < 3; a++)
    if (n[a] < 1) fail(C2GPU_ERR_ARG, w + ": grid dimensions must be positive");
  if (yt) {
    const auto &Y = ctx->yt;
    if (!Y.valid) fail(C2GPU_ERR_ARG, w + ": no c2gpu_yt call on this context has left its weights");
    if (Y.n[0] != n[0] || Y.n[1] != n[1] || Y.n[2] != n[2] || Y.nattr != nattr)
      fail(C2GPU_ERR_ARG, w + ": the grid and nattr must be those of the last c2gpu_yt call");
  }
  const int nlm = (lmax + 1) * (lmax + 1);
  const size_t nbins = (size_t)nattr * nlm;
  // bins per workgroup in LDS, or in rows of global memory when they do
  // not fit (fewer workgroups then, the rows at most about 1 GiB)
  const bool rows = 8 * nbins > 120 * 1024;
  const size_t lds = rows ? 0 : 8 * nbins;
  const uint64_t npts = (uint64_t)n[0] * n[1] * n[2];
  hipStream_t st = ctx->stream;
  MpGeo G{};
  G.n1 = n[0];
  G.n2 = n[1];
  G.n3 = n[2];
  G.lmax = lmax;
  G.nattr = nattr;
  for (int q = 0; q < 9; q++) G.x2c[q] = x2c[q];
  // crystal%shortest takes the nint fold alone when the cell is orthogonal
  // (isortho: a Wigner-Seitz cell of at most 6 facets, each across a single
  // lattice vector, tools@proc.f90:842-849), and searches neighbouring
  // translates otherwise
  {
    int vec[42];
    double area[14];
    const int nf = ws_facets(x2c, vec, area);
    G.ortho = nf <= 6;
    for (int q = 0; q < nf && G.ortho; q++) {
      int ones = 0, zeros = 0;
      for (int c = 0; c < 3; c++) {
        ones += std::abs(vec[3 * q + c]) == 1;
        zeros += vec[3 * q + c] == 0;
      }
      G.ortho = ones == 1 && zeros == 2;
    }
  }
  const int32_t *d_lab = label;
  const double *d_f = f;
  if (flags & C2GPU_PTR_DEVICE) {
    require_device_ptr(label, "label");
    require_device_ptr(f, "f");
  } else {
    int32_t *dl = buf<int32_t>(ctx, "mp_lab", npts);
    double *df = buf<double>(ctx, "mp_f", npts);
    C2_HIP(hipMemcpyAsync(dl, label, 4 * npts, hipMemcpyHostToDevice, st));
    C2_HIP(hipMemcpyAsync(df, f, 8 * npts, hipMemcpyHostToDevice, st));
    d_lab = dl;
    d_f = df;
  }
  double *d_x = buf<double>(ctx, "mp_xattr", 3 * (size_t)nattr);
  C2_HIP(hipMemcpyAsync(d_x, xattr, 24 * (size_t)nattr, hipMemcpyHostToDevice, st));
  double *d_out = buf<double>(ctx, "mp_out", (size_t)nattr * nlm);
  C2_HIP(hipMemsetAsync(d_out, 0, 8 * (size_t)nattr * nlm, st));
  unsigned gp = (unsigned)persistent_threads(ctx->device, MB) / MB;
  double *gbins = nullptr;
  if (rows) {
    const uint64_t fit = ((uint64_t)1 << 30) / (8 * nbins);
    gp = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(gp, fit));
    gbins = buf<double>(ctx, "mp_rows", (size_t)gp * nbins);
  }
  const unsigned g = std::min<unsigned>(grid_for((int64_t)npts, MB), gp);
  C2_LAUNCH(ctx, k_mpole, dim3(g), dim3(MB), lds, G, npts, d_lab, d_f, d_x, d_out, gbins);
  add_kernel_bytes(ctx, "k_mpole", 12.0 * (double)npts);
  if (yt && ctx->yt.nias) {
    const auto &Y = ctx->yt;
    const unsigned gs = std::min<unsigned>(grid_for((int64_t)Y.nias, MB), gp);
    C2_LAUNCH(ctx, k_mpole_yt_surface, dim3(gs), dim3(MB), lds, G, Y.nias, Y.ias, Y.woff, Y.wc, Y.wb, Y.ww, d_f,
              d_x, d_out, gbins);
    add_kernel_bytes(ctx, "k_mpole_yt_surface", 22.0 * (double)Y.nias + 12.0 * (double)Y.nent);
  }
  copy_to_host(ctx, mpole_out, d_out, 8 * (size_t)nattr * nlm);
  C2_HIP(hipStreamSynchronize(st));
}

}  // namespace

void mpole_run(c2gpu_ctx *ctx, const int32_t *label, const double *f, const int n[3], const double x2c[9],
               int nattr, const double *xattr, int lmax, int flags, double *mpole_out) {
  mpole_impl(ctx, "c2gpu_multipoles", false, label, f, n, x2c, nattr, xattr, lmax, flags, mpole_out);
}

void yt_mpole_run(c2gpu_ctx *ctx, const int32_t *label, const double *f, const int n[3], const double x2c[9],
                  int nattr, const double *xattr, int lmax, int flags, double *mpole_out) {
  mpole_impl(ctx, "c2gpu_yt_multipoles", true, label, f, n, x2c, nattr, xattr, lmax, flags, mpole_out);
}

}  // namespace c2
