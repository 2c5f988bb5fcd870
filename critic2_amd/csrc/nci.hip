// NCI reduced density gradient on the nodes of a periodic grid
// (nci@proc.f90:525-605), the FFT derivative grids of grid3%fft
// (grid3mod@proc.f90:1757-1872), the promolecular radial-table density
// (crystalmod@env.f90:622-753, grid1mod@proc.f90:86-137), and the synthetic
// Slater-sum grid used as test and benchmark input.
//
// Atom-sum kernels work on bricks of 8x8x4 grid nodes (one 256-thread
// workgroup per brick).  A brick's neighbour list holds the periodic atom
// images whose cutoff sphere can reach the brick; lists are built in two
// deterministic passes (count, scan, fill in image order), so every point
// sums its atoms in a fixed order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <hipfft/hipfft.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "nci_dev.hpp"

namespace c2 {

// ---------------------------------------------------------------- brick lists

struct Images {
  const double4 *pos;  // x, y, z (bohr), w = species index (as double)
  const double *rcut;  // per species: cutoff radius (bohr)
  int nimg;
  double brad;         // brick radius (bohr)
};

__device__ __forceinline__ bool image_hits_brick(const Geo &G, const Images &I, int b, int m) {
  const int b1 = b % G.nb1, b2 = (b / G.nb1) % G.nb2, b3 = b / (G.nb1 * G.nb2);
  double c[3];
  geo_cart(G, b1 * BX + 0.5 * (BX - 1), b2 * BY + 0.5 * (BY - 1), b3 * BZ + 0.5 * (BZ - 1), c);
  const double4 p = I.pos[m];
  const double dx = p.x - c[0], dy = p.y - c[1], dz = p.z - c[2];
  const double rr = I.rcut[(int)p.w] + I.brad;
  return dx * dx + dy * dy + dz * dz <= rr * rr;
}

// one wavefront per brick; count (out == nullptr) or fill in image order
__global__ void __launch_bounds__(64) k_brick_list(Geo G, Images I, uint32_t *__restrict__ cnt,
                                                   const uint64_t *__restrict__ off,
                                                   uint32_t *__restrict__ out) {
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  uint32_t total = 0;
  for (int m0 = 0; m0 < I.nimg; m0 += 64) {
    const int m = m0 + lane;
    const bool hit = (m < I.nimg) && image_hits_brick(G, I, b, m);
    const uint64_t mask = __ballot(hit);
    if (out && hit) {
      const uint32_t before = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
      out[off[b] + total + before] = (uint32_t)m;
    }
    total += (uint32_t)__popcll(mask);
  }
  if (!out && lane == 0) cnt[b] = total;
}

// ---------------------------------------------------------------- synthetic grid

__global__ void __launch_bounds__(256) k_slater(Geo G, const double4 *__restrict__ img,
                                                const double2 *__restrict__ zp,  // (pref, zeta) per image
                                                double rimg, const uint64_t *__restrict__ off,
                                                const uint32_t *__restrict__ lst,
                                                double *__restrict__ rho) {
  const int b = blockIdx.x;
  const int b1 = b % G.nb1, b2 = (b / G.nb1) % G.nb2, b3 = b / (G.nb1 * G.nb2);
  const int t = threadIdx.x;
  const int i = b1 * BX + (t % BX), j = b2 * BY + ((t / BX) % BY), k = b3 * BZ + t / (BX * BY);
  if (i >= G.n1 || j >= G.n2 || k >= G.n3) return;
  double xc[3];
  node_cart(G, i, j, k, xc);
  double acc = 0.0;
  for (uint64_t q = off[b]; q < off[b + 1]; q++) {
    const uint32_t m = lst[q];
    const double4 p = img[m];
    const double dx = xc[0] - p.x, dy = xc[1] - p.y, dz = xc[2] - p.z;
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    if (d > rimg) continue;
    const double2 c = zp[m];
    acc += c.x * exp(-c.y * d);
  }
  rho[(size_t)i + (size_t)G.n1 * ((size_t)j + (size_t)G.n2 * k)] = acc;
}

// ---------------------------------------------------------------- promolecular

struct RadTabD {  // FP64 tables, exact grid1%interp
  const double *r, *f, *fp, *fpp;
  int off, ngrid;
  double a, b, rmax;
};

struct PromolD {
  const double *r, *f, *fp, *fpp;  // concatenated tables
  const int *toff, *tng;           // per species offset / ngrid
  const double *ta, *tb, *trmax, *tcut;
};

// grid1%interp (grid1mod@proc.f90:86-137), FP64, same operation order
__device__ inline void interp_d(const PromolD &T, int s, double r0, double &f, double &fp,
                                double &fpp) {
  f = 0.0;
  fp = 0.0;
  fpp = 0.0;
  const int o = T.toff[s], ng = T.tng[s];
  const double *R = T.r + o;
  if (r0 >= T.trmax[s]) return;
  int ir;
  double r;
  if (r0 <= R[0]) {
    ir = 1;
    r = R[0];
  } else {
    ir = 1 + (int)floor(log(r0 / T.ta[s]) / T.tb[s]);
    r = r0;
  }
  int base = ir < 2 ? 2 : ir;
  base = base > ng - 2 ? ng - 2 : base;
  double rr[4], dr1[4], x[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    rr[i] = R[base - 2 + i];
    dr1[i] = r - rr[i];
#pragma unroll
    for (int j = 0; j < 4; j++) x[i][j] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < i; j++) {
      x[i][j] = 1.0 / (rr[i] - rr[j]);
      x[j][i] = -x[i][j];
    }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    double prod = 1.0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (i == j) continue;
      prod = prod * dr1[j] * x[i][j];
    }
    const int ii = o + base - 2 + i;
    f = f + T.f[ii] * prod;
    fp = fp + T.fp[ii] * prod;
    fpp = fpp + T.fpp[ii] * prod;
  }
}

// FP32 tables: per species and interval base (1-based, 2..ngrid-2) the cubic
// of each 4-node Lagrange interpolant in u = r - r(base-1), 12 coefficients
// padded to 16 floats.
struct PromolF {
  const float4 *coef;       // [toff[s] + base] * 4 float4
  const int *toff, *tng;
  const float *ta_inv_b;    // 1/b
  const float *tlog_a;      // log(a)
  const float *tr0;         // r(1)
  const float *trmax;
  const float *tnode;       // node r(base-1) per [toff+base]
};

// FP64 promolecular NCI (nci@proc.f90:565-604 with crystalmod@env.f90:622-753)
__global__ void __launch_bounds__(256) k_promol_d(Geo G, PromolD T, const double4 *__restrict__ img,
                                                  const uint64_t *__restrict__ off,
                                                  const uint32_t *__restrict__ lst, NciOut O) {
  const int b = blockIdx.x;
  const int b1 = b % G.nb1, b2 = (b / G.nb1) % G.nb2, b3 = b / (G.nb1 * G.nb2);
  const int t = threadIdx.x;
  const int i = b1 * BX + (t % BX), j = b2 * BY + ((t / BX) % BY), k = b3 * BZ + t / (BX * BY);
  if (i >= G.n1 || j >= G.n2 || k >= G.n3) return;
  double xc[3];
  node_cart(G, i, j, k, xc);
  double f = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
  double h00 = 0.0, h11 = 0.0, h22 = 0.0, h01 = 0.0, h02 = 0.0, h12 = 0.0;
  bool isnuc = false;
  for (uint64_t q = off[b]; q < off[b + 1]; q++) {
    const double4 p = img[lst[q]];
    const int s = (int)p.w;
    const double x0 = xc[0] - p.x, x1 = xc[1] - p.y, x2 = xc[2] - p.z;
    double r = sqrt(x0 * x0 + x1 * x1 + x2 * x2);
    if (r <= 1e-5) isnuc = true;  // identify_atom(distmax=1d-5), fieldmod@proc.f90:1148
    if (r > T.tcut[s]) continue;
    const double r1 = T.r[T.toff[s]];
    r = r > r1 ? r : r1;
    r = r > 1e-14 ? r : 1e-14;
    double rho, rhop, rhopp;
    interp_d(T, s, r, rho, rhop, rhopp);
    rho = rho > 0.0 ? rho : 0.0;
    f = f + rho;
    const double rinv1 = 1.0 / r;
    const double rinv1rp = rinv1 * rhop;
    g0 = g0 + x0 * rinv1rp;
    g1 = g1 + x1 * rinv1rp;
    g2 = g2 + x2 * rinv1rp;
    const double rfac = (rhopp - rinv1rp) * rinv1 * rinv1;
    h00 = h00 + rinv1rp + rfac * x0 * x0;
    h11 = h11 + rinv1rp + rfac * x1 * x1;
    h22 = h22 + rinv1rp + rfac * x2 * x2;
    h01 = h01 + rfac * x1 * x0;
    h02 = h02 + rfac * x2 * x0;
    h12 = h12 + rfac * x2 * x1;
  }
  if (isnuc) g0 = g1 = g2 = 0.0;
  const double gmod = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
  const double cst = 2.0 * pow(3.0 * PI_D * PI_D, 1.0 / 3.0);
  const double den = f > 1e-80 ? f : 1e-80;
  const double s = gmod / (cst * pow(den, 4.0 / 3.0));
  const double l2 = eig_mid_t<double>(h00, h11, h22, h01, h02, h12);
  const double a = fabs(f) * 100.0;
  const double crho = (l2 < 0.0 || (l2 == 0.0 && signbit(l2))) ? -a : a;
  const size_t o = out_index(G, O.zfast, i, j, k);
  store_out<double>(O, o, crho, s);
  if (O.rhoat) ((double *)O.rhoat)[o] = f;
}

// FP32 promolecular NCI at node (i,j,k) from the image list lst[qb..qe)
// (the atom sum of crystalmod@env.f90:622-753 and nci@proc.f90:565-599),
// with per-interval cubic tables.  The node-atom difference is taken in
// FP64 and then rounded: node and image coordinates reach tens of bohr,
// where an FP32 coordinate is off by up to 4e-6 bohr, a relative density
// error of about zeta * 4e-6 for an exponential tail; the rounded
// difference is off by 6e-8 of itself.
__device__ __forceinline__ void promol_f_node(const Geo &G, const PromolF &T, const double4 *__restrict__ img,
                                              const float *__restrict__ cut2, const uint32_t *__restrict__ lst,
                                              uint64_t qb, uint64_t qe, int i, int j, int k, const NciOut &O) {
  double xcd[3];
  node_cart(G, i, j, k, xcd);
  float f = 0.f, g0 = 0.f, g1 = 0.f, g2 = 0.f;
  float h00 = 0.f, h11 = 0.f, h22 = 0.f, h01 = 0.f, h02 = 0.f, h12 = 0.f;
  bool isnuc = false;
  for (uint64_t q = qb; q < qe; q++) {
    const double4 p = img[lst[q]];
    const int s = (int)p.w;
    const float x0 = (float)(xcd[0] - p.x), x1 = (float)(xcd[1] - p.y), x2 = (float)(xcd[2] - p.z);
    const float rsq = x0 * x0 + x1 * x1 + x2 * x2;
    if (rsq <= 1e-10f) isnuc = true;
    if (rsq > cut2[s]) continue;
    float r = sqrtf(rsq);
    if (r >= T.trmax[s]) continue;
    const float r1 = T.tr0[s];
    int ir;
    if (r <= r1) {
      ir = 1;
      r = r1;
    } else {
      ir = 1 + (int)floorf((__logf(r) - T.tlog_a[s]) * T.ta_inv_b[s]);
    }
    const int ng = T.tng[s];
    int base = ir < 2 ? 2 : ir;
    base = base > ng - 2 ? ng - 2 : base;
    const int ci = T.toff[s] + base;
    const float u = r - T.tnode[ci];
    const float4 cf = T.coef[4 * (size_t)ci + 0];  // f: c0..c3
    const float4 cp = T.coef[4 * (size_t)ci + 1];  // fp
    const float4 cq = T.coef[4 * (size_t)ci + 2];  // fpp
    float rho = ((cf.w * u + cf.z) * u + cf.y) * u + cf.x;
    const float rhop = ((cp.w * u + cp.z) * u + cp.y) * u + cp.x;
    const float rhopp = ((cq.w * u + cq.z) * u + cq.y) * u + cq.x;
    rho = rho > 0.f ? rho : 0.f;
    f += rho;
    const float rinv1 = 1.0f / r;
    const float rinv1rp = rinv1 * rhop;
    g0 += x0 * rinv1rp;
    g1 += x1 * rinv1rp;
    g2 += x2 * rinv1rp;
    const float rfac = (rhopp - rinv1rp) * rinv1 * rinv1;
    h00 += rinv1rp + rfac * x0 * x0;
    h11 += rinv1rp + rfac * x1 * x1;
    h22 += rinv1rp + rfac * x2 * x2;
    h01 += rfac * x1 * x0;
    h02 += rfac * x2 * x0;
    h12 += rfac * x2 * x1;
  }
  if (isnuc) g0 = g1 = g2 = 0.f;
  const float gmod = sqrtf(g0 * g0 + g1 * g1 + g2 * g2);
  // 2 (3 pi^2)^(1/3); max(rho, 1e-30f): 1d-80 is not representable in FP32
  const float cst = 6.18733545256027186f;
  const float den = f > 1e-30f ? f : 1e-30f;
  const float s = gmod / (cst * den * cbrtf(den));
  const float l2 = eig_mid_t<float>(h00, h11, h22, h01, h02, h12);
  const float a = fabsf(f) * 100.f;
  const float crho = (l2 < 0.f || (l2 == 0.f && signbit(l2))) ? -a : a;
  const size_t o = out_index(G, O.zfast, i, j, k);
  store_out<float>(O, o, crho, s);
  if (O.rhoat) ((float *)O.rhoat)[o] = f;
}

__global__ void __launch_bounds__(256) k_promol_f(Geo G, PromolF T, const double4 *__restrict__ img,
                                                  const float *__restrict__ cut2,
                                                  const uint64_t *__restrict__ off,
                                                  const uint32_t *__restrict__ lst, NciOut O) {
  const int b = blockIdx.x;
  const int b1 = b % G.nb1, b2 = (b / G.nb1) % G.nb2, b3 = b / (G.nb1 * G.nb2);
  const int t = threadIdx.x;
  const int i = b1 * BX + (t % BX), j = b2 * BY + ((t / BX) % BY), k = b3 * BZ + t / (BX * BY);
  if (i >= G.n1 || j >= G.n2 || k >= G.n3) return;
  promol_f_node(G, T, img, cut2, lst, off[b], off[b + 1], i, j, k, O);
}

// ---------------------------------------------------------------- FFT scheme

struct Recip {
  double bvec[9];  // bvec(:,l) column-major as grid3mod@proc.f90:1786-1789
  int n1, n2, n3;
  double invN;
};

__device__ __forceinline__ void vgc_of(const Recip &R, size_t l, double g[3]) {
  const int i = (int)(l % R.n1);
  const size_t rr = l / R.n1;
  const int j = (int)(rr % R.n2);
  const int k = (int)(rr / R.n2);
  // index range n/2-n+1 .. n/2 (Nyquist positive), grid3mod@proc.f90:1791-1797
  const int i1 = i <= R.n1 / 2 ? i : i - R.n1;
  const int i2 = j <= R.n2 / 2 ? j : j - R.n2;
  const int i3 = k <= R.n3 / 2 ? k : k - R.n3;
#pragma unroll
  for (int c = 0; c < 3; c++)
    g[c] = (double)i1 * R.bvec[c] + (double)i2 * R.bvec[c + 3] + (double)i3 * R.bvec[c + 6];
}

__global__ void k_real_to_complex(const double *__restrict__ f, double2 *__restrict__ z, size_t N) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N) z[p] = make_double2(f[p], 0.0);
}

// kind: 0,1,2 -> i*vgc_l (gradient component); 10+l -> -vgc_l^2; 20 -> -|vgc|^2
__global__ void k_spectral(const Recip R, const double2 *__restrict__ spec, double2 *__restrict__ w,
                           size_t N, int kind) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  double g[3];
  vgc_of(R, p, g);
  const double2 z = make_double2(spec[p].x * R.invN, spec[p].y * R.invN);
  double2 o;
  if (kind < 3) {
    const double v = g[kind];  // vgc * cmplx(-aimag(z), real(z))
    o = make_double2(v * -z.y, v * z.x);
  } else if (kind < 20) {
    const int c = kind - 10;
    const double v = -g[c] * g[c];
    o = make_double2(v * z.x, v * z.y);
  } else {
    const double v = -(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    o = make_double2(v * z.x, v * z.y);
  }
  w[p] = o;
}

__global__ void k_accum_sq(const double2 *__restrict__ w, double *__restrict__ acc, size_t N, int first) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const double x = w[p].x;
  acc[p] = first ? x * x : acc[p] + x * x;
}
__global__ void k_count_pos(const double2 *__restrict__ w, int8_t *__restrict__ npos, size_t N, int first) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int8_t v = w[p].x > 0.0 ? 1 : 0;
  npos[p] = first ? v : (int8_t)(npos[p] + v);
}
__global__ void k_real_part(const double2 *__restrict__ w, double *__restrict__ out, size_t N) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N) out[p] = w[p].x;
}
__global__ void k_sqrt(double *__restrict__ a, size_t N) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N) a[p] = sqrt(a[p]);
}

// nci@proc.f90:546-564 and :598-599 from node values
__global__ void k_nci_fft_point(Geo G, const double *__restrict__ rho, const double *__restrict__ gsq,
                                const int8_t *__restrict__ npos, NciOut O) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t N = (size_t)G.n1 * G.n2 * G.n3;
  if (p >= N) return;
  const int i = (int)(p % G.n1);
  const size_t rr = p / G.n1;
  const int j = (int)(rr % G.n2), k = (int)(rr / G.n2);
  const double f = rho[p];
  const double gmod = sqrt(gsq[p]);
  const double cst = 2.0 * pow(3.0 * PI_D * PI_D, 1.0 / 3.0);
  const double den = f > 1e-80 ? f : 1e-80;
  const double s = gmod / (cst * pow(den, 4.0 / 3.0));
  const double a = fabs(f) * 100.0;
  const double crho = (npos[p] >= 2) ? a : -a;
  store_out<double>(O, out_index(G, O.zfast, i, j, k), crho, s);
}

// ---------------------------------------------------------------- host side

// grid-interpolation scheme (ncigrid.hip, built without FP contraction)
void nci_grid_launch(c2gpu_ctx *ctx, const Geo &G, const double x2c[9], const double c2x[9],
                     const double *d_rho, const NciOut &O);

static Geo make_geo(const int n[3], const double x2c[9]) {
  Geo G{};
  G.n1 = n[0];
  G.n2 = n[1];
  G.n3 = n[2];
  G.nb1 = (n[0] + BX - 1) / BX;
  G.nb2 = (n[1] + BY - 1) / BY;
  G.nb3 = (n[2] + BZ - 1) / BZ;
  for (int q = 0; q < 9; q++) G.x2c[q] = x2c[q];
  return G;
}

// The nodes x0 + i xmat(:,1) + j xmat(:,2) + k xmat(:,3) of an nciplot cube.
static Geo make_geo_box(const c2gpu_nci_nodes &nd, const double x2c[9]) {
  Geo G = make_geo(nd.nstep, x2c);
  G.box = 1;
  for (int r = 0; r < 3; r++) G.x0[r] = nd.x0[r];
  for (int q = 0; q < 9; q++) G.xm[q] = nd.xmat[q];
  return G;
}

static double brick_radius(const Geo &G) {
  double h[3][3];
  const double e[3] = {0.5 * (BX - 1) / G.n1, 0.5 * (BY - 1) / G.n2, 0.5 * (BZ - 1) / G.n3};
  for (int a = 0; a < 3; a++)
    for (int r = 0; r < 3; r++) h[a][r] = G.box ? G.xm[r + 3 * a] * (0.5 * ((a == 0 ? BX : a == 1 ? BY : BZ) - 1))
                                                : G.x2c[r + 3 * a] * e[a];
  double best = 0;
  for (int s = 0; s < 8; s++) {
    double v[3];
    for (int r = 0; r < 3; r++)
      v[r] = ((s & 1) ? 1 : -1) * h[0][r] + ((s & 2) ? 1 : -1) * h[1][r] + ((s & 4) ? 1 : -1) * h[2][r];
    const double d = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    best = d > best ? d : best;
  }
  return best;
}

// All periodic images of the atoms that can lie within rc_max of a point of
// the cell: fractional coordinates within rc_max*|b_c| of [0,1] per axis.
// With lo/hi: the images that can reach a point whose fractional
// coordinates lie in [lo, hi] instead of the cell.
static void expand_images(const double x2c[9], int nat, const double *xat, const int *spc,
                          double rc_max, std::vector<double> &pos, std::vector<int> &isp,
                          std::vector<int> &iat, const double *lo = nullptr, const double *hi = nullptr) {
  // reciprocal vectors (rows of inverse of x2c)
  const double *a = x2c;
  double inv[9];
  const double det = a[0] * (a[4] * a[8] - a[7] * a[5]) - a[3] * (a[1] * a[8] - a[7] * a[2]) +
                     a[6] * (a[1] * a[5] - a[4] * a[2]);
  inv[0] = (a[4] * a[8] - a[7] * a[5]) / det;
  inv[3] = -(a[3] * a[8] - a[6] * a[5]) / det;
  inv[6] = (a[3] * a[7] - a[6] * a[4]) / det;
  inv[1] = -(a[1] * a[8] - a[7] * a[2]) / det;
  inv[4] = (a[0] * a[8] - a[6] * a[2]) / det;
  inv[7] = -(a[0] * a[7] - a[6] * a[1]) / det;
  inv[2] = (a[1] * a[5] - a[4] * a[2]) / det;
  inv[5] = -(a[0] * a[5] - a[3] * a[2]) / det;
  inv[8] = (a[0] * a[4] - a[3] * a[1]) / det;
  double delta[3], flo[3], fhi[3];
  int tlo[3], thi[3];
  for (int c = 0; c < 3; c++) {  // row c of inv = reciprocal vector b_c
    const double bl = std::sqrt(inv[c] * inv[c] + inv[c + 3] * inv[c + 3] + inv[c + 6] * inv[c + 6]);
    delta[c] = rc_max * bl;
    flo[c] = lo ? lo[c] : 0.0;
    fhi[c] = hi ? hi[c] : 1.0;
    tlo[c] = (int)std::floor(flo[c] - delta[c]) - 1;
    thi[c] = (int)std::ceil(fhi[c] + delta[c]) + 1;
  }
  for (int at = 0; at < nat; at++)
    for (int t0 = tlo[0]; t0 <= thi[0]; t0++)
      for (int t1 = tlo[1]; t1 <= thi[1]; t1++)
        for (int t2 = tlo[2]; t2 <= thi[2]; t2++) {
          double xf[3] = {xat[3 * at] - std::floor(xat[3 * at]) + t0,
                          xat[3 * at + 1] - std::floor(xat[3 * at + 1]) + t1,
                          xat[3 * at + 2] - std::floor(xat[3 * at + 2]) + t2};
          bool ok = true;
          for (int c = 0; c < 3; c++)
            if (xf[c] < flo[c] - delta[c] || xf[c] > fhi[c] + delta[c]) ok = false;
          if (!ok) continue;
          for (int r = 0; r < 3; r++)
            pos.push_back(x2c[r] * xf[0] + x2c[r + 3] * xf[1] + x2c[r + 6] * xf[2]);
          isp.push_back(spc[at]);
          iat.push_back(at);
        }
}

struct BrickLists {
  uint64_t *off;
  uint32_t *lst;
  uint64_t total;
};

static BrickLists build_bricks(c2gpu_ctx *ctx, const Geo &G, const double4 *d_img, const double *d_rcut,
                               int nimg, const std::string &tag) {
  const int nbk = G.nb1 * G.nb2 * G.nb3;
  Images I{d_img, d_rcut, nimg, brick_radius(G)};
  uint32_t *cnt = buf<uint32_t>(ctx, tag + "_bcnt", (size_t)nbk + 1);
  uint64_t *off = buf<uint64_t>(ctx, tag + "_boff", (size_t)nbk + 1);
  C2_HIP(hipMemsetAsync(cnt + nbk, 0, sizeof(uint32_t), ctx->stream));
  C2_LAUNCH(ctx, k_brick_list, dim3(nbk), dim3(64), 0, G, I, cnt, nullptr, nullptr);
  size_t tmp = 0;
  C2_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, cnt, off, nbk + 1, ctx->stream));
  void *t = buffer(ctx, tag + "_bscan", tmp);
  C2_HIP(hipcub::DeviceScan::ExclusiveSum(t, tmp, cnt, off, nbk + 1, ctx->stream));
  uint64_t total = 0;
  copy_to_host(ctx, &total, off + nbk, 8);
  C2_HIP(hipStreamSynchronize(ctx->stream));
  uint32_t *lst = buf<uint32_t>(ctx, tag + "_blst", total > 0 ? total : 1);
  C2_LAUNCH(ctx, k_brick_list, dim3(nbk), dim3(64), 0, G, I, cnt, off, lst);
  return BrickLists{off, lst, total};
}

void synth_slater_run(c2gpu_ctx *ctx, const double x2c[9], const int n[3], int nat, const double *xat,
                      const int *zat, double rimg, int flags, double *rho_out) {
  if (nat < 1) fail(C2GPU_ERR_ARG, "c2gpu_synth_slater: no atoms");
  for (int c = 0; c < 3; c++)
    if (n[c] < 1) fail(C2GPU_ERR_ARG, "c2gpu_synth_slater: bad grid");
  const Geo G = make_geo(n, x2c);
  std::vector<int> spc(nat), isp, iat;
  for (int a = 0; a < nat; a++) spc[a] = 0;
  std::vector<double> pos;
  expand_images(x2c, nat, xat, spc.data(), rimg, pos, isp, iat);
  const int nimg = (int)isp.size();
  std::vector<double4> h_img(nimg);
  std::vector<double2> h_zp(nimg);
  for (int m = 0; m < nimg; m++) {
    h_img[m] = make_double4(pos[3 * m], pos[3 * m + 1], pos[3 * m + 2], 0.0);
    const double z = (double)zat[iat[m]];
    const double ze = 2.0 * std::cbrt(z);
    h_zp[m] = make_double2(z * ze * ze * ze / (8.0 * PI_D), ze);
  }
  double4 *d_img = buf<double4>(ctx, "syn_img", nimg);
  double2 *d_zp = buf<double2>(ctx, "syn_zp", nimg);
  double *d_rc = buf<double>(ctx, "syn_rc", 1);
  C2_HIP(hipMemcpyAsync(d_img, h_img.data(), sizeof(double4) * nimg, hipMemcpyHostToDevice, ctx->stream));
  C2_HIP(hipMemcpyAsync(d_zp, h_zp.data(), sizeof(double2) * nimg, hipMemcpyHostToDevice, ctx->stream));
  C2_HIP(hipMemcpyAsync(d_rc, &rimg, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  BrickLists BL = build_bricks(ctx, G, d_img, d_rc, nimg, "syn");
  const size_t N = (size_t)n[0] * n[1] * n[2];
  double *d_out = rho_out;
  if (!(flags & C2GPU_PTR_DEVICE)) d_out = buf<double>(ctx, "syn_out", N);
  C2_LAUNCH(ctx, k_slater, dim3(G.nb1 * G.nb2 * G.nb3), dim3(256), 0, G, d_img, d_zp,
                     rimg, BL.off, BL.lst, d_out);
  if (!(flags & C2GPU_PTR_DEVICE))
    C2_HIP(hipMemcpyAsync(rho_out, d_out, 8 * N, hipMemcpyDeviceToHost, ctx->stream));
  C2_HIP(hipStreamSynchronize(ctx->stream));
}

// ---- promolecular tables

struct PromolDev {
  PromolD D;
  PromolF F;
  double4 *img;
  double *rcut;
  float *cut2f;
  int nimg;
};

// The radial tables on the device (FP64 nodes and values, FP32 cubic
// coefficients); returns the largest cutoff min(cutrad, rmax).
static double upload_tables(c2gpu_ctx *ctx, int ns, const c2gpu_radial_table *tables, PromolDev &P) {
  if (ns < 1 || !tables) fail(C2GPU_ERR_ARG, "c2gpu_nci: radial tables required");
  std::vector<int> toff(ns), tng(ns);
  int tot = 0;
  double rcmax = 0;
  std::vector<double> ta(ns), tb(ns), trmax(ns), tcut(ns);
  for (int s = 0; s < ns; s++) {
    const c2gpu_radial_table &t = tables[s];
    if (t.ngrid < 6 || !t.r || !t.f || !t.fp || !t.fpp) fail(C2GPU_ERR_ARG, "c2gpu_nci: bad radial table");
    toff[s] = tot;
    tng[s] = t.ngrid;
    tot += t.ngrid;
    ta[s] = t.a;
    tb[s] = t.b;
    trmax[s] = t.rmax;
    // list_near_atoms up2dsp = min(cutrad, rmax), crystalmod@env.f90:676-690
    tcut[s] = t.cutrad < t.rmax ? t.cutrad : t.rmax;
    rcmax = tcut[s] > rcmax ? tcut[s] : rcmax;
  }
  std::vector<double> r(tot), f(tot), fp(tot), fpp(tot);
  for (int s = 0; s < ns; s++) {
    const c2gpu_radial_table &t = tables[s];
    std::memcpy(&r[toff[s]], t.r, 8 * t.ngrid);
    std::memcpy(&f[toff[s]], t.f, 8 * t.ngrid);
    std::memcpy(&fp[toff[s]], t.fp, 8 * t.ngrid);
    std::memcpy(&fpp[toff[s]], t.fpp, 8 * t.ngrid);
  }
  // FP32 cubic coefficients per (species, base): expand the Lagrange basis
  // l_i(r) = prod_{j != i} (r - r_j)/(r_i - r_j) in u = r - r_{base-1}.
  std::vector<float> coef((size_t)tot * 16, 0.f), node(tot, 0.f);
  for (int s = 0; s < ns; s++) {
    const int ng = tng[s];
    for (int base = 2; base <= ng - 2; base++) {
      const int o = toff[s];
      double rr[4];
      for (int i = 0; i < 4; i++) rr[i] = r[o + base - 2 + i];
      const double u0 = rr[1];
      double pc[3][4] = {{0}};
      for (int i = 0; i < 4; i++) {
        // polynomial prod_{j != i} (u - (r_j - u0)) / (r_i - r_j)
        double poly[4] = {1, 0, 0, 0};
        double den = 1;
        for (int j = 0; j < 4; j++) {
          if (j == i) continue;
          const double root = rr[j] - u0;
          double np[4] = {0, 0, 0, 0};
          for (int d = 0; d < 3; d++) {
            np[d + 1] += poly[d];
            np[d] -= root * poly[d];
          }
          for (int d = 0; d < 4; d++) poly[d] = np[d];
          den *= (rr[i] - rr[j]);
        }
        const int ii = o + base - 2 + i;
        for (int d = 0; d < 4; d++) {
          pc[0][d] += f[ii] * poly[d] / den;
          pc[1][d] += fp[ii] * poly[d] / den;
          pc[2][d] += fpp[ii] * poly[d] / den;
        }
      }
      for (int tbl = 0; tbl < 3; tbl++)
        for (int d = 0; d < 4; d++) coef[(size_t)(o + base) * 16 + 4 * tbl + d] = (float)pc[tbl][d];
      node[o + base] = (float)u0;
    }
  }
  std::vector<float> cut2(ns), tainvb(ns), tloga(ns), tr0(ns), trmaxf(ns);
  for (int s = 0; s < ns; s++) {
    cut2[s] = (float)(tcut[s] * tcut[s]);
    tainvb[s] = (float)(1.0 / tb[s]);
    tloga[s] = (float)std::log(ta[s]);
    tr0[s] = (float)r[toff[s]];
    trmaxf[s] = (float)trmax[s];
  }
  auto up = [&](const std::string &name, const void *src, size_t bytes) {
    void *d = buffer(ctx, name, bytes);
    C2_HIP(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return d;
  };
  P.D.r = (const double *)up("pm_r", r.data(), 8 * tot);
  P.D.f = (const double *)up("pm_f", f.data(), 8 * tot);
  P.D.fp = (const double *)up("pm_fp", fp.data(), 8 * tot);
  P.D.fpp = (const double *)up("pm_fpp", fpp.data(), 8 * tot);
  P.D.toff = (const int *)up("pm_toff", toff.data(), 4 * ns);
  P.D.tng = (const int *)up("pm_tng", tng.data(), 4 * ns);
  P.D.ta = (const double *)up("pm_ta", ta.data(), 8 * ns);
  P.D.tb = (const double *)up("pm_tb", tb.data(), 8 * ns);
  P.D.trmax = (const double *)up("pm_trmax", trmax.data(), 8 * ns);
  P.D.tcut = (const double *)up("pm_tcut", tcut.data(), 8 * ns);
  P.F.coef = (const float4 *)up("pm_coef", coef.data(), 4 * coef.size());
  P.F.toff = P.D.toff;
  P.F.tng = P.D.tng;
  P.F.ta_inv_b = (const float *)up("pm_tainvb", tainvb.data(), 4 * ns);
  P.F.tlog_a = (const float *)up("pm_tloga", tloga.data(), 4 * ns);
  P.F.tr0 = (const float *)up("pm_tr0", tr0.data(), 4 * ns);
  P.F.trmax = (const float *)up("pm_trmaxf", trmaxf.data(), 4 * ns);
  P.F.tnode = (const float *)up("pm_node", node.data(), 4 * node.size());
  P.rcut = (double *)up("pm_rcut", tcut.data(), 8 * ns);
  P.cut2f = (float *)up("pm_cut2", cut2.data(), 4 * ns);
  C2_HIP(hipStreamSynchronize(ctx->stream));  // the host vectors end here
  return rcmax;
}

// Tables and the atom images of the cell that can reach the nodes (the cell,
// or the fractional range [lo, hi] of an nciplot cube).
static PromolDev upload_promol(c2gpu_ctx *ctx, const c2gpu_promol *pm, const double x2c[9],
                               const double *lo = nullptr, const double *hi = nullptr) {
  if (!pm || pm->nspecies < 1 || !pm->tables || pm->nat < 1 || !pm->xat || !pm->species)
    fail(C2GPU_ERR_ARG, "c2gpu_nci: promolecular tables required");
  for (int a = 0; a < pm->nat; a++)
    if (pm->species[a] < 0 || pm->species[a] >= pm->nspecies) fail(C2GPU_ERR_ARG, "c2gpu_nci: bad species index");
  PromolDev P;
  const double rcmax = upload_tables(ctx, pm->nspecies, pm->tables, P);
  std::vector<double> pos;
  std::vector<int> isp, iat;
  expand_images(x2c, pm->nat, pm->xat, pm->species, rcmax, pos, isp, iat, lo, hi);
  const int nimg = (int)isp.size();
  std::vector<double4> h_img(nimg > 0 ? nimg : 1);
  for (int m = 0; m < nimg; m++)
    h_img[m] = make_double4(pos[3 * m], pos[3 * m + 1], pos[3 * m + 2], (double)isp[m]);
  P.img = buf<double4>(ctx, "pm_img", h_img.size());
  C2_HIP(hipMemcpyAsync(P.img, h_img.data(), sizeof(double4) * h_img.size(), hipMemcpyHostToDevice, ctx->stream));
  C2_HIP(hipStreamSynchronize(ctx->stream));
  P.nimg = nimg;
  return P;
}

#define HIPFFT_CHECK(call)                                                     \
  do {                                                                         \
    hipfftResult r_ = (call);                                                  \
    if (r_ != HIPFFT_SUCCESS) fail(C2GPU_ERR_HIP, std::string(#call) + " failed"); \
  } while (0)

struct FftPlan {
  hipfftHandle h;
  explicit FftPlan(c2gpu_ctx *ctx, const int n[3]) {
    HIPFFT_CHECK(hipfftPlan3d(&h, n[2], n[1], n[0], HIPFFT_Z2Z));
    HIPFFT_CHECK(hipfftSetStream(h, ctx->stream));
  }
  ~FftPlan() { (void)hipfftDestroy(h); }
};

static Recip make_recip(const int n[3], const double x2c[9]) {
  // bvec(:,1) = cross(x2c(:,3), x2c(:,2)) etc. scaled by 2 pi/|det|
  auto cross = [](const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  Recip R;
  cross(x2c + 6, x2c + 3, R.bvec + 0);
  cross(x2c + 0, x2c + 6, R.bvec + 3);
  cross(x2c + 3, x2c + 0, R.bvec + 6);
  const double *m = x2c;  // det3 (tools_math@proc.f90 det3 expansion)
  const double det = m[0] * (m[4] * m[8] - m[7] * m[5]) + m[3] * (m[7] * m[2] - m[1] * m[8]) +
                     m[6] * (m[1] * m[5] - m[4] * m[2]);
  const double fac = 2.0 * PI_D / std::fabs(det);
  for (int q = 0; q < 9; q++) R.bvec[q] = fac * R.bvec[q];
  R.n1 = n[0];
  R.n2 = n[1];
  R.n3 = n[2];
  R.invN = 1.0 / ((double)n[0] * n[1] * n[2]);
  return R;
}

static const double *to_device_grid(c2gpu_ctx *ctx, const double *rho, size_t N, int flags,
                                    const char *name) {
  if (flags & C2GPU_PTR_DEVICE) return rho;
  double *d = buf<double>(ctx, name, N);
  C2_HIP(hipMemcpyAsync(d, rho, 8 * N, hipMemcpyHostToDevice, ctx->stream));
  return d;
}

void fft_deriv_run(c2gpu_ctx *ctx, const double *rho, const int n[3], const double x2c[9], int kind,
                   int flags, double *out) {
  for (int c = 0; c < 3; c++)
    if (n[c] < 1) fail(C2GPU_ERR_ARG, "c2gpu_fft_deriv: bad grid");
  if (kind < 0 || kind > 4) fail(C2GPU_ERR_ARG, "c2gpu_fft_deriv: bad kind");
  const size_t N = (size_t)n[0] * n[1] * n[2];
  hipStream_t st = ctx->stream;
  const double *d_rho = to_device_grid(ctx, rho, N, flags, "fft_rho");
  double2 *spec = buf<double2>(ctx, "fft_spec", N), *w = buf<double2>(ctx, "fft_work", N);
  double *d_out = (flags & C2GPU_PTR_DEVICE) ? out : buf<double>(ctx, "fft_out", N);
  FftPlan plan(ctx, n);
  const unsigned g = grid_for((int64_t)N, 256);
  const Recip R = make_recip(n, x2c);
  C2_LAUNCH(ctx, k_real_to_complex, dim3(g), dim3(256), 0, d_rho, spec, N);
  HIPFFT_CHECK(hipfftExecZ2Z(plan.h, (hipfftDoubleComplex *)spec, (hipfftDoubleComplex *)spec, HIPFFT_FORWARD));
  if (kind == C2GPU_FT_GRAD) {
    for (int c = 0; c < 3; c++) {
      C2_LAUNCH(ctx, k_spectral, dim3(g), dim3(256), 0, R, spec, w, N, c);
      HIPFFT_CHECK(hipfftExecZ2Z(plan.h, (hipfftDoubleComplex *)w, (hipfftDoubleComplex *)w, HIPFFT_BACKWARD));
      C2_LAUNCH(ctx, k_accum_sq, dim3(g), dim3(256), 0, w, d_out, N, c == 0);
    }
    C2_LAUNCH(ctx, k_sqrt, dim3(g), dim3(256), 0, d_out, N);
  } else {
    const int sk = (kind == C2GPU_FT_LAP) ? 20 : 10 + (kind - 1);
    C2_LAUNCH(ctx, k_spectral, dim3(g), dim3(256), 0, R, spec, w, N, sk);
    HIPFFT_CHECK(hipfftExecZ2Z(plan.h, (hipfftDoubleComplex *)w, (hipfftDoubleComplex *)w, HIPFFT_BACKWARD));
    C2_LAUNCH(ctx, k_real_part, dim3(g), dim3(256), 0, w, d_out, N);
  }
  if (!(flags & C2GPU_PTR_DEVICE)) C2_HIP(hipMemcpyAsync(out, d_out, 8 * N, hipMemcpyDeviceToHost, st));
  C2_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------- batches of cells

// A cell of a batch on the device: its geometry, its first brick in the
// batch's brick numbering, its images in the batch image array, and where its
// NCI outputs go.
struct BatchCell {
  Geo G;
  uint32_t brick0;
  uint32_t img0, nimg;
  double brad;
  NciOut O;
};

__device__ __forceinline__ int cell_of_brick(const BatchCell *__restrict__ cells, int ncell, uint32_t gb) {
  int lo = 0, hi = ncell - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cells[mid].brick0 <= gb)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// k_brick_list over the bricks of every cell of a batch (one wavefront per
// brick); list entries are batch image indices.
__global__ void __launch_bounds__(64) k_brick_list_batch(const BatchCell *__restrict__ cells, int ncell,
                                                         const double4 *__restrict__ img,
                                                         const double *__restrict__ rcut,
                                                         uint32_t *__restrict__ cnt, const uint64_t *__restrict__ off,
                                                         uint32_t *__restrict__ out) {
  const uint32_t gb = blockIdx.x;
  const BatchCell &C = cells[cell_of_brick(cells, ncell, gb)];
  const Images I{img + C.img0, rcut, (int)C.nimg, C.brad};
  const int b = (int)(gb - C.brick0);
  const int lane = threadIdx.x;
  uint32_t total = 0;
  for (int m0 = 0; m0 < I.nimg; m0 += 64) {
    const int m = m0 + lane;
    const bool hit = (m < I.nimg) && image_hits_brick(C.G, I, b, m);
    const uint64_t mask = __ballot(hit);
    if (out && hit)
      out[off[gb] + total + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = C.img0 + (uint32_t)m;
    total += (uint32_t)__popcll(mask);
  }
  if (!out && lane == 0) cnt[gb] = total;
}

__global__ void k_set_u32(uint32_t *p, uint32_t v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

// Persistent work-queue kernel over the bricks of a batch: each workgroup
// takes the next brick until the queue is drained; the last take of every
// workgroup returns a brick number past the end, so all of them exit.
__global__ void __launch_bounds__(256) k_promol_f_queue(const BatchCell *__restrict__ cells, int ncell,
                                                        uint32_t nbrick, PromolF T, const double4 *__restrict__ img,
                                                        const float *__restrict__ cut2,
                                                        const uint64_t *__restrict__ off,
                                                        const uint32_t *__restrict__ lst, uint32_t *queue) {
  __shared__ uint32_t next;
  for (;;) {
    __syncthreads();  // every thread has read the previous brick number
    if (threadIdx.x == 0) next = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t gb = next;
    if (gb >= nbrick) return;
    const BatchCell &C = cells[cell_of_brick(cells, ncell, gb)];
    const int b = (int)(gb - C.brick0);
    const int b1 = b % C.G.nb1, b2 = (b / C.G.nb1) % C.G.nb2, b3 = b / (C.G.nb1 * C.G.nb2);
    const int t = threadIdx.x;
    const int i = b1 * BX + (t % BX), j = b2 * BY + ((t / BX) % BY), k = b3 * BZ + t / (BX * BY);
    if (i < C.G.n1 && j < C.G.n2 && k < C.G.n3)
      promol_f_node(C.G, T, img, cut2, lst, off[gb], off[gb + 1], i, j, k, C.O);
  }
}

// FP32 promolecular NCI of every cell of a batch that asks for it (crho_out
// and cgrad_out set): one graph launch of the work-queue kernel.  The grid of
// that kernel (workgroups per CU) is tuned once per context on the first
// batch by timing the candidates.
void batch_nci_run(c2gpu_ctx *ctx, int ncell, const c2gpu_cell *cells, int nspecies,
                   const c2gpu_radial_table *tables, int flags, double *ms_out, int *wg_out) {
  hipStream_t st = ctx->stream;
  const bool dev = flags & C2GPU_PTR_DEVICE;
  std::vector<int> sel;
  for (int c = 0; c < ncell; c++)
    if (cells[c].crho_out && cells[c].cgrad_out) sel.push_back(c);
  if (ms_out) *ms_out = 0.0;
  if (sel.empty()) return;
  if (nspecies < 1 || !tables) fail(C2GPU_ERR_ARG, "c2gpu_batch_run: NCI asked for but no radial tables");
  const c2gpu_cell &c0 = cells[sel[0]];
  if (c0.nat < 1 || !c0.xat || !c0.species) fail(C2GPU_ERR_ARG, "c2gpu_batch_run: NCI cell without atoms");
  const c2gpu_promol pm0{nspecies, tables, c0.nat, c0.xat, c0.species};
  const PromolDev P = upload_promol(ctx, &pm0, c0.x2c);  // the tables (shared by all cells)
  double rcmax = 0;
  for (int s = 0; s < nspecies; s++) {
    const double rc = tables[s].cutrad < tables[s].rmax ? tables[s].cutrad : tables[s].rmax;
    rcmax = rc > rcmax ? rc : rcmax;
  }
  std::vector<BatchCell> hc;
  std::vector<double4> himg;
  uint64_t nbrick = 0;
  size_t nstage = 0;
  std::vector<size_t> stage_at;
  for (int c : sel) {
    const c2gpu_cell &C = cells[c];
    for (int q = 0; q < 3; q++)
      if (C.n[q] < 1) fail(C2GPU_ERR_ARG, "c2gpu_batch_run: bad grid");
    if (C.nat < 1 || !C.xat || !C.species) fail(C2GPU_ERR_ARG, "c2gpu_batch_run: NCI cell without atoms");
    for (int a = 0; a < C.nat; a++)
      if (C.species[a] < 0 || C.species[a] >= nspecies) fail(C2GPU_ERR_ARG, "c2gpu_batch_run: bad species");
    BatchCell B;
    B.G = make_geo(C.n, C.x2c);
    std::vector<double> pos;
    std::vector<int> isp, iat;
    expand_images(C.x2c, C.nat, C.xat, C.species, rcmax, pos, isp, iat);
    B.brick0 = (uint32_t)nbrick;
    B.img0 = (uint32_t)himg.size();
    B.nimg = (uint32_t)isp.size();
    B.brad = brick_radius(B.G);
    for (size_t m = 0; m < isp.size(); m++) {
      himg.push_back(make_double4(pos[3 * m], pos[3 * m + 1], pos[3 * m + 2], (double)isp[m]));
    }
    B.O.is_f32 = 1;
    B.O.zfast = 1;
    const size_t npt = (size_t)C.n[0] * C.n[1] * C.n[2];
    stage_at.push_back(nstage);
    if (dev) {
      B.O.crho = C.crho_out;
      B.O.cgrad = C.cgrad_out;
    } else {
      nstage += npt;
    }
    nbrick += (uint64_t)B.G.nb1 * B.G.nb2 * B.G.nb3;
    if (nbrick >= (1ull << 31)) fail(C2GPU_ERR_ARG, "c2gpu_batch_run: too many node bricks");
    hc.push_back(B);
  }
  float *stage = nullptr;
  if (!dev) {
    stage = buf<float>(ctx, "bq_stage", 2 * (nstage > 0 ? nstage : 1));
    for (size_t q = 0; q < hc.size(); q++) {
      hc[q].O.crho = stage + stage_at[q];
      hc[q].O.cgrad = stage + nstage + stage_at[q];
    }
  }
  const int ncs = (int)hc.size();
  BatchCell *d_cells = buf<BatchCell>(ctx, "bq_cells", hc.size());
  double4 *d_img = buf<double4>(ctx, "bq_img", himg.size() > 0 ? himg.size() : 1);
  C2_HIP(hipMemcpyAsync(d_cells, hc.data(), sizeof(BatchCell) * hc.size(), hipMemcpyHostToDevice, st));
  C2_HIP(hipMemcpyAsync(d_img, himg.data(), sizeof(double4) * himg.size(), hipMemcpyHostToDevice, st));
  uint32_t *cnt = buf<uint32_t>(ctx, "bq_bcnt", nbrick + 1);
  uint64_t *off = buf<uint64_t>(ctx, "bq_boff", nbrick + 1);
  C2_HIP(hipMemsetAsync(cnt + nbrick, 0, sizeof(uint32_t), st));
  C2_LAUNCH(ctx, k_brick_list_batch, dim3((unsigned)nbrick), dim3(64), 0, d_cells, ncs, d_img, P.rcut, cnt,
            (const uint64_t *)nullptr, (uint32_t *)nullptr);
  size_t tmp = 0;
  C2_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, cnt, off, (int)nbrick + 1, st));
  void *t = buffer(ctx, "bq_bscan", tmp);
  C2_HIP(hipcub::DeviceScan::ExclusiveSum(t, tmp, cnt, off, (int)nbrick + 1, st));
  uint64_t total = 0;
  copy_to_host(ctx, &total, off + nbrick, 8);
  C2_HIP(hipStreamSynchronize(st));
  uint32_t *lst = buf<uint32_t>(ctx, "bq_blst", total > 0 ? total : 1);
  C2_LAUNCH(ctx, k_brick_list_batch, dim3((unsigned)nbrick), dim3(64), 0, d_cells, ncs, d_img, P.rcut, cnt, off,
            lst);
  uint32_t *queue = buf<uint32_t>(ctx, "bq_queue", 1);
  int ncu = 256;
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device);
  // one graph launch with `wpc` workgroups per CU; returns its device time
  auto run = [&](int wpc) -> double {
    const unsigned nwg = (unsigned)(ncu * wpc);
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    C2_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, st, queue, 0u);  // no memset nodes in graphs
    hipLaunchKernelGGL(k_promol_f_queue, dim3(nwg), dim3(256), 0, st, d_cells, ncs, (uint32_t)nbrick, P.F,
                       (const double4 *)d_img, (const float *)P.cut2f, (const uint64_t *)off,
                       (const uint32_t *)lst, queue);
    C2_HIP(hipStreamEndCapture(st, &g));
    C2_HIP(hipGetLastError());
    C2_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    hipEvent_t a = event_get(ctx), b = event_get(ctx);
    C2_HIP(hipEventRecord(a, st));
    C2_HIP(hipGraphLaunch(ge, st));
    C2_HIP(hipEventRecord(b, st));
    C2_HIP(hipEventSynchronize(b));
    float ms = 0;
    C2_HIP(hipEventElapsedTime(&ms, a, b));
    ctx->evpool.push_back(a);
    ctx->evpool.push_back(b);
    (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
    return ms;
  };
  if (ctx->queue_wg_per_cu == 0) {
    // tune the persistent grid: the candidates write identical results
    double best = 1e300;
    for (int wpc : {2, 4, 8, 16}) {
      const double ms = run(wpc);
      if (ms < best) {
        best = ms;
        ctx->queue_wg_per_cu = wpc;
      }
    }
  }
  const double ms = run(ctx->queue_wg_per_cu);
  if (ctx->timing) {
    auto &k = ctx->ktime["k_promol_f_queue"];
    k.ms += ms;
    k.launches += 1;
    double pts = 0;
    for (const BatchCell &B : hc) pts += (double)B.G.n1 * B.G.n2 * B.G.n3;
    k.bytes += 8.0 * pts;  // FP32 crho + cgrad per node
  }
  if (ms_out) *ms_out = ms;
  if (wg_out) *wg_out = ncu * ctx->queue_wg_per_cu;
  if (!dev) {
    for (size_t q = 0; q < hc.size(); q++) {
      const c2gpu_cell &C = cells[sel[q]];
      const size_t npt = (size_t)C.n[0] * C.n[1] * C.n[2];
      C2_HIP(hipMemcpyAsync(C.crho_out, stage + stage_at[q], 4 * npt, hipMemcpyDeviceToHost, st));
      C2_HIP(hipMemcpyAsync(C.cgrad_out, stage + nstage + stage_at[q], 4 * npt, hipMemcpyDeviceToHost, st));
    }
  }
  C2_HIP(hipStreamSynchronize(st));
}

void nci_run(c2gpu_ctx *ctx, const double *rho, const int n[3], const double x2c[9], const double c2x[9],
             int scheme, int precision, const c2gpu_promol *promol, int zfast, int flags, void *crho_out,
             void *cgrad_out, int out_f32) {
  for (int c = 0; c < 3; c++)
    if (n[c] < 1) fail(C2GPU_ERR_ARG, "c2gpu_nci: bad grid");
  const size_t N = (size_t)n[0] * n[1] * n[2];
  hipStream_t st = ctx->stream;
  const Geo G = make_geo(n, x2c);
  const size_t osz = out_f32 ? 4 : 8;
  const bool dev = flags & C2GPU_PTR_DEVICE;
  NciOut O;
  O.is_f32 = out_f32;
  O.zfast = zfast ? 1 : 0;
  O.crho = dev ? crho_out : buffer(ctx, "nci_crho", osz * N);
  O.cgrad = dev ? cgrad_out : buffer(ctx, "nci_cgrad", osz * N);
  const unsigned g = grid_for((int64_t)N, 256);
  if (scheme == C2GPU_NCI_PROMOL) {
    PromolDev P = upload_promol(ctx, promol, x2c);
    BrickLists BL = build_bricks(ctx, G, P.img, P.rcut, P.nimg, "pm");
    const int nbk = G.nb1 * G.nb2 * G.nb3;
    if (precision == C2GPU_FP32)
      C2_LAUNCH(ctx, k_promol_f, dim3(nbk), dim3(256), 0, G, P.F, P.img, P.cut2f, BL.off, BL.lst, O);
    else
      C2_LAUNCH(ctx, k_promol_d, dim3(nbk), dim3(256), 0, G, P.D, P.img, BL.off, BL.lst, O);
  } else if (scheme == C2GPU_NCI_FFT) {
    if (precision != C2GPU_FP64) fail(C2GPU_ERR_UNSUPPORTED, "c2gpu_nci: FFT scheme runs in FP64");
    const double *d_rho = to_device_grid(ctx, rho, N, flags, "nci_rho");
    double2 *spec = buf<double2>(ctx, "fft_spec", N), *w = buf<double2>(ctx, "fft_work", N);
    double *gsq = buf<double>(ctx, "nci_gsq", N);
    int8_t *npos = buf<int8_t>(ctx, "nci_npos", N);
    FftPlan plan(ctx, n);
    const Recip R = make_recip(n, x2c);
    C2_LAUNCH(ctx, k_real_to_complex, dim3(g), dim3(256), 0, d_rho, spec, N);
    HIPFFT_CHECK(hipfftExecZ2Z(plan.h, (hipfftDoubleComplex *)spec, (hipfftDoubleComplex *)spec, HIPFFT_FORWARD));
    for (int c = 0; c < 3; c++) {  // |grad rho|^2 (ifformat_as_ft_grad)
      C2_LAUNCH(ctx, k_spectral, dim3(g), dim3(256), 0, R, spec, w, N, c);
      HIPFFT_CHECK(hipfftExecZ2Z(plan.h, (hipfftDoubleComplex *)w, (hipfftDoubleComplex *)w, HIPFFT_BACKWARD));
      C2_LAUNCH(ctx, k_accum_sq, dim3(g), dim3(256), 0, w, gsq, N, c == 0);
    }
    for (int c = 0; c < 3; c++) {  // xx, yy, zz signs
      C2_LAUNCH(ctx, k_spectral, dim3(g), dim3(256), 0, R, spec, w, N, 10 + c);
      HIPFFT_CHECK(hipfftExecZ2Z(plan.h, (hipfftDoubleComplex *)w, (hipfftDoubleComplex *)w, HIPFFT_BACKWARD));
      C2_LAUNCH(ctx, k_count_pos, dim3(g), dim3(256), 0, w, npos, N, c == 0);
    }
    if (out_f32) fail(C2GPU_ERR_UNSUPPORTED, "c2gpu_nci_f32: FFT scheme writes FP64");
    C2_LAUNCH(ctx, k_nci_fft_point, dim3(g), dim3(256), 0, G, d_rho, gsq, npos, O);
  } else if (scheme == C2GPU_NCI_GRID) {
    if (precision != C2GPU_FP64) fail(C2GPU_ERR_UNSUPPORTED, "c2gpu_nci: grid scheme runs in FP64");
    for (int c = 0; c < 3; c++)
      if (n[c] < 2) fail(C2GPU_ERR_ARG, "c2gpu_nci: grid scheme needs n >= 2 per axis");
    const double *d_rho = to_device_grid(ctx, rho, N, flags, "nci_rho");
    nci_grid_launch(ctx, G, x2c, c2x, d_rho, O);
  } else {
    fail(C2GPU_ERR_UNSUPPORTED, "c2gpu_nci: unknown scheme");
  }
  if (!dev) {
    C2_HIP(hipMemcpyAsync(crho_out, O.crho, osz * N, hipMemcpyDeviceToHost, st));
    C2_HIP(hipMemcpyAsync(cgrad_out, O.cgrad, osz * N, hipMemcpyDeviceToHost, st));
  }
  C2_HIP(hipStreamSynchronize(st));
}


// ---------------------------------------------------------------- nciplot cubes and fragments

static void inv3(const double a[9], double inv[9]) {
  const double det = a[0] * (a[4] * a[8] - a[7] * a[5]) - a[3] * (a[1] * a[8] - a[7] * a[2]) +
                     a[6] * (a[1] * a[5] - a[4] * a[2]);
  if (!(det != 0.0)) fail(C2GPU_ERR_ARG, "c2gpu_nci: singular cell");
  inv[0] = (a[4] * a[8] - a[7] * a[5]) / det;
  inv[3] = -(a[3] * a[8] - a[6] * a[5]) / det;
  inv[6] = (a[3] * a[7] - a[6] * a[4]) / det;
  inv[1] = -(a[1] * a[8] - a[7] * a[2]) / det;
  inv[4] = (a[0] * a[8] - a[6] * a[2]) / det;
  inv[7] = -(a[0] * a[7] - a[6] * a[1]) / det;
  inv[2] = (a[1] * a[5] - a[4] * a[2]) / det;
  inv[5] = -(a[0] * a[5] - a[3] * a[2]) / det;
  inv[8] = (a[0] * a[4] - a[3] * a[1]) / det;
}

// Fractional range [lo, hi] of the nodes of an nciplot cube (its corners).
static void box_frac_range(const Geo &G, const double x2c[9], double lo[3], double hi[3]) {
  double c2x[9];
  inv3(x2c, c2x);
  for (int c = 0; c < 3; c++) {
    lo[c] = 1e300;
    hi[c] = -1e300;
  }
  for (int q = 0; q < 8; q++) {
    const double u = (q & 1) ? G.n1 - 1 : 0, v = (q & 2) ? G.n2 - 1 : 0, w = (q & 4) ? G.n3 - 1 : 0;
    double x[3], f[3];
    for (int r = 0; r < 3; r++) x[r] = G.x0[r] + u * G.xm[r] + v * G.xm[r + 3] + w * G.xm[r + 6];
    for (int r = 0; r < 3; r++) f[r] = c2x[r] * x[0] + c2x[r + 3] * x[1] + c2x[r + 6] * x[2];
    for (int c = 0; c < 3; c++) {
      lo[c] = f[c] < lo[c] ? f[c] : lo[c];
      hi[c] = f[c] > hi[c] ? f[c] : hi[c];
    }
  }
}

static Geo checked_box(const c2gpu_nci_nodes *nd, const double x2c[9], const char *who) {
  if (!nd || !x2c) fail(C2GPU_ERR_ARG, std::string(who) + ": NULL argument");
  for (int c = 0; c < 3; c++)
    if (nd->nstep[c] < 1) fail(C2GPU_ERR_ARG, std::string(who) + ": nstep must be positive");
  if ((uint64_t)nd->nstep[0] * nd->nstep[1] * nd->nstep[2] >= (1ull << 32))
    fail(C2GPU_ERR_ARG, std::string(who) + ": fewer than 2^32 nodes");
  return make_geo_box(*nd, x2c);
}

void nci_promol_nodes_run(c2gpu_ctx *ctx, const c2gpu_nci_nodes *nd, const double x2c[9], const c2gpu_promol *pm,
                          int precision, int flags, void *crho_out, void *cgrad_out, void *rhoat_out) {
  const Geo G = checked_box(nd, x2c, "c2gpu_nci_promol_nodes");
  if (!crho_out || !cgrad_out) fail(C2GPU_ERR_ARG, "c2gpu_nci_promol_nodes: NULL output");
  if (precision != C2GPU_FP64 && precision != C2GPU_FP32)
    fail(C2GPU_ERR_ARG, "c2gpu_nci_promol_nodes: precision must be C2GPU_FP64 or C2GPU_FP32");
  const bool f32 = precision == C2GPU_FP32;
  const size_t N = (size_t)G.n1 * G.n2 * G.n3, osz = f32 ? 4 : 8;
  const bool dev = flags & C2GPU_PTR_DEVICE;
  if (dev) {
    require_device_ptr(crho_out, "crho_out");
    require_device_ptr(cgrad_out, "cgrad_out");
    if (rhoat_out) require_device_ptr(rhoat_out, "rhoat_out");
  }
  hipStream_t st = ctx->stream;
  double lo[3], hi[3];
  box_frac_range(G, x2c, lo, hi);
  PromolDev P = upload_promol(ctx, pm, x2c, lo, hi);
  BrickLists BL = build_bricks(ctx, G, P.img, P.rcut, P.nimg, "pm");
  NciOut O;
  O.is_f32 = f32 ? 1 : 0;
  O.zfast = 1;
  O.crho = dev ? crho_out : buffer(ctx, "nci_crho", osz * N);
  O.cgrad = dev ? cgrad_out : buffer(ctx, "nci_cgrad", osz * N);
  O.rhoat = rhoat_out ? (dev ? rhoat_out : buffer(ctx, "nci_rhoat", osz * N)) : nullptr;
  const int nbk = G.nb1 * G.nb2 * G.nb3;
  if (f32)
    C2_LAUNCH(ctx, k_promol_f, dim3(nbk), dim3(256), 0, G, P.F, P.img, P.cut2f, BL.off, BL.lst, O);
  else
    C2_LAUNCH(ctx, k_promol_d, dim3(nbk), dim3(256), 0, G, P.D, P.img, BL.off, BL.lst, O);
  if (!dev) {
    C2_HIP(hipMemcpyAsync(crho_out, O.crho, osz * N, hipMemcpyDeviceToHost, st));
    C2_HIP(hipMemcpyAsync(cgrad_out, O.cgrad, osz * N, hipMemcpyDeviceToHost, st));
    if (rhoat_out) C2_HIP(hipMemcpyAsync(rhoat_out, O.rhoat, osz * N, hipMemcpyDeviceToHost, st));
  }
  C2_HIP(hipStreamSynchronize(st));
}

// FP32 density of one table at r (the lookup of promol_f_node); 0 at and
// beyond rmax.
__device__ __forceinline__ float table_rho_f(const PromolF &T, int s, float r) {
  if (r >= T.trmax[s]) return 0.f;
  const float r1 = T.tr0[s];
  int ir;
  if (r <= r1) {
    ir = 1;
    r = r1;
  } else {
    ir = 1 + (int)floorf((__logf(r) - T.tlog_a[s]) * T.ta_inv_b[s]);
  }
  const int ng = T.tng[s];
  int base = ir < 2 ? 2 : ir;
  base = base > ng - 2 ? ng - 2 : base;
  const int ci = T.toff[s] + base;
  const float u = r - T.tnode[ci];
  const float4 cf = T.coef[4 * (size_t)ci + 0];
  const float rho = ((cf.w * u + cf.z) * u + cf.y) * u + cf.x;
  return rho > 0.f ? rho : 0.f;
}

// Per node and fragment, the density of the fragment's atoms alone
// (nci@proc.f90:580-596): atoms beyond rmax of their table skipped, the
// distance raised to r(1) and 1e-14, the interpolated density clipped at 0.
// out(p + f npts), p the z-fastest node index.
template <class T>
__global__ void __launch_bounds__(256) k_frag(Geo G, PromolD D, PromolF F, const double4 *__restrict__ fat,
                                              const int *__restrict__ fstart, int nfrag, uint64_t npts,
                                              T *__restrict__ out) {
  const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= npts) return;
  const int k = (int)(p % (uint64_t)G.n3);
  const uint64_t q = p / (uint64_t)G.n3;
  const int j = (int)(q % (uint64_t)G.n2), i = (int)(q / (uint64_t)G.n2);
  double xc[3];
  node_cart(G, i, j, k, xc);
  for (int f = 0; f < nfrag; f++) {
    T acc = T(0);
    for (int a = fstart[f]; a < fstart[f + 1]; a++) {
      const double4 at = fat[a];
      const int s = (int)at.w;
      if constexpr (sizeof(T) == 8) {
        const double x0 = xc[0] - at.x, x1 = xc[1] - at.y, x2 = xc[2] - at.z;
        double d = sqrt(x0 * x0 + x1 * x1 + x2 * x2);
        if (d > D.trmax[s]) continue;
        const double r1 = D.r[D.toff[s]];
        d = d > r1 ? d : r1;
        d = d > 1e-14 ? d : 1e-14;
        double rho, rhop, rhopp;
        interp_d(D, s, d, rho, rhop, rhopp);
        rho = rho > 0.0 ? rho : 0.0;
        acc = acc + rho;
      } else {
        const float x0 = (float)(xc[0] - at.x), x1 = (float)(xc[1] - at.y), x2 = (float)(xc[2] - at.z);
        acc += table_rho_f(F, s, sqrtf(x0 * x0 + x1 * x1 + x2 * x2));
      }
    }
    out[(uint64_t)f * npts + p] = acc;
  }
}

void nci_fragments_run(c2gpu_ctx *ctx, const c2gpu_nci_nodes *nd, int nspecies, const c2gpu_radial_table *tables,
                       int nfrag, const int *frag_nat, const double *xyz, const int *species, int precision,
                       int flags, void *rhofrag_out) {
  static const double ident[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  const Geo G = checked_box(nd, ident, "c2gpu_nci_fragments");
  if (nfrag < 1 || !frag_nat || !rhofrag_out) fail(C2GPU_ERR_ARG, "c2gpu_nci_fragments: no fragments");
  if (precision != C2GPU_FP64 && precision != C2GPU_FP32)
    fail(C2GPU_ERR_ARG, "c2gpu_nci_fragments: precision must be C2GPU_FP64 or C2GPU_FP32");
  std::vector<int> fstart(nfrag + 1, 0);
  for (int f = 0; f < nfrag; f++) {
    if (frag_nat[f] < 0 || (int64_t)fstart[f] + frag_nat[f] > (int64_t)INT32_MAX)
      fail(C2GPU_ERR_ARG, "c2gpu_nci_fragments: bad atom count");
    fstart[f + 1] = fstart[f] + frag_nat[f];
  }
  const int nat = fstart[nfrag];
  if (nat > 0 && (!xyz || !species)) fail(C2GPU_ERR_ARG, "c2gpu_nci_fragments: NULL atoms");
  std::vector<double4> h(nat > 0 ? nat : 1);
  for (int a = 0; a < nat; a++) {
    if (species[a] < 0 || species[a] >= nspecies) fail(C2GPU_ERR_ARG, "c2gpu_nci_fragments: bad species index");
    h[a] = make_double4(xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2], (double)species[a]);
  }
  PromolDev P;
  upload_tables(ctx, nspecies, tables, P);
  hipStream_t st = ctx->stream;
  double4 *d_at = buf<double4>(ctx, "frag_at", h.size());
  int *d_fs = buf<int>(ctx, "frag_start", fstart.size());
  C2_HIP(hipMemcpyAsync(d_at, h.data(), sizeof(double4) * h.size(), hipMemcpyHostToDevice, st));
  C2_HIP(hipMemcpyAsync(d_fs, fstart.data(), sizeof(int) * fstart.size(), hipMemcpyHostToDevice, st));
  const uint64_t npts = (uint64_t)G.n1 * G.n2 * G.n3;
  const size_t osz = precision == C2GPU_FP32 ? 4 : 8;
  const bool dev = flags & C2GPU_PTR_DEVICE;
  if (dev) require_device_ptr(rhofrag_out, "rhofrag_out");
  void *out = dev ? rhofrag_out : buffer(ctx, "frag_out", osz * npts * nfrag);
  const unsigned g = grid_for((int64_t)npts, 256);
  if (precision == C2GPU_FP32)
    C2_LAUNCH(ctx, k_frag<float>, dim3(g), dim3(256), 0, G, P.D, P.F, (const double4 *)d_at, (const int *)d_fs,
              nfrag, npts, (float *)out);
  else
    C2_LAUNCH(ctx, k_frag<double>, dim3(g), dim3(256), 0, G, P.D, P.F, (const double4 *)d_at, (const int *)d_fs,
              nfrag, npts, (double *)out);
  if (!dev) C2_HIP(hipMemcpyAsync(rhofrag_out, out, osz * npts * nfrag, hipMemcpyDeviceToHost, st));
  C2_HIP(hipStreamSynchronize(st));
}

}  // namespace c2
