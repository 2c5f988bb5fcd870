/* Oracle (test infrastructure only; never linked into the product):
 * restatement of the NCIPLOT grid-interpolation scheme,
 *   nci@proc.f90:565-571 (INTERPOLATION GRID, or a reference field with a
 *   core contribution), through field%grd (fieldmod@proc.f90:808-974),
 *   grid3%interp (grid3mod@proc.f90:1695-1752) and grinterp_tricubic
 *   (grid3mod@proc.f90:2649-2799), on the nodes of a periodic grid
 *   (nci@proc.f90:412-446 with nstep = grid%n).
 *
 * The 64x64 tricubic coefficient matrix of Lekien and Marsden (the table
 * c(64,64) of grid3mod@proc.f90:76) is the tensor product of the 1-D cubic
 * Hermite matrix: it is generated here from that identity and pinned against
 * golden vectors written from the reference's table. */
#include <math.h>
#include <string.h>

#include "c2oracle.h"

/* 1-D cubic Hermite: coefficient p of x^p from (v(0), v(1), v'(0), v'(1)) */
static const double HERM[4][4] = {
    {1, 0, 0, 0}, {0, 0, 1, 0}, {-3, 3, -2, -1}, {2, -2, 1, 1}};

/* c(l, m): l = 1 + p + 4q + 16r is the coefficient of x^p y^q z^r;
 * m = 8*type + corner with corner = cx + 2cy + 4cz and type 0..7 =
 * f, fx, fy, fz, fxy, fxz, fyz, fxyz (the b vector of
 * grid3mod@proc.f90:2687-2767). */
void oracle_tricubic_matrix(double c[64 * 64]) {
  static const int type_of[2][2][2] = {{{0, 3}, {2, 6}}, {{1, 5}, {4, 7}}}; /* [dx][dy][dz] */
  for (int l = 0; l < 64; l++) {
    const int p = l & 3, q = (l >> 2) & 3, r = l >> 4;
    for (int dx = 0; dx < 4; dx++)
      for (int dy = 0; dy < 4; dy++)
        for (int dz = 0; dz < 4; dz++) {
          const int corner = (dx & 1) + 2 * (dy & 1) + 4 * (dz & 1);
          const int type = type_of[dx >> 1][dy >> 1][dz >> 1];
          const int m = 8 * type + corner;
          c[l + 64 * m] = HERM[p][dx] * HERM[q][dy] * HERM[r][dz];
        }
  }
}

static int wrapi(long v, int n) {
  long m = v % n;
  return (int)(m < 0 ? m + n : m);
}

/* grinterp_tricubic (grid3mod@proc.f90:2649-2799): xi fractional in [0,1]
 * (after modulo); y, yp (fractional derivatives times n), ypp (column-major). */
void oracle_tricubic(const double *f, const int n[3], const double xi[3], double *y, double yp[3],
                     double ypp[9]) {
  static double C[64 * 64];
  static int cinit = 0;
  if (!cinit) {
    oracle_tricubic_matrix(C);
    cinit = 1;
  }
  int idx[3]; /* 0-based: grid_floor(f,xi) - 1 (grid3mod@proc.f90:3119-3136) */
  for (int c = 0; c < 3; c++) idx[c] = wrapi((long)floor(xi[c] * (double)n[c]), n[c]);
  double g[4][4][4]; /* g(-1:2,-1:2,-1:2) */
  for (int i = -1; i <= 2; i++)
    for (int j = -1; j <= 2; j++)
      for (int k = -1; k <= 2; k++) {
        const int a = wrapi(idx[0] + i, n[0]), b = wrapi(idx[1] + j, n[1]), cc = wrapi(idx[2] + k, n[2]);
        g[i + 1][j + 1][k + 1] = f[a + (size_t)n[0] * (b + (size_t)n[1] * cc)];
      }
#define G(i, j, k) g[(i) + 1][(j) + 1][(k) + 1]
  double b[64];
  /* f, fx, fy, fz, fxy, fxz, fyz, fxyz at the corners (cx,cy,cz) */
  for (int cz = 0; cz < 2; cz++)
    for (int cy = 0; cy < 2; cy++)
      for (int cx = 0; cx < 2; cx++) {
        const int m = cx + 2 * cy + 4 * cz;
        b[m] = G(cx, cy, cz);
        b[8 + m] = 0.5 * (G(cx + 1, cy, cz) - G(cx - 1, cy, cz));
        b[16 + m] = 0.5 * (G(cx, cy + 1, cz) - G(cx, cy - 1, cz));
        b[24 + m] = 0.5 * (G(cx, cy, cz + 1) - G(cx, cy, cz - 1));
        b[32 + m] = 0.25 * (G(cx + 1, cy + 1, cz) - G(cx - 1, cy + 1, cz) - G(cx + 1, cy - 1, cz) +
                            G(cx - 1, cy - 1, cz));
        b[40 + m] = 0.25 * (G(cx + 1, cy, cz + 1) - G(cx - 1, cy, cz + 1) - G(cx + 1, cy, cz - 1) +
                            G(cx - 1, cy, cz - 1));
        b[48 + m] = 0.25 * (G(cx, cy + 1, cz + 1) - G(cx, cy - 1, cz + 1) - G(cx, cy + 1, cz - 1) +
                            G(cx, cy - 1, cz - 1));
        b[56 + m] = 0.125 * (G(cx + 1, cy + 1, cz + 1) - G(cx - 1, cy + 1, cz + 1) - G(cx + 1, cy - 1, cz + 1) +
                             G(cx - 1, cy - 1, cz + 1) - G(cx + 1, cy + 1, cz - 1) + G(cx - 1, cy + 1, cz - 1) +
                             G(cx + 1, cy - 1, cz - 1) - G(cx - 1, cy - 1, cz - 1));
      }
#undef G
  double a[64];
  for (int l = 0; l < 64; l++) {
    double s = 0.0;
    for (int m = 0; m < 64; m++) s = s + C[l + 64 * m] * b[m];
    a[l] = s;
  }
  double x[3];
  for (int c = 0; c < 3; c++) x[c] = xi[c] * (double)n[c] - (double)idx[c];
  double aa[4], aax[4], aay[4], aaxy[4], aaxx[4], aayy[4];
  int l = 0;
  for (int k = 0; k < 4; k++) {
    double bb[4], bbx[4], bbxx[4];
    for (int j = 0; j < 4; j++) {
      bb[j] = a[l] + x[0] * (a[l + 1] + x[0] * (a[l + 2] + x[0] * a[l + 3]));
      bbx[j] = a[l + 1] + x[0] * (2.0 * a[l + 2] + x[0] * 3.0 * a[l + 3]);
      bbxx[j] = 2.0 * a[l + 2] + 6.0 * x[0] * a[l + 3];
      l += 4;
    }
    aa[k] = bb[0] + x[1] * (bb[1] + x[1] * (bb[2] + x[1] * bb[3]));
    aax[k] = bbx[0] + x[1] * (bbx[1] + x[1] * (bbx[2] + x[1] * bbx[3]));
    aay[k] = bb[1] + x[1] * (2.0 * bb[2] + x[1] * 3.0 * bb[3]);
    aaxy[k] = bbx[1] + x[1] * (2.0 * bbx[2] + x[1] * 3.0 * bbx[3]);
    aaxx[k] = bbxx[0] + x[1] * (bbxx[1] + x[1] * (bbxx[2] + x[1] * bbxx[3]));
    aayy[k] = 2.0 * bb[2] + 6.0 * x[1] * bb[3];
  }
  const double z = x[2];
  *y = aa[0] + z * (aa[1] + z * (aa[2] + z * aa[3]));
  yp[0] = aax[0] + z * (aax[1] + z * (aax[2] + z * aax[3]));
  yp[1] = aay[0] + z * (aay[1] + z * (aay[2] + z * aay[3]));
  yp[2] = aa[1] + z * (2.0 * aa[2] + z * 3.0 * aa[3]);
  double h[3][3];
  h[0][0] = aaxx[0] + z * (aaxx[1] + z * (aaxx[2] + z * aaxx[3]));
  h[0][1] = aaxy[0] + z * (aaxy[1] + z * (aaxy[2] + z * aaxy[3]));
  h[0][2] = aax[1] + z * (2.0 * aax[2] + z * 3.0 * aax[3]);
  h[1][1] = aayy[0] + z * (aayy[1] + z * (aayy[2] + z * aayy[3]));
  h[1][2] = aay[1] + z * (2.0 * aay[2] + z * 3.0 * aay[3]);
  h[2][2] = 2.0 * aa[2] + 6.0 * z * aa[3];
  for (int i = 0; i < 3; i++) {
    yp[i] = yp[i] * (double)n[i];
    for (int j = i; j < 3; j++) {
      h[i][j] = h[i][j] * (double)n[i] * (double)n[j];
      h[j][i] = h[i][j];
    }
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) ypp[i + 3 * j] = h[i][j];
}

/* NCIPLOT grid scheme on every node of a periodic grid (nci@proc.f90:543-571,
 * :598-599). x2c: cell matrix (column-major); c2x: its inverse as critic2's
 * matinv computes it (used both as crystal%m_c2x and grid%c2xl).  Outputs
 * z-fastest, (0:n3-1,0:n2-1,0:n1-1). */
void oracle_nci_grid(const double *f, const int n[3], const double x2c[9], const double c2x[9],
                     double *crho, double *cgrad) {
  double xm[9]; /* xmat(:,c) = x2c(e_c) / nstep(c) (nci@proc.f90:418-428) */
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) xm[r + 3 * c] = x2c[r + 3 * c] / (double)n[c];
  for (int i = 0; i < n[0]; i++)
    for (int j = 0; j < n[1]; j++)
      for (int k = 0; k < n[2]; k++) {
        double xc[3], wx[3];
        for (int r = 0; r < 3; r++)
          xc[r] = ((0.0 + (double)i * xm[r]) + (double)j * xm[r + 3]) + (double)k * xm[r + 6];
        for (int r = 0; r < 3; r++) {
          double s = 0.0;
          for (int c = 0; c < 3; c++) s = s + c2x[r + 3 * c] * xc[c];
          wx[r] = s;
        }
        /* field%grd: fold into the cell outside a 1d-4 border
         * (fieldmod@proc.f90:922-928); grid3%interp: modulo(xi,1d0) */
        for (int r = 0; r < 3; r++) {
          if (wx[r] < -1e-4 || wx[r] > 1.0 + 1e-4) wx[r] = wx[r] - floor(wx[r]);
          wx[r] = wx[r] - floor(wx[r]);
        }
        double y, yp[3], ypp[9];
        oracle_tricubic(f, n, wx, &y, yp, ypp);
        /* yp = matmul(transpose(c2xl),yp); ypp = c2xl^T ypp c2xl */
        double g[3], t[9], hc[9];
        for (int r = 0; r < 3; r++) {
          double s = 0.0;
          for (int q = 0; q < 3; q++) s = s + c2x[q + 3 * r] * yp[q];
          g[r] = s;
        }
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) {
            double s = 0.0;
            for (int q = 0; q < 3; q++) s = s + c2x[q + 3 * r] * ypp[q + 3 * c];
            t[r + 3 * c] = s;
          }
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) {
            double s = 0.0;
            for (int q = 0; q < 3; q++) s = s + t[r + 3 * q] * c2x[q + 3 * c];
            hc[r + 3 * c] = s;
          }
        const double gmod = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        const size_t o = (size_t)k + (size_t)n[2] * ((size_t)j + (size_t)n[1] * (size_t)i);
        oracle_nci_point(y, gmod, hc, 1, &crho[o], &cgrad[o]);
      }
}
