/*
 * nci.c -- CPU restatement of critic2's NCI per-point quantities, the
 * promolecular radial-table density and the synthetic Slater-sum grid.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle); never part of the product path.
 *
 *   grid1 table build      grid1mod@proc.f90:206-331 (read_critic, 6-point FD)
 *   grid1%interp           grid1mod@proc.f90:86-137 (4-point Lagrange, log grid)
 *   promolecular_atom      crystalmod@env.f90:622-753
 *   fragment densities     nci@proc.f90:580-596
 *   NCI point (s, sign)    nci@proc.f90:91, :543-605
 */
#include "c2oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double PI = 3.14159265358979323846264338327950288;

/* 6-point derivative formulas, grid1mod@proc.f90:25-38 */
static const int noef[3][6] = {{0, 1, 2, 3, 4, 5}, {-2, -1, 0, 1, 2, 3}, {-5, -4, -3, -2, -1, 0}};
static const double coef1[3][6] = {{-274, 600, -600, 400, -150, 24},
                                   {6, -60, -40, 120, -30, 4},
                                   {-24, 150, -400, 600, -600, 274}};
static const double coef2[3][6] = {{225, -770, 1070, -780, 305, -50},
                                   {-5, 80, -150, 80, -5, 0},
                                   {-50, 305, -780, 1070, -770, 225}};

void oracle_grid1_build(int ngrid, const double *r, const double *rr0, double *f,
                        double *fp, double *fpp, double *a, double *b) {
  const double fac1 = 1.0 / 120.0, fac2 = 2.0 / 120.0;
  *a = r[0];
  *b = log(r[1] / r[0]);
  for (int i = 0; i < ngrid; i++) {
    int ic;
    /* Fortran i is 1-based: i <= 2 -> ic 1; i >= ngrid-2 -> ic 3; else 2 */
    int i1 = i + 1;
    if (i1 <= 2) ic = 0;
    else if (i1 >= ngrid - 2) ic = 2;
    else ic = 1;
    double d1 = 0.0, d2 = 0.0;
    for (int j = 0; j < 6; j++) {
      d1 = d1 + coef1[ic][j] * rr0[i + noef[ic][j]];
      d2 = d2 + coef2[ic][j] * rr0[i + noef[ic][j]];
    }
    d1 = d1 * fac1;
    d2 = d2 * fac2;
    double rv = r[i];
    double r1 = 1.0 / rv, r2 = r1 * r1, r3 = r2 * r1, r4 = r3 * r1;
    double delta = 1.0 / *b, delta2 = delta * delta;
    f[i] = rr0[i] * r2;
    fp[i] = (d1 * delta - 2.0 * rr0[i]) * r3;
    fpp[i] = (d2 * delta2 - 5.0 * d1 * delta + 6.0 * rr0[i]) * r4;
  }
  for (int i = 0; i < ngrid; i++) {
    f[i] = f[i] / (4.0 * PI);
    fp[i] = fp[i] / (4.0 * PI);
    fpp[i] = fpp[i] / (4.0 * PI);
  }
}

void oracle_grid1_interp(const orc_rgrid *g, double r0, double *f, double *fp,
                         double *fpp) {
  *f = 0.0; *fp = 0.0; *fpp = 0.0;
  if (r0 >= g->rmax) return;
  int ir;
  double r;
  if (r0 <= g->r[0]) {
    ir = 1;
    r = g->r[0];
  } else {
    ir = 1 + (int)floor(log(r0 / g->a) / g->b);
    r = r0;
  }
  double rr[4], dr1[4], x1dr12[4][4];
  memset(x1dr12, 0, sizeof(x1dr12));
  int base = ir;
  if (base < 2) base = 2;
  if (base > g->ngrid - 2) base = g->ngrid - 2;
  for (int i = 0; i < 4; i++) {
    int ii = base - 2 + (i + 1); /* 1-based */
    rr[i] = g->r[ii - 1];
    dr1[i] = r - rr[i];
    for (int j = 0; j < i; j++) {
      x1dr12[i][j] = 1.0 / (rr[i] - rr[j]);
      x1dr12[j][i] = -x1dr12[i][j];
    }
  }
  for (int i = 0; i < 4; i++) {
    int ii = base - 2 + (i + 1);
    double prod = 1.0;
    for (int j = 0; j < 4; j++) {
      if (i == j) continue;
      prod = prod * dr1[j] * x1dr12[i][j];
    }
    *f = *f + g->f[ii - 1] * prod;
    *fp = *fp + g->fp[ii - 1] * prod;
    *fpp = *fpp + g->fpp[ii - 1] * prod;
  }
}

void oracle_promol(const double xc[3], int nimg, const double *ximg,
                   const int *zidx, const orc_rgrid *tabs, const double *cut,
                   int nder, double *f, double g[3], double h[9]) {
  *f = 0.0;
  for (int c = 0; c < 3; c++) g[c] = 0.0;
  for (int c = 0; c < 9; c++) h[c] = 0.0;
  for (int ii = 0; ii < nimg; ii++) {
    const orc_rgrid *gr = &tabs[zidx[ii]];
    double xx[3] = {xc[0] - ximg[3 * ii], xc[1] - ximg[3 * ii + 1], xc[2] - ximg[3 * ii + 2]};
    double r = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
    double rc = cut[zidx[ii]];
    if (gr->rmax < rc) rc = gr->rmax;
    if (r > rc) continue;
    if (r < gr->r[0]) r = gr->r[0];
    if (r < 1e-14) r = 1e-14;
    double rho, rhop, rhopp;
    oracle_grid1_interp(gr, r, &rho, &rhop, &rhopp);
    if (rho < 0.0) rho = 0.0;
    *f = *f + rho;
    if (nder < 1) continue;
    double rinv1 = 1.0 / r;
    double rinv1rp = rinv1 * rhop;
    for (int c = 0; c < 3; c++) g[c] = g[c] + xx[c] * rinv1rp;
    if (nder < 2) continue;
    double rfac = (rhopp - rinv1rp) * rinv1 * rinv1;
    for (int j = 0; j < 3; j++) {
      h[j + 3 * j] = h[j + 3 * j] + rinv1rp + rfac * xx[j] * xx[j];
      for (int k = 0; k < j; k++) {
        double radd = rfac * xx[j] * xx[k];
        h[j + 3 * k] += radd;
        h[k + 3 * j] += radd;
      }
    }
  }
}

void oracle_frag_density(const double xc[3], int nfrag, const int *fstart,
                         const double *xyz, const int *zidx,
                         const orc_rgrid *tabs, double *out) {
  for (int ifr = 0; ifr < nfrag; ifr++) {
    double acc = 0.0;
    for (int iat = fstart[ifr]; iat < fstart[ifr + 1]; iat++) {
      const orc_rgrid *g = &tabs[zidx[iat]];
      double xd[3];
      for (int c = 0; c < 3; c++) xd[c] = xc[c] - xyz[3 * iat + c];
      double dist = sqrt(xd[0] * xd[0] + xd[1] * xd[1] + xd[2] * xd[2]);
      if (dist > g->rmax) continue;
      if (dist < g->r[0]) dist = g->r[0];
      if (dist < 1e-14) dist = 1e-14;
      double rrho, rrho1, rrho2;
      oracle_grid1_interp(g, dist, &rrho, &rrho1, &rrho2);
      if (rrho < 0.0) rrho = 0.0;
      acc = acc + rrho;
    }
    out[ifr] = acc;
  }
}

/* Eigenvalues of a symmetric 3x3 matrix by cyclic Jacobi rotations, sorted
 * ascending (the order dsyev returns). */
static void eig3(const double hin[9], double ev[3]) {
  double a[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) a[i][j] = hin[i + 3 * j];
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
    double dia = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]);
    if (off <= 1e-300 || off <= 1e-17 * dia) break;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        if (a[p][q] == 0.0) continue;
        double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; k++) {
          double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; k++) {
          double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
      }
  }
  ev[0] = a[0][0]; ev[1] = a[1][1]; ev[2] = a[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2 - i; j++)
      if (ev[j] > ev[j + 1]) { double t = ev[j]; ev[j] = ev[j + 1]; ev[j + 1] = t; }
}

void oracle_nci_point(double rho, double gmod, const double e[9],
                      int from_full_hessian, double *crho, double *s) {
  const double cst = 2.0 * pow(3.0 * PI * PI, 1.0 / 3.0);
  const double fthirds = 4.0 / 3.0;
  double den = rho > 1e-80 ? rho : 1e-80; /* max(res%f, vsmall), param.F90 */
  *s = gmod / (cst * pow(den, fthirds));
  double l2;
  if (from_full_hessian) {
    double ev[3];
    eig3(e, ev);
    l2 = ev[1];
  } else {
    int npos = (e[0] > 0.0) + (e[1] > 0.0) + (e[2] > 0.0);
    l2 = (npos >= 2) ? 1.0 : -1.0;
  }
  /* Fortran sign(a,b): |a| with the sign of b (b = +0 gives +|a|) */
  double a = fabs(rho);
  *crho = (l2 < 0.0 || (l2 == 0.0 && signbit(l2))) ? -a * 100.0 : a * 100.0;
}

void oracle_slater_grid(const double x2c[9], const int n[3], int nat,
                        const double *xat, const int *zat, double rimg,
                        double *out) {
  /* image range per axis: enough translations to cover rimg */
  int lim[3];
  for (int c = 0; c < 3; c++) {
    double len = sqrt(x2c[0 + 3 * c] * x2c[0 + 3 * c] + x2c[1 + 3 * c] * x2c[1 + 3 * c] +
                      x2c[2 + 3 * c] * x2c[2 + 3 * c]);
    lim[c] = (int)ceil(rimg / len) + 1;
  }
  /* expand images */
  int cap = nat * (2 * lim[0] + 1) * (2 * lim[1] + 1) * (2 * lim[2] + 1);
  double *pos = (double *)malloc(sizeof(double) * 3 * cap);
  double *pref = (double *)malloc(sizeof(double) * cap);
  double *zeta = (double *)malloc(sizeof(double) * cap);
  int m = 0;
  for (int a = 0; a < nat; a++) {
    double z = (double)zat[a];
    double ze = 2.0 * cbrt(z);
    for (int i = -lim[0]; i <= lim[0]; i++)
      for (int j = -lim[1]; j <= lim[1]; j++)
        for (int k = -lim[2]; k <= lim[2]; k++) {
          double xf[3] = {xat[3 * a] + i, xat[3 * a + 1] + j, xat[3 * a + 2] + k};
          for (int r = 0; r < 3; r++)
            pos[3 * m + r] = x2c[r] * xf[0] + x2c[r + 3] * xf[1] + x2c[r + 6] * xf[2];
          pref[m] = z * ze * ze * ze / (8.0 * PI);
          zeta[m] = ze;
          m++;
        }
  }
  for (int k = 0; k < n[2]; k++)
    for (int j = 0; j < n[1]; j++)
      for (int i = 0; i < n[0]; i++) {
        double xf[3] = {(double)i / n[0], (double)j / n[1], (double)k / n[2]};
        double xc[3];
        for (int r = 0; r < 3; r++) xc[r] = x2c[r] * xf[0] + x2c[r + 3] * xf[1] + x2c[r + 6] * xf[2];
        double acc = 0.0;
        for (int q = 0; q < m; q++) {
          /* images: nearest copy of each atom relative to the cell, then all
           * translations whose distance to the point is within rimg */
          double dx = xc[0] - pos[3 * q], dy = xc[1] - pos[3 * q + 1], dz = xc[2] - pos[3 * q + 2];
          double d = sqrt(dx * dx + dy * dy + dz * dz);
          if (d > rimg) continue;
          acc += pref[q] * exp(-zeta[q] * d);
        }
        out[(int64_t)i + (int64_t)n[0] * ((int64_t)j + (int64_t)n[1] * k)] = acc;
      }
  free(pos); free(pref); free(zeta);
}

/* The promolecular branch of nciplot's per-point loop over many nodes,
 * parallel as the reference's (nci@proc.f90:543, !$omp parallel do
 * schedule(dynamic)): per node rho, |grad rho| (zero within 1e-5 bohr of a
 * nucleus, fieldmod@proc.f90:1146-1155), s and sign(rho,lambda2)*100 from the
 * full Hessian.  out: 4 doubles per node (rho, |grad|, s, crho).  The CPU
 * baseline leg times it. */
void oracle_promol_nci_batch(int64_t npts, const double *pts, int nimg, const double *ximg,
                             const int *zidx, const orc_rgrid *tabs, const double *cut,
                             int full_hessian, double *out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t q = 0; q < npts; q++) {
    const double *x = pts + 3 * q;
    double f, g[3], h[9];
    oracle_promol(x, nimg, ximg, zidx, tabs, cut, 2, &f, g, h);
    double dmin = 1e300;
    for (int ii = 0; ii < nimg; ii++) {
      double d0 = x[0] - ximg[3 * ii], d1 = x[1] - ximg[3 * ii + 1], d2 = x[2] - ximg[3 * ii + 2];
      double d = d0 * d0 + d1 * d1 + d2 * d2;
      if (d < dmin) dmin = d;
    }
    double gm = (sqrt(dmin) <= 1e-5) ? 0.0 : sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    double cr, s;
    oracle_nci_point(f, gm, h, full_hessian, &cr, &s);
    out[4 * q] = f;
    out[4 * q + 1] = gm;
    out[4 * q + 2] = s;
    out[4 * q + 3] = cr;
  }
}
