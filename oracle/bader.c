/*
 * bader.c -- CPU restatement of critic2's near-grid Bader integration.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  Loaded by tests/, smoke() and the
 * cpu_baseline leg of bench.py as the checker; never shipped as the product.
 *
 * Follows critic2's src/bader@proc.f90 statement by statement:
 *   bader_integrate       bader@proc.f90:80-234
 *   refine_edge           bader@proc.f90:300-422
 *   max_neargrid          bader@proc.f90:427-450
 *   step_neargrid         bader@proc.f90:455-494   (dr is SAVE state)
 *   step_ongrid           bader@proc.f90:500-527
 *   rho_grad_dir          bader@proc.f90:532-567
 *   is_max                bader@proc.f90:571-597
 *   pbc/rho_val/volnum_val bader@proc.f90:601-660
 *   assign_surrounding_pts/known_volnum_ongrid bader@proc.f90:665-726
 *   is_vol_edge           bader@proc.f90:730-752
 *   reassign_volnum_ongrid2 bader@proc.f90:756-771
 *   matinv (dgetrf2/dgetri/dtrti2) tools_math@proc.f90:1203-1235,
 *     src/oldlibs/dgetrf2.f, dgetri.f, dtrti2.f
 *   identify_atom/nearest_atom crystalmod@env.f90:516-614
 *   are_lclose/shortest   crystalmod@proc.f90:1054-1133
 *
 * Indices are 0-based; the iteration order (k innermost, then j, then i) and
 * every comparison, rounding and accumulation order is kept.  Compile without
 * FP contraction (-ffp-contract=off): the reference build emits no FMA.
 *
 * Pinning (tests/test_oracle_bader_pins.py, tests/test_oracle_ref.py,
 * tests/test_oracle.py): car2lat equals the reference's own LAPACK
 * (src/oldlibs dgetrf/dgetri, built unchanged into oracle/_ref) bit for bit;
 * hand-traced tiny grids decide labels through each order-dependent rule:
 * the carried dr, nint at +-0.5, the known == 2 stop, relabelling by later
 * paths, step_ongrid's first-maximum tie rule, refine pass 1 reassignment
 * and pass 2's known(pt)-after-loop read (through the number of points it
 * re-ascends); every mutation switch below fails one of them.  Also: one
 * maximum gives one basin, counts sum to N and populations to the grid
 * sum, the attractor rules follow hand-worked cases.  The reference's own
 * Bader golden outputs need input grids its checkout does not hold.
 *
 * qcksort_r8 (the YT ranking): pinned by two hand-traced arrays (insertion
 * sort; one partition) and sortedness of large tied arrays
 * (tests/test_yt_cpu.py); its pivot sequence past the first partition is
 * parity unpinned (the reference's tools module does not build here).
 */
#include "c2oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

typedef struct {
  const double *f;
  int n[3];
  int64_t N;
  double car2lat[9];
  double lid[27];
  int32_t *volnum;
  int32_t *known;
  int *path; /* 3*cap */
  int64_t pcap;
  int64_t pnum;
  double dr[3];
  int nbasin;
  int64_t nsteps;
} bstate;

/* Mutation switches (tests only): each bit replaces one rule of the
 * reference by a plausible mistake, so that a test can show that the pins
 * of the Bader restatement fail on it.  0 (the default) is the reference. */
static int orc_mut = 0;
void oracle_set_mutation(int mask) { orc_mut = mask; }

/* Points re-ascended by each refine_edge call of the last run (the negative
 * points of its reassign loop, bader@proc.f90:391-417); tests only. */
#define ORC_MAXPASS 64
static int64_t orc_reascend[ORC_MAXPASS];
static int orc_npass = 0;
int oracle_refine_counts(int64_t *out, int max) {
  int m = orc_npass < max ? orc_npass : max;
  for (int q = 0; q < m; q++) out[q] = orc_reascend[q];
  return orc_npass;
}
int oracle_get_mutation(void) { return orc_mut; }

/* Reassign paths (all passes of the last run) that ended on another point
 * which was negative when its pass's reassign loop began; tests only. */
static int64_t orc_negend = 0;
int64_t oracle_refine_negend(void) {
  const int64_t v = orc_negend;
  orc_negend = 0;
  return v;
}

static inline int wrap(int p, int n) {
  /* pbc: repeated +n / -n until 1..n (0..n-1 here) */
  while (p < 0) p += n;
  while (p >= n) p -= n;
  return p;
}
static inline int64_t lin(const bstate *s, int i, int j, int k) {
  return (int64_t)i + (int64_t)s->n[0] * ((int64_t)j + (int64_t)s->n[1] * k);
}
static inline double rho_val(const bstate *s, int i, int j, int k) {
  return s->f[lin(s, wrap(i, s->n[0]), wrap(j, s->n[1]), wrap(k, s->n[2]))];
}
static inline int32_t volnum_val(const bstate *s, int i, int j, int k) {
  return s->volnum[lin(s, wrap(i, s->n[0]), wrap(j, s->n[1]), wrap(k, s->n[2]))];
}
/* Fortran nint: round half away from 0 (ORC_MUT_RINT: half to even) */
static inline long fnint(double x) { return (orc_mut & ORC_MUT_RINT) ? lrint(x) : lround(x); }

/* rho_grad_dir, bader@proc.f90:532-567 */
static void rho_grad_dir(const bstate *s, const int p[3], double res[3]) {
  int p1 = p[0], p2 = p[1], p3 = p[2];
  double rho000 = rho_val(s, p1, p2, p3);
  double rho001 = rho_val(s, p1, p2, p3 + 1);
  double rho010 = rho_val(s, p1, p2 + 1, p3);
  double rho100 = rho_val(s, p1 + 1, p2, p3);
  double rho00_1 = rho_val(s, p1, p2, p3 - 1);
  double rho_100 = rho_val(s, p1 - 1, p2, p3);
  double rho0_10 = rho_val(s, p1, p2 - 1, p3);
  double gl[3], gc[3];
  gl[0] = (rho100 - rho_100) / 2.0;
  gl[1] = (rho010 - rho0_10) / 2.0;
  gl[2] = (rho001 - rho00_1) / 2.0;
  if (rho100 < rho000 && rho_100 < rho000) gl[0] = 0.0;
  if (rho010 < rho000 && rho0_10 < rho000) gl[1] = 0.0;
  if (rho001 < rho000 && rho00_1 < rho000) gl[2] = 0.0;
  /* rho_grad_car = matmul(rho_grad_lat, car2lat): gc(j) = sum_i gl(i)*c(i,j) */
  const double *c = s->car2lat;
  for (int j = 0; j < 3; j++) {
    double acc = 0.0;
    for (int i = 0; i < 3; i++) acc = acc + gl[i] * c[i + 3 * j];
    gc[j] = acc;
  }
  /* res = matmul(car2lat, rho_grad_car): res(i) = sum_j c(i,j)*gc(j) */
  for (int i = 0; i < 3; i++) {
    double acc = 0.0;
    for (int j = 0; j < 3; j++) acc = acc + c[i + 3 * j] * gc[j];
    res[i] = acc;
  }
}

/* is_max, bader@proc.f90:571-597 */
static int is_max(const bstate *s, const int p[3]) {
  double rho = rho_val(s, p[0], p[1], p[2]);
  for (int d1 = -1; d1 <= 1; d1++)
    for (int d2 = -1; d2 <= 1; d2++)
      for (int d3 = -1; d3 <= 1; d3++)
        if (rho_val(s, p[0] + d1, p[1] + d2, p[2] + d3) > rho) return 0;
  return 1;
}

/* step_ongrid, bader@proc.f90:500-527 */
static void step_ongrid(const bstate *s, int p[3]) {
  int pm[3] = {p[0], p[1], p[2]};
  double rho_ctr = rho_val(s, p[0], p[1], p[2]);
  double rho_max = rho_ctr;
  for (int d1 = -1; d1 <= 1; d1++)
    for (int d2 = -1; d2 <= 1; d2++)
      for (int d3 = -1; d3 <= 1; d3++) {
        int pt0 = p[0] + d1, pt1 = p[1] + d2, pt2 = p[2] + d3;
        double rho_tmp = rho_val(s, pt0, pt1, pt2);
        rho_tmp = rho_ctr + (rho_tmp - rho_ctr) * s->lid[(d1 + 1) + 3 * (d2 + 1) + 9 * (d3 + 1)];
        if (rho_tmp > rho_max || ((orc_mut & ORC_MUT_ONGRID_GE) && rho_tmp == rho_max && rho_tmp > rho_ctr)) {
          rho_max = rho_tmp;
          pm[0] = pt0; pm[1] = pt1; pm[2] = pt2;
        }
      }
  p[0] = wrap(pm[0], s->n[0]);
  p[1] = wrap(pm[1], s->n[1]);
  p[2] = wrap(pm[2], s->n[2]);
}

/* step_neargrid, bader@proc.f90:455-494 */
static void step_neargrid(bstate *s, int p[3]) {
  double g[3];
  int pm[3];
  s->nsteps++;
  if (s->pnum == 1 || (orc_mut & ORC_MUT_DR_RESET)) { s->dr[0] = s->dr[1] = s->dr[2] = 0.0; }
  rho_grad_dir(s, p, g);
  double gmax = fabs(g[0]);
  if (fabs(g[1]) > gmax) gmax = fabs(g[1]);
  if (fabs(g[2]) > gmax) gmax = fabs(g[2]);
  if (gmax < 1e-30) {
    s->dr[0] = s->dr[1] = s->dr[2] = 0.0;
    if (is_max(s, p)) return;
    pm[0] = p[0]; pm[1] = p[1]; pm[2] = p[2];
    step_ongrid(s, pm);
  } else {
    double coeff = 1.0 / gmax;
    for (int c = 0; c < 3; c++) {
      g[c] = coeff * g[c];
      long ng = fnint(g[c]);
      pm[c] = p[c] + (int)ng;
      s->dr[c] = (s->dr[c] + g[c]) - (double)ng;
      long nd = fnint(s->dr[c]);
      pm[c] = pm[c] + (int)nd;
      s->dr[c] = s->dr[c] - (double)nd;
    }
  }
  s->known[lin(s, p[0], p[1], p[2])] = 1;
  pm[0] = wrap(pm[0], s->n[0]);
  pm[1] = wrap(pm[1], s->n[1]);
  pm[2] = wrap(pm[2], s->n[2]);
  if (s->known[lin(s, pm[0], pm[1], pm[2])] == 1 && !(orc_mut & ORC_MUT_NO_LOOP_CHECK)) {
    pm[0] = p[0]; pm[1] = p[1]; pm[2] = p[2];
    step_ongrid(s, pm);
    s->dr[0] = s->dr[1] = s->dr[2] = 0.0;
  }
  p[0] = pm[0]; p[1] = pm[1]; p[2] = pm[2];
}

static void path_push(bstate *s, const int p[3]) {
  if (s->pnum >= s->pcap) {
    s->pcap = 2 * (s->pnum + 1);
    s->path = (int *)realloc(s->path, sizeof(int) * 3 * s->pcap);
  }
  s->path[3 * s->pnum + 0] = p[0];
  s->path[3 * s->pnum + 1] = p[1];
  s->path[3 * s->pnum + 2] = p[2];
  s->pnum++;
}

/* max_neargrid, bader@proc.f90:427-450 */
static void max_neargrid(bstate *s, int p[3]) {
  s->pnum = 0;
  path_push(s, p); /* pnum = 1 */
  for (;;) {
    step_neargrid(s, p);
    const int *last = &s->path[3 * (s->pnum - 1)];
    if (p[0] == last[0] && p[1] == last[1] && p[2] == last[2]) break;
    path_push(s, p);
    if (s->known[lin(s, p[0], p[1], p[2])] == 2 && !(orc_mut & ORC_MUT_NO_KNOWN2_STOP)) break;
    if (orc_mut && s->pnum > 4 * s->N + 16) break; /* a mutant may cycle: stop it */
  }
}

/* known_volnum_ongrid, bader@proc.f90:705-726 */
static void known_volnum_ongrid(bstate *s, const int p[3]) {
  int p1 = p[0], p2 = p[1], p3 = p[2];
  int32_t v = volnum_val(s, p1, p2, p3);
  if (v <= 0) return;
  if (volnum_val(s, p1, p2, p3 + 1) != v) return;
  if (volnum_val(s, p1, p2, p3 - 1) != v) return;
  if (volnum_val(s, p1, p2 + 1, p3) != v) return;
  if (volnum_val(s, p1, p2 - 1, p3) != v) return;
  if (volnum_val(s, p1 + 1, p2, p3) != v) return;
  if (volnum_val(s, p1 - 1, p2, p3) != v) return;
  s->known[lin(s, p1, p2, p3)] = 2;
}

/* assign_surrounding_pts, bader@proc.f90:665-700 */
static void assign_surrounding_pts(bstate *s, const int p[3]) {
  static const int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0},
                                {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  for (int t = 0; t < 6; t++) {
    int pt[3] = {wrap(p[0] + off[t][0], s->n[0]), wrap(p[1] + off[t][1], s->n[1]),
                 wrap(p[2] + off[t][2], s->n[2])};
    if (s->known[lin(s, pt[0], pt[1], pt[2])] != 2) known_volnum_ongrid(s, pt);
  }
}

/* is_vol_edge, bader@proc.f90:730-752 */
static int is_vol_edge(const bstate *s, const int p[3]) {
  int32_t v = s->volnum[lin(s, p[0], p[1], p[2])];
  for (int d1 = -1; d1 <= 1; d1++)
    for (int d2 = -1; d2 <= 1; d2++)
      for (int d3 = -1; d3 <= 1; d3++) {
        int32_t vn = volnum_val(s, p[0] + d1, p[1] + d2, p[2] + d3);
        if (abs(vn) != abs(v)) return 1;
      }
  return 0;
}

/* reassign_volnum_ongrid2, bader@proc.f90:756-771 */
static void reassign_volnum_ongrid2(bstate *s, const int p[3]) {
  for (int d1 = -1; d1 <= 1; d1++)
    for (int d2 = -1; d2 <= 1; d2++)
      for (int d3 = -1; d3 <= 1; d3++)
        s->known[lin(s, wrap(p[0] + d1, s->n[0]), wrap(p[1] + d2, s->n[1]),
                     wrap(p[2] + d3, s->n[2]))] = 0;
}

/* refine_edge, bader@proc.f90:300-422.  Returns num_reassign, or -1 on the
 * "should be no new maxima in edge refinement" fatal error. */
static int64_t refine_edge(bstate *s, int ref_itrs) {
  const int n1 = s->n[0], n2 = s->n[1], n3 = s->n[2];
  int p[3];
  if (ref_itrs == 1) {
    for (int i = 0; i < n1; i++)
      for (int j = 0; j < n2; j++)
        for (int k = 0; k < n3; k++) {
          int64_t id = lin(s, i, j, k);
          p[0] = i; p[1] = j; p[2] = k;
          if (s->volnum[id] == s->nbasin + 1) continue;
          if (is_vol_edge(s, p) && !is_max(s, p)) {
            s->volnum[id] = -s->volnum[id];
            s->known[id] = 0;
            reassign_volnum_ongrid2(s, p);
          }
        }
  } else {
    for (int i = 0; i < n1; i++)
      for (int j = 0; j < n2; j++)
        for (int k = 0; k < n3; k++) {
          int64_t id = lin(s, i, j, k);
          p[0] = i; p[1] = j; p[2] = k;
          if (s->volnum[id] == s->nbasin + 1) continue;
          if (s->volnum[id] < 0 && s->known[id] != -1) {
            int pt[3] = {0, 0, 0};
            for (int d1 = -1; d1 <= 1; d1++)
              for (int d2 = -1; d2 <= 1; d2++)
                for (int d3 = -1; d3 <= 1; d3++) {
                  pt[0] = wrap(i + d1, n1); pt[1] = wrap(j + d2, n2); pt[2] = wrap(k + d3, n3);
                  int64_t it = lin(s, pt[0], pt[1], pt[2]);
                  if (s->volnum[it] == s->nbasin + 1) continue;
                  if (!is_max(s, pt)) {
                    if (s->volnum[it] > 0) {
                      s->volnum[it] = -s->volnum[it];
                      s->known[it] = -1;
                    } else if (s->volnum[it] < 0 && s->known[it] == 0) {
                      s->known[it] = -2;
                    }
                  }
                }
            /* pt keeps its last value, p + (1,1,1) wrapped (bader@proc.f90:359) */
            if (orc_mut & ORC_MUT_NO_PT_QUIRK) { pt[0] = i; pt[1] = j; pt[2] = k; }
            if (s->known[lin(s, pt[0], pt[1], pt[2])] != -2)
              s->volnum[id] = abs(s->volnum[id]);
          }
        }
    for (int i = 0; i < n1; i++)
      for (int j = 0; j < n2; j++)
        for (int k = 0; k < n3; k++) {
          if (s->volnum[lin(s, i, j, k)] < 0) {
            for (int d1 = -1; d1 <= 1; d1++)
              for (int d2 = -1; d2 <= 1; d2++)
                for (int d3 = -1; d3 <= 1; d3++) {
                  int64_t it = lin(s, wrap(i + d1, n1), wrap(j + d2, n2), wrap(k + d3, n3));
                  if (s->known[it] == 2) s->known[it] = 0;
                }
          }
        }
  }

  int64_t num_reassign = 0;
  {
    int64_t neg = 0;
    for (int64_t q = 0; q < s->N; q++) neg += s->volnum[q] < 0;
    if (ref_itrs <= ORC_MAXPASS) orc_reascend[ref_itrs - 1] = neg;
    orc_npass = ref_itrs;
  }
  /* the signs at the start of the reassign loop, for the count of paths
   * that end on another point that was negative then (tests only) */
  uint8_t *neg0 = (uint8_t *)malloc((size_t)s->N);
  for (int64_t q = 0; q < s->N; q++) neg0[q] = s->volnum[q] < 0;
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n2; j++)
      for (int k = 0; k < n3; k++) {
        int64_t id = lin(s, i, j, k);
        int32_t bvolnum = s->volnum[id];
        if (bvolnum < 0) {
          p[0] = i; p[1] = j; p[2] = k;
          max_neargrid(s, p);
          const int64_t ie = lin(s, p[0], p[1], p[2]);
          if (ie != id && neg0[ie]) orc_negend++;
          int32_t pv = s->volnum[ie];
          if (pv < 0 || pv > s->nbasin) {
            free(neg0);
            return -1;
          }
          s->volnum[id] = pv;
          if (abs(bvolnum) != pv) {
            num_reassign++;
            s->volnum[id] = -pv;
          }
          for (int64_t l = 0; l < s->pnum; l++) {
            int64_t it = lin(s, s->path[3 * l], s->path[3 * l + 1], s->path[3 * l + 2]);
            if (s->known[it] != 2) s->known[it] = 0;
          }
        }
      }
  free(neg0);
  return num_reassign;
}

/* ---- crystal geometry used by attractor assignment ---- */

static void x2c_apply(const double m[9], const double x[3], double out[3]) {
  /* matmul(m_x2c, x): out(i) = sum_j m(i,j)*x(j) */
  for (int i = 0; i < 3; i++) {
    double acc = 0.0;
    for (int j = 0; j < 3; j++) acc = acc + m[i + 3 * j] * x[j];
    out[i] = acc;
  }
}

/* Shortest lattice-translated distance of a fractional difference vector.
 * Brute force over translations -2..2 after folding by nint (exact minimum
 * for the reduced cells used here). */
static double lattice_dist(const double x2c[9], const double xd[3]) {
  double x0[3];
  for (int c = 0; c < 3; c++) x0[c] = xd[c] - (double)lround(xd[c]);
  double best = DBL_MAX;
  for (int a = -2; a <= 2; a++)
    for (int b = -2; b <= 2; b++)
      for (int c = -2; c <= 2; c++) {
        double xt[3] = {x0[0] + a, x0[1] + b, x0[2] + c}, xc[3];
        x2c_apply(x2c, xt, xc);
        double d = sqrt(xc[0] * xc[0] + xc[1] * xc[1] + xc[2] * xc[2]);
        if (d < best) best = d;
      }
  return best;
}

/* identify_atom(x, icrd_crys, distmax): nearest atom id (1-based) within
 * distmax, first one in list order on exact ties; 0 if none. */
static int identify_atom(const orc_geom *g, const double x[3], double distmax) {
  int best = 0;
  double bd = DBL_MAX;
  const int natoms = g->nat_atoms > 0 && g->nat_atoms < g->nat ? g->nat_atoms : g->nat;
  for (int a = 0; a < natoms; a++) {
    double xd[3] = {x[0] - g->xat[3 * a], x[1] - g->xat[3 * a + 1], x[2] - g->xat[3 * a + 2]};
    double d = lattice_dist(g->x2c, xd);
    if (d < bd) { bd = d; best = a + 1; }
  }
  if (best == 0) return 0;
  if (distmax < 1.0) {
    return (bd <= distmax) ? best : 0; /* list_near_atoms up2d */
  }
  return (bd > distmax) ? 0 : best;
}

static int are_lclose(const orc_geom *g, const double x0[3], const double x1[3], double eps) {
  double xd[3] = {x0[0] - x1[0], x0[1] - x1[1], x0[2] - x1[2]};
  return lattice_dist(g->x2c, xd) < eps;
}

/* ---- 3x3 matinv: dgetrf2 + dgetri (dtrti2 + unblocked loop) ---- */
#define A_(r, c) a[(r) + 3 * (c)]
static void matinv3(double *a) {
  int ipiv[3];
  /* dgetrf2 recursion for 3x3: n1=1 panel, then 2x2 (n1=1), then 1x1 */
  /* panel column 0, rows 0..2 */
  {
    int ip = 0;
    double mx = fabs(A_(0, 0));
    for (int r = 1; r < 3; r++)
      if (fabs(A_(r, 0)) > mx) { mx = fabs(A_(r, 0)); ip = r; }
    ipiv[0] = ip;
    if (A_(ip, 0) != 0.0) {
      if (ip != 0) { double t = A_(0, 0); A_(0, 0) = A_(ip, 0); A_(ip, 0) = t; }
      if (fabs(A_(0, 0)) >= DBL_MIN) {
        double rc = 1.0 / A_(0, 0);
        for (int r = 1; r < 3; r++) A_(r, 0) = rc * A_(r, 0);
      } else {
        for (int r = 1; r < 3; r++) A_(r, 0) = A_(r, 0) / A_(0, 0);
      }
    }
    /* dlaswp on columns 1..2 */
    if (ip != 0)
      for (int c = 1; c < 3; c++) { double t = A_(0, c); A_(0, c) = A_(ip, c); A_(ip, c) = t; }
    /* dgemm: A22 -= A21*A12 (k=1): C(i,j) = C(i,j) + (-B(0,j))*A(i,0) */
    for (int c = 1; c < 3; c++) {
      double temp = -A_(0, c);
      for (int r = 1; r < 3; r++) A_(r, c) = A_(r, c) + temp * A_(r, 0);
    }
  }
  /* 2x2 sub-block rows/cols 1..2 */
  {
    int ip = 1;
    double mx = fabs(A_(1, 1));
    if (fabs(A_(2, 1)) > mx) { ip = 2; }
    ipiv[1] = ip;
    if (A_(ip, 1) != 0.0) {
      if (ip != 1) { double t = A_(1, 1); A_(1, 1) = A_(ip, 1); A_(ip, 1) = t; }
      if (fabs(A_(1, 1)) >= DBL_MIN) {
        double rc = 1.0 / A_(1, 1);
        A_(2, 1) = rc * A_(2, 1);
      } else {
        A_(2, 1) = A_(2, 1) / A_(1, 1);
      }
    }
    if (ip != 1) { double t = A_(1, 2); A_(1, 2) = A_(ip, 2); A_(ip, 2) = t; }
    double temp = -A_(1, 2);
    A_(2, 2) = A_(2, 2) + temp * A_(2, 1);
    ipiv[2] = 2;
    /* dlaswp of the 2x2 pivots applied to column 0 (left of the sub-block) */
    if (ip != 1) { double t = A_(1, 0); A_(1, 0) = A_(ip, 0); A_(ip, 0) = t; }
  }
  /* dtrti2 upper non-unit */
  for (int j = 0; j < 3; j++) {
    A_(j, j) = 1.0 / A_(j, j);
    double ajj = -A_(j, j);
    /* dtrmv upper, no-trans, non-unit, order j, x = A(0:j-1, j) */
    for (int jj = 0; jj < j; jj++) {
      if (A_(jj, j) != 0.0) {
        double temp = A_(jj, j);
        for (int i = 0; i < jj; i++) A_(i, j) = A_(i, j) + temp * A_(i, jj);
        A_(jj, j) = A_(jj, j) * A_(jj, jj);
      }
    }
    for (int i = 0; i < j; i++) A_(i, j) = ajj * A_(i, j);
  }
  /* dgetri unblocked: solve inv(A)*L = inv(U) */
  double work[3];
  for (int j = 2; j >= 0; j--) {
    for (int i = j + 1; i < 3; i++) { work[i] = A_(i, j); A_(i, j) = 0.0; }
    if (j < 2) {
      /* dgemv N: y = y + (-1)*A(:,j+1:n)*work(j+1:n), column by column */
      for (int jj = j + 1; jj < 3; jj++) {
        double temp = -work[jj];
        for (int i = 0; i < 3; i++) A_(i, j) = A_(i, j) + temp * A_(i, jj);
      }
    }
  }
  for (int j = 1; j >= 0; j--) {
    int jp = ipiv[j];
    if (jp != j)
      for (int r = 0; r < 3; r++) { double t = A_(r, j); A_(r, j) = A_(r, jp); A_(r, jp) = t; }
  }
}
#undef A_

void oracle_matinv3(double a[9]) { matinv3(a); }

void oracle_bader_metrics(const double x2c[9], const int n[3], double lat2car[9],
                          double car2lat[9], double lat_i_dist[27]) {
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) lat2car[r + 3 * c] = x2c[r + 3 * c] / (double)n[c];
  memcpy(car2lat, lat2car, 9 * sizeof(double));
  matinv3(car2lat);
  for (int i = -1; i <= 1; i++)
    for (int j = -1; j <= 1; j++)
      for (int k = -1; k <= 1; k++) {
        double dlat[3] = {(double)i, (double)j, (double)k}, dcar[3];
        x2c_apply(lat2car, dlat, dcar);
        double d = sqrt(dcar[0] * dcar[0] + dcar[1] * dcar[1] + dcar[2] * dcar[2]);
        int id = (i + 1) + 3 * (j + 1) + 9 * (k + 1);
        lat_i_dist[id] = (i == 0 && j == 0 && k == 0) ? 0.0 : 1.0 / d;
      }
}

/* First sweep of bader_integrate, bader@proc.f90:109-215. */
static int first_sweep(bstate *s, const orc_geom *g, double **xattr, int *xcap) {
  const int n1 = s->n[0], n2 = s->n[1], n3 = s->n[2];
  int atexist = (g->mode != ORC_ATTR_NOATOMS);
  double ratom = (g->mode == ORC_ATTR_NEAREST_ATOM) ? 1e40 : g->ratom;
  s->nbasin = 0;
  if (atexist) {
    s->nbasin = g->nat;
    for (int a = 0; a < g->nat; a++)
      for (int c = 0; c < 3; c++) (*xattr)[3 * a + c] = g->xat[3 * a + c];
  }
  int p[3];
  int64_t nstarts = 0;
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n2; j++)
      for (int k = 0; k < n3; k++) {
        if (s->volnum[lin(s, i, j, k)] != 0) continue;
        nstarts++;
        p[0] = i; p[1] = j; p[2] = k;
        max_neargrid(s, p);
        int32_t pv = s->volnum[lin(s, p[0], p[1], p[2])];
        if (pv == 0) {
          double dv[3] = {(double)p[0] / (double)n1, (double)p[1] / (double)n2,
                          (double)p[2] / (double)n3};
          int assigned = 0;
          if (atexist) {
            int nid = identify_atom(g, dv, ratom);
            if (nid > 0) { pv = nid; assigned = 1; }
          }
          if (!assigned && ratom > 1e-40) { /* vsmall = 1d-40 (param.F90) */
            for (int l = 0; l < s->nbasin; l++)
              if (are_lclose(g, dv, &(*xattr)[3 * l], ratom)) { pv = l + 1; assigned = 1; break; }
          }
          if (!assigned) {
            /* DISCARD: the expression at x = s%c%x2c(dv), evaluated each
             * time a path ends at this (still unlabelled) maximum */
            int ok = 1;
            if (g->discard) {
              double xc[3];
              x2c_apply(g->x2c, dv, xc);
              ok = !g->discard(xc, g->discard_user);
            }
            if (ok) {
              s->nbasin++;
              if (s->nbasin > *xcap) {
                *xcap = 2 * s->nbasin;
                *xattr = (double *)realloc(*xattr, sizeof(double) * 3 * (*xcap));
              }
              pv = s->nbasin;
              for (int c = 0; c < 3; c++) (*xattr)[3 * (s->nbasin - 1) + c] = dv[c];
            }
          }
        }
        for (int64_t l = 0; l < s->pnum; l++) {
          int pt[3] = {s->path[3 * l], s->path[3 * l + 1], s->path[3 * l + 2]};
          int64_t it = lin(s, pt[0], pt[1], pt[2]);
          if (s->volnum[it] != -1 && !((orc_mut & ORC_MUT_KEEP_LABEL) && s->volnum[it] != 0)) s->volnum[it] = pv;
          if (s->known[it] != 2) s->known[it] = 0;
          assign_surrounding_pts(s, pt);
        }
      }
  return (int)(nstarts > 0x7fffffff ? 0x7fffffff : nstarts);
}

static int run_bader(const double *f, const int n[3], const double car2lat[9],
                     const double lat_i_dist[27], const orc_geom *geom,
                     int32_t *volnum_out, int32_t *known_out, double *xattr_out,
                     int xattr_cap, orc_bader_info *info, int do_refine) {
  bstate s;
  memset(&s, 0, sizeof(s));
  s.f = f;
  s.n[0] = n[0]; s.n[1] = n[1]; s.n[2] = n[2];
  s.N = (int64_t)n[0] * n[1] * n[2];
  memcpy(s.car2lat, car2lat, sizeof(s.car2lat));
  memcpy(s.lid, lat_i_dist, sizeof(s.lid));
  s.volnum = volnum_out;
  memset(s.volnum, 0, sizeof(int32_t) * s.N);
  s.known = (int32_t *)calloc((size_t)s.N, sizeof(int32_t));
  s.pcap = 16;
  s.path = (int *)malloc(sizeof(int) * 3 * s.pcap);
  int xcap = geom->nat > 4 ? geom->nat : 4;
  double *xattr = (double *)malloc(sizeof(double) * 3 * xcap);

  int64_t nstarts = 0;
  orc_npass = 0;
  {
    /* count starts inside first_sweep through nsteps bookkeeping */
    nstarts = first_sweep(&s, geom, &xattr, &xcap);
  }
  int64_t sweep_steps = s.nsteps;
  int status = 0, iters = 0;
  if (do_refine && !(orc_mut & ORC_MUT_NO_REFINE)) {
    int ref_itrs = 1;
    for (;;) {
      int64_t nr = refine_edge(&s, ref_itrs);
      iters++;
      if (nr < 0) { status = 1; break; }
      if (nr == 0) break;
      ref_itrs++;
    }
  }
  if (known_out) memcpy(known_out, s.known, sizeof(int32_t) * s.N);
  if (info) {
    info->nattr = s.nbasin;
    info->refine_iters = iters;
    info->nstarts = nstarts;
    info->nsteps = sweep_steps;
    info->status = status;
    info->nmax_found = 0;
  }
  int ncopy = s.nbasin < xattr_cap ? s.nbasin : xattr_cap;
  if (xattr_out) memcpy(xattr_out, xattr, sizeof(double) * 3 * ncopy);
  free(xattr);
  free(s.known);
  free(s.path);
  return status;
}

int oracle_bader(const double *f, const int n[3], const double car2lat[9],
                 const double lat_i_dist[27], const orc_geom *geom,
                 int32_t *volnum_out, double *xattr_out, int xattr_cap,
                 orc_bader_info *info) {
  return run_bader(f, n, car2lat, lat_i_dist, geom, volnum_out, NULL, xattr_out,
                   xattr_cap, info, 1);
}

int oracle_bader_sweep_only(const double *f, const int n[3],
                            const double car2lat[9], const double lat_i_dist[27],
                            const orc_geom *geom, int32_t *volnum_out,
                            int32_t *known_out, double *xattr_out, int xattr_cap,
                            orc_bader_info *info) {
  return run_bader(f, n, car2lat, lat_i_dist, geom, volnum_out, known_out,
                   xattr_out, xattr_cap, info, 0);
}

void oracle_basin_sums(const int32_t *idg, int64_t npts, int nattr, int nfields,
                       const double *const *fields, int64_t *counts, double *sums) {
  for (int a = 0; a < nattr; a++) counts[a] = 0;
  for (int64_t q = 0; q < (int64_t)nattr * nfields; q++) sums[q] = 0.0;
  /* count(idg == i) and sum(fint, idg == i), one sequential pass per
   * attractor in memory order (integration@proc.f90:1215, :1296); the pass
   * order over attractors does not change any individual sum. */
  for (int64_t p = 0; p < npts; p++) {
    int32_t a = idg[p];
    if (a < 1 || a > nattr) continue;
    counts[a - 1]++;
    for (int fi = 0; fi < nfields; fi++) sums[(int64_t)fi * nattr + (a - 1)] += fields[fi][p];
  }
}

void oracle_relabel(int32_t *idg, int64_t npts, const int32_t *map, int nmap) {
  for (int64_t p = 0; p < npts; p++) {
    int32_t a = idg[p];
    if (a >= 1 && a <= nmap) idg[p] = map[a - 1];
  }
}

/* Step table of c2gpu_bader_step_pointers, from the restated routines:
 * step_ongrid target (p if is_max), first step_neargrid target of a fresh
 * path from p (dr = 0, only p known), and dr after that step. */
void oracle_step_pointers(const double *f, const int n[3], const double car2lat[9],
                          const double lat_i_dist[27], int64_t *ongrid, int64_t *near0,
                          double *dr1) {
  bstate s;
  memset(&s, 0, sizeof(s));
  s.f = f;
  s.n[0] = n[0]; s.n[1] = n[1]; s.n[2] = n[2];
  s.N = (int64_t)n[0] * n[1] * n[2];
  memcpy(s.car2lat, car2lat, sizeof(s.car2lat));
  memcpy(s.lid, lat_i_dist, sizeof(s.lid));
  s.known = (int32_t *)calloc((size_t)s.N, sizeof(int32_t));
  s.volnum = (int32_t *)calloc((size_t)s.N, sizeof(int32_t));
  for (int k = 0; k < n[2]; k++)
    for (int j = 0; j < n[1]; j++)
      for (int i = 0; i < n[0]; i++) {
        int64_t id = lin(&s, i, j, k);
        int p[3] = {i, j, k};
        if (is_max(&s, p)) ongrid[id] = id;
        else { step_ongrid(&s, p); ongrid[id] = lin(&s, p[0], p[1], p[2]); }
        p[0] = i; p[1] = j; p[2] = k;
        s.pnum = 1;
        step_neargrid(&s, p);
        near0[id] = lin(&s, p[0], p[1], p[2]);
        dr1[3 * id] = s.dr[0]; dr1[3 * id + 1] = s.dr[1]; dr1[3 * id + 2] = s.dr[2];
        s.known[id] = 0;
      }
  free(s.known);
  free(s.volnum);
}

/* ==================================================================== */
/* Yu-Trinkle integration: yt_integrate (yt@proc.f90:38-224) and        */
/* yt_weights (yt@proc.f90:399-531) with the sums of intgrid_fields     */
/* (integration@proc.f90:1208-1218, :1289-1300) for imtype_yt.          */
/* ==================================================================== */

/* f: the grid (x fastest).  vec (3 per vector, grid steps) and area: the
 * Voronoi-relevant vectors of the grid lattice and their facet areas
 * (grid3%vec, grid3%area).  Points are ranked by oracle_qcksort_r8, the
 * reference's sort, so equal values take its order.  Returns 0, or 2 if the
 * attractor list overflows xcap.  idg: the basin of each point in grid
 * order (0 on the interatomic surface; critic2 stores ibasin in sorted
 * order, yt@proc.f90:218).  vol[a] = sum of the weights of basin a+1,
 * sums[q*nattr + a] = sum of w*fields[q]; both unscaled (critic2
 * multiplies by Omega/N). */
/* qcksort_r8 (tools@proc.f90:83-168), the reference's quicksort: iord(first:last)
 * (1-based positions in the reference; 0-based here, as are the element
 * indices in iord) sorted by arr ascending.  The pivot position comes from a
 * linear congruential sequence (fx) advanced once per partition; the
 * partition moves elements through a hole, right scan then left scan;
 * partitions below m = 7 elements are insertion-sorted.  The arrangement of
 * equal keys is this sequence's, which is what yt_integrate's ranks depend
 * on. */
void oracle_qcksort_r8(const double *arr, int64_t *iord, int64_t first, int64_t last) {
  const int64_t m = 7;
  const int nstack = 50;
  const double fm = 7875.0, fa = 211.0, fc = 1663.0, fmi = 1.2698413e-4;
  int64_t istack[50];
  int jstack = 0;
  int64_t l = first, ir = last, i, j, na, iq;
  double fx = 0.0, a;
L10:
  if (ir - l < m) {
    for (j = l + 1; j <= ir; j++) {
      na = iord[j];
      a = arr[na];
      for (i = j - 1; i >= first; i--) {
        if (arr[iord[i]] <= a) goto L12;
        iord[i + 1] = iord[i];
      }
      i = first - 1;
    L12:
      iord[i + 1] = na;
    }
    if (jstack == 0) return;
    ir = istack[jstack - 1];
    l = istack[jstack - 2];
    jstack -= 2;
  } else {
    i = l;
    j = ir;
    fx = fmod(fx * fa + fc, fm);
    iq = l + (int64_t)((double)(ir - l + 1) * (fx * fmi));
    na = iord[iq];
    a = arr[na];
    iord[iq] = iord[l];
  L20:
  L21:
    if (j >= first) {
      if (a < arr[iord[j]]) {
        j--;
        goto L21;
      }
    }
    if (j <= i) {
      iord[i] = na;
      goto L30;
    }
    iord[i] = iord[j];
    i++;
  L22:
    if (i <= last) {
      if (a > arr[iord[i]]) {
        i++;
        goto L22;
      }
    }
    if (j <= i) {
      iord[j] = na;
      i = j;
      goto L30;
    }
    iord[j] = iord[i];
    j--;
    goto L20;
  L30:
    jstack += 2;
    if (jstack > nstack) return; /* the reference stops here ("Increase nstack") */
    if (ir - i >= i - l) {
      istack[jstack - 1] = ir;
      istack[jstack - 2] = i + 1;
      ir = i - 1;
    } else {
      istack[jstack - 1] = i - 1;
      istack[jstack - 2] = l;
      l = i + 1;
    }
  }
  goto L10;
}

int oracle_yt(const double *f, const int n[3], const orc_geom *g, int nvec, const int *vec, const double *area,
              int nfields, const double *const *fields, int32_t *idg, int *nattr_out, double *xattr, int xcap,
              double *vol, double *sums, double *wout, int wcap) {
  const int64_t nn = (int64_t)n[0] * n[1] * n[2];
  int64_t *io = malloc(sizeof(int64_t) * nn), *iio = malloc(sizeof(int64_t) * nn);
  int32_t *ibasin = calloc(nn, sizeof(int32_t)); /* by sorted position, as in yt_integrate */
  int *nlo = calloc(nn, sizeof(int));
  int64_t *inear = calloc(nn * (size_t)nvec, sizeof(int64_t));
  double *fnear = calloc(nn * (size_t)nvec, sizeof(double));
  int rc = 0;
  for (int64_t i = 0; i < nn; i++) io[i] = i;
  oracle_qcksort_r8(f, io, 0, nn - 1); /* from smaller to larger (yt@proc.f90:85-93) */
  for (int64_t i = 0; i < nn; i++) iio[io[i]] = i;
#define YT_NB(i, k)                                                                           \
  (wrap((int)((i) % n[0]) + vec[3 * (k)], n[0]) +                                             \
   n[0] * (wrap((int)(((i) / n[0]) % n[1]) + vec[3 * (k) + 1], n[1]) +                        \
           (int64_t)n[1] * wrap((int)((i) / ((int64_t)n[0] * n[1])) + vec[3 * (k) + 2], n[2])))
  int nattr = g->mode != ORC_ATTR_NOATOMS ? g->nat : 0;
  if (nattr > xcap) rc = 2;
  for (int a = 0; a < nattr && !rc; a++)
    for (int c = 0; c < 3; c++) xattr[3 * a + c] = g->xat[3 * a + c];
  const double ratom = g->mode == ORC_ATTR_NEAREST_ATOM ? 1e40 : g->ratom;
  const double vsmall = 1e-40;
  int64_t ihi[64];
  double chi[64];
  /* run over the points in order of decreasing density (yt@proc.f90:100-181) */
  for (int64_t ii = nn - 1; ii >= 0 && !rc; ii--) {
    const int64_t i = io[ii];
    int nhi = 0;
    double csum = 0.0;
    for (int k = 0; k < nvec; k++) {
      const int64_t j = YT_NB(i, k);
      const int64_t jj = iio[j];
      if (jj > ii) {
        ihi[nhi] = jj;
        chi[nhi] = fmax(area[k] * (f[j] - f[i]), vsmall);
        csum = csum + chi[nhi];
        nhi++;
      }
    }
    nlo[ii] = 0;
    if (nhi == 0) { /* a local maximum */
      const double dv[3] = {(double)(i % n[0]) / n[0], (double)((i / n[0]) % n[1]) / n[1],
                            (double)(i / ((int64_t)n[0] * n[1])) / n[2]};
      int assigned = 0;
      if (g->mode != ORC_ATTR_NOATOMS) {
        const int nid = identify_atom(g, dv, ratom);
        if (nid > 0) {
          ibasin[ii] = nid;
          assigned = 1;
        }
      }
      if (!assigned && ratom > vsmall) {
        for (int k = 0; k < nattr; k++)
          if (are_lclose(g, dv, xattr + 3 * k, ratom)) {
            ibasin[ii] = k + 1;
            assigned = 1;
            break;
          }
      }
      if (!assigned) {
        if (nattr + 1 > xcap) {
          rc = 2;
          break;
        }
        for (int c = 0; c < 3; c++) xattr[3 * nattr + c] = dv[c];
        nattr++;
        ibasin[ii] = nattr;
      }
    } else {
      int isias = ibasin[ihi[0]] == 0;
      for (int k = 0; k < nhi; k++) isias = isias || (ibasin[ihi[k]] != ibasin[ihi[0]]);
      if (!isias) {
        ibasin[ii] = ibasin[ihi[0]];
      } else {
        ibasin[ii] = 0;
        for (int k = 0; k < nhi; k++) {
          const int64_t kk = ihi[k];
          inear[(size_t)nvec * kk + nlo[kk]] = ii;
          fnear[(size_t)nvec * kk + nlo[kk]] = chi[k] / fmax(csum, vsmall);
          nlo[kk]++;
        }
      }
    }
  }
#undef YT_NB
  if (!rc) {
    for (int64_t i = 0; i < nn; i++) idg[i] = ibasin[iio[i]];
    /* weights of each basin (yt_weights) and their sums (intgrid_fields) */
    double *waux = malloc(sizeof(double) * nn);
    for (int m = 1; m <= nattr; m++) {
      for (int64_t j = 0; j < nn; j++) waux[j] = ibasin[j] == m ? 1.0 : 0.0;
      for (int64_t j = nn - 1; j >= 0; j--)
        if (fabs(waux[j]) > 0.0)
          for (int k = 0; k < nlo[j]; k++) {
            const int64_t t = inear[(size_t)nvec * j + k];
            waux[t] = waux[t] + fnear[(size_t)nvec * j + k] * waux[j];
          }
      double sv = 0.0;
      for (int64_t p = 0; p < nn; p++) sv += waux[iio[p]];
      vol[m - 1] = sv;
      if (wout && m <= wcap)  /* w(i,j,k) = waux(iio(p)), grid order */
        for (int64_t p = 0; p < nn; p++) wout[(size_t)(m - 1) * nn + p] = waux[iio[p]];
      for (int q = 0; q < nfields; q++) {
        double s = 0.0;
        for (int64_t p = 0; p < nn; p++) s += waux[iio[p]] * fields[q][p];
        sums[(size_t)q * nattr + (m - 1)] = s;
      }
    }
    free(waux);
  }
  *nattr_out = nattr;
  free(io);
  free(iio);
  free(ibasin);
  free(nlo);
  free(inear);
  free(fnear);
  return rc;
}

/* ==================================================================== */
/* Basin multipoles (integration@proc.f90:1340-1362, Bader branch):     */
/* shortest (crystalmod@proc.f90:1054-1085), tosphere and genylm /      */
/* genrlm_real (tools_math@proc.f90:273-406), in the reference's route  */
/* through spherical angles.                                            */
/* ==================================================================== */

#include <complex.h>

/* genylm: spherical harmonics with the Condon-Shortley phase, packed
 * j = l(l+1)+m (0-based), Masters and Richards-Dinger recursion. */
static void genylm(int lmax, const double tp[2], double complex *ylm) {
  const double fourpi = 12.566370614359172954;
  double x[64];
  double complex z[64];
  ylm[0] = 0.28209479177387814347;
  if (lmax == 0) return;
  const double sn = sin(tp[0]), cs = cos(tp[0]);
  for (int m = 1; m <= lmax; m++) {
    const double t1 = (double)m * tp[1];
    z[m] = cos(t1) + I * sin(t1);
  }
  for (int l = 1; l <= lmax; l++) {
    x[l] = (l % 2 == 0) ? 1.0 : -1.0;
    double dx = 0.0;
    for (int m = l; m >= 1; m--) {
      const double t1 = sqrt((double)((l + m) * (l - m + 1)));
      x[m - 1] = -(sn * dx + (double)(2 * m) * cs * x[m]) / t1;
      dx = sn * x[m] * t1;
    }
    double t1 = sn, sum = 0.0;
    for (int m = 1; m <= l; m++) {
      x[m] = t1 * x[m];
      sum = sum + x[m] * x[m];
      t1 = t1 * sn;
    }
    sum = 2.0 * sum + x[0] * x[0];
    t1 = sqrt((double)(2 * l + 1) / (fourpi * sum));
    int lm1 = l * (l + 1), lm2 = lm1;
    ylm[lm1] = t1 * x[0];
    for (int m = 1; m <= l; m++) {
      lm1++;
      lm2--;
      ylm[lm1] = t1 * x[m] * z[m];
      ylm[lm2] = conj(ylm[lm1]);
      if (m % 2 != 0) ylm[lm2] = -ylm[lm2];
    }
  }
}

/* genrlm_real: real regular solid harmonics from (r, theta, phi). */
void oracle_genrlm_real(int lmax, double r, const double tp[2], double *rrlm) {
  double complex rlm[(16 + 1) * (16 + 1)];
  const double pi = 3.14159265358979323846, sh = 1.0 / sqrt(2.0);
  genylm(lmax, tp, rlm);
  for (int l = 0; l <= lmax; l++) {
    const int imin = l * l, imax = (l + 1) * (l + 1);
    const double sc = sqrt(4.0 * pi / (2 * l + 1)) * pow(r, l);
    for (int q = imin; q < imax; q++) rlm[q] = rlm[q] * sc;
    rrlm[imin + l] = creal(rlm[imin + l]);
    for (int m = 1; m <= l; m++) {
      const int ip = imin + l + m, im = imin + l - m;
      const double iphas = (m % 2 == 0) ? 1.0 : -1.0;
      const double complex a = iphas * rlm[ip] + rlm[im];
      const double complex b = -iphas * I * rlm[ip] + I * rlm[im];
      rrlm[im] = sh * creal(a);
      rrlm[ip] = sh * creal(b);
    }
  }
}

/* tosphere */
void oracle_tosphere(const double v[3], double *r, double tp[2]) {
  const double eps = 1e-14, pi = 3.14159265358979323846;
  *r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  tp[0] = tp[1] = 0.0;
  if (*r > eps) {
    const double t1 = v[2] / *r;
    if (t1 >= 1.0) tp[0] = 0.0;
    else if (t1 <= -1.0) tp[0] = pi;
    else tp[0] = acos(t1);
    if (fabs(v[0]) > eps || fabs(v[1]) > eps) tp[1] = atan2(v[1], v[0]);
  }
}

/* mpole[a * nlm + q], a the basin (0-based), unscaled (critic2 multiplies
 * by Omega/N).  shortest: nint fold, then the shortest translate over
 * -2..2 (first strictly shorter). */
/* The Cartesian shortest lattice translate of p/n - xattr (crystal%shortest,
 * crystalmod@proc.f90:1054-1085), by the nint fold and a search over the
 * neighbouring translates. */
/* crystal%isortho for the cells used here: the lattice vectors are
 * mutually orthogonal (then the Wigner-Seitz cell is the box of 6 facets
 * across +-a, +-b, +-c that tools@proc.f90:842-849 tests for), to 1e-12 of
 * the product of their lengths. */
static int cell_is_ortho(const double x2c[9]) {
  const double *a = x2c, *b = x2c + 3, *c = x2c + 6;
  const double la = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  const double lb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  const double lc = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  const double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  const double ac = a[0] * c[0] + a[1] * c[1] + a[2] * c[2];
  const double bc = b[0] * c[0] + b[1] * c[1] + b[2] * c[2];
  return fabs(ab) <= 1e-12 * la * lb && fabs(ac) <= 1e-12 * la * lc && fabs(bc) <= 1e-12 * lb * lc;
}

/* crystal%shortest (crystalmod@proc.f90:1054-1085) of p/n - xa, Cartesian:
 * the nint fold; for a non-orthogonal cell then the first strictly shorter
 * of the translates x0 + lattice vector (Cartesian addition, as the
 * reference adds ws_ineighc), over the translates -2..2. */
static void shortest_offset(const int n[3], const double x2c[9], int64_t i, int64_t j, int64_t k,
                            const double *xa, double bv[3]) {
  double dv[3] = {(double)i / n[0] - xa[0], (double)j / n[1] - xa[1], (double)k / n[2] - xa[2]};
  for (int c = 0; c < 3; c++) dv[c] = dv[c] - (double)lround(dv[c]);
  double v[3];
  x2c_apply(x2c, dv, v);
  double best = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  bv[0] = v[0]; bv[1] = v[1]; bv[2] = v[2];
  if (cell_is_ortho(x2c)) return;
  for (int a = -2; a <= 2; a++)
    for (int b = -2; b <= 2; b++)
      for (int c = -2; c <= 2; c++) {
        double t[3];
        for (int e = 0; e < 3; e++) t[e] = v[e] + x2c[e] * a + x2c[e + 3] * b + x2c[e + 6] * c;
        const double d = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
        if (d < best) {
          best = d;
          bv[0] = t[0];
          bv[1] = t[1];
          bv[2] = t[2];
        }
      }
}

void oracle_multipoles(const int32_t *idg, const double *f, const int n[3], const double x2c[9], int nattr,
                       const double *xattr, int lmax, double *mpole) {
  const int nlm = (lmax + 1) * (lmax + 1);
  double rrlm[17 * 17];
  for (int q = 0; q < nattr * nlm; q++) mpole[q] = 0.0;
  for (int64_t k = 0; k < n[2]; k++)
    for (int64_t j = 0; j < n[1]; j++)
      for (int64_t i = 0; i < n[0]; i++) {
        const int64_t p = i + n[0] * (j + n[1] * k);
        const int ix = idg[p];
        if (ix < 1 || ix > nattr) continue;
        double bv[3];
        shortest_offset(n, x2c, i, j, k, &xattr[3 * (ix - 1)], bv);
        double r, tp[2];
        oracle_tosphere(bv, &r, tp);
        oracle_genrlm_real(lmax, r, tp, rrlm);
        for (int q = 0; q < nlm; q++) mpole[(ix - 1) * nlm + q] += rrlm[q] * f[p];
      }
}

/* The YT branch (integration@proc.f90:1314-1336): per basin m, the weights
 * w(m*N + p) of yt_weights (1 on its interior, the propagated surface
 * weight elsewhere); points with |w| < 1d-15 are skipped. */
void oracle_yt_multipoles(const double *w, const double *f, const int n[3], const double x2c[9], int nattr,
                          const double *xattr, int lmax, double *mpole) {
  const int nlm = (lmax + 1) * (lmax + 1);
  const int64_t N = (int64_t)n[0] * n[1] * n[2];
  double rrlm[17 * 17];
  for (int q = 0; q < nattr * nlm; q++) mpole[q] = 0.0;
  for (int m = 0; m < nattr; m++)
    for (int64_t i = 0; i < n[0]; i++)
      for (int64_t j = 0; j < n[1]; j++)
        for (int64_t k = 0; k < n[2]; k++) {
          const int64_t p = i + n[0] * (j + n[1] * k);
          const double wp = w[(size_t)m * N + p];
          if (fabs(wp) < 1e-15) continue;
          double bv[3];
          shortest_offset(n, x2c, i, j, k, &xattr[3 * m], bv);
          double r, tp[2];
          oracle_tosphere(bv, &r, tp);
          oracle_genrlm_real(lmax, r, tp, rrlm);
          for (int q = 0; q < nlm; q++) mpole[m * nlm + q] += rrlm[q] * f[p] * wp;
        }
}
