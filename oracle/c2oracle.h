/*
 * c2oracle.h -- CPU restatement of critic2's on-grid Bader and NCI path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in critic2_amd/.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  It is never the thing measured
 * as the GPU result and never a fallback for the product path.
 *
 * Conventions (same as critic2):
 *   - grids are Fortran column-major f(n1,n2,n3): index = i + n1*(j + n2*k),
 *     i (x) fastest, all indices 0-based here (critic2 uses 1-based).
 *   - 3x3 matrices are column-major: m[r + 3*c] == m(r+1,c+1) in Fortran.
 *   - lat_i_dist is indexed [(d1+1) + 3*(d2+1) + 9*(d3+1)], d in {-1,0,1}.
 */
#ifndef C2ORACLE_H
#define C2ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Attractor-assignment modes of BADER (integration@proc.f90:236-248). */
enum {
  ORC_ATTR_NEAREST_ATOM = 0, /* default (nonnm): atexist, ratom = 1d40     */
  ORC_ATTR_RATOM = 1,        /* NNM / RATOM r: atexist, ratom = r           */
  ORC_ATTR_NOATOMS = 2       /* NOATOMS: no atoms, maxima merged by ratom   */
};

typedef struct {
  double x2c[9];      /* lattice vectors as columns, bohr                   */
  int nat;            /* atoms in the cell (complete list, critic2 order)   */
  const double *xat;  /* fractional atom coordinates, 3*nat                 */
  int mode;           /* ORC_ATTR_*                                         */
  double ratom;       /* bohr                                               */
  /* DISCARD (bader@proc.f90:185-191): nonzero discards the new maximum at
   * Cartesian x; NULL: no expression */
  int (*discard)(const double x[3], void *user);
  void *discard_user;
  /* with atexist the known attractors are the field's critical points
   * cpcel(1:nat) (bader@proc.f90:109-117); the first nat_atoms are the
   * atoms identify_atom looks at (<= 0: all nat) */
  int nat_atoms;
} orc_geom;

/* Mutation switches of the Bader restatement (tests only; 0 = the
 * reference's rules).  Each replaces one rule by a plausible mistake:
 *   DR_RESET       dr reset at every step_neargrid, not only at a path start
 *   NO_KNOWN2_STOP max_neargrid does not stop at a known == 2 point
 *   RINT           nint rounds half to even instead of half away from zero
 *   NO_PT_QUIRK    refine pass >= 2 reads known(p), not known(pt) after the loop
 *   NO_REFINE      no refine_edge passes
 *   ONGRID_GE      step_ongrid takes the last of equal maxima (>= not >)
 *   NO_LOOP_CHECK  step_neargrid does not fall back on an on-grid step when
 *                  the next point is already on the path (known == 1)
 *   KEEP_LABEL     a committed path does not relabel already labelled points */
enum {
  ORC_MUT_DR_RESET = 1, ORC_MUT_NO_KNOWN2_STOP = 2, ORC_MUT_RINT = 4, ORC_MUT_NO_PT_QUIRK = 8,
  ORC_MUT_NO_REFINE = 16, ORC_MUT_ONGRID_GE = 32, ORC_MUT_NO_LOOP_CHECK = 64, ORC_MUT_KEEP_LABEL = 128
};
void oracle_set_mutation(int mask);
/* Number of points each refine_edge call of the last Bader run re-ascended
 * (its negative points); returns the number of calls. */
int oracle_refine_counts(int64_t *out, int max);
int oracle_get_mutation(void);

/* Outputs of oracle_bader. */
typedef struct {
  int nattr;          /* number of attractors (bas%nattr after bader)       */
  int nmax_found;     /* number of distinct grid maxima reached             */
  int refine_iters;   /* number of refine_edge calls                        */
  int64_t nstarts;    /* first-sweep paths started                          */
  int64_t nsteps;     /* first-sweep step_neargrid calls                    */
  int status;         /* 0 ok; 1 new maximum during refinement             */
} orc_bader_info;

/* bader_integrate (bader@proc.f90:80-234).  volnum_out: N int32 labels,
 * 1-based attractor ids.  xattr_out: capacity 3*xattr_cap fractional
 * attractor positions (discovery order, atoms first when atexist).
 * lat2car/car2lat/lat_i_dist are passed in so the caller fixes their bits. */
int oracle_bader(const double *f, const int n[3], const double car2lat[9],
                 const double lat_i_dist[27], const orc_geom *geom,
                 int32_t *volnum_out, double *xattr_out, int xattr_cap,
                 orc_bader_info *info);

/* Same, but stop after the first sweep (no refine_edge). Test hook. */
int oracle_bader_sweep_only(const double *f, const int n[3],
                            const double car2lat[9], const double lat_i_dist[27],
                            const orc_geom *geom, int32_t *volnum_out,
                            int32_t *known_out, double *xattr_out, int xattr_cap,
                            orc_bader_info *info);

/* lat2car(:,i) = x2c(:,i)/n(i); car2lat = inverse (LU with partial pivoting
 * and the dgetri column order); lat_i_dist as bader@proc.f90:124-145. */
void oracle_bader_metrics(const double x2c[9], const int n[3], double lat2car[9],
                          double car2lat[9], double lat_i_dist[27]);

/* Per-basin sums (integration@proc.f90:1208-1218, :1289-1300):
 * counts[a] = count(idg == a+1); sums[fi*nattr + a] = sum(field_fi, idg==a+1)
 * summed in column (memory) order, sequentially. */
void oracle_basin_sums(const int32_t *idg, int64_t npts, int nattr, int nfields,
                       const double *const *fields, int64_t *counts, double *sums);

/* Step table (see c2gpu_bader_step_pointers). */
void oracle_step_pointers(const double *f, const int n[3], const double car2lat[9],
                          const double lat_i_dist[27], int64_t *ongrid, int64_t *near0,
                          double *dr1);

/* Relabel: idg = map[idg-1] for idg >= 1 (int_reorder_gridout remap). */
void oracle_relabel(int32_t *idg, int64_t npts, const int32_t *map, int nmap);

/* Promolecular radial tables (grid1mod@proc.f90:206-331 semantics):
 * r[i] = a*exp(b*i), tables f, fp, fpp already built. */
typedef struct {
  int ngrid;
  double a, b, rmax;
  const double *r, *f, *fp, *fpp;
} orc_rgrid;

/* grid1%interp (grid1mod@proc.f90:86-137). */
void oracle_grid1_interp(const orc_rgrid *g, double r0, double *f, double *fp,
                         double *fpp);

/* Build f, fp, fpp from rr0 = sum(occ*psi^2) on the log grid
 * (grid1mod@proc.f90:281-313); ngrid is the already-truncated count. */
void oracle_grid1_build(int ngrid, const double *r, const double *rr0,
                        double *f, double *fp, double *fpp, double *a,
                        double *b);

/* Promolecular density, gradient, Hessian at Cartesian point xc
 * (crystalmod@env.f90:622-753), atoms listed with Cartesian positions of all
 * images already expanded (the caller chooses the image set). Each atom uses
 * table tab[zidx[i]] and cutoff cut[zidx[i]]. */
void oracle_promol(const double xc[3], int nimg, const double *ximg,
                   const int *zidx, const orc_rgrid *tabs, const double *cut,
                   int nder, double *f, double g[3], double h[9]);
/* oracle_promol + oracle_nci_point over npts nodes, OpenMP schedule(dynamic)
 * like nciplot's loop; out: (rho, |grad|, s, crho) per node. */
void oracle_promol_nci_batch(int64_t npts, const double *pts, int nimg, const double *ximg,
                             const int *zidx, const orc_rgrid *tabs, const double *cut,
                             int full_hessian, double *out);

/* Fragment densities at Cartesian point xc (nci@proc.f90:580-596): for
 * each of nfrag fragments (fstart[f] .. fstart[f+1]-1 index its atoms,
 * Cartesian positions xyz, table zidx), the sum over its atoms within rmax
 * of max(interp(max(d, r(1), 1e-14)), 0). */
void oracle_frag_density(const double xc[3], int nfrag, const int *fstart,
                         const double *xyz, const int *zidx,
                         const orc_rgrid *tabs, double *out);

/* NCI per-point quantities from node values (nci@proc.f90:543-605):
 * s = gmod / (const * max(rho,1d-80)^(4/3)); crho = sign(rho, l2)*100. */
void oracle_nci_point(double rho, double gmod, const double ehess_or_h[9],
                      int from_full_hessian, double *crho, double *s);

/* In-place inverse of a column-major 3x3 matrix in the operation order of
 * critic2's matinv (LAPACK dgetrf2 + dgetri). */
void oracle_matinv3(double a[9]);

/* Tricubic interpolation matrix c(64,64) of grid3mod@proc.f90:76,
 * column-major. */
void oracle_tricubic_matrix(double c[64 * 64]);

/* grinterp_tricubic (grid3mod@proc.f90:2649-2799) at fractional xi. */
void oracle_tricubic(const double *f, const int n[3], const double xi[3], double *y,
                     double yp[3], double ypp[9]);

/* NCIPLOT grid-interpolation scheme on the nodes of a periodic grid
 * (nci@proc.f90:565-571); outputs z-fastest. */
void oracle_nci_grid(const double *f, const int n[3], const double x2c[9], const double c2x[9],
                     double *crho, double *cgrad);

/* Synthetic Slater-sum density (FP64) on an n1*n2*n3 grid, x fastest.
 * rho = sum_{atoms,images} Z*(zeta^3/(8 pi))*exp(-zeta*|r-R|),
 * zeta = 2*Z^(1/3) bohr^-1, images within rimg bohr. */
void oracle_slater_grid(const double x2c[9], const int n[3], int nat,
                        const double *xat, const int *zat, double rimg,
                        double *out);

/* Yu-Trinkle integration (yt@proc.f90:38-224, :399-531; the imtype_yt sums
 * of intgrid_fields).  vec: 3*nvec grid steps, area: nvec facet areas of the
 * grid lattice's Wigner-Seitz cell.  idg in grid order; vol and sums
 * unscaled; sums[q*nattr + a].  Points are ranked by oracle_qcksort_r8, so
 * equal values take the reference sort's order.  Returns 0, or 2 (more than
 * xcap attractors). */
/* Reassign paths of the last run's refinement that ended on another point
 * negative at the start of its pass (resets the count; tests only). */
int64_t oracle_refine_negend(void);

/* qcksort_r8 (tools@proc.f90:83-168): iord[first..last] (0-based positions,
 * element indices) sorted by arr ascending, equal keys in that quicksort's
 * arrangement. */
void oracle_qcksort_r8(const double *arr, int64_t *iord, int64_t first, int64_t last);

int oracle_yt(const double *f, const int n[3], const orc_geom *g, int nvec, const int *vec, const double *area,
              int nfields, const double *const *fields, int32_t *idg, int *nattr_out, double *xattr, int xcap,
              double *vol, double *sums, double *wout, int wcap);
/* YT multipoles (integration@proc.f90:1314-1336) from the per-basin weights
 * wout of oracle_yt (basin-major, grid order). */
void oracle_yt_multipoles(const double *w, const double *f, const int n[3], const double x2c[9], int nattr,
                          const double *xattr, int lmax, double *mpole);

/* Basin multipoles (integration@proc.f90:1340-1362): mpole[a*nlm + q],
 * nlm = (lmax+1)^2, unscaled.  genrlm_real / tosphere restated for tests. */
void oracle_multipoles(const int32_t *idg, const double *f, const int n[3], const double x2c[9], int nattr,
                       const double *xattr, int lmax, double *mpole);
void oracle_genrlm_real(int lmax, double r, const double tp[2], double *rrlm);
void oracle_tosphere(const double v[3], double *r, double tp[2]);

/* Cube readers (grid3mod@proc.f90:512-635); f == NULL: header only.
 * step: column i = the file's axis-i vector.  Return 0, or 1 on a read
 * error. */
int oracle_read_cube(const char *path, int *nat, int *ismo, int n[3], double x0[3], double step[9], double *f);
int oracle_read_bincube(const char *path, int *nat, int *ismo, int n[3], double x0[3], double step[9],
                        double *f);

#ifdef __cplusplus
}
#endif
#endif
