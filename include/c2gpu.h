/*
 * c2gpu.h -- C ABI of the MI355X (gfx950) implementation of critic2's
 * on-grid density-analysis path: near-grid Bader basin labelling, per-basin
 * integration, and the NCI reduced density gradient.
 *
 * All entry points return an int status (C2GPU_OK == 0).  A thin Fortran
 * shim turns a non-zero status into ferror(...,faterr), which keeps the
 * keyword-level behaviour of critic2 (tools_io@proc.F90:1573-1618).
 *
 * Array conventions follow critic2:
 *   - 3-D grids are Fortran column-major f(n1,n2,n3): x (n1) fastest.
 *   - 3x3 matrices are column-major, m[r + 3*c] == m(r+1,c+1).
 *   - lat_i_dist is lat_i_dist(d1,d2,d3), d in -1..1, stored
 *     [(d1+1) + 3*(d2+1) + 9*(d3+1)] (bader@proc.f90:62-63).
 * Array arguments are host memory unless flags has C2GPU_PTR_DEVICE, in
 * which case the grid/label/field/output arrays are device pointers on the
 * context's device (small parameter arrays are always host memory).
 *
 * Contexts are independent and reentrant (critic2's bader module keeps
 * module-global state; this library keeps everything in the context).
 * Calls on one context are serialised on the context's stream.
 */
#ifndef C2GPU_H
#define C2GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  C2GPU_OK = 0,
  C2GPU_ERR_ARG = 1,          /* invalid argument                          */
  C2GPU_ERR_HIP = 2,          /* HIP runtime error                         */
  C2GPU_ERR_ALLOC = 3,        /* device allocation failed                  */
  C2GPU_ERR_UNSUPPORTED = 4,  /* option not implemented on the GPU path    */
  C2GPU_ERR_NOCONVERGE = 5,   /* first-sweep fixpoint did not converge     */
  C2GPU_ERR_NEWMAX = 6,       /* bader@proc.f90:400-402: new maximum in
                                 edge refinement (fatal in critic2 too)    */
  C2GPU_ERR_PATHLEN = 7,      /* an ascent path exceeded the path limit    */
  C2GPU_ERR_DEGENERATE = 8,   /* plateau case whose serial outcome depends
                                 on visiting order inside one sweep        */
  C2GPU_ERR_FORMAT = 9        /* malformed or truncated input file (the
                                 "Error reading file" of grid3mod@proc.f90:
                                 527-571)                                  */
};

enum { C2GPU_PTR_DEVICE = 1 };

typedef struct c2gpu_ctx c2gpu_ctx;

/* Context management.  device: HIP device ordinal. */
int c2gpu_ctx_create(c2gpu_ctx **ctx, int device);
int c2gpu_ctx_destroy(c2gpu_ctx *ctx);
/* Message of the last failing call on this context ("" if none). */
const char *c2gpu_ctx_error(const c2gpu_ctx *ctx);
/* Use an existing hipStream_t (NULL: the context's own stream). */
int c2gpu_ctx_set_stream(c2gpu_ctx *ctx, void *hip_stream);
/* Wait for all work queued on the context's stream. */
int c2gpu_ctx_sync(c2gpu_ctx *ctx);

/* Device memory for callers that keep grids resident on the GPU (no
 * counterpart in critic2, whose arrays live in host memory).  kind for
 * c2gpu_memcpy: 1 host->device, 2 device->host, 3 device->device. */
int c2gpu_dev_alloc(c2gpu_ctx *ctx, uint64_t bytes, void **ptr);
int c2gpu_dev_free(c2gpu_ctx *ctx, void *ptr);
int c2gpu_memcpy(c2gpu_ctx *ctx, void *dst, const void *src, uint64_t bytes, int kind);

/* Per-kernel device timing, the counterpart of critic2's tictac wall-clock
 * stamps (tools_io@proc.F90:1646) at kernel granularity: while enabled, each
 * kernel launch is bracketed by HIP events on the context's stream and its
 * time accumulated under the kernel's name.  c2gpu_ctx_timing resets the
 * table.  c2gpu_ctx_kernel_time returns the accumulated milliseconds and
 * launch count of one kernel (0 and 0 if it never ran); names_out receives
 * the comma-separated kernel names (truncated to len bytes). */
int c2gpu_ctx_timing(c2gpu_ctx *ctx, int enable);
int c2gpu_ctx_kernel_time(c2gpu_ctx *ctx, const char *kernel, double *ms, int64_t *launches);
int c2gpu_ctx_kernel_names(c2gpu_ctx *ctx, char *names_out, uint64_t len);
/* Algorithmic bytes a kernel had to move while timing was enabled (each
 * array element it must read or write, counted once; a read-modify-write
 * counts twice), accumulated over its launches; 0 for kernels that do not
 * report.  No counterpart in critic2: the measurement side of the timing
 * table, used for the HBM roofline. */
int c2gpu_ctx_kernel_bytes(c2gpu_ctx *ctx, const char *kernel, double *bytes);

/* ------------------------------------------------------------------ */
/* Bader (near-grid, Henkelman et al.)                                 */
/* ------------------------------------------------------------------ */

/* Grid metrics of bader_integrate (bader@proc.f90:123-145):
 * lat2car(:,i) = x2c(:,i)/n(i); car2lat = matinv(lat2car) with the
 * dgetrf/dgetri operation order (tools_math@proc.f90:1203-1235);
 * lat_i_dist(d) = 1/|lat2car*d| (0 at d = 0). */
int c2gpu_bader_metrics(const double x2c[9], const int n[3], double lat2car[9],
                        double car2lat[9], double lat_i_dist[27]);

/* Per-point, order-free step table (step_neargrid/step_ongrid/is_max of
 * bader@proc.f90:455-597 evaluated from every grid point with dr = 0):
 *   ongrid_next[p]    step_ongrid target of p, or p itself if is_max(p);
 *   neargrid_next0[p] target of the first step_neargrid call from p on a
 *                     fresh path (known all 0, dr = 0), p if no move;
 *   dr1_out[3*p+c]    dr after that step (0 when an on-grid step was taken).
 * Indices are 0-based linear grid indices.  Any output may be NULL. */
int c2gpu_bader_step_pointers(c2gpu_ctx *ctx, const double *rho, const int n[3],
                              const double car2lat[9], const double lat_i_dist[27],
                              int flags, int64_t *ongrid_next,
                              int64_t *neargrid_next0, double *dr1_out);

/* Attractor assignment options of BADER (integration@proc.f90:236-248). */
enum {
  C2GPU_ATTR_NEAREST_ATOM = 0, /* default: atexist, ratom = 1d40 (nonnm)   */
  C2GPU_ATTR_RATOM = 1,        /* NNM / RATOM r                            */
  C2GPU_ATTR_NOATOMS = 2       /* NOATOMS                                  */
};
/* BADER's DISCARD expression (bader@proc.f90:185-191) as a caller
 * callback: given the Cartesian position x (bohr, x2c times the fractional
 * position of the grid maximum), return nonzero to discard the maximum (the
 * expression's value is not below 1d-30 in magnitude).  It is consulted only
 * for maxima that would become new attractors (not an atom within ratom nor
 * an earlier attractor within ratom); a discarded maximum gets no attractor
 * and the points whose paths end there keep label 0, as in the reference.
 * It may be called once or more per maximum and must depend on x only. */
typedef int (*c2gpu_discard_fn)(const double x[3], void *user);

/* The attractors known before the sweep are the field's critical points,
 * cpcel(1:ncpcel) (bader@proc.f90:109-117): the atoms of the complete cell
 * list first, then any other critical points an earlier run added (known
 * non-nuclear maxima, or the bond/ring/cage points of a critical-point
 * search).  A new maximum takes the nearest ATOM within ratom
 * (identify_atom: only the first nat_atoms entries), else the first entry or
 * earlier new attractor within ratom (are_lclose over all of them). */
typedef struct {
  int mode;           /* C2GPU_ATTR_*                                      */
  double ratom;       /* bohr (ignored for NEAREST_ATOM)                   */
  double x2c[9];      /* lattice vectors as columns (bohr), crystal m_x2c  */
  int nat;            /* entries of xat: ncpcel (the atoms when no other
                         critical points are known)                        */
  const double *xat;  /* fractional cpcel coordinates, 3*nat, host memory  */
  c2gpu_discard_fn discard; /* DISCARD filter, or NULL (none)              */
  void *discard_user;       /* passed to discard                           */
  int nat_atoms;      /* leading entries of xat that are atoms (ncel);
                         <= 0: all nat entries are atoms                   */
} c2gpu_attractor_opts;

typedef struct {
  int nattr;            /* bas%nattr on return                            */
  int nmaxima;          /* distinct grid maxima reached by the sweep      */
  int sweep_iters;      /* fixpoint iterations of the first sweep         */
  int refine_iters;     /* refine_edge passes (bader@proc.f90:218-224)    */
  int64_t nstarts;      /* paths started by the first sweep               */
  int64_t traj_points;  /* stored free-trajectory points                  */
  int64_t edge_points;  /* points re-ascended in the first refine pass    */
  int64_t events_total; /* sweep events processed, summed over iterations */
  double ms_traj, ms_sweep, ms_refine, ms_total; /* device time per phase */
} c2gpu_bader_stats;

/* bader_integrate(s,bas,iref) (bader@proc.f90:80-234; interface
 * bader.f90:48-53).  label_out receives bas%idg (1-based attractor ids,
 * before int_reorder_gridout); xattr_out (host, 3*xattr_cap) receives
 * bas%xattr in fractional coordinates; *nattr receives bas%nattr.
 * car2lat and lat_i_dist are taken as given so that the caller fixes their
 * bits (critic2 computes them with LAPACK).  stats may be NULL.
 * Attractors of new maxima follow bader@proc.f90:159-199 in every mode: the
 * nearest atom (default), or an atom within ratom, else an earlier attractor
 * within ratom, else a new one (NNM/RATOM: atoms first; NOATOMS: none), the
 * new ones numbered in the serial sweep's discovery order; with a discard
 * callback (DISCARD), a maximum it rejects gets no attractor and its basin
 * keeps label 0.  If *nattr exceeds xattr_cap only the first xattr_cap
 * positions are written (call again with a larger xattr_out).  Limits: every grid edge >= 3 points, fewer than 2^32 - 1
 * points, and ascent paths of at most 2048 points (a longer one returns
 * C2GPU_ERR_PATHLEN; the reference's path array grows without bound). */
int c2gpu_bader(c2gpu_ctx *ctx, const double *rho, const int n[3],
                const double car2lat[9], const double lat_i_dist[27],
                const c2gpu_attractor_opts *opts, int flags, int32_t *label_out,
                int *nattr, double *xattr_out, int xattr_cap,
                c2gpu_bader_stats *stats);

/* where (idg == i) idg = map(i) for i = 1..nmap (the idg update of
 * int_reorder_gridout, integration@proc.f90:1113-1122). */
int c2gpu_relabel(c2gpu_ctx *ctx, int64_t npts, const int32_t *map, int nmap,
                  int flags, int32_t *label);

/* Per-basin sums (intgrid_fields, integration@proc.f90:1208-1218 and
 * :1289-1300) without the Omega/N factor, which stays with the caller:
 *   counts_out[a]          = count(idg == a+1)
 *   sums_out[k*nattr + a]  = sum(field_k, idg == a+1)
 * fields: nfields pointers to N-point FP64 grids.  Deterministic: the
 * summation order is fixed for a given grid size.  sums_out/counts_out are
 * host memory. */
int c2gpu_basin_sums(c2gpu_ctx *ctx, const int32_t *label, int64_t npts, int nattr,
                     int nfields, const double *const *fields, int flags,
                     double *sums_out, int64_t *counts_out);

/* No counterpart in the reference: c2gpu_basin_sums for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It calls no hipMalloc, hipFree or other
 * allocating routine: its block partials live in a fixed device scratch of
 * 2048 x 2048 doubles (32 MiB, held by every device the library runs on
 * until the process ends).  label, the nfields field grids, sums_out
 * (nfields*nattr doubles, as c2gpu_basin_sums) and counts_out (nattr int64,
 * may be NULL when nfields > 0) are device memory.  The work is enqueued on
 * the context's stream and the call returns without waiting; the results
 * equal c2gpu_basin_sums's bit for bit.  Limits: nattr*(nfields+1) <= 2048
 * bins (as c2gpu_basin_sums), nfields <= 4; larger requests return
 * C2GPU_ERR_UNSUPPORTED.  Every call on one device shares that device's
 * scratch: calls that overlap in time on one GPU (two streams, two threads)
 * race and must be serialised by the caller. */
int c2gpu_basin_sums_noalloc(c2gpu_ctx *ctx, const int32_t *label, int64_t npts, int nattr,
                             int nfields, const double *const *fields,
                             double *sums_out, int64_t *counts_out);

/* ------------------------------------------------------------------ */
/* FFT derivative grids and NCI                                        */
/* ------------------------------------------------------------------ */

enum {
  C2GPU_FT_GRAD = 0, /* |grad f| (ifformat_as_ft_grad)                    */
  C2GPU_FT_XX = 1,   /* d2f/dx2 (ifformat_as_ft_xx)                       */
  C2GPU_FT_YY = 2,
  C2GPU_FT_ZZ = 3,
  C2GPU_FT_LAP = 4   /* Laplacian (ifformat_as_ft_lap)                    */
};

/* grid3%fft(fnew, fold, iff) (grid3mod@proc.f90:1757-1872): spectral
 * derivative of a periodic grid, Nyquist components at +n/2. */
int c2gpu_fft_deriv(c2gpu_ctx *ctx, const double *rho, const int n[3],
                    const double x2c[9], int kind, int flags, double *out);

/* Radial density tables of the promolecular density (grid1 type,
 * grid1mod.f90:26-48): nodes r (a = r(1), b = log(r(2)/r(1)), the index
 * formula of grid1%interp assumes r_i = a*exp(b*(i-1))), f = rho,
 * fp = rho', fpp = rho'', rmax = r(ngrid).  cutrad: cutoff radius of the
 * species (global.f90:53, bohr). */
typedef struct {
  int ngrid;
  double a, b, rmax, cutrad;
  const double *r, *f, *fp, *fpp; /* host arrays, ngrid each               */
} c2gpu_radial_table;

typedef struct {
  int nspecies;
  const c2gpu_radial_table *tables; /* nspecies tables                    */
  int nat;                          /* atoms in the cell                   */
  const double *xat;                /* fractional coordinates, 3*nat       */
  const int *species;               /* table index of each atom            */
} c2gpu_promol;

enum { C2GPU_NCI_FFT = 0, C2GPU_NCI_GRID = 1, C2GPU_NCI_PROMOL = 2 };
enum { C2GPU_FP64 = 0, C2GPU_FP32 = 1 };

/* The NCIPLOT per-point block (nci@proc.f90:525-605) on the nodes of a
 * periodic grid:
 *   cgrad = |grad rho| / (2 (3 pi^2)^(1/3) max(rho, vsmall)^(4/3))
 *   crho  = sign(rho, lambda2) * 100
 * scheme C2GPU_NCI_FFT: rho is the loaded FP64 grid; derivatives by FFT and
 *   the lambda2 sign from count(diag Hessian > 0) >= 2 (nci@proc.f90:546-564).
 * scheme C2GPU_NCI_GRID: rho is the loaded FP64 grid (INTERPOLATION GRID, or
 *   a field with a core contribution); density, gradient and Hessian from the
 *   grid's tricubic interpolant (grid3mod@proc.f90:2649-2799) at the node as
 *   field%grd locates it, lambda2 from the Hessian eigenvalues
 *   (nci@proc.f90:565-571). FP64 only; n >= 2 per axis.
 * scheme C2GPU_NCI_PROMOL: rho is ignored; density, gradient and Hessian from
 *   the promolecular radial tables (crystalmod@env.f90:622-753), lambda2 from
 *   the Hessian eigenvalues (nci@proc.f90:565-571); precision selects FP64 or
 *   FP32 arithmetic.
 * crho_out/cgrad_out are written z-fastest, (0:n3-1,0:n2-1,0:n1-1) as in
 * nci@proc.f90:502-505, unless zfast == 0 (then x-fastest like the input). */
int c2gpu_nci(c2gpu_ctx *ctx, const double *rho, const int n[3], const double x2c[9],
              int scheme, int precision, const c2gpu_promol *promol, int zfast,
              int flags, double *crho_out, double *cgrad_out);

/* FP32-output variant of c2gpu_nci (same arguments, float outputs). */
int c2gpu_nci_f32(c2gpu_ctx *ctx, const double *rho, const int n[3], const double x2c[9],
                  int scheme, int precision, const c2gpu_promol *promol, int zfast,
                  int flags, float *crho_out, float *cgrad_out);

/* The nodes of an nciplot cube (nci@proc.f90:412-446, :548):
 *   x = x0 + i*xmat(:,1) + j*xmat(:,2) + k*xmat(:,3),  0 <= i < nstep[0] ...
 * Cartesian bohr, xmat column-major.  The periodic cell: x0 = 0 and
 * xmat(:,c) = a_c / nstep[c] (:413-426); a CUBE or fragment box: x0 its
 * lower corner and xmat = diag(xinc) (:427-443). */
typedef struct {
  int nstep[3];
  double x0[3];
  double xmat[9];
} c2gpu_nci_nodes;

/* The promolecular NCI block (nci@proc.f90:548-604 with the promolecular
 * field, crystalmod@env.f90:622-753) on the nodes of an nciplot cube, in
 * the cell or a box: crho = sign(rho, lambda2) * 100, cgrad = s, and, when
 * rhoat_out is not NULL, rhoat = rho, the promolecular density that
 * dopromol stores (nci@proc.f90:574-577, 600).  x2c: the cell, whose lattice
 * repeats the atoms of promol.  precision: C2GPU_FP64 (double outputs) or
 * C2GPU_FP32 (float outputs).  Outputs z-fastest,
 * (0:nstep[2]-1, 0:nstep[1]-1, 0:nstep[0]-1), as nci@proc.f90:502-505. */
int c2gpu_nci_promol_nodes(c2gpu_ctx *ctx, const c2gpu_nci_nodes *nodes, const double x2c[9],
                           const c2gpu_promol *promol, int precision, int flags, void *crho_out,
                           void *cgrad_out, void *rhoat_out);

/* NCIPLOT fragments (FRAGMENT ... ENDFRAGMENT, nci@proc.f90:580-596): per
 * node and fragment, the promolecular density of the fragment's atoms alone,
 * at the positions given (no lattice repeats): the sum over its atoms within
 * rmax of their table of max(interp(max(d, r(1), 1e-14)), 0).
 * frag_nat[f]: atoms of fragment f; xyz: Cartesian bohr, 3 per atom,
 * fragments one after the other; species: radial-table index per atom.
 * rhofrag_out(k, j, i, f): z fastest, fragment slowest, nstep product times
 * nfrag values of the precision's type (C2GPU_FP64: double, C2GPU_FP32:
 * float). */
int c2gpu_nci_fragments(c2gpu_ctx *ctx, const c2gpu_nci_nodes *nodes, int nspecies,
                        const c2gpu_radial_table *tables, int nfrag, const int *frag_nat,
                        const double *xyz, const int *species, int precision, int flags,
                        void *rhofrag_out);

/* Options of nciplot's cutoff pass (nci@proc.f90:612-643; keywords
 * CUTOFFS rhocut dimcut, CUTPLOT rhoplot, SRHORANGE, ONLYNEG, RHO_VOID). */
typedef struct {
  double rhocut, dimcut;  /* .dat rows: |rho| < rhocut and s < dimcut       */
  double rhoplot;         /* s <- 100 where |rho| > rhoplot                 */
  double srhorange[2];    /* s <- 100 where sign(l2) rho is outside          */
  int onlyneg;            /* s <- 100 where sign(l2) rho > 0                 */
  double rho_void;        /* > 0: only nodes with promolecular density below
                             it count; the void sums are returned           */
} c2gpu_nci_cut;

/* nciplot's cutoff pass (nci@proc.f90:612-654) without fragments over crho
 * (sign(l2) rho * 100) and cgrad (s) as c2gpu_nci returns them, z fastest:
 * the .dat rows '(1p,E15.7,E15.7)' rho, s of the selected nodes in node
 * order (31 bytes each) into dat, then cgrad <- 100 where masked (in place).
 * rhoat: promolecular density (needed with rho_void > 0).  precision:
 * C2GPU_FP64 or C2GPU_FP32, the type of the three arrays.  *len receives
 * the .dat byte count; dat NULL only queries it (nothing is modified).
 * sums (may be NULL): unscaled void charge, promolecular charge and node
 * count (critic2 multiplies them by Omega/N).  With C2GPU_PTR_DEVICE the
 * arrays and dat are device memory. */
int c2gpu_nci_cutoff(c2gpu_ctx *ctx, const void *crho, void *cgrad, const void *rhoat, const int n[3],
                     int precision, const c2gpu_nci_cut *opts, int flags, char *dat, uint64_t cap,
                     uint64_t *len, double sums[3]);

/* The cutoff pass with fragments (nci@proc.f90:622-626): as
 * c2gpu_nci_cutoff, where a node counts as intermolecular only when
 *   S = sum_f rhofrag(f) >= rhoparam2 * rhoat  and  rhofrag(f) <= S * rhoparam
 * for every fragment (keywords RHOPARAM2, RHOPARAM; critic2's defaults are
 * 0.75 and 0.95), and with RHO_VOID also rhoat < rho_void.  rhofrag: nfrag
 * arrays of the nodes, one after the other, as c2gpu_nci_fragments writes
 * them; rhoat is required.  nfrag = 0 is c2gpu_nci_cutoff. */
int c2gpu_nci_cutoff_frag(c2gpu_ctx *ctx, const void *crho, void *cgrad, const void *rhoat,
                          const void *rhofrag, int nfrag, double rhoparam, double rhoparam2,
                          const int n[3], int precision, const c2gpu_nci_cut *opts, int flags,
                          char *dat, uint64_t cap, uint64_t *len, double sums[3]);

/* ------------------------------------------------------------------ */
/* Batches of cells                                                    */
/* ------------------------------------------------------------------ */

/* One periodic cell of a batch.  Inputs: the FP64 grid (x fastest), the
 * cell, the grid metrics as c2gpu_bader_metrics returns them, and the atoms
 * (fractional coordinates; species = radial-table index for NCI).  Outputs
 * (NULL skips that part for this cell): Bader labels (N), the attractor
 * count and positions, per-basin point counts and sums of rho (xattr_cap
 * entries each, host memory), and the promolecular NCI signed density x100
 * and reduced gradient in FP32, z-fastest as nci@proc.f90:502-505. */
typedef struct {
  const double *rho;
  int n[3];
  double x2c[9];
  double car2lat[9];
  double lat_i_dist[27];
  int nat;
  const double *xat;
  const int *species;
  int32_t *label_out;
  int *nattr_out;
  double *xattr_out;
  int xattr_cap;
  int64_t *counts_out;
  double *sums_out;
  float *crho_out, *cgrad_out;
} c2gpu_cell;

typedef struct {
  double ms_total;      /* wall time of the call                          */
  double ms_nci;        /* device time of the NCI work-queue launch        */
  int64_t points;       /* grid points of all cells                       */
  int nstreams;         /* concurrent Bader workers used                   */
  int queue_workgroups; /* workgroups of the NCI work-queue kernel         */
} c2gpu_batch_stats;

/* Bader (bader_integrate + the Volume/Pop sums of intgrid_fields) and the
 * promolecular NCI block over many independent cells in one call; no
 * counterpart in critic2, which runs them one system at a time.  Results per
 * cell equal c2gpu_bader / c2gpu_basin_sums / c2gpu_nci_f32 (FP32 promol) on
 * that cell alone.  NCI: every cell's 8x8x4 node bricks go through one
 * persistent work-queue kernel launched as a HIP graph.  Bader: cells are
 * taken from a queue by nstreams workers (0: 8), each on its own stream.
 * attr_mode/ratom as in c2gpu_attractor_opts.  tables: the radial tables
 * (shared by all cells; may be NULL when no cell asks for NCI).  With
 * C2GPU_PTR_DEVICE, rho, label_out, crho_out and cgrad_out are device
 * memory; the other outputs are always host memory.  stats may be NULL. */
int c2gpu_batch_run(c2gpu_ctx *ctx, int ncell, const c2gpu_cell *cells, int attr_mode, double ratom,
                    int nspecies, const c2gpu_radial_table *tables, int nstreams, int flags,
                    c2gpu_batch_stats *stats);

/* Basin multipoles (INTEGRABLE f MULTIPOLE lmax, integration@proc.f90:
 * 1302-1362, the Bader branch): mpole_out[a*(lmax+1)^2 + q] = sum over the
 * points of basin a+1 of rrlm(q) f, rrlm the real regular solid harmonics
 * of genrlm_real (tools_math@proc.f90:273-306; order C00, C11, C10, S11,
 * C22, C21, C20, S21, S22, ...) at the shortest lattice translate of
 * p/n - xattr(a) (Cartesian).  Unscaled (critic2 multiplies by Omega/N);
 * the layout is critic2's mpole((lmax+1)**2, nattr).  xattr: fractional,
 * 3 per attractor, host memory.  0 <= lmax <= 10.  With C2GPU_PTR_DEVICE
 * label and f are device memory. */
int c2gpu_multipoles(c2gpu_ctx *ctx, const int32_t *label, const double *f, const int n[3], const double x2c[9],
                     int nattr, const double *xattr, int lmax, int flags, double *mpole_out);

/* ------------------------------------------------------------------ */
/* Yu-Trinkle integration                                               */
/* ------------------------------------------------------------------ */

/* Facets of the Wigner-Seitz cell of the lattice with vectors m (columns):
 * the Voronoi-relevant lattice vectors (3 integers each, lattice
 * coordinates) and facet areas; for the grid lattice x2c(:,i)/n(i) this is
 * grid3%vec / grid3%area (wscell, tools@proc.f90:679-800).  *nvec <= 14.
 * Host only. */
int c2gpu_ws_facets(const double m[9], int *nvec, int vec[42], double area[14]);

typedef struct {
  int nattr;          /* attractors                                        */
  int nmaxima;        /* grid maxima                                       */
  int64_t nias;       /* points on the interatomic surfaces (basin 0)      */
  int label_passes;   /* propagation passes for the labels                 */
  int weight_passes;  /* propagation passes for the surface weights        */
  int64_t wide;       /* surface points where more than 32 basins meet     */
  double ms_total;
  int ranked;         /* 1: equal values met, points ranked by the
                         reference's quicksort on the host                 */
} c2gpu_yt_stats;

/* yt_integrate (yt@proc.f90:38-224) with the weights of yt_weights
 * (:399-531) summed as intgrid_fields does for imtype_yt
 * (integration@proc.f90:1208-1218, :1289-1300).  Points are ranked by rho;
 * a point with no higher neighbour among the grid's Voronoi-relevant
 * neighbours is a maximum (attractor rules as in c2gpu_bader); a point whose
 * higher neighbours all lie in one basin joins it; the others (the
 * interatomic surfaces, label 0) split their weight over the basins of
 * their higher neighbours in proportion to facet area times rho difference.
 * label_out: basin per point in grid order (x fastest; critic2 keeps it in
 * rank order, yt@proc.f90:218).  vol_out[a] = sum of basin a+1's weights,
 * sums_out[q*nattr + a] = sum of weights times fields[q] (both unscaled:
 * critic2 multiplies by Omega/N).  Where equal values meet (neighbouring
 * points, the higher neighbours of one point, or maxima), their order is the
 * one critic2's qcksort gives them (tools@proc.f90:83-168): the points are
 * then ranked by a restatement of that quicksort on the host (seconds for
 * 16 M points) and the ranks decide every comparison.  At most 13653
 * attractors (C2GPU_ERR_UNSUPPORTED beyond).  rho, fields and label_out are
 * device memory with C2GPU_PTR_DEVICE. */
int c2gpu_yt(c2gpu_ctx *ctx, const double *rho, const int n[3], const double x2c[9],
             const c2gpu_attractor_opts *opts, int nfields, const double *const *fields, int flags,
             int32_t *label_out, int *nattr_out, double *xattr_out, int xattr_cap, double *vol_out,
             double *sums_out, c2gpu_yt_stats *stats);

/* The YT checkpoint update of int_reorder_gridout (integration@proc.f90:
 * 1125-1158): the surface weights kept by the last c2gpu_yt call change
 * basin ids through map (ibasin2 = assigned(ibasin)); nmap must be that
 * call's nattr ("inconsistent number of attractors in yt checkpoint"
 * otherwise, C2GPU_ERR_ARG) and every map entry lie in 1..nattr_new.
 * Weights of merged basins add, as yt_weights over the remapped checkpoint
 * gives them.  The labels are remapped by c2gpu_relabel with the same map. */
int c2gpu_yt_remap(c2gpu_ctx *ctx, const int32_t *map, int nmap, int nattr_new);

/* Basin multipoles of a YT partition (INTEGRABLE f MULTIPOLE lmax,
 * integration@proc.f90:1314-1336, the imtype_yt branch):
 * mpole_out[a*(lmax+1)^2 + q] = sum over the points p with |w_a(p)| >= 1e-15
 * of rrlm(q) f(p) w_a(p), w_a = yt_weights' weight of basin a+1 (1 at its
 * interior points, the surface weight elsewhere), rrlm as in
 * c2gpu_multipoles.  The surface weights are those of the last c2gpu_yt
 * call on this context, which keeps them as critic2 keeps its YT checkpoint
 * for later yt_weights calls (after c2gpu_yt_remap, with their new ids);
 * label is that call's label_out (relabelled the same way), and n and nattr
 * must be that call's, or c2gpu_yt_remap's nattr_new (else C2GPU_ERR_ARG).
 * A surface point's entries are skipped one by one below 1e-15, where the
 * reference tests their sum for merged basins (a difference below 1e-15 of
 * the point's contribution).  Every basin is computed:
 * critic2 leaves those whose attractor it does not integrate
 * (docelatom) at zero.  Flags, layout and units as c2gpu_multipoles. */
int c2gpu_yt_multipoles(c2gpu_ctx *ctx, const int32_t *label, const double *f, const int n[3],
                        const double x2c[9], int nattr, const double *xattr, int lmax, int flags,
                        double *mpole_out);

/* ------------------------------------------------------------------ */
/* Gaussian cube input and output                                       */
/* ------------------------------------------------------------------ */

/* What the header of a cube holds.  x0 and step are as in the file
 * (Cartesian bohr); critic2 then sets x2cl(:,i) = step(:,i)*n(i)
 * (grid3mod@proc.f90:546-547) and converts x0 to crystallographic
 * coordinates with the caller's cell (:540-541). */
typedef struct {
  int nat;               /* number of atom records (|natoms| of the file)  */
  int ismo;              /* natoms < 0: orbital cube, one more record       */
  int n[3];              /* grid points along each axis                     */
  double x0[3];          /* origin                                          */
  double step[9];        /* step(:,i) = the file's axis-i vector, col-major */
  uint64_t data_offset;  /* byte offset of the grid values (text), or of the
                            grid record's first marker (binary)             */
} c2gpu_cube_header;

/* Header only (the host reads it; no device work).  binary: 0 text cube
 * (grid3mod@proc.f90:533-551), 1 binary cube (:592-609). */
int c2gpu_cube_read_header(c2gpu_ctx *ctx, const char *path, int binary, c2gpu_cube_header *hdr);

/* grid3%read_cube (grid3mod@proc.f90:512-572): the text cube at path, its
 * grid values converted on the GPU with correct rounding (the value the
 * Fortran list-directed read gives) into f(n1,n2,n3), x fastest (the file
 * holds them z fastest).  f: n1*n2*n3 doubles, device memory with
 * C2GPU_PTR_DEVICE.  Values past n1*n2*n3 are ignored; fewer values, a
 * non-numeric token, or a list-directed form other than blank-separated
 * constants (r*c repeats, commas, slashes) give C2GPU_ERR_FORMAT. */
int c2gpu_cube_read(c2gpu_ctx *ctx, const char *path, int flags, c2gpu_cube_header *hdr, double *f);

/* The same over the bytes of a text cube already in memory: text holds len
 * bytes (device memory with C2GPU_PTR_DEVICE, as is f then).  f NULL: only
 * hdr is filled. */
int c2gpu_cube_parse(c2gpu_ctx *ctx, const char *text, uint64_t len, int flags, c2gpu_cube_header *hdr,
                     double *f);

/* grid3%read_bincube (grid3mod@proc.f90:575-635): unformatted sequential
 * records with 4-byte markers (subrecords of records over 2 GiB followed):
 * (natoms, x0), (n, step), one record per atom, then f in Fortran order,
 * streamed to f (device memory with C2GPU_PTR_DEVICE). */
int c2gpu_bincube_read(c2gpu_ctx *ctx, const char *path, int flags, c2gpu_cube_header *hdr, double *f);

/* Cube body styles: the Fortran edit descriptor of each grid value.
 * Values are correctly rounded (nearest, ties to even on the exact binary
 * value), negative values carry '-', fields too narrow hold asterisks, as
 * gfortran's formatted WRITE gives them. */
enum {
  C2GPU_CUBE_E22_14 = 0,   /* '(6(" ",E22.14E3))': writegrid_cube with
                              precisecube, the default
                              (crystalmod@write.f90:3563)              */
  C2GPU_CUBE_1P_E12_5 = 1, /* '(1p,6(" ",E12.5E3))': writegrid_cube
                              without precisecube (:3565)              */
  C2GPU_CUBE_1P_E13_5 = 2  /* '(6(" ",1p,e13.5e3))': nciplot's
                              write_cube_body (nci@proc.f90:927)       */
};

/* The body text of a cube: for each (i,j) in file order, the n3 values
 * f(i,j,1..n3) in records of six, a record closed early also holding the
 * format's next " " (crystalmod@write.f90:3557-3568, nci@proc.f90:925-929).
 * zfast: 0 f is x fastest (critic2's grids), 1 f is already z fastest
 * (c2gpu_nci's zfast outputs).  *len receives the byte count; text NULL
 * queries it, else text must hold cap >= *len bytes.  With
 * C2GPU_PTR_DEVICE f and text are device memory. */
int c2gpu_cube_format(c2gpu_ctx *ctx, const double *f, const int n[3], int style, int zfast, int flags,
                      char *text, uint64_t cap, uint64_t *len);

/* grid3 writer writegrid_cube (crystalmod@write.f90:3476-3570), text form:
 * two "critic2-cube" titles, the header from hdr (nat, x0, n, step) and the
 * atoms (zat, xat: Cartesian bohr, as the caller shifted them), then the
 * body of f (x fastest; NULL writes the header only, critic2's onlyheader).
 * style: C2GPU_CUBE_E22_14 (precisecube header and body) or
 * C2GPU_CUBE_1P_E12_5 (the F12.6 header and 1p body). */
int c2gpu_cube_write(c2gpu_ctx *ctx, const char *path, const c2gpu_cube_header *hdr, const int *zat,
                     const double *xat, const double *f, int style, int flags);

/* ------------------------------------------------------------------ */
/* Synthetic input (no counterpart in critic2: test and benchmark input) */
/* ------------------------------------------------------------------ */

/* rho = sum_{atoms,images within rimg} Z zeta^3/(8 pi) exp(-zeta |r-R|),
 * zeta = 2 Z^(1/3), evaluated in FP64 on the grid nodes x = (i/n1,j/n2,k/n3). */
int c2gpu_synth_slater(c2gpu_ctx *ctx, const double x2c[9], const int n[3], int nat,
                       const double *xat, const int *zat, double rimg, int flags,
                       double *rho_out);

#ifdef __cplusplus
}
#endif
#endif
