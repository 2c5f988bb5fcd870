// extern "C" entry points of libc2gpu (declared in include/c2gpu.h) and the
// host-side pieces of the path: grid metrics (bader@proc.f90:123-145 with a
// 3x3 restatement of LAPACK dgetrf2/dgetri) and the attractor assignment of
// new maxima (bader@proc.f90:159-199).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>

#include "common.hpp"

namespace c2 {
void bader_run(c2gpu_ctx *, const double *, const int *, const double *, const double *,
               const c2gpu_attractor_opts *, int, int32_t *, int *, double *, int, c2gpu_bader_stats *);
void step_pointers_run(c2gpu_ctx *, const double *, const int *, const double *, const double *, int,
                       int64_t *, int64_t *, double *);
void noalloc_init(int);
void basin_sums_noalloc_run(c2gpu_ctx *, const int32_t *, int64_t, int, int, const double *const *, double *,
                            int64_t *);
void basin_sums_run(c2gpu_ctx *, const int32_t *, int64_t, int, int, const double *const *, int,
                    double *, int64_t *);
void relabel_run(c2gpu_ctx *, int64_t, const int32_t *, int, int, int32_t *);
void fft_deriv_run(c2gpu_ctx *, const double *, const int *, const double *, int, int, double *);
void nci_run(c2gpu_ctx *, const double *, const int *, const double *, const double *, int, int,
             const c2gpu_promol *, int, int, void *, void *, int);
void batch_run(c2gpu_ctx *, int, const c2gpu_cell *, int, double, int, const c2gpu_radial_table *, int, int,
               c2gpu_batch_stats *);
void destroy_workers(c2gpu_ctx *);
void cube_header(const char *, c2gpu_cube_header *, int);
void cube_parse_run(c2gpu_ctx *, const char *, uint64_t, int, c2gpu_cube_header *, double *);
void cube_read_run(c2gpu_ctx *, const char *, int, c2gpu_cube_header *, double *);
void bincube_read_run(c2gpu_ctx *, const char *, int, c2gpu_cube_header *, double *);
void nci_cutoff_run(c2gpu_ctx *, const void *, void *, const void *, const void *, int, double, double, const int *,
                    int, const c2gpu_nci_cut *, int, char *, uint64_t, uint64_t *, double *);
void nci_promol_nodes_run(c2gpu_ctx *, const c2gpu_nci_nodes *, const double *, const c2gpu_promol *, int, int,
                          void *, void *, void *);
void nci_fragments_run(c2gpu_ctx *, const c2gpu_nci_nodes *, int, const c2gpu_radial_table *, int, const int *,
                       const double *, const int *, int, int, void *);
int ws_facets(const double m[9], int vec[3 * 14], double area[14]);
void yt_run(c2gpu_ctx *, const double *, const int *, const double *, const c2gpu_attractor_opts *, int,
            const double *const *, int, int32_t *, int *, double *, int, double *, double *, c2gpu_yt_stats *);
void mpole_run(c2gpu_ctx *, const int32_t *, const double *, const int *, const double *, int, const double *, int,
               int, double *);
void yt_mpole_run(c2gpu_ctx *, const int32_t *, const double *, const int *, const double *, int, const double *, int,
               int, double *);
void yt_remap_run(c2gpu_ctx *, const int32_t *, int, int);
void cube_format_run(c2gpu_ctx *, const double *, const int *, int, int, int, char *, uint64_t, uint64_t *);
void cube_write_run(c2gpu_ctx *, const char *, const c2gpu_cube_header *, const int *, const double *,
                    const double *, int, int);
void synth_slater_run(c2gpu_ctx *, const double *, const int *, int, const double *, const int *,
                      double, int, double *);

// matmul(m, x) for a column-major 3x3 m
static void mat3_apply(const double m[9], const double x[3], double out[3]) {
  for (int i = 0; i < 3; i++) {
    double acc = 0.0;
    for (int j = 0; j < 3; j++) acc = acc + m[i + 3 * j] * x[j];
    out[i] = acc;
  }
}

// Inverse of a 3x3 column-major matrix with the operation order of the
// reference LAPACK: dgetrf2 (recursive, n1 = 1 split, reciprocal-scaled
// pivot column), dtrti2 (upper inverse), dgetri's unblocked solve and the
// final column interchanges (src/oldlibs/dgetrf2.f, dgetri.f, dtrti2.f).
static bool inv3_lapack(double *a) {
#define M(r, c) a[(r) + 3 * (c)]
  int ipiv[3];
  for (int col = 0; col < 2; col++) {
    int ip = col;
    double mx = std::fabs(M(col, col));
    for (int r = col + 1; r < 3; r++)
      if (std::fabs(M(r, col)) > mx) {
        mx = std::fabs(M(r, col));
        ip = r;
      }
    ipiv[col] = ip;
    if (M(ip, col) == 0.0) return false;
    if (ip != col) std::swap(M(col, col), M(ip, col));
    if (std::fabs(M(col, col)) >= DBL_MIN) {
      const double rc = 1.0 / M(col, col);
      for (int r = col + 1; r < 3; r++) M(r, col) = rc * M(r, col);
    } else {
      for (int r = col + 1; r < 3; r++) M(r, col) = M(r, col) / M(col, col);
    }
    if (ip != col)
      for (int c = col + 1; c < 3; c++) std::swap(M(col, c), M(ip, c));
    for (int c = col + 1; c < 3; c++) {
      const double temp = -M(col, c);
      for (int r = col + 1; r < 3; r++) M(r, c) = M(r, c) + temp * M(r, col);
    }
    if (col == 1 && ip != 1) std::swap(M(1, 0), M(ip, 0));
  }
  ipiv[2] = 2;
  if (M(2, 2) == 0.0) return false;
  for (int j = 0; j < 3; j++) {  // dtrti2
    M(j, j) = 1.0 / M(j, j);
    const double ajj = -M(j, j);
    for (int jj = 0; jj < j; jj++) {
      if (M(jj, j) != 0.0) {
        const double temp = M(jj, j);
        for (int i = 0; i < jj; i++) M(i, j) = M(i, j) + temp * M(i, jj);
        M(jj, j) = M(jj, j) * M(jj, jj);
      }
    }
    for (int i = 0; i < j; i++) M(i, j) = ajj * M(i, j);
  }
  double work[3];
  for (int j = 2; j >= 0; j--) {  // dgetri, NB >= N branch
    for (int i = j + 1; i < 3; i++) {
      work[i] = M(i, j);
      M(i, j) = 0.0;
    }
    for (int jj = j + 1; jj < 3; jj++) {
      const double temp = -work[jj];
      for (int i = 0; i < 3; i++) M(i, j) = M(i, j) + temp * M(i, jj);
    }
  }
  for (int j = 1; j >= 0; j--)
    if (ipiv[j] != j)
      for (int r = 0; r < 3; r++) std::swap(M(r, j), M(r, ipiv[j]));
#undef M
  return true;
}

// Shortest lattice-translated length of a fractional difference (Cartesian
// norm), by search over translations -2..2 after folding.
static double lattice_dist(const double x2c[9], const double xd[3]) {
  double x0[3];
  for (int c = 0; c < 3; c++) x0[c] = xd[c] - std::round(xd[c]);
  double best = DBL_MAX;
  for (int a = -2; a <= 2; a++)
    for (int b = -2; b <= 2; b++)
      for (int c = -2; c <= 2; c++) {
        const double xt[3] = {x0[0] + a, x0[1] + b, x0[2] + c};
        double xc[3];
        mat3_apply(x2c, xt, xc);
        const double d = std::sqrt(xc[0] * xc[0] + xc[1] * xc[1] + xc[2] * xc[2]);
        if (d < best) best = d;
      }
  return best;
}

// the atoms among the known attractors (the leading nat_atoms entries)
static int atom_count(const c2gpu_attractor_opts *o) {
  return o->nat_atoms > 0 && o->nat_atoms < o->nat ? o->nat_atoms : o->nat;
}

int nearest_atom_label(const c2gpu_attractor_opts *o, const double x[3]) {
  int best = 0;
  double bd = DBL_MAX;
  const int natoms = atom_count(o);
  for (int a = 0; a < natoms; a++) {
    const double xd[3] = {x[0] - o->xat[3 * a], x[1] - o->xat[3 * a + 1], x[2] - o->xat[3 * a + 2]};
    const double d = lattice_dist(o->x2c, xd);
    if (d < bd) {
      bd = d;
      best = a + 1;
    }
  }
  return best;
}

// identify_atom(x, icrd_crys, distmax) (crystalmod@env.f90:593-614 through
// nearest_atom, :516-563): the nearest atom, if within distmax (atoms up to
// distmax inclusive when distmax < 1 bohr, else the nearest unless farther
// than distmax); 0 if none.
static int identify_atom_within(const c2gpu_attractor_opts *o, const double x[3], double distmax) {
  int best = 0;
  double bd = DBL_MAX;
  const int natoms = atom_count(o);
  for (int a = 0; a < natoms; a++) {
    const double xd[3] = {x[0] - o->xat[3 * a], x[1] - o->xat[3 * a + 1], x[2] - o->xat[3 * a + 2]};
    const double d = lattice_dist(o->x2c, xd);
    if (d < bd) {
      bd = d;
      best = a + 1;
    }
  }
  if (best == 0) return 0;
  if (distmax < 1.0) return bd <= distmax ? best : 0;
  return bd > distmax ? 0 : best;
}

// Attractor of each new maximum, taken in discovery order
// (bader@proc.f90:159-199): an atom within ratom (if atoms exist), else a
// known attractor within ratom (are_lclose, crystalmod@proc.f90:1118-1133),
// else, unless the DISCARD callback rejects it (:185-191, at the Cartesian
// position x2c*dv), a new attractor numbered nbasin+1 at the maximum; a
// rejected maximum gets label 0.  dv: fractional positions of the maxima;
// order: their discovery order; discarded: per maximum, -1 until the
// callback was asked, then its answer (kept across calls, so the callback
// runs at most once per maximum).  Fills lab (per maximum) and xattr (3 per
// attractor); returns the attractor count.
int attractor_rule(const c2gpu_attractor_opts *o, uint32_t nmax, const uint32_t *order, const double *dv,
                   int32_t *lab, std::vector<double> &xattr, std::vector<int8_t> &discarded) {
  if (discarded.size() != nmax) discarded.assign(nmax, -1);
  const bool atexist = o->mode != C2GPU_ATTR_NOATOMS;
  const double ratom = o->mode == C2GPU_ATTR_NEAREST_ATOM ? 1e40 : o->ratom;
  xattr.clear();
  int nbasin = 0;
  if (atexist) {
    nbasin = o->nat;
    xattr.assign(o->xat, o->xat + 3 * (size_t)o->nat);
  }
  for (uint32_t q = 0; q < nmax; q++) {
    const uint32_t m = order[q];
    const double *x = dv + 3 * (size_t)m;
    int pv = 0;
    if (atexist) pv = identify_atom_within(o, x, ratom);
    if (pv == 0 && ratom > 1e-40) {  // vsmall (param.F90)
      for (int l = 0; l < nbasin; l++) {
        const double xd[3] = {x[0] - xattr[3 * l], x[1] - xattr[3 * l + 1], x[2] - xattr[3 * l + 2]};
        if (lattice_dist(o->x2c, xd) < ratom) {
          pv = l + 1;
          break;
        }
      }
    }
    if (pv == 0 && o->discard) {
      if (discarded[m] < 0) {
        double xc[3];
        mat3_apply(o->x2c, x, xc);  // x = s%c%x2c(dv)
        discarded[m] = o->discard(xc, o->discard_user) != 0;
      }
      if (discarded[m]) {
        lab[m] = 0;
        continue;
      }
    }
    if (pv == 0) {
      nbasin++;
      xattr.insert(xattr.end(), x, x + 3);
      pv = nbasin;
    }
    lab[m] = pv;
  }
  // The reference evaluates DISCARD each time a path ends at a maximum that
  // is still unlabelled: a discarded maximum within ratom of an attractor
  // created after it would join that attractor on its later visits (a label
  // that changes during the sweep).  Refused rather than approximated.
  if (o->discard && ratom > 1e-40) {
    const int first_new = atexist ? o->nat : 0;
    for (uint32_t q = 0; q < nmax; q++) {
      const uint32_t m = order[q];
      if (lab[m] != 0) continue;
      const double *x = dv + 3 * (size_t)m;
      for (int l = first_new; l < nbasin; l++) {
        const double xd[3] = {x[0] - xattr[3 * l], x[1] - xattr[3 * l + 1], x[2] - xattr[3 * l + 2]};
        if (lattice_dist(o->x2c, xd) < ratom)
          fail(C2GPU_ERR_UNSUPPORTED,
               "c2gpu_bader: a discarded maximum lies within ratom of an attractor found after it (the "
               "reference would assign its later paths to that attractor)");
      }
    }
  }
  return nbasin;
}

}  // namespace c2

using namespace c2;

extern "C" {

int c2gpu_ctx_create(c2gpu_ctx **ctx, int device) {
  if (!ctx) return C2GPU_ERR_ARG;
  *ctx = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return C2GPU_ERR_HIP;
  c2gpu_ctx *c = new c2gpu_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return C2GPU_ERR_HIP;
  }
  c->own_stream = true;
  // the device address of this device's scratch of c2gpu_basin_sums_noalloc
  try {
    c2::noalloc_init(device);
  } catch (const c2::Error &) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return C2GPU_ERR_HIP;
  }
  *ctx = c;
  return C2GPU_OK;
}

int c2gpu_ctx_destroy(c2gpu_ctx *ctx) {
  if (!ctx) return C2GPU_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  destroy_workers(ctx);
  release_all(ctx);
  if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return C2GPU_OK;
}

const char *c2gpu_ctx_error(const c2gpu_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int c2gpu_ctx_set_stream(c2gpu_ctx *ctx, void *hip_stream) {
  return guarded(ctx, [&] {
    if (ctx->own_stream && ctx->stream) {
      C2_HIP(hipStreamSynchronize(ctx->stream));
      C2_HIP(hipStreamDestroy(ctx->stream));
    }
    if (hip_stream) {
      ctx->stream = (hipStream_t)hip_stream;
      ctx->own_stream = false;
    } else {
      C2_HIP(hipStreamCreate(&ctx->stream));
      ctx->own_stream = true;
    }
  });
}

int c2gpu_ctx_sync(c2gpu_ctx *ctx) {
  return guarded(ctx, [&] { C2_HIP(hipStreamSynchronize(ctx->stream)); });
}

int c2gpu_dev_alloc(c2gpu_ctx *ctx, uint64_t bytes, void **ptr) {
  return guarded(ctx, [&] {
    if (!ptr) fail(C2GPU_ERR_ARG, "null argument");
    *ptr = nullptr;
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) fail(C2GPU_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
  });
}

int c2gpu_dev_free(c2gpu_ctx *ctx, void *ptr) {
  return guarded(ctx, [&] {
    C2_HIP(hipStreamSynchronize(ctx->stream));
    if (ptr) C2_HIP(hipFree(ptr));
  });
}

int c2gpu_memcpy(c2gpu_ctx *ctx, void *dst, const void *src, uint64_t bytes, int kind) {
  return guarded(ctx, [&] {
    hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                      : kind == 2 ? hipMemcpyDeviceToHost
                      : kind == 3 ? hipMemcpyDeviceToDevice
                                  : hipMemcpyDefault;
    if (kind < 1 || kind > 3) fail(C2GPU_ERR_ARG, "c2gpu_memcpy: bad kind");
    C2_HIP(hipMemcpyAsync(dst, src, bytes, k, ctx->stream));
    C2_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int c2gpu_ctx_timing(c2gpu_ctx *ctx, int enable) {
  return guarded(ctx, [&] {
    flush_timing(ctx);
    ctx->ktime.clear();
    ctx->timing = enable != 0;
  });
}

int c2gpu_ctx_kernel_time(c2gpu_ctx *ctx, const char *kernel, double *ms, int64_t *launches) {
  return guarded(ctx, [&] {
    if (!kernel) fail(C2GPU_ERR_ARG, "null kernel name");
    auto it = ctx->ktime.find(kernel);
    if (ms) *ms = it == ctx->ktime.end() ? 0.0 : it->second.ms;
    if (launches) *launches = it == ctx->ktime.end() ? 0 : it->second.launches;
  });
}

int c2gpu_ctx_kernel_bytes(c2gpu_ctx *ctx, const char *kernel, double *bytes) {
  return guarded(ctx, [&] {
    if (!kernel || !bytes) fail(C2GPU_ERR_ARG, "null argument");
    auto it = ctx->ktime.find(kernel);
    *bytes = it == ctx->ktime.end() ? 0.0 : it->second.bytes;
  });
}

int c2gpu_ctx_kernel_names(c2gpu_ctx *ctx, char *names_out, uint64_t len) {
  return guarded(ctx, [&] {
    if (!names_out || len == 0) fail(C2GPU_ERR_ARG, "null buffer");
    std::string s;
    for (auto &kv : ctx->ktime) {
      if (!s.empty()) s += ",";
      s += kv.first;
    }
    const size_t m = s.size() < len - 1 ? s.size() : (size_t)len - 1;
    std::memcpy(names_out, s.data(), m);
    names_out[m] = '\0';
  });
}

int c2gpu_bader_metrics(const double x2c[9], const int n[3], double lat2car[9], double car2lat[9],
                        double lat_i_dist[27]) {
  if (!x2c || !n || !lat2car || !car2lat || !lat_i_dist) return C2GPU_ERR_ARG;
  for (int c = 0; c < 3; c++)
    if (n[c] < 1) return C2GPU_ERR_ARG;
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) lat2car[r + 3 * c] = x2c[r + 3 * c] / (double)n[c];
  std::memcpy(car2lat, lat2car, 9 * sizeof(double));
  if (!inv3_lapack(car2lat)) return C2GPU_ERR_ARG;
  for (int i = -1; i <= 1; i++)
    for (int j = -1; j <= 1; j++)
      for (int k = -1; k <= 1; k++) {
        const double dlat[3] = {(double)i, (double)j, (double)k};
        double dcar[3];
        mat3_apply(lat2car, dlat, dcar);
        const double d = std::sqrt(dcar[0] * dcar[0] + dcar[1] * dcar[1] + dcar[2] * dcar[2]);
        lat_i_dist[(i + 1) + 3 * (j + 1) + 9 * (k + 1)] = (i == 0 && j == 0 && k == 0) ? 0.0 : 1.0 / d;
      }
  return C2GPU_OK;
}

int c2gpu_bader_step_pointers(c2gpu_ctx *ctx, const double *rho, const int n[3], const double car2lat[9],
                              const double lat_i_dist[27], int flags, int64_t *ongrid_next,
                              int64_t *neargrid_next0, double *dr1_out) {
  return guarded(ctx, [&] {
    if (!rho || !n || !car2lat || !lat_i_dist) fail(C2GPU_ERR_ARG, "null argument");
    if (flags & C2GPU_PTR_DEVICE) {
      require_device_ptr(rho, "rho");
      if (ongrid_next) require_device_ptr(ongrid_next, "ongrid_next");
      if (neargrid_next0) require_device_ptr(neargrid_next0, "neargrid_next0");
      if (dr1_out) require_device_ptr(dr1_out, "dr1_out");
    }
    step_pointers_run(ctx, rho, n, car2lat, lat_i_dist, flags, ongrid_next, neargrid_next0, dr1_out);
  });
}

int c2gpu_bader(c2gpu_ctx *ctx, const double *rho, const int n[3], const double car2lat[9],
                const double lat_i_dist[27], const c2gpu_attractor_opts *opts, int flags, int32_t *label_out,
                int *nattr, double *xattr_out, int xattr_cap, c2gpu_bader_stats *stats) {
  return guarded(ctx, [&] {
    if (!rho || !n || !car2lat || !lat_i_dist || !label_out) fail(C2GPU_ERR_ARG, "null argument");
    if (flags & C2GPU_PTR_DEVICE) {
      require_device_ptr(rho, "rho");
      require_device_ptr(label_out, "label_out");
    }
    bader_run(ctx, rho, n, car2lat, lat_i_dist, opts, flags, label_out, nattr, xattr_out, xattr_cap, stats);
  });
}

int c2gpu_relabel(c2gpu_ctx *ctx, int64_t npts, const int32_t *map, int nmap, int flags, int32_t *label) {
  return guarded(ctx, [&] {
    if (!map || !label) fail(C2GPU_ERR_ARG, "null argument");
    if (flags & C2GPU_PTR_DEVICE) require_device_ptr(label, "label");
    relabel_run(ctx, npts, map, nmap, flags, label);
  });
}

int c2gpu_basin_sums(c2gpu_ctx *ctx, const int32_t *label, int64_t npts, int nattr, int nfields,
                     const double *const *fields, int flags, double *sums_out, int64_t *counts_out) {
  return guarded(ctx, [&] {
    if (!label || (nfields > 0 && !fields)) fail(C2GPU_ERR_ARG, "null argument");
    if (flags & C2GPU_PTR_DEVICE) {
      require_device_ptr(label, "label");
      for (int q = 0; q < nfields; q++) require_device_ptr(fields[q], "fields[k]");
    }
    basin_sums_run(ctx, label, npts, nattr, nfields, fields, flags, sums_out, counts_out);
  });
}

int c2gpu_basin_sums_noalloc(c2gpu_ctx *ctx, const int32_t *label, int64_t npts, int nattr, int nfields,
                             const double *const *fields, double *sums_out, int64_t *counts_out) {
  return guarded(ctx, [&] {
    if (!label || (nfields > 0 && (!fields || !sums_out)) || (nfields == 0 && !counts_out))
      fail(C2GPU_ERR_ARG, "null argument");
    require_device_ptr(label, "label");
    for (int q = 0; q < nfields; q++) require_device_ptr(fields[q], "fields[k]");
    if (nfields > 0) require_device_ptr(sums_out, "sums_out");
    if (counts_out) require_device_ptr(counts_out, "counts_out");
    basin_sums_noalloc_run(ctx, label, npts, nattr, nfields, fields, sums_out, counts_out);
  });
}

int c2gpu_fft_deriv(c2gpu_ctx *ctx, const double *rho, const int n[3], const double x2c[9], int kind,
                    int flags, double *out) {
  return guarded(ctx, [&] {
    if (!rho || !n || !x2c || !out) fail(C2GPU_ERR_ARG, "null argument");
    if (flags & C2GPU_PTR_DEVICE) {
      require_device_ptr(rho, "rho");
      require_device_ptr(out, "out");
    }
    fft_deriv_run(ctx, rho, n, x2c, kind, flags, out);
  });
}

static void nci_checked(c2gpu_ctx *ctx, const double *rho, const int n[3], const double x2c[9], int scheme,
                        int precision, const c2gpu_promol *promol, int zfast, int flags, void *crho_out,
                        void *cgrad_out, int out_f32) {
  if (!n || !x2c || !crho_out || !cgrad_out) fail(C2GPU_ERR_ARG, "null argument");
  const bool grid_in = (scheme == C2GPU_NCI_FFT || scheme == C2GPU_NCI_GRID);
  if (grid_in && !rho) fail(C2GPU_ERR_ARG, "null grid");
  if (scheme == C2GPU_NCI_PROMOL && !promol) fail(C2GPU_ERR_ARG, "null promolecular tables");
  if (flags & C2GPU_PTR_DEVICE) {
    if (grid_in) require_device_ptr(rho, "rho");
    require_device_ptr(crho_out, "crho_out");
    require_device_ptr(cgrad_out, "cgrad_out");
  }
  // crystal%m_c2x and the grid's c2xl: critic2's matinv of the cell matrix
  double c2x[9];
  std::memcpy(c2x, x2c, sizeof(c2x));
  if (!inv3_lapack(c2x)) fail(C2GPU_ERR_ARG, "c2gpu_nci: singular cell matrix");
  nci_run(ctx, rho, n, x2c, c2x, scheme, precision, promol, zfast, flags, crho_out, cgrad_out, out_f32);
}

int c2gpu_nci(c2gpu_ctx *ctx, const double *rho, const int n[3], const double x2c[9], int scheme,
              int precision, const c2gpu_promol *promol, int zfast, int flags, double *crho_out,
              double *cgrad_out) {
  return guarded(ctx, [&] {
    nci_checked(ctx, rho, n, x2c, scheme, precision, promol, zfast, flags, crho_out, cgrad_out, 0);
  });
}

int c2gpu_nci_f32(c2gpu_ctx *ctx, const double *rho, const int n[3], const double x2c[9], int scheme,
                  int precision, const c2gpu_promol *promol, int zfast, int flags, float *crho_out,
                  float *cgrad_out) {
  return guarded(ctx, [&] {
    nci_checked(ctx, rho, n, x2c, scheme, precision, promol, zfast, flags, crho_out, cgrad_out, 1);
  });
}

int c2gpu_batch_run(c2gpu_ctx *ctx, int ncell, const c2gpu_cell *cells, int attr_mode, double ratom,
                    int nspecies, const c2gpu_radial_table *tables, int nstreams, int flags,
                    c2gpu_batch_stats *stats) {
  return guarded(ctx, [&] {
    batch_run(ctx, ncell, cells, attr_mode, ratom, nspecies, tables, nstreams, flags, stats);
  });
}

int c2gpu_synth_slater(c2gpu_ctx *ctx, const double x2c[9], const int n[3], int nat, const double *xat,
                       const int *zat, double rimg, int flags, double *rho_out) {
  return guarded(ctx, [&] {
    if (!x2c || !n || !xat || !zat || !rho_out) fail(C2GPU_ERR_ARG, "null argument");
    if (flags & C2GPU_PTR_DEVICE) require_device_ptr(rho_out, "rho_out");
    synth_slater_run(ctx, x2c, n, nat, xat, zat, rimg, flags, rho_out);
  });
}

int c2gpu_cube_read_header(c2gpu_ctx *ctx, const char *path, int binary, c2gpu_cube_header *hdr) {
  return guarded(ctx, [&] {
    if (!hdr) fail(C2GPU_ERR_ARG, "hdr is NULL");
    memset(hdr, 0, sizeof(*hdr));
    cube_header(path, hdr, binary);
  });
}

int c2gpu_cube_read(c2gpu_ctx *ctx, const char *path, int flags, c2gpu_cube_header *hdr, double *f) {
  return guarded(ctx, [&] { cube_read_run(ctx, path, flags, hdr, f); });
}

int c2gpu_cube_parse(c2gpu_ctx *ctx, const char *text, uint64_t len, int flags, c2gpu_cube_header *hdr,
                     double *f) {
  return guarded(ctx, [&] { cube_parse_run(ctx, text, len, flags, hdr, f); });
}

int c2gpu_bincube_read(c2gpu_ctx *ctx, const char *path, int flags, c2gpu_cube_header *hdr, double *f) {
  return guarded(ctx, [&] { bincube_read_run(ctx, path, flags, hdr, f); });
}

int c2gpu_cube_format(c2gpu_ctx *ctx, const double *f, const int n[3], int style, int zfast, int flags,
                      char *text, uint64_t cap, uint64_t *len) {
  return guarded(ctx, [&] { cube_format_run(ctx, f, n, style, zfast, flags, text, cap, len); });
}

int c2gpu_cube_write(c2gpu_ctx *ctx, const char *path, const c2gpu_cube_header *hdr, const int *zat,
                     const double *xat, const double *f, int style, int flags) {
  return guarded(ctx, [&] { cube_write_run(ctx, path, hdr, zat, xat, f, style, flags); });
}

int c2gpu_nci_cutoff(c2gpu_ctx *ctx, const void *crho, void *cgrad, const void *rhoat, const int n[3],
                     int precision, const c2gpu_nci_cut *opts, int flags, char *dat, uint64_t cap,
                     uint64_t *len, double sums[3]) {
  return guarded(ctx, [&] {
    nci_cutoff_run(ctx, crho, cgrad, rhoat, nullptr, 0, 0.0, 0.0, n, precision, opts, flags, dat, cap, len, sums);
  });
}

int c2gpu_nci_cutoff_frag(c2gpu_ctx *ctx, const void *crho, void *cgrad, const void *rhoat, const void *rhofrag,
                          int nfrag, double rhoparam, double rhoparam2, const int n[3], int precision,
                          const c2gpu_nci_cut *opts, int flags, char *dat, uint64_t cap, uint64_t *len,
                          double sums[3]) {
  return guarded(ctx, [&] {
    nci_cutoff_run(ctx, crho, cgrad, rhoat, rhofrag, nfrag, rhoparam, rhoparam2, n, precision, opts, flags, dat,
                   cap, len, sums);
  });
}

int c2gpu_nci_promol_nodes(c2gpu_ctx *ctx, const c2gpu_nci_nodes *nodes, const double x2c[9],
                           const c2gpu_promol *promol, int precision, int flags, void *crho_out, void *cgrad_out,
                           void *rhoat_out) {
  return guarded(ctx, [&] {
    nci_promol_nodes_run(ctx, nodes, x2c, promol, precision, flags, crho_out, cgrad_out, rhoat_out);
  });
}

int c2gpu_nci_fragments(c2gpu_ctx *ctx, const c2gpu_nci_nodes *nodes, int nspecies, const c2gpu_radial_table *tables,
                        int nfrag, const int *frag_nat, const double *xyz, const int *species, int precision,
                        int flags, void *rhofrag_out) {
  return guarded(ctx, [&] {
    nci_fragments_run(ctx, nodes, nspecies, tables, nfrag, frag_nat, xyz, species, precision, flags, rhofrag_out);
  });
}

int c2gpu_ws_facets(const double m[9], int *nvec, int vec[42], double area[14]) {
  if (!m || !nvec || !vec || !area) return C2GPU_ERR_ARG;
  try {
    *nvec = ws_facets(m, vec, area);
    return C2GPU_OK;
  } catch (const Error &e) {
    return e.code;
  }
}

int c2gpu_yt(c2gpu_ctx *ctx, const double *rho, const int n[3], const double x2c[9],
             const c2gpu_attractor_opts *opts, int nfields, const double *const *fields, int flags,
             int32_t *label_out, int *nattr_out, double *xattr_out, int xattr_cap, double *vol_out,
             double *sums_out, c2gpu_yt_stats *stats) {
  return guarded(ctx, [&] {
    yt_run(ctx, rho, n, x2c, opts, nfields, fields, flags, label_out, nattr_out, xattr_out, xattr_cap, vol_out,
           sums_out, stats);
  });
}

int c2gpu_multipoles(c2gpu_ctx *ctx, const int32_t *label, const double *f, const int n[3], const double x2c[9],
                     int nattr, const double *xattr, int lmax, int flags, double *mpole_out) {
  return guarded(ctx, [&] { mpole_run(ctx, label, f, n, x2c, nattr, xattr, lmax, flags, mpole_out); });
}

int c2gpu_yt_remap(c2gpu_ctx *ctx, const int32_t *map, int nmap, int nattr_new) {
  return guarded(ctx, [&] { yt_remap_run(ctx, map, nmap, nattr_new); });
}

int c2gpu_yt_multipoles(c2gpu_ctx *ctx, const int32_t *label, const double *f, const int n[3], const double x2c[9],
                        int nattr, const double *xattr, int lmax, int flags, double *mpole_out) {
  return guarded(ctx, [&] { yt_mpole_run(ctx, label, f, n, x2c, nattr, xattr, lmax, flags, mpole_out); });
}

}  // extern "C"
