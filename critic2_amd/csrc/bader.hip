// Near-grid Bader basin labelling on gfx950, exact to critic2's serial
// bader_integrate (bader@proc.f90:80-234): free ascent trajectories of every
// grid point, the first sweep as a parallel fixpoint (sweep.hip), the edge
// refinement (refine.hip) and the host-side attractor assignment.
//
// Compiled with -ffp-contract=off: ascent arithmetic follows the reference
// FP64 operation order exactly.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "bader_dev.hpp"
#include "common.hpp"
#include "devutil.hpp"

namespace c2 {

struct SweepResult {
  int iters;
  int64_t events_total;
  int64_t nstarts;
  const int32_t *ell;  // the converged story's committed path lengths (device)
};
SweepResult sweep_run(c2gpu_ctx *ctx, const BParams &P, uint32_t N, const uint64_t *toff, const uint32_t *tp,
                      const int32_t *maxlab, int32_t *V, int8_t *known);
struct RefineResult {
  int iters;
  int64_t edge_points;
};
RefineResult refine_run(c2gpu_ctx *ctx, const BParams &P, uint32_t N, int nbasin, int32_t *&V, int32_t *&V2,
                        int8_t *&known, int8_t *&known2, uint32_t *scratch, int nth, int maxlen);

static constexpr int MAX_PATH = 2048;  // positions are stamped in 11 bits

struct Never {
  __device__ bool operator()(uint32_t) const { return false; }
};
struct NoVisit {
  __device__ void operator()(uint32_t) const {}
};

// ---------------------------------------------------------------- trajectories

// Free trajectory (no known==2 stops) of every start, walked once into the
// thread's scratch (through an LDS ring) and appended to a pool through a
// bump allocator; pos[s] / len[s] locate it.  A pool overflow only counts
// the needed size.  Persistent, TRAJ_B threads per workgroup.
//
// Walk order: a wavefront takes a row of 64 neighbouring starts along x
// from a queue, and the lanes walk side by side.  The rows are ordered in
// blocks of 64 x 64 x 64 points (blocks x fastest), within a block in
// 4 x 4 groups of rows, so the walks in flight start in a region about
// 128 x 64 x 64 points wide whose ascents re-read the direction records of
// one neighbourhood.  At 512^3 this takes the walk from 389 ms (each thread
// striding over the grid in grid order) to 261 ms, at 768^3 from 1627 ms
// to 1228 ms; block edges of 16 to 128 points and row groups of 2 to 16
// measured the same.  (Letting each lane take its next start as soon as
// its own path ends was measured 2.4x slower: the lanes drift apart on the
// grid and the path copies interleave.)
constexpr int TRAJ_B = 256;
struct TrajOrder {
  uint32_t nbi, nbj, nbk;  // blocks along x, y, z
  uint32_t bj, bk;         // block edge along y and z (points); along x 64
  uint32_t g;              // rows per group edge (g x g rows a group)
  uint32_t nrows;          // nbi * nbj * nbk * bj * bk
};
__device__ __forceinline__ bool traj_row(const BParams &P, const TrajOrder &T, uint32_t r, int &i0, int &j,
                                         int &k) {
  const uint32_t per = T.bj * T.bk;
  const uint32_t b = r / per, q = r % per;
  const uint32_t bi = b % T.nbi, bj = (b / T.nbi) % T.nbj, bk = b / (T.nbi * T.nbj);
  const uint32_t gg = T.g * T.g, t = q / gg, u = q % gg, gj = T.bj / T.g;
  i0 = (int)(bi * 64);
  j = (int)(bj * T.bj + (t % gj) * T.g + u % T.g);
  k = (int)(bk * T.bk + (t / gj) * T.g + u / T.g);
  return j < P.n2 && k < P.n3;
}
__global__ void __launch_bounds__(TRAJ_B) k_traj(BParams P, TrajOrder T, unsigned *row_next, uint32_t *scratch,
                                                       int maxlen, uint32_t *__restrict__ pool,
                                                       unsigned long long pool_cap, unsigned long long *pool_top,
                                                       uint64_t *__restrict__ pos, uint32_t *__restrict__ len,
                                                       uint32_t *__restrict__ tpk, int *err) {
  __shared__ uint32_t ring[RING * TRAJ_B];
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const RingPath path{ring + threadIdx.x, scratch + (size_t)tid * maxlen, TRAJ_B};
  for (;;) {
    unsigned r = 0;
    if (lane == 0) r = atomicAdd(row_next, 1u);
    r = __shfl(r, 0, 64);
    if (r >= T.nrows) break;  // every lane of the wavefront leaves together
    int i0, j, k;
    if (!traj_row(P, T, r, i0, j, k)) continue;  // uniform over the wavefront
    const int i = i0 + lane;
    const bool live = i < P.n1;
    const uint32_t s = live ? lin3(P, i, j, k) : 0u;
    int L = 0;
    if (live) {
      L = walk_path(P, s, path, maxlen, Never{}, NoVisit{});
      if (L < 0) {
        atomicOr(err, 1);
        L = 1;
      }
    }
    // the wavefront's paths side by side in the pool: one reservation per
    // wavefront instead of one per lane on the shared counter
    uint32_t incl = (uint32_t)L;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t v = __shfl_up(incl, o, 64);
      if (lane >= o) incl += v;
    }
    unsigned long long base = 0;
    if (lane == 63) base = atomicAdd(pool_top, (unsigned long long)incl);
    base = __shfl(base, 63, 64);
    if (!live) continue;
    const unsigned long long at = base + (incl - (uint32_t)L);
    len[s] = (uint32_t)L;
    pos[s] = at;
    if (at + (unsigned long long)L <= pool_cap)
      for (int q = 0; q < L; q++) pool[at + q] = path.get(q, L);
    if (tpk)
      for (int q = 1; q <= TPK && q < L; q++) tpk[(size_t)(q - 1) * P.npts + s] = path.get(q, L);
  }
}

// The step direction table of BParams::G, one record per point.
__global__ void k_step_dirs(BParams P, uint32_t N, double4 *__restrict__ G) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  int p[3];
  coords(P, s, p[0], p[1], p[2]);
  double g[3];
  G[s] = step_dir(P, p, g) ? make_double4(g[0], g[1], g[2], P.f[s]) : make_double4(0.0, 0.0, 0.0, P.f[s]);
}

// Pool -> CSR in start order (toff = exclusive scan of len): REORDER_LANES
// lanes per path, so each copy moves whole 64-byte runs.
constexpr int REORDER_LANES = 16;
__global__ void k_traj_reorder(uint32_t N, const uint64_t *__restrict__ pos, const uint64_t *__restrict__ toff,
                               const uint32_t *__restrict__ pool, uint32_t *__restrict__ tp) {
  // grid-stride: N * REORDER_LANES lanes exceed one dispatch's 2^32 work-items on large grids
  const uint64_t nlanes = (uint64_t)N * REORDER_LANES;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < nlanes;
       g += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t s = g / REORDER_LANES;
    const uint32_t r = (uint32_t)(g % REORDER_LANES);
    const uint64_t a = pos[s], b = toff[s], n = toff[s + 1] - b;
    for (uint64_t q = r; q < n; q += REORDER_LANES) tp[b + q] = pool[a + q];
  }
}

__global__ void k_mark_ends(uint32_t N, const uint64_t *__restrict__ toff, const uint32_t *__restrict__ tp,
                            uint8_t *__restrict__ isend) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) isend[tp[toff[s + 1] - 1]] = 1;
}

__global__ void k_scatter_labels(uint32_t n, const uint32_t *__restrict__ idx, const int32_t *__restrict__ lab,
                                 int32_t *__restrict__ out) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[idx[q]] = lab[q];
}

// Earliest start (sweep index) whose path ends at each point: over the free
// trajectories (ell == null) or the committed paths of a story.  A path that
// reaches a maximum ends there, so at a maximum this is the first start to
// reach it: its discovery (bader@proc.f90:159-199).
__global__ void k_first_reach(BParams P, uint32_t N, const uint64_t *__restrict__ toff,
                              const uint32_t *__restrict__ tp, const int32_t *__restrict__ ell,
                              uint32_t *__restrict__ first) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const int e = ell ? ell[s] : (int)(toff[s + 1] - toff[s]);
  if (e <= 0) return;
  atomicMin(&first[tp[toff[s] + e - 1]], (uint32_t)sweep_index(P, s));
}

__global__ void k_gather_u32(uint32_t n, const uint32_t *__restrict__ idx, const uint32_t *__restrict__ src,
                             uint32_t *__restrict__ out) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[q] = src[idx[q]];
}

// ---------------------------------------------------------------- step table

__global__ void k_step_pointers(BParams P, uint32_t N, int64_t *ongrid, int64_t *near0, double *dr1) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  int p[3];
  coords(P, s, p[0], p[1], p[2]);
  if (ongrid) {
    int q[3] = {p[0], p[1], p[2]};
    if (!is_max(P, p[0], p[1], p[2])) step_ongrid(P, q);
    ongrid[s] = (int64_t)lin3(P, q[0], q[1], q[2]);
  }
  if (near0 || dr1) {
    double dr[3] = {0.0, 0.0, 0.0};
    int pm[3];
    int64_t tgt = s;
    const double4 t = P.G ? P.G[s] : make_double4(0.0, 0.0, 0.0, 0.0);
    if (step_near_candidate(P, p, t, dr, pm)) {
      if (pm[0] == p[0] && pm[1] == p[1] && pm[2] == p[2]) {  // known(pm)==1 (pm == p)
        step_ongrid(P, pm);
        dr[0] = dr[1] = dr[2] = 0.0;
      }
      tgt = (int64_t)lin3(P, pm[0], pm[1], pm[2]);
    }
    if (near0) near0[s] = tgt;
    if (dr1) {
      dr1[3 * (size_t)s + 0] = dr[0];
      dr1[3 * (size_t)s + 1] = dr[1];
      dr1[3 * (size_t)s + 2] = dr[2];
    }
  }
}

// ---------------------------------------------------------------- host side

static BParams make_params(const double *d_rho, const int n[3], const double car2lat[9],
                           const double lat_i_dist[27]) {
  BParams P;
  P.f = d_rho;
  P.n1 = n[0];
  P.n2 = n[1];
  P.n3 = n[2];
  P.inv1 = 1.0 / n[0];
  P.inv2 = 1.0 / n[1];
  P.inv3 = 1.0 / n[2];
  for (int q = 0; q < 9; q++) P.c[q] = car2lat[q];
  for (int q = 0; q < 27; q++) P.lid[q] = lat_i_dist[q];
  P.G = nullptr;
  P.tpk = nullptr;
  P.npts = (uint32_t)((uint64_t)n[0] * n[1] * n[2]);
  return P;
}

// Attractor of a new maximum in the default mode (bader@proc.f90:159-199):
// identify_atom with ratom = 1d40, the nearest atom of the cell list.
int nearest_atom_label(const c2gpu_attractor_opts *o, const double x[3]);
// Attractors of maxima taken in discovery order (bader@proc.f90:159-199).
int attractor_rule(const c2gpu_attractor_opts *o, uint32_t nmax, const uint32_t *order, const double *dv,
                   int32_t *lab, std::vector<double> &xattr, std::vector<int8_t> &discarded);

void bader_run(c2gpu_ctx *ctx, const double *rho, const int n[3], const double car2lat[9],
               const double lat_i_dist[27], const c2gpu_attractor_opts *opts, int flags, int32_t *label_out,
               int *nattr, double *xattr_out, int xattr_cap, c2gpu_bader_stats *stats) {
  if (!opts) fail(C2GPU_ERR_ARG, "c2gpu_bader: opts is NULL");
  if (opts->mode != C2GPU_ATTR_NEAREST_ATOM && opts->mode != C2GPU_ATTR_RATOM && opts->mode != C2GPU_ATTR_NOATOMS)
    fail(C2GPU_ERR_ARG, "c2gpu_bader: unknown attractor mode");
  const bool ordered = opts->mode != C2GPU_ATTR_NEAREST_ATOM;  // labels depend on discovery order
  if (opts->mode != C2GPU_ATTR_NOATOMS && (opts->nat < 1 || !opts->xat))
    fail(C2GPU_ERR_ARG, "c2gpu_bader: no atoms");
  if (opts->nat > 0 && !opts->xat) fail(C2GPU_ERR_ARG, "c2gpu_bader: null atom positions");
  if (opts->nat >= (1 << 20)) fail(C2GPU_ERR_ARG, "c2gpu_bader: too many atoms");
  for (int c = 0; c < 3; c++)
    if (n[c] < 3) fail(C2GPU_ERR_ARG, "c2gpu_bader: grid edges must be >= 3");
  const uint64_t N64 = (uint64_t)n[0] * n[1] * n[2];
  if (N64 >= (1ull << 32) - 1) fail(C2GPU_ERR_ARG, "c2gpu_bader: grid too large");
  if (N64 >= (1ull << 34)) fail(C2GPU_ERR_ARG, "c2gpu_bader: grid too large for event time stamps");
  const uint32_t N = (uint32_t)N64;
  hipStream_t st = ctx->stream;
  c2gpu_bader_stats S;
  std::memset(&S, 0, sizeof(S));
  Timer ttot(st);

  const double *d_rho = rho;
  if (!(flags & C2GPU_PTR_DEVICE)) {
    double *d = buf<double>(ctx, "bader_rho", N);
    C2_HIP(hipMemcpyAsync(d, rho, sizeof(double) * N, hipMemcpyHostToDevice, st));
    d_rho = d;
  }
  BParams P = make_params(d_rho, n, car2lat, lat_i_dist);
  const int B = 256;
  const unsigned gN = grid_for(N, B);
  int *d_err = buf<int>(ctx, "bader_err", 1);
  C2_HIP(hipMemsetAsync(d_err, 0, sizeof(int), st));

  // 1) free trajectories: one walk into a pool, then CSR in start order
  Timer ttraj(st);
  {  // every point's step direction, once, for the walks of this call
    double4 *G = buf<double4>(ctx, "bader_gdir", N);
    C2_LAUNCH(ctx, k_step_dirs, dim3(gN), dim3(B), 0, P, N, G);
    add_kernel_bytes(ctx, "k_step_dirs", 40.0 * (double)N);
    P.G = G;
  }
  const int nth = persistent_threads(ctx->device, B);
  uint32_t *scratch = buf<uint32_t>(ctx, "bader_scratch", (size_t)nth * MAX_PATH);
  uint32_t *d_len = buf<uint32_t>(ctx, "bader_len", (size_t)N + 1);
  uint64_t *d_pos = buf<uint64_t>(ctx, "bader_pos", N);
  uint64_t *d_toff = buf<uint64_t>(ctx, "bader_toff", (size_t)N + 1);
  unsigned long long *d_top = buf<unsigned long long>(ctx, "bader_pooltop", 1);
  C2_HIP(hipMemsetAsync(d_len + N, 0, sizeof(uint32_t), st));
  // the walk pool lives in the arena that later holds the sweep's event
  // lists (the walk is over once the CSR copy is made), so large cells do
  // not hold both or allocate either anew on every call
  auto &poolbuf = ctx->bufs["bader_arena"];
  // the walk pool: 48 points per start, or up to 80 when the device can hold
  // that pool and the CSR copy of it side by side (long ascents in large
  // cells would otherwise need a second walk)
  uint64_t cap = (uint64_t)N * 48;
  if (poolbuf.ptr && poolbuf.bytes / 4 > cap) cap = poolbuf.bytes / 4;
  {
    // keep room for the sweep's per-point state (about 112 bytes a point)
    // unless an earlier call left it allocated
    auto lq = ctx->bufs.find("sweep_longq");
    const bool have_sweep = lq != ctx->bufs.end() && lq->second.ptr && lq->second.bytes >= 4ull * N;
    const uint64_t spare = (8ull << 30) + (have_sweep ? 0 : 112ull * N), most = (uint64_t)N * 80;
    const uint64_t own = poolbuf.ptr ? poolbuf.bytes : 0;
    auto tp = ctx->bufs.find("bader_tp");
    const uint64_t tp_own = (tp != ctx->bufs.end() && tp->second.ptr) ? tp->second.bytes : 0;
    size_t fr = 0, tot = 0;
    C2_HIP(hipMemGetInfo(&fr, &tot));
    // pool (4 B a point) now, and the CSR copy (at most as large) next to it
    const uint64_t room = fr + own > spare ? fr + own - spare : 0;
    uint64_t afford = (room + tp_own) / 8;
    if (afford > room / 4) afford = room / 4;
    const uint64_t want = afford < most ? afford : most;
    if (want > cap) cap = want;
  }
  // the first TPK positions of every trajectory, position-major, when the
  // device has room for them beside the largest pool and CSR copy planned
  // and the sweep's state (4 * TPK bytes a point; large cells go without)
  uint32_t *d_tpk = nullptr;
  {
    auto tk = ctx->bufs.find("bader_tpk");
    const uint64_t need = 4ull * TPK * N, own = (tk != ctx->bufs.end() && tk->second.ptr) ? tk->second.bytes : 0;
    const uint64_t arena = poolbuf.ptr ? poolbuf.bytes : 0;
    size_t fr = 0, tot = 0;
    C2_HIP(hipMemGetInfo(&fr, &tot));
    const uint64_t rest = 8ull * cap + (8ull << 30) + 112ull * N;  // pool + CSR, spare, sweep state
    if (own >= need || fr + own + arena > need + rest) d_tpk = buf<uint32_t>(ctx, "bader_tpk", (size_t)TPK * N);
  }
  uint64_t ntraj = 0;
  for (int attempt = 0; attempt < 2; attempt++) {
    uint32_t *pool = buf<uint32_t>(ctx, "bader_arena", cap);
    cap = ctx->bufs["bader_arena"].bytes / 4;
    C2_HIP(hipMemsetAsync(d_top, 0, 8, st));
    TrajOrder T;
    T.bj = T.bk = 64;
    T.g = 4;
    T.nbi = (uint32_t)((n[0] + 63) / 64);
    T.nbj = (uint32_t)((n[1] + T.bj - 1) / T.bj);
    T.nbk = (uint32_t)((n[2] + T.bk - 1) / T.bk);
    const uint64_t rows = (uint64_t)T.nbi * T.nbj * T.nbk * T.bj * T.bk;
    if (rows >= (1ull << 32)) fail(C2GPU_ERR_ARG, "c2gpu_bader: grid too large for the walk order");
    T.nrows = (uint32_t)rows;
    unsigned *d_row = buf<unsigned>(ctx, "bader_rownext", 1);
    C2_HIP(hipMemsetAsync(d_row, 0, sizeof(unsigned), st));
    C2_LAUNCH(ctx, k_traj, dim3(nth / TRAJ_B), dim3(TRAJ_B), 0, P, T, d_row, scratch, MAX_PATH, pool,
              (unsigned long long)cap, d_top, d_pos, d_len, d_tpk, d_err);
    ntraj = read_scalar(ctx, d_top);
    if (ntraj <= cap) break;
    cap = ntraj;  // exact size now known: walk again into a pool that fits
  }
  if (read_scalar(ctx, d_err) & 1) fail(C2GPU_ERR_PATHLEN, "c2gpu_bader: ascent path longer than 2048 points");
  P.tpk = d_tpk;
  exclusive_scan(ctx, "len", d_len, d_toff, (int64_t)N + 1);
  uint32_t *d_tp = buf<uint32_t>(ctx, "bader_tp", ntraj);
  const unsigned gR = (unsigned)std::min<int64_t>(((int64_t)N * REORDER_LANES + B - 1) / B, int64_t(1) << 20);
  C2_LAUNCH(ctx, k_traj_reorder, dim3(gR), dim3(B), 0, N, d_pos, d_toff,
            (const uint32_t *)ctx->bufs["bader_arena"].ptr, d_tp);
  S.traj_points = (int64_t)ntraj;
  if (ctx->bufs["bader_gdir"].bytes > (8ull << 30)) {
    // a large direction table leaves the event lists short of room: the
    // refinement's few walks compute their directions in place instead
    release(ctx, "bader_gdir");
    P.G = nullptr;
  }

  // endpoints (grid maxima and no-move points) -> attractors, on the host
  uint8_t *d_isend = buf<uint8_t>(ctx, "bader_isend", N);
  C2_HIP(hipMemsetAsync(d_isend, 0, N, st));
  C2_LAUNCH(ctx, k_mark_ends, dim3(gN), dim3(B), 0, N, d_toff, d_tp, d_isend);
  uint32_t *d_ends = buf<uint32_t>(ctx, "bader_ends", N);
  const uint32_t nends = compact_flags(ctx, "ends", d_isend, N, d_ends);
  std::vector<uint32_t> h_ends(nends);
  copy_to_host(ctx, h_ends.data(), d_ends, sizeof(uint32_t) * nends);
  C2_HIP(hipStreamSynchronize(st));
  // fractional position of each maximum, dv = (p-1)/n (bader@proc.f90:161)
  std::vector<double> h_dv(3 * (size_t)nends);
  for (uint32_t q = 0; q < nends; q++) {
    const uint32_t l = h_ends[q];
    const int i = (int)(l % (uint32_t)n[0]);
    const int j = (int)((l / (uint32_t)n[0]) % (uint32_t)n[1]);
    const int k = (int)(l / ((uint32_t)n[0] * (uint32_t)n[1]));
    h_dv[3 * q] = (double)i / n[0];
    h_dv[3 * q + 1] = (double)j / n[1];
    h_dv[3 * q + 2] = (double)k / n[2];
  }
  std::vector<int32_t> h_lab(nends);
  std::vector<double> h_xattr;
  std::vector<int8_t> h_discarded;  // DISCARD answers per maximum, asked once
  int nbasin = opts->nat;
  uint32_t *d_first = nullptr, *d_key = nullptr;
  // discovery order of the maxima from the earliest start whose path ends at
  // each (free trajectories first; then the committed paths of the story)
  auto discovery_labels = [&](const int32_t *ell) {
    C2_HIP(hipMemsetAsync(d_first, 0xff, sizeof(uint32_t) * N, st));
    C2_LAUNCH(ctx, k_first_reach, dim3(gN), dim3(B), 0, P, N, d_toff, d_tp, ell, d_first);
    std::vector<uint32_t> key(nends), order(nends);
    if (nends) {
      C2_LAUNCH(ctx, k_gather_u32, dim3(grid_for(nends, B)), dim3(B), 0, nends, d_ends, d_first, d_key);
      copy_to_host(ctx, key.data(), d_key, sizeof(uint32_t) * nends);
    }
    C2_HIP(hipStreamSynchronize(st));
    for (uint32_t q = 0; q < nends; q++) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
    std::vector<int32_t> lab(nends);
    nbasin = attractor_rule(opts, nends, order.data(), h_dv.data(), lab.data(), h_xattr, h_discarded);
    if (nbasin >= (1 << 20)) fail(C2GPU_ERR_ARG, "c2gpu_bader: too many attractors");
    return lab;
  };
  if (ordered) {
    d_first = buf<uint32_t>(ctx, "bader_first", N);
    d_key = buf<uint32_t>(ctx, "bader_firstkey", nends > 0 ? nends : 1);
    h_lab = discovery_labels(nullptr);
  } else {
    for (uint32_t q = 0; q < nends; q++) h_lab[q] = nearest_atom_label(opts, &h_dv[3 * q]);
    h_xattr.assign(opts->xat, opts->xat + 3 * (size_t)opts->nat);
  }
  S.nmaxima = (int)nends;
  int32_t *d_maxlab = buf<int32_t>(ctx, "bader_maxlab", N);
  int32_t *d_endlab = buf<int32_t>(ctx, "bader_endlab", nends > 0 ? nends : 1);
  auto upload_labels = [&]() {
    C2_HIP(hipMemcpyAsync(d_endlab, h_lab.data(), sizeof(int32_t) * nends, hipMemcpyHostToDevice, st));
    if (nends)
      C2_LAUNCH(ctx, k_scatter_labels, dim3(grid_for(nends, B)), dim3(B), 0, nends, d_ends, d_endlab, d_maxlab);
  };
  upload_labels();
  S.ms_traj = ttraj.stop();

  // 2) first sweep (bader@proc.f90:151-215).  With the NNM/RATOM/NOATOMS
  // rules a maximum's label depends on the order in which the sweep reaches
  // the maxima; labels and story are made consistent by iteration: the serial
  // sweep is the unique story whose own discovery order reproduces the labels
  // it was computed with.
  Timer tsweep(st);
  int32_t *V = buf<int32_t>(ctx, "bader_V", N), *V2 = buf<int32_t>(ctx, "bader_V2", N);
  int8_t *known = buf<int8_t>(ctx, "bader_known", N), *known2 = buf<int8_t>(ctx, "bader_known2", N);
  SweepResult SR = sweep_run(ctx, P, N, d_toff, d_tp, d_maxlab, V, known);
  for (int outer = 0; ordered; outer++) {
    std::vector<int32_t> lab = discovery_labels(SR.ell);
    if (lab == h_lab) break;
    if (outer >= 8) fail(C2GPU_ERR_NOCONVERGE, "c2gpu_bader: attractor discovery order did not settle");
    h_lab.swap(lab);
    upload_labels();
    SR = sweep_run(ctx, P, N, d_toff, d_tp, d_maxlab, V, known);
  }
  S.sweep_iters = SR.iters;
  S.events_total = SR.events_total;
  S.nstarts = SR.nstarts;
  S.ms_sweep = tsweep.stop();

  // 3) refinement (bader@proc.f90:217-224)
  Timer tref(st);
  const RefineResult RR = refine_run(ctx, P, N, nbasin, V, V2, known, known2, scratch, nth, MAX_PATH);
  S.refine_iters = RR.iters;
  S.edge_points = RR.edge_points;
  S.ms_refine = tref.stop();

  if (flags & C2GPU_PTR_DEVICE)
    C2_HIP(hipMemcpyAsync(label_out, V, sizeof(int32_t) * N, hipMemcpyDeviceToDevice, st));
  else
    C2_HIP(hipMemcpyAsync(label_out, V, sizeof(int32_t) * N, hipMemcpyDeviceToHost, st));
  C2_HIP(hipStreamSynchronize(st));
  S.nattr = nbasin;
  if (nattr) *nattr = nbasin;
  if (xattr_out)
    for (int a = 0; a < nbasin && a < xattr_cap; a++)
      for (int c = 0; c < 3; c++) xattr_out[3 * a + c] = h_xattr[3 * (size_t)a + c];
  S.ms_total = ttot.stop();
  if (stats) *stats = S;
}

void step_pointers_run(c2gpu_ctx *ctx, const double *rho, const int n[3], const double car2lat[9],
                       const double lat_i_dist[27], int flags, int64_t *ongrid, int64_t *near0,
                       double *dr1) {
  for (int c = 0; c < 3; c++)
    if (n[c] < 3) fail(C2GPU_ERR_ARG, "c2gpu_bader_step_pointers: grid edges must be >= 3");
  const uint64_t N64 = (uint64_t)n[0] * n[1] * n[2];
  if (N64 >= (1ull << 32) - 1) fail(C2GPU_ERR_ARG, "grid too large");
  const uint32_t N = (uint32_t)N64;
  hipStream_t st = ctx->stream;
  const bool dev = flags & C2GPU_PTR_DEVICE;
  const double *d_rho = rho;
  if (!dev) {
    double *d = buf<double>(ctx, "sp_rho", N);
    C2_HIP(hipMemcpyAsync(d, rho, sizeof(double) * N, hipMemcpyHostToDevice, st));
    d_rho = d;
  }
  int64_t *d_on = ongrid, *d_ne = near0;
  double *d_dr = dr1;
  if (!dev) {
    d_on = ongrid ? buf<int64_t>(ctx, "sp_on", N) : nullptr;
    d_ne = near0 ? buf<int64_t>(ctx, "sp_ne", N) : nullptr;
    d_dr = dr1 ? buf<double>(ctx, "sp_dr", 3 * (size_t)N) : nullptr;
  }
  BParams P = make_params(d_rho, n, car2lat, lat_i_dist);
  C2_LAUNCH(ctx, k_step_pointers, dim3(grid_for(N, 256)), dim3(256), 0, P, N, d_on, d_ne, d_dr);
  if (!dev) {
    if (ongrid) C2_HIP(hipMemcpyAsync(ongrid, d_on, 8 * (size_t)N, hipMemcpyDeviceToHost, st));
    if (near0) C2_HIP(hipMemcpyAsync(near0, d_ne, 8 * (size_t)N, hipMemcpyDeviceToHost, st));
    if (dr1) C2_HIP(hipMemcpyAsync(dr1, d_dr, 24 * (size_t)N, hipMemcpyDeviceToHost, st));
  }
  C2_HIP(hipStreamSynchronize(st));
}

}  // namespace c2