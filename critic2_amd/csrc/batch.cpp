// c2gpu_batch_run: Bader labels, per-basin Volume/Pop sums and the
// promolecular NCI block over many independent cells in one call.
//
// NCI: the bricks of every cell feed one persistent work-queue kernel
// (batch_nci_run, launched as a HIP graph on the caller's stream).
// Bader: the exact sweep of each cell is a host-driven fixpoint, so cells are
// taken from a shared queue by worker threads, each driving its own context
// and stream; small cells then overlap on the device instead of running one
// after another.  Results per cell are those of the single-cell entries.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "common.hpp"

namespace c2 {

void bader_run(c2gpu_ctx *, const double *, const int *, const double *, const double *,
               const c2gpu_attractor_opts *, int, int32_t *, int *, double *, int, c2gpu_bader_stats *);
void basin_sums_run(c2gpu_ctx *, const int32_t *, int64_t, int, int, const double *const *, int, double *,
                    int64_t *);
void batch_nci_run(c2gpu_ctx *, int, const c2gpu_cell *, int, const c2gpu_radial_table *, int, double *, int *);

// Worker contexts of a batch (created once, kept by the parent context).
static void ensure_workers(c2gpu_ctx *ctx, int nw) {
  while ((int)ctx->subs.size() < nw) {
    c2gpu_ctx *c = new c2gpu_ctx();
    c->device = ctx->device;
    if (hipStreamCreate(&c->stream) != hipSuccess) {
      delete c;
      fail(C2GPU_ERR_HIP, "c2gpu_batch_run: cannot create a worker stream");
    }
    c->own_stream = true;
    ctx->subs.push_back(c);
  }
}

// One cell's Bader labels and sums on a worker context.
static void bader_cell(c2gpu_ctx *w, const c2gpu_cell &C, int attr_mode, double ratom, int flags) {
  const bool dev = flags & C2GPU_PTR_DEVICE;
  const size_t N = (size_t)C.n[0] * C.n[1] * C.n[2];
  if (!C.rho) fail(C2GPU_ERR_ARG, "c2gpu_batch_run: cell without a grid");
  const double *d_rho = C.rho;
  if (!dev) {
    double *d = buf<double>(w, "batch_rho", N);
    C2_HIP(hipMemcpyAsync(d, C.rho, 8 * N, hipMemcpyHostToDevice, w->stream));
    d_rho = d;
  }
  int32_t *d_lab = dev ? C.label_out : buf<int32_t>(w, "batch_lab", N);
  c2gpu_attractor_opts o;
  std::memset(&o, 0, sizeof(o));
  o.mode = attr_mode;
  o.ratom = ratom;
  std::memcpy(o.x2c, C.x2c, sizeof(o.x2c));
  o.nat = C.nat;
  o.xat = C.xat;
  int nattr = 0;
  bader_run(w, d_rho, C.n, C.car2lat, C.lat_i_dist, &o, C2GPU_PTR_DEVICE, d_lab, &nattr, C.xattr_out, C.xattr_cap,
            nullptr);
  if (C.nattr_out) *C.nattr_out = nattr;
  if (C.counts_out || C.sums_out) {
    if (nattr > C.xattr_cap) fail(C2GPU_ERR_ARG, "c2gpu_batch_run: more attractors than xattr_cap");
    std::vector<double> sums(nattr > 0 ? nattr : 1);
    std::vector<int64_t> counts(nattr > 0 ? nattr : 1);
    const double *fields[1] = {d_rho};
    basin_sums_run(w, d_lab, (int64_t)N, nattr, 1, fields, C2GPU_PTR_DEVICE, sums.data(), counts.data());
    if (C.sums_out) std::memcpy(C.sums_out, sums.data(), sizeof(double) * nattr);
    if (C.counts_out) std::memcpy(C.counts_out, counts.data(), sizeof(int64_t) * nattr);
  }
  if (!dev) C2_HIP(hipMemcpyAsync(C.label_out, d_lab, 4 * N, hipMemcpyDeviceToHost, w->stream));
  C2_HIP(hipStreamSynchronize(w->stream));
}

void batch_run(c2gpu_ctx *ctx, int ncell, const c2gpu_cell *cells, int attr_mode, double ratom, int nspecies,
               const c2gpu_radial_table *tables, int nstreams, int flags, c2gpu_batch_stats *stats) {
  const auto t0 = std::chrono::steady_clock::now();
  if (ncell < 0 || (ncell > 0 && !cells)) fail(C2GPU_ERR_ARG, "c2gpu_batch_run: bad cell list");
  const int nw = nstreams > 0 ? nstreams : 8;
  if (nw > 64) fail(C2GPU_ERR_ARG, "c2gpu_batch_run: at most 64 streams");
  int64_t points = 0;
  bool any_bader = false;
  for (int c = 0; c < ncell; c++) {
    points += (int64_t)cells[c].n[0] * cells[c].n[1] * cells[c].n[2];
    if (cells[c].label_out) any_bader = true;
    if (flags & C2GPU_PTR_DEVICE) {
      if (cells[c].label_out) {
        require_device_ptr(cells[c].rho, "rho");
        require_device_ptr(cells[c].label_out, "label_out");
      }
      if (cells[c].crho_out) require_device_ptr(cells[c].crho_out, "crho_out");
      if (cells[c].cgrad_out) require_device_ptr(cells[c].cgrad_out, "cgrad_out");
    }
  }
  // Bader workers: a shared queue of cells; the first error stops the rest
  std::atomic<int> next{0};
  std::atomic<bool> failed{false};
  std::mutex emu;
  Error first{C2GPU_OK, ""};
  std::vector<std::thread> pool;
  if (any_bader) {
    ensure_workers(ctx, nw);
    for (int q = 0; q < nw; q++)
      pool.emplace_back([&, q]() {
        c2gpu_ctx *w = ctx->subs[q];
        try {
          C2_HIP(hipSetDevice(w->device));
          for (;;) {
            const int c = next.fetch_add(1);
            if (c >= ncell || failed.load()) break;
            if (cells[c].label_out) bader_cell(w, cells[c], attr_mode, ratom, flags);
          }
        } catch (const Error &e) {
          std::lock_guard<std::mutex> g(emu);
          if (!failed.exchange(true)) first = e;
        }
      });
  }
  // NCI of all cells on the caller's stream while the workers run
  double ms_nci = 0;
  int qwg = 0;
  try {
    batch_nci_run(ctx, ncell, cells, nspecies, tables, flags, &ms_nci, &qwg);
  } catch (const Error &e) {
    std::lock_guard<std::mutex> g(emu);
    if (!failed.exchange(true)) first = e;
  }
  for (auto &th : pool) th.join();
  if (failed.load()) throw first;
  if (stats) {
    stats->ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats->ms_nci = ms_nci;
    stats->points = points;
    stats->nstreams = any_bader ? nw : 0;
    stats->queue_workgroups = qwg;
  }
}

void destroy_workers(c2gpu_ctx *ctx) {
  for (c2gpu_ctx *w : ctx->subs) {
    (void)hipStreamSynchronize(w->stream);
    release_all(w);
    for (hipEvent_t e : w->evpool) (void)hipEventDestroy(e);
    if (w->own_stream && w->stream) (void)hipStreamDestroy(w->stream);
    delete w;
  }
  ctx->subs.clear();
}

}  // namespace c2
