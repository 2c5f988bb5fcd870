// Per-basin integration (intgrid_fields, integration@proc.f90:1195-1301) and
// the attractor relabel of int_reorder_gridout (:1113-1122).
//
// The reference does one masked full-grid pass per attractor (O(nattr*N)).
// Here one streaming pass reads each label and field value once: every
// wavefront takes 64 consecutive points, forms runs of equal labels with a
// segmented scan over the lanes (labels are spatially coherent, so runs are
// long), and the run tails add into the workgroup's LDS bins one after the
// other.  Block partials go to HBM and a second kernel adds them in block
// order, so the result is bitwise reproducible for a given grid size.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "common.hpp"

namespace c2 {

static constexpr int SUM_WAVES = 4;          // wavefronts per workgroup

struct FieldPtrs {
  const double *p[4];
};

// One wavefront's 64-point tile added into its bins wb[q*nattr + a]
// (q = 0 count, q > 0 field q-1): runs of equal labels found by a segmented
// scan over the lanes, each run's total added by its tail in lane order; a
// tile of one label (the common case, labels being spatially coherent) is a
// plain reduction.  The order of the additions depends only on the grid.
__device__ __forceinline__ void tile_sums(const int32_t *__restrict__ lab, uint64_t t, uint64_t b1, int nattr, int nf,
                                          const FieldPtrs &F, double *wb) {
  const int lane = threadIdx.x & 63;
  const uint64_t p = t + lane;
  const bool in = p < b1;
  const int32_t l = in ? lab[p] : -1;
  const int32_t l0 = __shfl(l, 0, 64);
  if (__all(in && l == l0)) {
    if (l0 < 1 || l0 > nattr) return;
    if (lane == 0) wb[l0 - 1] += 64.0;
    for (int q = 1; q <= nf; q++) {
      double v = F.p[q - 1][p];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      if (lane == 0) wb[(size_t)q * nattr + (l0 - 1)] += v;
    }
    return;
  }
  {
    const int32_t lprev = __shfl_up(l, 1, 64);
    const int32_t lnext = __shfl_down(l, 1, 64);
    const bool head = (lane == 0) || (l != lprev);
    const bool tail = (lane == 63) || (l != lnext);
    const bool valid = in && l >= 1 && l <= nattr;
    // segmented inclusive scans (count, then each field)
    for (int q = 0; q <= nf; q++) {
      double v = valid ? (q == 0 ? 1.0 : F.p[q - 1][p]) : 0.0;
      bool f = head;
      for (int off = 1; off < 64; off <<= 1) {
        const double ov = __shfl_up(v, off, 64);
        const int of = __shfl_up((int)f, off, 64);
        if (lane >= off && !f) {
          v = ov + v;
          f = f || of;
        }
      }
      // run tails add their run total, lane order (deterministic)
      uint64_t tails = __ballot(tail && valid);
      while (tails) {
        const int src = __ffsll((long long)tails) - 1;
        tails &= tails - 1;
        const double rv = __shfl(v, src, 64);
        const int32_t rl = __shfl(l, src, 64);
        if (lane == 0) wb[(size_t)q * nattr + (rl - 1)] += rv;
      }
    }
  }
}

// Bins layout in LDS and in the partials: [q][a], q = 0 count, q>0 field q-1.
__global__ void __launch_bounds__(64 * SUM_WAVES)
    k_basin_partials(const int32_t *__restrict__ lab, uint64_t npts, int nattr, int nf,
                     FieldPtrs F, uint64_t chunk, double *__restrict__ partial) {
  extern __shared__ double bins[];  // SUM_WAVES * (nf+1) * nattr
  const int wave = threadIdx.x >> 6;
  const int nb = (nf + 1) * nattr;
  for (int q = threadIdx.x; q < nb * SUM_WAVES; q += blockDim.x) bins[q] = 0.0;
  __syncthreads();
  double *wb = bins + (size_t)wave * nb;
  const uint64_t b0 = (uint64_t)blockIdx.x * chunk;
  uint64_t b1 = b0 + chunk;
  if (b1 > npts) b1 = npts;
  // each wave handles 64-point tiles, tiles interleaved over the waves
  for (uint64_t t = b0 + (uint64_t)wave * 64; t < b1; t += 64 * SUM_WAVES) tile_sums(lab, t, b1, nattr, nf, F, wb);
  __syncthreads();
  // fold the waves in wave order and write the block partial
  for (int q = threadIdx.x; q < nb; q += blockDim.x) {
    double s = 0.0;
    for (int w = 0; w < SUM_WAVES; w++) s += bins[(size_t)w * nb + q];
    partial[(size_t)blockIdx.x * nb + q] = s;
  }
}

// The same with each wavefront's bins in its own row of global memory
// (zeroed beforehand; rows in block-then-wave order): for more bins than
// the LDS holds.
__global__ void __launch_bounds__(64 * SUM_WAVES)
    k_basin_partials_rows(const int32_t *__restrict__ lab, uint64_t npts, int nattr, int nf, FieldPtrs F,
                          uint64_t chunk, double *__restrict__ rows) {
  const int wave = threadIdx.x >> 6;
  const int nb = (nf + 1) * nattr;
  double *wb = rows + ((size_t)blockIdx.x * SUM_WAVES + wave) * nb;
  const uint64_t b0 = (uint64_t)blockIdx.x * chunk;
  uint64_t b1 = b0 + chunk;
  if (b1 > npts) b1 = npts;
  for (uint64_t t = b0 + (uint64_t)wave * 64; t < b1; t += 64 * SUM_WAVES) tile_sums(lab, t, b1, nattr, nf, F, wb);
}

__global__ void k_basin_reduce(const double *__restrict__ partial, int nblocks, int nb,
                               double *__restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nb) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; b++) s += partial[(size_t)b * nb + q];
  out[q] = s;
}

__global__ void k_relabel(int32_t *__restrict__ lab, uint64_t npts, const int32_t *__restrict__ map,
                          int nmap) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const int32_t a = lab[p];
  if (a >= 1 && a <= nmap) lab[p] = map[a - 1];
}

void basin_sums_run(c2gpu_ctx *ctx, const int32_t *label, int64_t npts, int nattr, int nfields,
                    const double *const *fields, int flags, double *sums_out,
                    int64_t *counts_out) {
  if (npts <= 0 || nattr <= 0 || nfields < 0 || nfields > 4)
    fail(C2GPU_ERR_ARG, "c2gpu_basin_sums: need npts > 0, nattr > 0, 0 <= nfields <= 4");
  // more bins than the LDS holds go to per-wavefront rows in global memory
  const bool rows = (int64_t)nattr * (nfields + 1) * SUM_WAVES * 8 > 65536;
  hipStream_t st = ctx->stream;
  const bool dev = flags & C2GPU_PTR_DEVICE;
  const int32_t *d_lab = label;
  FieldPtrs F{};
  if (!dev) {
    int32_t *d = buf<int32_t>(ctx, "sum_lab", npts);
    C2_HIP(hipMemcpyAsync(d, label, 4 * (size_t)npts, hipMemcpyHostToDevice, st));
    d_lab = d;
    for (int q = 0; q < nfields; q++) {
      double *df = buf<double>(ctx, "sum_f" + std::to_string(q), npts);
      C2_HIP(hipMemcpyAsync(df, fields[q], 8 * (size_t)npts, hipMemcpyHostToDevice, st));
      F.p[q] = df;
    }
  } else {
    for (int q = 0; q < nfields; q++) F.p[q] = fields[q];
  }
  const int nb = (nfields + 1) * nattr;
  // chunk per block: multiple of 64*SUM_WAVES, about 2048 blocks in total
  const uint64_t tile = 64 * SUM_WAVES;
  uint64_t chunk = ((uint64_t)npts + 2047) / 2048;
  chunk = ((chunk + tile - 1) / tile) * tile;
  if (chunk < tile) chunk = tile;
  int nblocks = (int)(((uint64_t)npts + chunk - 1) / chunk);
  double *d_out = buf<double>(ctx, "sum_out", nb);
  if (!rows) {
    double *partial = buf<double>(ctx, "sum_partial", (size_t)nblocks * nb);
    const size_t lds = sizeof(double) * (size_t)nb * SUM_WAVES;
    C2_LAUNCH(ctx, k_basin_partials, dim3(nblocks), dim3(64 * SUM_WAVES), lds, d_lab, (uint64_t)npts, nattr, nfields,
              F, chunk, partial);
    C2_LAUNCH(ctx, k_basin_reduce, dim3(grid_for(nb, 256)), dim3(256), 0, partial, nblocks, nb, d_out);
  } else {
    // fewer, longer chunks keep the rows (8 bytes x bins x 4 per block) to
    // at most about 1 GiB
    const uint64_t row_bytes = 8ull * nb * SUM_WAVES;
    uint64_t maxb = ((uint64_t)1 << 30) / row_bytes;
    if (maxb < 1) maxb = 1;
    if ((uint64_t)nblocks > maxb) {
      chunk = ((uint64_t)npts + maxb - 1) / maxb;
      chunk = ((chunk + tile - 1) / tile) * tile;
      nblocks = (int)(((uint64_t)npts + chunk - 1) / chunk);
    }
    double *rowsb = buf<double>(ctx, "sum_rows", (size_t)nblocks * SUM_WAVES * nb);
    C2_HIP(hipMemsetAsync(rowsb, 0, sizeof(double) * (size_t)nblocks * SUM_WAVES * nb, st));
    C2_LAUNCH(ctx, k_basin_partials_rows, dim3(nblocks), dim3(64 * SUM_WAVES), 0, d_lab, (uint64_t)npts, nattr,
              nfields, F, chunk, rowsb);
    C2_LAUNCH(ctx, k_basin_reduce, dim3(grid_for(nb, 256)), dim3(256), 0, rowsb, nblocks * SUM_WAVES, nb, d_out);
  }
  std::vector<double> h(nb);
  copy_to_host(ctx, h.data(), d_out, 8 * (size_t)nb);
  C2_HIP(hipStreamSynchronize(st));
  if (counts_out)
    for (int a = 0; a < nattr; a++) counts_out[a] = (int64_t)h[a];
  if (sums_out)
    for (int q = 0; q < nfields; q++)
      for (int a = 0; a < nattr; a++) sums_out[(size_t)q * nattr + a] = h[(size_t)(q + 1) * nattr + a];
}

// ---- the per-basin sums without allocation (c2gpu_basin_sums_noalloc)
//
// The block partials live in a file-scope device array instead of a
// context buffer: every device the library runs on holds its own copy (32
// MiB) until the process ends.  Host code never passes the array's name to a
// kernel (in host code it names a host-side stand-in); it passes the
// device address hipGetSymbolAddress returned for the current device, kept
// per device from the first context created on it.
static constexpr size_t NOALLOC_PARTIALS = (size_t)2048 * 2048;  // 2048 blocks x 2048 bins (doubles)
static constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[NOALLOC_PARTIALS];
static double *scratch_dev[NOALLOC_MAX_DEVICES];
static std::mutex scratch_mu;

// Called with `device` current (c2gpu_ctx_create, or the entry's first call
// on a device): records the device address of this device's scratch.
void noalloc_init(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) fail(C2GPU_ERR_ARG, "noalloc scratch: device index out of range");
  std::lock_guard<std::mutex> g(scratch_mu);
  if (scratch_dev[device]) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    fail(C2GPU_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  scratch_dev[device] = static_cast<double *>(p);
}

// The block partials in bin order, added in block order (as k_basin_reduce),
// written straight to the caller's outputs: counts (bins 0..nattr-1) as
// int64, then each field's sums.
__global__ void k_basin_reduce_out(const double *__restrict__ partial, int nblocks, int nattr, int nf,
                                   double *__restrict__ sums_out, int64_t *__restrict__ counts_out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int nb = (nf + 1) * nattr;
  if (q >= nb) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; b++) s += partial[(size_t)b * nb + q];
  if (q < nattr) {
    if (counts_out) counts_out[q] = (int64_t)s;
  } else {
    sums_out[q - nattr] = s;
  }
}

void basin_sums_noalloc_run(c2gpu_ctx *ctx, const int32_t *label, int64_t npts, int nattr, int nfields,
                            const double *const *fields, double *sums_out, int64_t *counts_out) {
  if (npts <= 0 || nattr <= 0 || nfields < 0 || nfields > 4)
    fail(C2GPU_ERR_ARG, "c2gpu_basin_sums_noalloc: need npts > 0, nattr > 0, 0 <= nfields <= 4");
  if ((int64_t)nattr * (nfields + 1) * SUM_WAVES * 8 > 65536)
    fail(C2GPU_ERR_UNSUPPORTED, "c2gpu_basin_sums_noalloc: nattr*(nfields+1) above 2048 bins");
  FieldPtrs F{};
  for (int q = 0; q < nfields; q++) F.p[q] = fields[q];
  const int nb = (nfields + 1) * nattr;
  // the chunking of basin_sums_run: identical partials, identical sums
  const uint64_t tile = 64 * SUM_WAVES;
  uint64_t chunk = ((uint64_t)npts + 2047) / 2048;
  chunk = ((chunk + tile - 1) / tile) * tile;
  if (chunk < tile) chunk = tile;
  const int nblocks = (int)(((uint64_t)npts + chunk - 1) / chunk);
  if ((size_t)nblocks * (size_t)nb > NOALLOC_PARTIALS)
    fail(C2GPU_ERR_UNSUPPORTED, "c2gpu_basin_sums_noalloc: " + std::to_string(nblocks) + " blocks x " +
                                    std::to_string(nb) + " bins exceed the scratch of " +
                                    std::to_string(NOALLOC_PARTIALS) + " partials");
  if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES) fail(C2GPU_ERR_ARG, "device index out of range");
  double *partial = nullptr;
  {
    std::lock_guard<std::mutex> g(scratch_mu);
    partial = scratch_dev[ctx->device];
  }
  if (!partial) {
    noalloc_init(ctx->device);
    partial = scratch_dev[ctx->device];
  }
  const size_t lds = sizeof(double) * (size_t)nb * SUM_WAVES;
  // plain launches: no timing events (creating one may allocate)
  hipLaunchKernelGGL(k_basin_partials, dim3(nblocks), dim3(64 * SUM_WAVES), lds, ctx->stream, label,
                     (uint64_t)npts, nattr, nfields, F, chunk, partial);
  C2_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_basin_reduce_out, dim3((nb + 255) / 256), dim3(256), 0, ctx->stream, partial, nblocks,
                     nattr, nfields, sums_out, counts_out);
  C2_HIP(hipGetLastError());
}

void relabel_run(c2gpu_ctx *ctx, int64_t npts, const int32_t *map, int nmap, int flags,
                 int32_t *label) {
  if (npts <= 0 || nmap < 0) fail(C2GPU_ERR_ARG, "c2gpu_relabel: bad sizes");
  hipStream_t st = ctx->stream;
  int32_t *d_map = buf<int32_t>(ctx, "relabel_map", nmap > 0 ? nmap : 1);
  C2_HIP(hipMemcpyAsync(d_map, map, 4 * (size_t)nmap, hipMemcpyHostToDevice, st));
  int32_t *d_lab = label;
  if (!(flags & C2GPU_PTR_DEVICE)) {
    d_lab = buf<int32_t>(ctx, "relabel_lab", npts);
    C2_HIP(hipMemcpyAsync(d_lab, label, 4 * (size_t)npts, hipMemcpyHostToDevice, st));
  }
  C2_LAUNCH(ctx, k_relabel, dim3(grid_for(npts, 256)), dim3(256), 0, d_lab,
                     (uint64_t)npts, d_map, nmap);
  if (!(flags & C2GPU_PTR_DEVICE))
    C2_HIP(hipMemcpyAsync(label, d_lab, 4 * (size_t)npts, hipMemcpyDeviceToHost, st));
  C2_HIP(hipStreamSynchronize(st));
}

}  // namespace c2
