// nciplot's cutoff pass and .dat rows (nci@proc.f90:612-654) on the GPU.
//
// Over crho (sign(lambda2) rho * 100) and cgrad (s) in critic2's z-fastest
// node order, the loop order of the reference:
//   - a node is written to the .dat file as '(1p,E15.7,E15.7)' crho/100, s
//     when |crho| < rhocut*100, s < dimcut and the node counts as
//     intermolecular (with RHO_VOID: promolecular density below it);
//   - then s <- 100 where |crho| > rhoplot*100, crho is outside
//     srhorange*100, the node is not intermolecular, or (ONLYNEG) crho > 0;
//   - with RHO_VOID the void charge, promolecular charge and volume are
//     summed (returned unscaled; critic2 multiplies by Omega/N).
// With fragments a node is intermolecular only when the fragment densities
// sum to at least rhoparam2 times the promolecular density and none exceeds
// rhoparam times that sum (:622-626).
// Rows keep the node order: k_cut_mark flags the written nodes (and sums the
// void), a compaction lists them in node order, k_cut_rows formats row r at
// 31 r bytes, and k_cut_mask then applies the masks (the reference writes a
// row before it masks the node).
#include <hip/hip_runtime.h>

#include <string>

#include "common.hpp"
#include "devutil.hpp"
#include "format.hpp"
#include "pow5tab.hpp"

namespace c2 {

namespace {

constexpr int ROW = 31;  // two E15.7 fields and a newline
constexpr int CB = 256;

// Fragment arrays of the cutoff pass: rf[p + f npts], f < nfrag.
template <class T>
struct Frags {
  const T *rf;
  int nfrag;
  double rhoparam, rhoparam2;
};

// sum(rhofrag) >= rhoparam2 * rhoat .and. all(rhofrag <= sum(rhofrag) * rhoparam),
// the sum in fragment order as Fortran's sum intrinsic forms it
template <class T>
__device__ __forceinline__ bool frag_inter(const Frags<T> &F, uint64_t npts, uint64_t p, double ra) {
  if (F.nfrag <= 0) return true;
  double S = 0.0;
  for (int f = 0; f < F.nfrag; f++) S = S + (double)F.rf[(uint64_t)f * npts + p];
  bool ok = S >= F.rhoparam2 * ra;
  const double lim = S * F.rhoparam;
  for (int f = 0; f < F.nfrag; f++) ok = ok && (double)F.rf[(uint64_t)f * npts + p] <= lim;
  return ok;
}

template <class T>
__global__ void __launch_bounds__(CB) k_cut_mark(uint64_t npts, const T *__restrict__ crho,
                                                 const T *__restrict__ cgrad, const T *__restrict__ rhoat,
                                                 Frags<T> F, c2gpu_nci_cut o, uint8_t *__restrict__ flag,
                                                 double *__restrict__ part) {
  const uint64_t p = (uint64_t)blockIdx.x * CB + threadIdx.x;
  double q = 0, qp = 0, v = 0;
  if (p < npts) {
    const double c = (double)crho[p];
    const double g = (double)cgrad[p];
    bool inter = F.nfrag > 0 ? frag_inter(F, npts, p, (double)rhoat[p]) : true;
    if (o.rho_void > 0) {
      const double ra = (double)rhoat[p];
      inter = inter && ra < o.rho_void;
      if (ra < o.rho_void) {  // the void sums ignore the fragments
        q = fabs(c) / 100.0;
        qp = ra;
        v = 1;
      }
    }
    flag[p] = (fabs(c) < o.rhocut * 100.0 && g < o.dimcut && inter) ? 1 : 0;
  }
  if (o.rho_void > 0) {  // block partial sums, combined on the host in block order
    __shared__ double red[3][CB / 64];
    for (int off = 32; off > 0; off >>= 1) {
      q += __shfl_xor(q, off);
      qp += __shfl_xor(qp, off);
      v += __shfl_xor(v, off);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      red[0][w] = q;
      red[1][w] = qp;
      red[2][w] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
      double t = 0;
      for (int i = 0; i < CB / 64; i++) t += red[threadIdx.x][i];
      part[3 * (uint64_t)blockIdx.x + threadIdx.x] = t;
    }
  }
}

template <class T>
__global__ void __launch_bounds__(CB) k_cut_mask(uint64_t npts, const T *__restrict__ crho, T *__restrict__ cgrad,
                                                 const T *__restrict__ rhoat, Frags<T> F, c2gpu_nci_cut o) {
  const uint64_t p = (uint64_t)blockIdx.x * CB + threadIdx.x;
  if (p >= npts) return;
  const double c = (double)crho[p];
  const bool inter = (F.nfrag <= 0 || frag_inter(F, npts, p, (double)rhoat[p])) &&
                     (!(o.rho_void > 0) || (double)rhoat[p] < o.rho_void);
  bool mask = fabs(c) > o.rhoplot * 100.0 || !inter;
  mask = mask || c < o.srhorange[0] * 100.0 || c > o.srhorange[1] * 100.0 || !inter;
  mask = mask || (c > 0 && o.onlyneg);
  if (mask) cgrad[p] = (T)100.0;
}

// One row per written node: idx[r] is the node of row r.
template <class T>
__global__ void __launch_bounds__(CB) k_cut_rows(uint32_t nrow, const uint32_t *__restrict__ idx,
                                                 const T *__restrict__ crho, const T *__restrict__ cgrad,
                                                 const uint64_t *pow5, const int16_t *pow5s, char *__restrict__ text,
                                                 unsigned long long *__restrict__ bad) {
  const uint32_t r = blockIdx.x * CB + threadIdx.x;
  if (r >= nrow) return;
  const uint32_t p = idx[r];
  char a[16], b[16];
  const int ra = fmt_value_s<true, FMT_1P_E15_7>((double)crho[p] / 100.0, pow5, pow5s, a);
  const int rb = fmt_value_s<true, FMT_1P_E15_7>((double)cgrad[p], pow5, pow5s, b);
  if (ra | rb) atomicAdd(bad, 1ull);
  char *o = text + (uint64_t)r * ROW;
  for (int c = 0; c < 15; c++) o[c] = a[c];
  for (int c = 0; c < 15; c++) o[15 + c] = b[c];
  o[30] = '\n';
}

// crho, cgrad, rhoat: device arrays of npts nodes.  dat: device, cap bytes,
// or NULL for the size only (no masks, no sums).
template <class T>
void cut_run(c2gpu_ctx *ctx, uint64_t npts, const T *crho, T *cgrad, const T *rhoat, const Frags<T> &F,
             const c2gpu_nci_cut &o, char *dat, uint64_t cap, uint64_t *len, double *sums) {
  uint8_t *flag = buf<uint8_t>(ctx, "ncicut_flag", npts);
  const unsigned nb = grid_for((int64_t)npts, CB);
  double *part = buf<double>(ctx, "ncicut_part", 3 * (uint64_t)nb);
  const bool use_at = o.rho_void > 0 || F.nfrag > 0;
  C2_LAUNCH(ctx, k_cut_mark<T>, dim3(nb), dim3(CB), 0, npts, crho, cgrad, rhoat, F, o, flag, part);
  add_kernel_bytes(ctx, "k_cut_mark<T>",
                   (double)npts * (2 * sizeof(T) + 1 + (use_at ? sizeof(T) : 0) + F.nfrag * sizeof(T)));
  uint32_t *idx = buf<uint32_t>(ctx, "ncicut_idx", npts);
  const uint32_t nrow = compact_flags(ctx, "ncicut", flag, (uint32_t)npts, idx);
  *len = (uint64_t)nrow * ROW;
  if (!dat) return;
  if (cap < *len)
    fail(C2GPU_ERR_ARG, "nci cutoff: dat buffer holds " + std::to_string(cap) + " bytes, need " +
                            std::to_string(*len));
  const uint64_t *p5;
  const int16_t *p5s;
  pow5_device(ctx, p5, p5s);
  unsigned long long *bad = buf<unsigned long long>(ctx, "ncicut_bad", 1);
  C2_HIP(hipMemsetAsync(bad, 0, 8, ctx->stream));
  if (nrow) {
    C2_LAUNCH(ctx, k_cut_rows<T>, dim3(grid_for(nrow, CB)), dim3(CB), 0, nrow, idx, crho, cgrad, p5, p5s, dat,
              bad);
    add_kernel_bytes(ctx, "k_cut_rows<T>", (double)nrow * (4 + 2 * sizeof(T) + ROW));
  }
  C2_LAUNCH(ctx, k_cut_mask<T>, dim3(nb), dim3(CB), 0, npts, crho, cgrad, rhoat, F, o);
  add_kernel_bytes(ctx, "k_cut_mask<T>",
                   (double)npts * (2 * sizeof(T) + (use_at ? sizeof(T) : 0) + F.nfrag * sizeof(T)));
  if (read_scalar(ctx, bad)) fail(C2GPU_ERR_HIP, "nci cutoff: value formatting overflowed its big integers");
  if (sums) {
    sums[0] = sums[1] = sums[2] = 0;
    if (o.rho_void > 0) {
      std::vector<double> h(3 * (size_t)nb);
      copy_to_host(ctx, h.data(), part, h.size() * 8);
      C2_HIP(hipStreamSynchronize(ctx->stream));
      for (unsigned b = 0; b < nb; b++)
        for (int c = 0; c < 3; c++) sums[c] += h[3 * (size_t)b + c];
    }
  }
}

template <class T>
void cut_entry(c2gpu_ctx *ctx, const void *crho_, void *cgrad_, const void *rhoat_, const void *rhofrag_, int nfrag,
               double rhoparam, double rhoparam2, uint64_t npts, const c2gpu_nci_cut &o, int flags, char *dat,
               uint64_t cap, uint64_t *len, double *sums) {
  const bool dev = flags & C2GPU_PTR_DEVICE;
  const T *crho = static_cast<const T *>(crho_);
  T *cgrad = static_cast<T *>(cgrad_);
  const T *rhoat = static_cast<const T *>(rhoat_);
  Frags<T> F{static_cast<const T *>(rhofrag_), nfrag, rhoparam, rhoparam2};
  const bool use_at = o.rho_void > 0 || nfrag > 0;
  if (dev) {
    require_device_ptr(crho, "crho");
    require_device_ptr(cgrad, "cgrad");
    if (use_at) require_device_ptr(rhoat, "rhoat");
    if (nfrag > 0) require_device_ptr(F.rf, "rhofrag");
    if (dat) require_device_ptr(dat, "dat");
    cut_run<T>(ctx, npts, crho, cgrad, rhoat, F, o, dat, cap, len, sums);
  } else {
    T *dc = buf<T>(ctx, "ncicut_crho", npts);
    T *dg = buf<T>(ctx, "ncicut_cgrad", npts);
    T *da = nullptr;
    C2_HIP(hipMemcpyAsync(dc, crho, npts * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    C2_HIP(hipMemcpyAsync(dg, cgrad, npts * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    if (use_at) {
      da = buf<T>(ctx, "ncicut_rhoat", npts);
      C2_HIP(hipMemcpyAsync(da, rhoat, npts * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    }
    if (nfrag > 0) {
      T *df = buf<T>(ctx, "ncicut_rhofrag", npts * (uint64_t)nfrag);
      C2_HIP(hipMemcpyAsync(df, F.rf, npts * (uint64_t)nfrag * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
      F.rf = df;
    }
    char *dd = nullptr;
    if (dat) {
      const uint64_t c = cap < npts * (uint64_t)ROW ? cap : npts * (uint64_t)ROW;
      dd = buf<char>(ctx, "ncicut_dat", c > 0 ? c : 1);
      cut_run<T>(ctx, npts, dc, dg, da, F, o, dd, c, len, sums);
      if (*len) C2_HIP(hipMemcpyAsync(dat, dd, *len, hipMemcpyDeviceToHost, ctx->stream));
      C2_HIP(hipMemcpyAsync(cgrad, dg, npts * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    } else {
      cut_run<T>(ctx, npts, dc, dg, da, F, o, nullptr, 0, len, sums);
    }
  }
  C2_HIP(hipStreamSynchronize(ctx->stream));
}

}  // namespace

void nci_cutoff_run(c2gpu_ctx *ctx, const void *crho, void *cgrad, const void *rhoat, const void *rhofrag, int nfrag,
                    double rhoparam, double rhoparam2, const int n[3], int precision, const c2gpu_nci_cut *o,
                    int flags, char *dat, uint64_t cap, uint64_t *len, double *sums) {
  if (!crho || !cgrad || !n || !o || !len) fail(C2GPU_ERR_ARG, "nci cutoff: NULL argument");
  if (o->rho_void > 0 && !rhoat) fail(C2GPU_ERR_ARG, "nci cutoff: rho_void needs rhoat");
  if (nfrag < 0) fail(C2GPU_ERR_ARG, "nci cutoff: negative fragment count");
  if (nfrag > 0 && (!rhofrag || !rhoat)) fail(C2GPU_ERR_ARG, "nci cutoff: fragments need rhofrag and rhoat");
  for (int a = 0; a < 3; a++)
    if (n[a] < 1) fail(C2GPU_ERR_ARG, "nci cutoff: grid dimensions must be positive");
  const uint64_t npts = (uint64_t)n[0] * n[1] * n[2];
  if (npts >= (1ull << 32)) fail(C2GPU_ERR_ARG, "nci cutoff: grids need fewer than 2^32 nodes");
  if (precision == C2GPU_FP64)
    cut_entry<double>(ctx, crho, cgrad, rhoat, rhofrag, nfrag, rhoparam, rhoparam2, npts, *o, flags, dat, cap, len,
                      sums);
  else if (precision == C2GPU_FP32)
    cut_entry<float>(ctx, crho, cgrad, rhoat, rhofrag, nfrag, rhoparam, rhoparam2, npts, *o, flags, dat, cap, len,
                     sums);
  else
    fail(C2GPU_ERR_ARG, "nci cutoff: precision must be C2GPU_FP64 or C2GPU_FP32");
}

}  // namespace c2
