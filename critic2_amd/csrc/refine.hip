// Edge refinement of bader_integrate (refine_edge, bader@proc.f90:300-422),
// order-free formulation.
//
// The reassign sweeps read only state frozen during the sweep (known == 2
// marks and the labels at their ends), so every point to reassign ascends
// independently.  The pass >= 2 marking loop (:329-366) depends on the
// visiting order only through local interactions, and reduces to closed-form
// per-point rules (k_refn_mark).
#include <hip/hip_runtime.h>

#include <string>

#include "bader_dev.hpp"
#include "common.hpp"
#include "devutil.hpp"

namespace c2 {

constexpr int REF_SLOTS = 64;

__global__ void k_ref1_mark(BParams P, uint32_t N, const int32_t *__restrict__ V,
                            uint8_t *__restrict__ E, unsigned long long *nedge) {
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  int mark = 0;
  if (y < N) {
    int i, j, k;
    coords(P, y, i, j, k);
    const int32_t v = abs(V[y]);
    bool edge = false;
    for (int d1 = -1; d1 <= 1 && !edge; d1++)
      for (int d2 = -1; d2 <= 1 && !edge; d2++)
        for (int d3 = -1; d3 <= 1 && !edge; d3++)
          if (abs(V[lin3(P, wrap1(i + d1, P.n1), wrap1(j + d2, P.n2), wrap1(k + d3, P.n3))]) != v)
            edge = true;
    mark = edge && !is_max(P, i, j, k);
    E[y] = (uint8_t)mark;
  }
  uint64_t m = mark;
  for (int off = 32; off > 0; off >>= 1) m += __shfl_down(m, off, 64);
  // spread over REF_SLOTS counters (one address taking an atomic from most
  // wavefronts of the grid serialises them); the host sums them
  if ((threadIdx.x & 63) == 0 && m)
    atomicAdd(nedge + ((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (REF_SLOTS - 1)),
              (unsigned long long)m);
}

// volnum = -volnum on marked points; known = 0 on their 27-neighbourhoods
// (reassign_volnum_ongrid2, :756-771).
__global__ void k_ref1_apply(BParams P, uint32_t N, const uint8_t *__restrict__ E,
                             int32_t *__restrict__ V, int8_t *__restrict__ known) {
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= N) return;
  int i, j, k;
  coords(P, y, i, j, k);
  if (E[y]) V[y] = -V[y];
  bool near = false;
  for (int d1 = -1; d1 <= 1 && !near; d1++)
    for (int d2 = -1; d2 <= 1 && !near; d2++)
      for (int d3 = -1; d3 <= 1 && !near; d3++)
        if (E[lin3(P, wrap1(i + d1, P.n1), wrap1(j + d2, P.n2), wrap1(k + d3, P.n3))]) near = true;
  if (near) known[y] = 0;
}

// Pass >=2 marking (bader@proc.f90:329-366) as closed-form rules.  Sources
// are the points negative at pass start.  A positive non-maximum within one
// cell of a source is marked (-volnum, known -1).  A source keeps its sign
// unless the point p+(1,1,1) read after its loop (:359) is not -2: that
// point is -2 exactly when it is a source processed later, or an earlier
// source that itself kept its sign.  A source whose sign was restored and
// that a later source touches is marked again (known -1).
__global__ void k_refn_mark(BParams P, uint32_t N, const int32_t *__restrict__ V,
                            const int8_t *__restrict__ known, int32_t *__restrict__ V2,
                            int8_t *__restrict__ known2) {
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= N) return;
  int i, j, k;
  coords(P, y, i, j, k);
  const int32_t v = V[y];
  int32_t vo = v;
  int8_t ko = known[y];
  if (v > 0) {
    if (!is_max(P, i, j, k)) {
      bool near = false;
      for (int d1 = -1; d1 <= 1 && !near; d1++)
        for (int d2 = -1; d2 <= 1 && !near; d2++)
          for (int d3 = -1; d3 <= 1 && !near; d3++)
            if (V[lin3(P, wrap1(i + d1, P.n1), wrap1(j + d2, P.n2), wrap1(k + d3, P.n3))] < 0)
              near = true;
      if (near) {
        vo = -v;
        ko = -1;
      }
    }
  } else if (v < 0) {
    // restore(y): follow p -> p+(1,1,1) through earlier sources
    uint32_t cur = y;
    int ci = i, cj = j, ck = k;
    bool restore = true;
    for (int guard = 0; guard < 64; guard++) {
      const int zi = wrap1(ci + 1, P.n1), zj = wrap1(cj + 1, P.n2), zk = wrap1(ck + 1, P.n3);
      const uint32_t z = lin3(P, zi, zj, zk);
      if (V[z] >= 0) {
        restore = true;
        break;
      }
      if (sweep_index(P, z) > sweep_index(P, cur)) {
        restore = false;
        break;
      }
      cur = z;
      ci = zi;
      cj = zj;
      ck = zk;
      if (z == y) {  // cycle only on degenerate 1-wide grids
        restore = false;
        break;
      }
    }
    if (!restore) {
      ko = -2;
    } else {
      const uint64_t sy = sweep_index(P, y);
      bool later = false;
      for (int d1 = -1; d1 <= 1 && !later; d1++)
        for (int d2 = -1; d2 <= 1 && !later; d2++)
          for (int d3 = -1; d3 <= 1 && !later; d3++) {
            const uint32_t z = lin3(P, wrap1(i + d1, P.n1), wrap1(j + d2, P.n2), wrap1(k + d3, P.n3));
            if (z != y && V[z] < 0 && sweep_index(P, z) > sy) later = true;
          }
      if (later) {
        ko = -1;
      } else {
        vo = -v;
        ko = -2;
      }
    }
  }
  V2[y] = vo;
  known2[y] = ko;
}

// known==2 -> 0 around every negative point (bader@proc.f90:368-388).
__global__ void k_refn_clear(BParams P, uint32_t N, const int32_t *__restrict__ V,
                             int8_t *__restrict__ known) {
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= N || known[y] != 2) return;
  int i, j, k;
  coords(P, y, i, j, k);
  for (int d1 = -1; d1 <= 1; d1++)
    for (int d2 = -1; d2 <= 1; d2++)
      for (int d3 = -1; d3 <= 1; d3++)
        if (V[lin3(P, wrap1(i + d1, P.n1), wrap1(j + d2, P.n2), wrap1(k + d3, P.n3))] < 0) {
          known[y] = 0;
          return;
        }
}

// Reassign sweep (bader@proc.f90:391-417): every negative point ascends
// against the frozen known==2 set; the label read at the path end is frozen
// during the sweep, so the points are independent.
// The points to reassign (negative) listed in grid order, so that every
// lane of the reassign sweep walks.
__global__ void k_neg_flags(uint32_t N, const int32_t *__restrict__ V, uint8_t *__restrict__ flag) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) flag[s] = V[s] < 0;
}

__global__ void k_reassign(BParams P, uint32_t nneg, const uint32_t *__restrict__ neg, const int32_t *__restrict__ V,
                           int32_t *__restrict__ Vout, int8_t *__restrict__ known,
                           uint32_t *scratch, int maxlen, int nbasin,
                           unsigned long long *nre, int *err) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nth = gridDim.x * blockDim.x;
  uint32_t *path = scratch + (size_t)tid * maxlen;
  uint64_t mine = 0;
  for (uint32_t t = tid; t < nneg; t += nth) {
    const uint32_t s = neg[t];
    const int32_t b = V[s];
    int L = walk_path(
        P, s, FlatPath{path}, maxlen, [&](uint32_t x) { return known[x] == 2; },
        [&](uint32_t x) {
          if (known[x] != 2) known[x] = 0;
        });
    if (L < 0) {
      atomicOr(err, 1);
      Vout[s] = b;
      continue;
    }
    const uint32_t e = path[L - 1];
    const int32_t pv = V[e];
    if (pv < 0 || pv > nbasin) {
      // the serial sweep stops here (:400-402).  An end e != s that is
      // negative cannot occur: a walk ends only at a grid maximum (a point
      // that is not is_max always has a strictly higher neighbour, so
      // step_ongrid moves, and a near-grid step back onto the path falls
      // back to step_ongrid) or at a known == 2 point, and the marking never
      // makes a maximum negative nor leaves a negative point known == 2.
      // Such an end would make the result depend on the serial order, so it
      // is reported as an internal inconsistency rather than followed.
      atomicOr(err, (pv < 0 && e != s) ? 8 : 4);
      Vout[s] = b;
      continue;
    }
    if (abs(b) != pv) {
      Vout[s] = -pv;
      mine++;
    } else {
      Vout[s] = pv;
    }
  }
  for (int off = 32; off > 0; off >>= 1) mine += __shfl_down(mine, off, 64);
  if ((threadIdx.x & 63) == 0 && mine) atomicAdd(nre, (unsigned long long)mine);
}

struct RefineResult {
  int iters;
  int64_t edge_points;
};

// V: labels after the first sweep (overwritten with the refined labels),
// known: known==2 marks; scratch: per-thread path buffers of maxlen points
// for nth threads.
RefineResult refine_run(c2gpu_ctx *ctx, const BParams &P, uint32_t N, int nbasin, int32_t *&V,
                        int32_t *&V2, int8_t *&known, int8_t *&known2, uint32_t *scratch, int nth,
                        int maxlen) {
  const int B = 256;
  const unsigned gN = grid_for(N, B);
  uint8_t *E = buf<uint8_t>(ctx, "bader_edge", N);
  unsigned long long *d_cnt = buf<unsigned long long>(ctx, "refine_counters", 4);
  unsigned long long *d_edge = buf<unsigned long long>(ctx, "refine_edges", REF_SLOTS);
  uint8_t *negf = buf<uint8_t>(ctx, "refine_negf", N);
  uint32_t *negl = buf<uint32_t>(ctx, "refine_negl", N);
  int *d_err = buf<int>(ctx, "refine_err", 1);
  C2_HIP(hipMemsetAsync(d_err, 0, sizeof(int), ctx->stream));
  RefineResult R{0, 0};
  int iter = 1;
  for (;; iter++) {
    if (iter > 100000) fail(C2GPU_ERR_NOCONVERGE, "c2gpu_bader: refinement did not converge");
    C2_HIP(hipMemsetAsync(d_cnt, 0, 4 * sizeof(unsigned long long), ctx->stream));
    if (iter == 1) {
      C2_HIP(hipMemsetAsync(d_edge, 0, REF_SLOTS * sizeof(unsigned long long), ctx->stream));
      C2_LAUNCH(ctx, k_ref1_mark, dim3(gN), dim3(B), 0, P, N, V, E, d_edge);
      C2_LAUNCH(ctx, k_ref1_apply, dim3(gN), dim3(B), 0, P, N, E, V, known);
      unsigned long long h[REF_SLOTS];
      copy_to_host(ctx, h, d_edge, sizeof(h));
      R.edge_points = 0;
      for (int q = 0; q < REF_SLOTS; q++) R.edge_points += (int64_t)h[q];
    } else {
      C2_LAUNCH(ctx, k_refn_mark, dim3(gN), dim3(B), 0, P, N, V, known, V2, known2);
      std::swap(V, V2);
      std::swap(known, known2);
      C2_LAUNCH(ctx, k_refn_clear, dim3(gN), dim3(B), 0, P, N, V, known);
    }
    // every point keeps its label but the negative ones, which the sweep
    // below reassigns from the listed points only
    C2_LAUNCH(ctx, k_neg_flags, dim3(gN), dim3(B), 0, N, V, negf);
    const uint32_t nneg = compact_flags(ctx, "refneg", negf, N, negl);
    C2_HIP(hipMemcpyAsync(V2, V, 4 * (size_t)N, hipMemcpyDeviceToDevice, ctx->stream));
    C2_LAUNCH(ctx, k_reassign, dim3(nth / B), dim3(B), 0, P, nneg, negl, V, V2, known, scratch, maxlen, nbasin,
              d_cnt, d_err);
    std::swap(V, V2);
    const unsigned long long nre = read_scalar(ctx, d_cnt);
    const int e = read_scalar(ctx, d_err);
    if (e & 1)
      fail(C2GPU_ERR_PATHLEN, "c2gpu_bader: refinement path longer than " + std::to_string(maxlen) + " points");
    if (e & 4) fail(C2GPU_ERR_NEWMAX, "refine_edge: should be no new maxima in edge refinement");
    if (e & 8) fail(C2GPU_ERR_DEGENERATE, "c2gpu_bader: internal inconsistency: a refinement path ends on a point being reassigned");
    if (nre == 0) break;
  }
  R.iters = iter;
  return R;
}

}  // namespace c2
