// Near-grid Bader basin labelling on gfx950, exact to critic2's serial
// bader_integrate (bader@proc.f90:80-234).
//
// The reference labels points in one serial sweep (do i; do j; do k) in
// which every path stops at points already marked known==2 and relabels the
// points it passes; the outcome depends on the visiting order.  Here the
// sweep is computed as the fixpoint of a parallel (Jacobi) iteration over the
// sweep "story" -- for every start point: is it a start, how long is its
// committed path, which label does it carry.  Given a story, the state the
// serial sweep would have at any moment is rebuilt from time-stamped events
// (time = sweep index << 10 | position on the path), and every decision is
// recomputed in parallel.  The serial recurrence has exactly one solution
// and any fixpoint of the iteration satisfies it, so the converged story is
// the serial sweep, bit for bit.  The refinement passes (refine_edge,
// :300-422) are order-free once rewritten: reassign sweeps read only frozen
// state, and the pass>=2 marking loop reduces to closed-form per-point rules.
//
// Compiled with -ffp-contract=off: ascent arithmetic follows the reference
// FP64 operation order exactly.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "bader_dev.hpp"
#include "common.hpp"

namespace c2 {

static constexpr uint64_t KT_INF = ~0ull;
static constexpr int POS_BITS = 10;    // position on a path: < 1024
static constexpr int LAB_BITS = 20;    // label: < 2^20
static constexpr int MAX_PATH = 1 << POS_BITS;
static constexpr int SORT_INLINE = 48; // event lists longer than this go to k_ev_sort_long

__device__ __forceinline__ uint64_t ev_time(uint64_t e) { return e >> LAB_BITS; }
__device__ __forceinline__ int32_t ev_lab(uint64_t e) { return (int32_t)(e & ((1u << LAB_BITS) - 1)); }

// Label of point y just after all events with time <= tau (inclusive) or
// time < tau (exclusive); 0 if unlabeled.  Events of y are time-sorted.
__device__ __forceinline__ int32_t v_at(const uint64_t *__restrict__ ev,
                                        const uint64_t *__restrict__ eoff, uint32_t y,
                                        uint64_t tau, bool inclusive) {
  // binary search for the first event after the cut (times are unique)
  uint64_t lo = eoff[y], hi = eoff[y + 1];
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    const uint64_t t = ev_time(ev[mid]);
    if (inclusive ? (t <= tau) : (t < tau))
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo > eoff[y]) ? ev_lab(ev[lo - 1]) : 0;
}

struct Never {
  __device__ bool operator()(uint32_t) const { return false; }
};
struct NoVisit {
  __device__ void operator()(uint32_t) const {}
};

// ---------------------------------------------------------------- trajectories

__global__ void k_traj_len(BParams P, uint32_t N, uint32_t *scratch, int maxlen,
                           uint32_t *__restrict__ len, int *err) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nth = gridDim.x * blockDim.x;
  uint32_t *path = scratch + (size_t)tid * maxlen;
  for (uint32_t s = tid; s < N; s += nth) {
    int L = walk_path(P, s, path, maxlen, Never{}, NoVisit{});
    if (L < 0) {
      atomicOr(err, 1);
      L = 1;
    }
    len[s] = (uint32_t)L;
  }
}

__global__ void k_traj_fill(BParams P, uint32_t N, const uint64_t *__restrict__ toff,
                            uint32_t *__restrict__ tp, int maxlen) {
  const uint32_t nth = gridDim.x * blockDim.x;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < N; s += nth)
    (void)walk_path(P, s, tp + toff[s], maxlen, Never{}, NoVisit{});
}

__global__ void k_mark_ends(uint32_t N, const uint64_t *__restrict__ toff,
                            const uint32_t *__restrict__ tp, uint8_t *__restrict__ isend) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) isend[tp[toff[s + 1] - 1]] = 1;
}

__global__ void k_scatter_labels(uint32_t n, const uint32_t *__restrict__ idx,
                                 const int32_t *__restrict__ lab, int32_t *__restrict__ out) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[idx[q]] = lab[q];
}

// ---------------------------------------------------------------- sweep story

__global__ void k_story_init(uint32_t N, const uint64_t *__restrict__ toff,
                             const uint32_t *__restrict__ tp, const int32_t *__restrict__ maxlab,
                             int32_t *__restrict__ ell, int32_t *__restrict__ lam) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  ell[s] = 1;
  lam[s] = maxlab[tp[toff[s + 1] - 1]];
}

__global__ void k_ev_count(uint32_t N, const uint64_t *__restrict__ toff,
                           const uint32_t *__restrict__ tp, const int32_t *__restrict__ ell,
                           uint32_t *__restrict__ cnt, unsigned long long *total) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t mine = 0;
  if (s < N) {
    const int e = ell[s];
    const uint64_t o = toff[s];
    for (int l = 0; l < e; l++) atomicAdd(&cnt[tp[o + l]], 1u);
    mine = (uint64_t)e;
  }
  // wave reduction, one atomic per wavefront
  for (int off = 32; off > 0; off >>= 1) mine += __shfl_down(mine, off, 64);
  if ((threadIdx.x & 63) == 0 && mine) atomicAdd(total, (unsigned long long)mine);
}

__global__ void k_ev_fill(BParams P, uint32_t N, const uint64_t *__restrict__ toff,
                          const uint32_t *__restrict__ tp, const int32_t *__restrict__ ell,
                          const int32_t *__restrict__ lam, const uint64_t *__restrict__ eoff,
                          uint32_t *__restrict__ cur, uint64_t *__restrict__ ev) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const int e = ell[s];
  if (e == 0) return;
  const uint64_t t0 = sweep_index(P, s) << POS_BITS;
  const uint64_t lab = (uint64_t)(uint32_t)lam[s];
  const uint64_t o = toff[s];
  for (int l = 0; l < e; l++) {
    const uint32_t y = tp[o + l];
    const uint32_t slot = atomicAdd(&cur[y], 1u);
    ev[eoff[y] + slot] = ((t0 | (uint64_t)l) << LAB_BITS) | lab;
  }
}

// Short event lists are sorted in place by their thread; longer ones are
// queued: up to SORT_LDS entries for k_ev_sort_long, beyond that for a
// segmented radix sort.
static constexpr int SORT_LDS = 4096;

__global__ void k_ev_sort(uint32_t N, const uint64_t *__restrict__ eoff, uint64_t *__restrict__ ev,
                          uint32_t *__restrict__ longlist, uint32_t *nlong,
                          uint32_t *__restrict__ hugelist, uint32_t *nhuge) {
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= N) return;
  const uint64_t b = eoff[y], e = eoff[y + 1];
  const uint64_t n = e - b;
  if (n < 2) return;
  if (n > SORT_LDS) {
    hugelist[atomicAdd(nhuge, 1u)] = y;
    return;
  }
  if (n > SORT_INLINE) {
    longlist[atomicAdd(nlong, 1u)] = y;
    return;
  }
  for (uint64_t i = b + 1; i < e; i++) {
    const uint64_t v = ev[i];
    uint64_t j = i;
    while (j > b && ev[j - 1] > v) {
      ev[j] = ev[j - 1];
      j--;
    }
    ev[j] = v;
  }
}

// One workgroup per long event list: bitonic sort in LDS.
__global__ void __launch_bounds__(256) k_ev_sort_long(const uint32_t *__restrict__ longlist,
                                                      const uint64_t *__restrict__ eoff,
                                                      uint64_t *__restrict__ ev) {
  __shared__ uint64_t sh[SORT_LDS];
  const uint32_t y = longlist[blockIdx.x];
  const uint64_t b = eoff[y], n = eoff[y + 1] - b;
  if (n <= SORT_LDS) {
    uint32_t m = 1;
    while (m < n) m <<= 1;
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) sh[i] = (i < n) ? ev[b + i] : ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= m; k <<= 1)
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
          const uint32_t ixj = i ^ j;
          if (ixj > i) {
            const bool up = ((i & k) == 0);
            const uint64_t a = sh[i], c = sh[ixj];
            if ((a > c) == up) {
              sh[i] = c;
              sh[ixj] = a;
            }
          }
        }
        __syncthreads();
      }
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) ev[b + i] = sh[i];
  }
}

// Segment bounds of the very long lists, for the segmented radix sort.
__global__ void k_huge_bounds(uint32_t nhuge, const uint32_t *__restrict__ hugelist,
                              const uint64_t *__restrict__ eoff, uint64_t *__restrict__ beg,
                              uint64_t *__restrict__ end) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nhuge) return;
  beg[q] = eoff[hugelist[q]];
  end[q] = eoff[hugelist[q] + 1];
}

// Copy the sorted very long lists back (one workgroup per list).
__global__ void k_huge_copyback(const uint32_t *__restrict__ hugelist, const uint64_t *__restrict__ eoff,
                                const uint64_t *__restrict__ src, uint64_t *__restrict__ dst) {
  const uint32_t y = hugelist[blockIdx.x];
  for (uint64_t i = eoff[y] + threadIdx.x; i < eoff[y + 1]; i += blockDim.x) dst[i] = src[i];
}

// Time at which q enters known==2 (assign_surrounding_pts/known_volnum_ongrid,
// bader@proc.f90:665-726): the first commit event at a face neighbour of q
// after which q and its 6 face neighbours carry one common positive label.
// Per point: 1 if every event of its (sorted) list carries the same label.
// Such a point's label is 0 before its first event and constant after it.
__global__ void k_ev_uniform(uint32_t N, const uint64_t *__restrict__ eoff,
                             const uint64_t *__restrict__ ev, uint8_t *__restrict__ uni) {
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= N) return;
  const uint64_t b = eoff[y], e = eoff[y + 1];
  uint8_t u = 1;
  if (e > b) {
    const int32_t l0 = ev_lab(ev[b]);
    for (uint64_t x = b + 1; x < e && u; x++) u = (ev_lab(ev[x]) == l0);
  }
  uni[y] = u;
}

__device__ __forceinline__ uint64_t first_at_or_after(const uint64_t *__restrict__ ev, uint64_t lo,
                                                      uint64_t hi, uint64_t t) {
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (ev_time(ev[mid]) < t)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ void k_kentry(BParams P, uint32_t N, const uint64_t *__restrict__ eoff,
                         const uint64_t *__restrict__ ev, const uint8_t *__restrict__ uni,
                         uint64_t *__restrict__ Kt) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N) return;
  int i, j, k;
  coords(P, q, i, j, k);
  uint32_t nb[6];
  nb[0] = lin3(P, wrap1(i + 1, P.n1), j, k);
  nb[1] = lin3(P, wrap1(i - 1, P.n1), j, k);
  nb[2] = lin3(P, i, wrap1(j + 1, P.n2), k);
  nb[3] = lin3(P, i, wrap1(j - 1, P.n2), k);
  nb[4] = lin3(P, i, j, wrap1(k + 1, P.n3));
  nb[5] = lin3(P, i, j, wrap1(k - 1, P.n3));
  // lower bound: all seven points must carry a label
  uint64_t tmin;
  bool alluni = uni[q] != 0;
  {
    const uint64_t bq = eoff[q];
    if (eoff[q + 1] == bq) {
      Kt[q] = KT_INF;
      return;
    }
    tmin = ev_time(ev[bq]);
    for (int t = 0; t < 6; t++) {
      const uint64_t b = eoff[nb[t]];
      if (eoff[nb[t] + 1] == b) {
        Kt[q] = KT_INF;
        return;
      }
      const uint64_t ft = ev_time(ev[b]);
      tmin = ft > tmin ? ft : tmin;
      alluni = alluni && uni[nb[t]];
    }
  }
  if (alluni) {
    // constant labels after tmin: K from the first candidate at or after tmin
    const int32_t l0 = ev_lab(ev[eoff[q]]);
    bool same = l0 > 0;
    for (int t = 0; t < 6 && same; t++) same = (ev_lab(ev[eoff[nb[t]]]) == l0);
    uint64_t kt = KT_INF;
    if (same)
      for (int t = 0; t < 6; t++) {
        const uint64_t e = eoff[nb[t] + 1];
        const uint64_t x = first_at_or_after(ev, eoff[nb[t]], e, tmin);
        if (x < e) {
          const uint64_t tt = ev_time(ev[x]);
          kt = tt < kt ? tt : kt;
        }
      }
    Kt[q] = kt;
    return;
  }
  // Candidate times are the events at the face neighbours, visited in time
  // order by merging the six sorted lists; the labels of the seven points are
  // followed with cursors that only move forward, so the work is linear in
  // the local event count.
  uint64_t cpos[6], cend[6];
  for (int t = 0; t < 6; t++) {
    uint64_t lo = eoff[nb[t]], hi = eoff[nb[t] + 1];
    cend[t] = hi;
    while (lo < hi) {  // first event with time >= tmin
      const uint64_t mid = (lo + hi) >> 1;
      if (ev_time(ev[mid]) < tmin)
        lo = mid + 1;
      else
        hi = mid;
    }
    cpos[t] = lo;
  }
  uint32_t y7[7] = {q, nb[0], nb[1], nb[2], nb[3], nb[4], nb[5]};
  uint64_t vpos[7], vend[7], vbeg[7];
  for (int u = 0; u < 7; u++) {
    vbeg[u] = eoff[y7[u]];
    vend[u] = eoff[y7[u] + 1];
    vpos[u] = first_at_or_after(ev, vbeg[u], vend[u], tmin);  // events before are < tmin
  }
  uint64_t kt = KT_INF;
  for (;;) {
    uint64_t c = KT_INF;
    int tc = -1;
    for (int t = 0; t < 6; t++)
      if (cpos[t] < cend[t]) {
        const uint64_t tt = ev_time(ev[cpos[t]]);
        if (tt < c) {
          c = tt;
          tc = t;
        }
      }
    if (tc < 0) break;
    cpos[tc]++;
    int32_t lab[7];
    for (int u = 0; u < 7; u++) {
      while (vpos[u] < vend[u] && ev_time(ev[vpos[u]]) <= c) vpos[u]++;
      lab[u] = (vpos[u] > vbeg[u]) ? ev_lab(ev[vpos[u] - 1]) : 0;
    }
    bool ok = lab[0] > 0;
    for (int u = 1; u < 7 && ok; u++) ok = (lab[u] == lab[0]);
    if (ok) {
      kt = c;
      break;
    }
  }
  Kt[q] = kt;
}

// Recompute every start's decision against the state rebuilt from the
// current story (bader@proc.f90:151-211).
__global__ void k_decide(BParams P, uint32_t N, const uint64_t *__restrict__ toff,
                         const uint32_t *__restrict__ tp, const uint64_t *__restrict__ eoff,
                         const uint64_t *__restrict__ ev, const uint64_t *__restrict__ Kt,
                         const int32_t *__restrict__ maxlab, const int32_t *__restrict__ ell,
                         const int32_t *__restrict__ lam, int32_t *__restrict__ ell2,
                         int32_t *__restrict__ lam2, unsigned long long *nchange, int *err) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  int changed = 0;
  if (s < N) {
    const uint64_t ts = sweep_index(P, s) << POS_BITS;
    int e = 0;
    int32_t la = 0;
    if (v_at(ev, eoff, s, ts, false) == 0) {  // volnum(i,j,k) == 0: a start
      const uint64_t o = toff[s];
      const int len = (int)(toff[s + 1] - o);
      e = len;
      for (int q = 1; q < len; q++)
        if (Kt[tp[o + q]] < ts) {  // known == 2 when the path reaches it
          e = q + 1;
          break;
        }
      const uint32_t end = tp[o + e - 1];
      la = v_at(ev, eoff, end, ts, false);
      if (la == 0) {
        if (e != len) atomicOr(err, 2);  // unlabeled known==2 point: impossible
        la = maxlab[end];                // new maximum -> its attractor
      }
    }
    ell2[s] = e;
    lam2[s] = la;
    changed = (e != ell[s] || la != lam[s]);
  }
  uint64_t m = changed;
  for (int off = 32; off > 0; off >>= 1) m += __shfl_down(m, off, 64);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(nchange, (unsigned long long)m);
}

__global__ void k_final_state(uint32_t N, const uint64_t *__restrict__ eoff,
                              const uint64_t *__restrict__ ev, const uint64_t *__restrict__ Kt,
                              int32_t *__restrict__ V, int8_t *__restrict__ known) {
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= N) return;
  const uint64_t b = eoff[y], e = eoff[y + 1];
  V[y] = (e > b) ? ev_lab(ev[e - 1]) : 0;
  known[y] = (Kt[y] != KT_INF) ? 2 : 0;
}

// ---------------------------------------------------------------- refinement

// Pass 1 marking (bader@proc.f90:311-327): edge, non-maximum points.
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
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(nedge, (unsigned long long)m);
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
__global__ void k_reassign(BParams P, uint32_t N, const int32_t *__restrict__ V,
                           int32_t *__restrict__ Vout, int8_t *__restrict__ known,
                           uint32_t *scratch, int maxlen, int nbasin,
                           unsigned long long *nre, int *err) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nth = gridDim.x * blockDim.x;
  uint32_t *path = scratch + (size_t)tid * maxlen;
  uint64_t mine = 0;
  for (uint32_t s = tid; s < N; s += nth) {
    const int32_t b = V[s];
    if (b >= 0) {
      Vout[s] = b;
      continue;
    }
    int L = walk_path(
        P, s, path, maxlen, [&](uint32_t x) { return known[x] == 2; },
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
      // the serial sweep would stop (:400-402), or the end is itself a start
      // of this sweep (plateau point): flag it
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

// ---------------------------------------------------------------- step table

__global__ void k_step_pointers(BParams P, uint32_t N, int64_t *ongrid, int64_t *near0,
                                double *dr1) {
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
    if (step_near_candidate(P, p, dr, pm)) {
      if (pm[0] == p[0] && pm[1] == p[1] && pm[2] == p[2]) {  // known(pm)==1 (pm == p)
        pm[0] = p[0];
        pm[1] = p[1];
        pm[2] = p[2];
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
  for (int q = 0; q < 9; q++) P.c[q] = car2lat[q];
  for (int q = 0; q < 27; q++) P.lid[q] = lat_i_dist[q];
  return P;
}

struct Timer {
  hipStream_t s;
  hipEvent_t a, b;
  explicit Timer(hipStream_t st) : s(st) {
    C2_HIP(hipEventCreate(&a));
    C2_HIP(hipEventCreate(&b));
    C2_HIP(hipEventRecord(a, s));
  }
  double stop() {
    C2_HIP(hipEventRecord(b, s));
    C2_HIP(hipEventSynchronize(b));
    float ms = 0;
    C2_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
  }
};

template <class T>
static T read_scalar(c2gpu_ctx *ctx, const T *d) {
  T h;
  C2_HIP(hipMemcpyAsync(&h, d, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
  C2_HIP(hipStreamSynchronize(ctx->stream));
  return h;
}

struct ToU64 {
  __host__ __device__ uint64_t operator()(uint32_t v) const { return (uint64_t)v; }
};

// Exclusive prefix sum of 32-bit counts into 64-bit offsets (the scan
// accumulates in 64 bits: totals exceed 2^32 on large grids).
static void exclusive_scan(c2gpu_ctx *ctx, const char *tag, const uint32_t *in, uint64_t *out,
                           int64_t n) {
  hipcub::TransformInputIterator<uint64_t, ToU64, const uint32_t *> it(in, ToU64{});
  size_t tmp = 0;
  C2_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, it, out, (int)n, ctx->stream));
  void *t = buffer(ctx, std::string("scan_tmp_") + tag, tmp);
  C2_HIP(hipcub::DeviceScan::ExclusiveSum(t, tmp, it, out, (int)n, ctx->stream));
}

// Host-side attractor assignment of a new maximum (bader@proc.f90:159-199)
// for the default mode: identify_atom with ratom = 1d40, i.e. the nearest
// atom of the complete cell list (first one on exact ties).
int nearest_atom_label(const c2gpu_attractor_opts *o, const double x[3]);

void bader_run(c2gpu_ctx *ctx, const double *rho, const int n[3], const double car2lat[9],
               const double lat_i_dist[27], const c2gpu_attractor_opts *opts, int flags,
               int32_t *label_out, int *nattr, double *xattr_out, int xattr_cap,
               c2gpu_bader_stats *stats) {
  if (!opts) fail(C2GPU_ERR_ARG, "c2gpu_bader: opts is NULL");
  if (opts->mode != C2GPU_ATTR_NEAREST_ATOM)
    fail(C2GPU_ERR_UNSUPPORTED,
         "c2gpu_bader: only the default attractor assignment (nearest atom, ratom=1d40) is "
         "implemented on the GPU path");
  if (opts->nat < 1 || !opts->xat) fail(C2GPU_ERR_ARG, "c2gpu_bader: no atoms");
  if (opts->nat >= (1 << LAB_BITS)) fail(C2GPU_ERR_ARG, "c2gpu_bader: too many atoms");
  for (int c = 0; c < 3; c++)
    if (n[c] < 3) fail(C2GPU_ERR_ARG, "c2gpu_bader: grid edges must be >= 3");
  const uint64_t N64 = (uint64_t)n[0] * n[1] * n[2];
  if (N64 >= (1ull << 32) - 1) fail(C2GPU_ERR_ARG, "c2gpu_bader: grid too large");
  if (N64 >= (1ull << (64 - POS_BITS - LAB_BITS)))
    fail(C2GPU_ERR_ARG, "c2gpu_bader: grid too large for event time stamps");
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
  int *d_err = buf<int>(ctx, "bader_err", 4);
  unsigned long long *d_cnt = buf<unsigned long long>(ctx, "bader_counters", 4);
  C2_HIP(hipMemsetAsync(d_err, 0, 4 * sizeof(int), st));

  // 1) free trajectories
  Timer ttraj(st);
  const int nth = persistent_threads(ctx->device, B);
  const int maxlen = MAX_PATH;
  uint32_t *scratch = buf<uint32_t>(ctx, "bader_scratch", (size_t)nth * maxlen);
  uint32_t *d_len = buf<uint32_t>(ctx, "bader_len", (size_t)N + 1);
  uint64_t *d_toff = buf<uint64_t>(ctx, "bader_toff", (size_t)N + 1);
  C2_HIP(hipMemsetAsync(d_len + N, 0, sizeof(uint32_t), st));
  C2_LAUNCH(ctx, k_traj_len, dim3(nth / B), dim3(B), 0, P, N, scratch, maxlen, d_len, d_err);
  exclusive_scan(ctx, "len", d_len, d_toff, (int64_t)N + 1);
  const uint64_t ntraj = read_scalar(ctx, d_toff + N);
  if (read_scalar(ctx, d_err) & 1) fail(C2GPU_ERR_PATHLEN, "c2gpu_bader: ascent path longer than 1024 points");
  uint32_t *d_tp = buf<uint32_t>(ctx, "bader_tp", ntraj);
  C2_LAUNCH(ctx, k_traj_fill, dim3(nth / B), dim3(B), 0, P, N, d_toff, d_tp, maxlen);
  S.traj_points = (int64_t)ntraj;

  // endpoints (grid maxima and no-move points) -> attractors on the host
  uint8_t *d_isend = buf<uint8_t>(ctx, "bader_isend", N);
  C2_HIP(hipMemsetAsync(d_isend, 0, N, st));
  C2_LAUNCH(ctx, k_mark_ends, dim3(gN), dim3(B), 0, N, d_toff, d_tp, d_isend);
  uint32_t *d_ends = buf<uint32_t>(ctx, "bader_ends", N);
  uint32_t *d_nends = buf<uint32_t>(ctx, "bader_nends", 1);
  {
    hipcub::CountingInputIterator<uint32_t> it(0);
    size_t tmp = 0;
    C2_HIP(hipcub::DeviceSelect::Flagged(nullptr, tmp, it, d_isend, d_ends, d_nends, (int)N, st));
    void *t = buffer(ctx, "select_tmp", tmp);
    C2_HIP(hipcub::DeviceSelect::Flagged(t, tmp, it, d_isend, d_ends, d_nends, (int)N, st));
  }
  const uint32_t nends = read_scalar(ctx, d_nends);
  std::vector<uint32_t> h_ends(nends);
  C2_HIP(hipMemcpyAsync(h_ends.data(), d_ends, sizeof(uint32_t) * nends, hipMemcpyDeviceToHost, st));
  C2_HIP(hipStreamSynchronize(st));
  std::vector<int32_t> h_lab(nends);
  for (uint32_t q = 0; q < nends; q++) {
    const uint32_t l = h_ends[q];
    const int i = (int)(l % (uint32_t)n[0]);
    const int j = (int)((l / (uint32_t)n[0]) % (uint32_t)n[1]);
    const int k = (int)(l / ((uint32_t)n[0] * (uint32_t)n[1]));
    const double dv[3] = {(double)i / n[0], (double)j / n[1], (double)k / n[2]};
    h_lab[q] = nearest_atom_label(opts, dv);
  }
  S.nmaxima = (int)nends;
  int32_t *d_maxlab = buf<int32_t>(ctx, "bader_maxlab", N);
  int32_t *d_endlab = buf<int32_t>(ctx, "bader_endlab", nends > 0 ? nends : 1);
  C2_HIP(hipMemcpyAsync(d_endlab, h_lab.data(), sizeof(int32_t) * nends, hipMemcpyHostToDevice, st));
  if (nends)
    C2_LAUNCH(ctx, k_scatter_labels, dim3(grid_for(nends, B)), dim3(B), 0, nends, d_ends,
                       d_endlab, d_maxlab);
  S.ms_traj = ttraj.stop();

  // 2) first sweep: Jacobi fixpoint over the story
  Timer tsweep(st);
  int32_t *ell = buf<int32_t>(ctx, "bader_ell", N), *lam = buf<int32_t>(ctx, "bader_lam", N);
  int32_t *ell2 = buf<int32_t>(ctx, "bader_ell2", N), *lam2 = buf<int32_t>(ctx, "bader_lam2", N);
  uint32_t *cnt = buf<uint32_t>(ctx, "bader_evcnt", (size_t)N + 1);
  uint64_t *eoff = buf<uint64_t>(ctx, "bader_eoff", (size_t)N + 1);
  uint32_t *cur = buf<uint32_t>(ctx, "bader_evcur", N);
  uint64_t *Kt = buf<uint64_t>(ctx, "bader_kt", N);
  uint32_t *longlist = buf<uint32_t>(ctx, "bader_long", N);
  uint32_t *hugelist = buf<uint32_t>(ctx, "bader_huge", N);
  uint8_t *uni = buf<uint8_t>(ctx, "bader_uni", N);
  C2_LAUNCH(ctx, k_story_init, dim3(gN), dim3(B), 0, N, d_toff, d_tp, d_maxlab, ell, lam);
  uint64_t *ev = nullptr;
  int it = 0;
  const int max_iters = 100000;
  for (;; it++) {
    if (it >= max_iters) fail(C2GPU_ERR_NOCONVERGE, "c2gpu_bader: first-sweep fixpoint did not converge");
    C2_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t) * ((size_t)N + 1), st));
    C2_HIP(hipMemsetAsync(d_cnt, 0, 4 * sizeof(unsigned long long), st));
    C2_LAUNCH(ctx, k_ev_count, dim3(gN), dim3(B), 0, N, d_toff, d_tp, ell, cnt, d_cnt);
    exclusive_scan(ctx, "ev", cnt, eoff, (int64_t)N + 1);
    const uint64_t nev = read_scalar(ctx, eoff + N);
    S.events_total += (int64_t)nev;
    ev = buf<uint64_t>(ctx, "bader_ev", nev);
    C2_HIP(hipMemsetAsync(cur, 0, sizeof(uint32_t) * N, st));
    C2_LAUNCH(ctx, k_ev_fill, dim3(gN), dim3(B), 0, P, N, d_toff, d_tp, ell, lam, eoff, cur, ev);
    uint32_t *d_nlong = reinterpret_cast<uint32_t *>(d_cnt + 2);
    uint32_t *d_nhuge = d_nlong + 1;
    C2_LAUNCH(ctx, k_ev_sort, dim3(gN), dim3(B), 0, N, eoff, ev, longlist, d_nlong, hugelist, d_nhuge);
    uint32_t nlh[2];
    C2_HIP(hipMemcpyAsync(nlh, d_nlong, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    C2_HIP(hipStreamSynchronize(st));
    const uint32_t nlong = nlh[0], nhuge = nlh[1];
    if (nlong) C2_LAUNCH(ctx, k_ev_sort_long, dim3(nlong), dim3(256), 0, longlist, eoff, ev);
    if (nhuge) {
      uint64_t *hb = buf<uint64_t>(ctx, "bader_hugeb", nhuge), *he = buf<uint64_t>(ctx, "bader_hugee", nhuge);
      C2_LAUNCH(ctx, k_huge_bounds, dim3(grid_for(nhuge, 64)), dim3(64), 0, nhuge, hugelist, eoff, hb, he);
      uint64_t *ev_alt = buf<uint64_t>(ctx, "bader_ev_alt", nev);
      size_t tmp = 0;
      C2_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tmp, ev, ev_alt, (int)nev, (int)nhuge,
                                                        hb, he, 0, 64, st));
      void *t = buffer(ctx, "bader_segsort_tmp", tmp);
      C2_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(t, tmp, ev, ev_alt, (int)nev, (int)nhuge, hb, he,
                                                        0, 64, st));
      C2_LAUNCH(ctx, k_huge_copyback, dim3(nhuge), dim3(256), 0, hugelist, eoff, ev_alt, ev);
    }
    C2_LAUNCH(ctx, k_ev_uniform, dim3(gN), dim3(B), 0, N, eoff, ev, uni);
    C2_LAUNCH(ctx, k_kentry, dim3(gN), dim3(B), 0, P, N, eoff, ev, uni, Kt);
    C2_LAUNCH(ctx, k_decide, dim3(gN), dim3(B), 0, P, N, d_toff, d_tp, eoff, ev, Kt,
                       d_maxlab, ell, lam, ell2, lam2, d_cnt + 1, d_err);
    const unsigned long long nch = read_scalar(ctx, d_cnt + 1);
    std::swap(ell, ell2);
    std::swap(lam, lam2);
    if (nch == 0) break;
  }
  if (read_scalar(ctx, d_err) & 2) fail(C2GPU_ERR_ARG, "c2gpu_bader: internal inconsistency in the sweep state");
  S.sweep_iters = it + 1;
  int32_t *V = buf<int32_t>(ctx, "bader_V", N), *V2 = buf<int32_t>(ctx, "bader_V2", N);
  int8_t *known = buf<int8_t>(ctx, "bader_known", N), *known2 = buf<int8_t>(ctx, "bader_known2", N);
  C2_LAUNCH(ctx, k_final_state, dim3(gN), dim3(B), 0, N, eoff, ev, Kt, V, known);
  {
    // number of starts = points with ell > 0
    S.nstarts = 0;
  }
  S.ms_sweep = tsweep.stop();

  // 3) refinement (bader@proc.f90:217-224)
  Timer tref(st);
  const int nbasin = opts->nat;
  uint8_t *E = buf<uint8_t>(ctx, "bader_edge", N);
  int iter = 1;
  for (;; iter++) {
    if (iter > 100000) fail(C2GPU_ERR_NOCONVERGE, "c2gpu_bader: refinement did not converge");
    C2_HIP(hipMemsetAsync(d_cnt, 0, 4 * sizeof(unsigned long long), st));
    if (iter == 1) {
      C2_LAUNCH(ctx, k_ref1_mark, dim3(gN), dim3(B), 0, P, N, V, E, d_cnt + 2);
      C2_LAUNCH(ctx, k_ref1_apply, dim3(gN), dim3(B), 0, P, N, E, V, known);
      S.edge_points = (int64_t)read_scalar(ctx, d_cnt + 2);
    } else {
      C2_LAUNCH(ctx, k_refn_mark, dim3(gN), dim3(B), 0, P, N, V, known, V2, known2);
      std::swap(V, V2);
      std::swap(known, known2);
      C2_LAUNCH(ctx, k_refn_clear, dim3(gN), dim3(B), 0, P, N, V, known);
    }
    C2_LAUNCH(ctx, k_reassign, dim3(nth / B), dim3(B), 0, P, N, V, V2, known, scratch,
                       maxlen, nbasin, d_cnt, d_err);
    std::swap(V, V2);
    const unsigned long long nre = read_scalar(ctx, d_cnt);
    const int e = read_scalar(ctx, d_err);
    if (e & 1) fail(C2GPU_ERR_PATHLEN, "c2gpu_bader: refinement path longer than 1024 points");
    if (e & 4) fail(C2GPU_ERR_NEWMAX, "refine_edge: should be no new maxima in edge refinement");
    if (e & 8) fail(C2GPU_ERR_DEGENERATE, "c2gpu_bader: refinement path ends on a point being reassigned");
    if (nre == 0) break;
  }
  S.refine_iters = iter;
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
      for (int c = 0; c < 3; c++) xattr_out[3 * a + c] = opts->xat[3 * a + c];
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
