// First sweep of bader_integrate (bader@proc.f90:151-215) as a parallel
// fixpoint, exact to the serial loop.
//
// The serial sweep visits points in the order do i; do j; do k.  A point whose
// volnum is still 0 starts a path (max_neargrid); the path is the point's free
// near-grid trajectory truncated at the first point already marked
// known == 2; every point of the path receives the path's label (the label of
// the end point, or the attractor of a new maximum), and assign_surrounding_pts
// marks known == 2 the face neighbours whose six face neighbours now share
// their label.  Which points start paths, where paths stop and which labels
// they carry depend on the visiting order.
//
// Here that history is a "story": for every point s, ell[s] (0 if s is not a
// start, else the length of its committed path) and lam[s] (its label).
// Every commit is an event stamped time = (sweep index << 10 | position on
// the path); per point, the events form a time-sorted list from which the
// state of the serial sweep at any moment is read back: volnum(y) at time t is
// the label of y's last event before t, and the time y enters known == 2 is
// the first commit at a face neighbour of y after which y and its six face
// neighbours carry one common positive label.  Recomputing every decision
// against the state its story implies is one Jacobi step.  The serial
// recurrence has exactly one solution and every fixpoint satisfies it, so the
// converged story is the serial sweep.
//
// After the first (full) step, steps are incremental: only the event lists of
// points covered by a changed path are rebuilt, only the known==2 times of
// their neighbourhoods recomputed, and only starts whose inputs changed are
// decided again.  A start's inputs are the events at the points of its
// committed path (its own start, the points where it could be truncated, its
// end), so the starts to redo are found from those points' event lists.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <vector>

#include "bader_dev.hpp"
#include "common.hpp"
#include "devutil.hpp"

namespace c2 {

static constexpr uint64_t KT_INF = ~0ull;
static constexpr int POS_BITS = 10;     // position on a path: < 1024
static constexpr int LAB_BITS = 20;     // label: < 2^20
static constexpr int SORT_INLINE = 48;  // longer event lists are sorted by workgroups
static constexpr int SORT_LDS = 4096;   // longer still: segmented radix sort

__device__ __forceinline__ uint64_t ev_time(uint64_t e) { return e >> LAB_BITS; }
__device__ __forceinline__ int32_t ev_lab(uint64_t e) { return (int32_t)(e & ((1u << LAB_BITS) - 1)); }
__device__ __forceinline__ uint64_t make_event(uint64_t t0, int l, int32_t lab) {
  return ((t0 | (uint64_t)l) << LAB_BITS) | (uint64_t)(uint32_t)lab;
}
// the start point (linear index) whose path produced an event
__device__ __forceinline__ uint32_t event_start(const BParams &P, uint64_t e) {
  const uint64_t sg = ev_time(e) >> POS_BITS;
  const uint32_t k = (uint32_t)(sg % (uint64_t)P.n3);
  const uint64_t r = sg / (uint64_t)P.n3;
  const uint32_t j = (uint32_t)(r % (uint64_t)P.n2);
  const uint32_t i = (uint32_t)(r / (uint64_t)P.n2);
  return lin3(P, (int)i, (int)j, (int)k);
}

// Label of y just after the events with time <= tau (inclusive) or < tau.
__device__ __forceinline__ int32_t v_at(const uint64_t *__restrict__ ev, const uint64_t *__restrict__ eoff,
                                        uint32_t y, uint64_t tau, bool inclusive) {
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

__device__ __forceinline__ void wave_add(unsigned long long *dst, uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

// ---------------------------------------------------------------- event lists

__global__ void k_story_init(uint32_t N, const uint64_t *__restrict__ toff, const uint32_t *__restrict__ tp,
                             const int32_t *__restrict__ maxlab, int32_t *__restrict__ ell,
                             int32_t *__restrict__ lam) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  ell[s] = 1;
  lam[s] = maxlab[tp[toff[s + 1] - 1]];
}

__global__ void k_ev_count(uint32_t N, const uint64_t *__restrict__ toff, const uint32_t *__restrict__ tp,
                           const int32_t *__restrict__ ell, uint32_t *__restrict__ cnt) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const int e = ell[s];
  const uint64_t o = toff[s];
  for (int l = 0; l < e; l++) atomicAdd(&cnt[tp[o + l]], 1u);
}

// Writes the events of the starts in `list` (all starts when list == null).
__global__ void k_ev_fill(BParams P, uint32_t n, const uint32_t *__restrict__ list,
                          const uint64_t *__restrict__ toff, const uint32_t *__restrict__ tp,
                          const int32_t *__restrict__ ell, const int32_t *__restrict__ lam,
                          const uint64_t *__restrict__ eoff, uint32_t *__restrict__ cur,
                          uint64_t *__restrict__ ev, int *err) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t s = list ? list[t] : t;
  const int e = ell[s];
  if (e == 0) return;
  const uint64_t t0 = sweep_index(P, s) << POS_BITS;
  const int32_t lab = lam[s];
  const uint64_t o = toff[s];
  for (int l = 0; l < e; l++) {
    const uint32_t y = tp[o + l];
    const uint32_t slot = atomicAdd(&cur[y], 1u);
    if (eoff[y] + slot < eoff[y + 1])
      ev[eoff[y] + slot] = make_event(t0, l, lab);
    else
      atomicOr(err, 4);  // count/fill mismatch (never expected): refuse to write
  }
}

__device__ __forceinline__ bool list_uniform(const uint64_t *__restrict__ ev, uint64_t b, uint64_t e) {
  if (e <= b) return true;
  const int32_t l0 = ev_lab(ev[b]);
  for (uint64_t x = b + 1; x < e; x++)
    if (ev_lab(ev[x]) != l0) return false;
  return true;
}

// Sorts short lists in place and records whether each list carries a single
// label; longer lists are queued (flag == null: every point, else the points
// with flag set).
__global__ void k_ev_sort(uint32_t N, const uint8_t *__restrict__ flag, const uint64_t *__restrict__ eoff,
                          uint64_t *__restrict__ ev, uint8_t *__restrict__ uni,
                          uint32_t *__restrict__ longlist, uint32_t *nlong, uint32_t *__restrict__ hugelist,
                          uint32_t *nhuge) {
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= N) return;
  if (flag && !flag[y]) return;
  const uint64_t b = eoff[y], e = eoff[y + 1];
  const uint64_t n = e - b;
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
  uni[y] = list_uniform(ev, b, e);
}

// One workgroup per long list: bitonic sort in LDS.
__global__ void __launch_bounds__(256) k_ev_sort_long(const uint32_t *__restrict__ longlist,
                                                      const uint64_t *__restrict__ eoff,
                                                      uint64_t *__restrict__ ev, uint8_t *__restrict__ uni) {
  __shared__ uint64_t sh[SORT_LDS];
  __shared__ int notuni;
  const uint32_t y = longlist[blockIdx.x];
  const uint64_t b = eoff[y], n = eoff[y + 1] - b;
  uint32_t m = 1;
  while (m < n) m <<= 1;
  for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) sh[i] = (i < n) ? ev[b + i] : ~0ull;
  if (threadIdx.x == 0) notuni = 0;
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
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    ev[b + i] = sh[i];
    if (ev_lab(sh[i]) != ev_lab(sh[0])) notuni = 1;
  }
  __syncthreads();
  if (threadIdx.x == 0) uni[y] = notuni ? 0 : 1;
}

__global__ void k_huge_bounds(uint32_t nhuge, const uint32_t *__restrict__ hugelist,
                              const uint64_t *__restrict__ eoff, uint64_t *__restrict__ beg,
                              uint64_t *__restrict__ end) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nhuge) return;
  beg[q] = eoff[hugelist[q]];
  end[q] = eoff[hugelist[q] + 1];
}

// Copies a sorted very long list back and records its single-label flag.
__global__ void k_huge_copyback(const uint32_t *__restrict__ hugelist, const uint64_t *__restrict__ eoff,
                                const uint64_t *__restrict__ src, uint64_t *__restrict__ dst,
                                uint8_t *__restrict__ uni) {
  __shared__ int notuni;
  const uint32_t y = hugelist[blockIdx.x];
  const uint64_t b = eoff[y], e = eoff[y + 1];
  if (threadIdx.x == 0) notuni = 0;
  __syncthreads();
  for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
    dst[i] = src[i];
    if (ev_lab(src[i]) != ev_lab(src[b])) notuni = 1;
  }
  __syncthreads();
  if (threadIdx.x == 0) uni[y] = notuni ? 0 : 1;
}

// ---------------------------------------------------------------- known == 2

// Time at which q enters known == 2 (assign_surrounding_pts and
// known_volnum_ongrid, bader@proc.f90:665-726).
__device__ uint64_t kentry_point(const BParams &P, uint32_t q, const uint64_t *__restrict__ eoff,
                                 const uint64_t *__restrict__ ev, const uint8_t *__restrict__ uni) {
  int i, j, k;
  coords(P, q, i, j, k);
  uint32_t nb[6];
  nb[0] = lin3(P, wrap1(i + 1, P.n1), j, k);
  nb[1] = lin3(P, wrap1(i - 1, P.n1), j, k);
  nb[2] = lin3(P, i, wrap1(j + 1, P.n2), k);
  nb[3] = lin3(P, i, wrap1(j - 1, P.n2), k);
  nb[4] = lin3(P, i, j, wrap1(k + 1, P.n3));
  nb[5] = lin3(P, i, j, wrap1(k - 1, P.n3));
  // all seven points must carry a label
  if (eoff[q + 1] == eoff[q]) return KT_INF;
  uint64_t tmin = ev_time(ev[eoff[q]]);
  bool alluni = uni[q] != 0;
  for (int t = 0; t < 6; t++) {
    const uint64_t b = eoff[nb[t]];
    if (eoff[nb[t] + 1] == b) return KT_INF;
    const uint64_t ft = ev_time(ev[b]);
    tmin = ft > tmin ? ft : tmin;
    alluni = alluni && uni[nb[t]];
  }
  if (alluni) {
    // labels are constant after tmin: the first candidate at or after it
    const int32_t l0 = ev_lab(ev[eoff[q]]);
    if (l0 <= 0) return KT_INF;
    for (int t = 0; t < 6; t++)
      if (ev_lab(ev[eoff[nb[t]]]) != l0) return KT_INF;
    uint64_t kt = KT_INF;
    for (int t = 0; t < 6; t++) {
      const uint64_t e = eoff[nb[t] + 1];
      const uint64_t x = first_at_or_after(ev, eoff[nb[t]], e, tmin);
      if (x < e) {
        const uint64_t tt = ev_time(ev[x]);
        kt = tt < kt ? tt : kt;
      }
    }
    return kt;
  }
  // general case: merge the six neighbour lists in time order; the labels of
  // the seven points are followed with forward-only cursors
  uint64_t cpos[6], cend[6];
  for (int t = 0; t < 6; t++) {
    cend[t] = eoff[nb[t] + 1];
    cpos[t] = first_at_or_after(ev, eoff[nb[t]], cend[t], tmin);
  }
  const uint32_t y7[7] = {q, nb[0], nb[1], nb[2], nb[3], nb[4], nb[5]};
  uint64_t vpos[7], vend[7], vbeg[7];
  for (int u = 0; u < 7; u++) {
    vbeg[u] = eoff[y7[u]];
    vend[u] = eoff[y7[u] + 1];
    vpos[u] = first_at_or_after(ev, vbeg[u], vend[u], tmin);
  }
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
    if (tc < 0) return KT_INF;
    cpos[tc]++;
    int32_t lab[7];
    for (int u = 0; u < 7; u++) {
      while (vpos[u] < vend[u] && ev_time(ev[vpos[u]]) <= c) vpos[u]++;
      lab[u] = (vpos[u] > vbeg[u]) ? ev_lab(ev[vpos[u] - 1]) : 0;
    }
    bool ok = lab[0] > 0;
    for (int u = 1; u < 7 && ok; u++) ok = (lab[u] == lab[0]);
    if (ok) return c;
  }
}

__global__ void k_kentry_all(BParams P, uint32_t N, const uint64_t *__restrict__ eoff,
                             const uint64_t *__restrict__ ev, const uint8_t *__restrict__ uni,
                             uint64_t *__restrict__ Kt) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < N) Kt[q] = kentry_point(P, q, eoff, ev, uni);
}

// Recomputes Kt where the seven-point neighbourhood holds a rebuilt list.
__global__ void k_kentry_dirty(BParams P, uint32_t N, const uint64_t *__restrict__ eoff,
                               const uint64_t *__restrict__ ev, const uint8_t *__restrict__ uni,
                               const uint8_t *__restrict__ evdirty, uint64_t *__restrict__ Kt,
                               uint8_t *__restrict__ ktchg) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N) return;
  int i, j, k;
  coords(P, q, i, j, k);
  const bool dirty = evdirty[q] || evdirty[lin3(P, wrap1(i + 1, P.n1), j, k)] ||
                     evdirty[lin3(P, wrap1(i - 1, P.n1), j, k)] ||
                     evdirty[lin3(P, i, wrap1(j + 1, P.n2), k)] ||
                     evdirty[lin3(P, i, wrap1(j - 1, P.n2), k)] ||
                     evdirty[lin3(P, i, j, wrap1(k + 1, P.n3))] || evdirty[lin3(P, i, j, wrap1(k - 1, P.n3))];
  if (!dirty) return;
  const uint64_t kt = kentry_point(P, q, eoff, ev, uni);
  if (kt != Kt[q]) {
    Kt[q] = kt;
    ktchg[q] = 1;
  }
}

// ---------------------------------------------------------------- decisions

// One start's decision against the state of the current story
// (bader@proc.f90:154-211).
__device__ void decide_start(const BParams &P, uint32_t s, const uint64_t *__restrict__ toff,
                             const uint32_t *__restrict__ tp, const uint64_t *__restrict__ eoff,
                             const uint64_t *__restrict__ ev, const uint64_t *__restrict__ Kt,
                             const int32_t *__restrict__ maxlab, int &e, int32_t &la, int *err) {
  const uint64_t ts = sweep_index(P, s) << POS_BITS;
  e = 0;
  la = 0;
  if (v_at(ev, eoff, s, ts, false) != 0) return;  // volnum /= 0: not a start
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
    if (e != len) atomicOr(err, 2);  // an unlabeled known==2 point cannot exist
    la = maxlab[end];                // new maximum -> its attractor
  }
}

// Decides the starts in `list` (every point when list == null); changed
// decisions are flagged and appended to `chlist`.
__global__ void k_decide(BParams P, uint32_t n, const uint32_t *__restrict__ list,
                         const uint64_t *__restrict__ toff, const uint32_t *__restrict__ tp,
                         const uint64_t *__restrict__ eoff, const uint64_t *__restrict__ ev,
                         const uint64_t *__restrict__ Kt, const int32_t *__restrict__ maxlab,
                         const int32_t *__restrict__ ell, const int32_t *__restrict__ lam,
                         int32_t *__restrict__ ell2, int32_t *__restrict__ lam2, uint8_t *__restrict__ chg,
                         uint32_t *__restrict__ chlist, uint32_t *nch, int *err) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t s = list ? list[t] : t;
  int e;
  int32_t la;
  decide_start(P, s, toff, tp, eoff, ev, Kt, maxlab, e, la, err);
  ell2[s] = e;
  lam2[s] = la;
  if (e != ell[s] || la != lam[s]) {
    chg[s] = 1;
    chlist[atomicAdd(nch, 1u)] = s;
  }
}

// ---------------------------------------------------------------- incremental update

// Count changes per point from the changed starts' old and new paths.
__global__ void k_delta(uint32_t nc, const uint32_t *__restrict__ chlist, const uint64_t *__restrict__ toff,
                        const uint32_t *__restrict__ tp, const int32_t *__restrict__ ell,
                        const int32_t *__restrict__ ell2, int32_t *__restrict__ delta,
                        uint8_t *__restrict__ evdirty) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nc) return;
  const uint32_t s = chlist[t];
  const uint64_t o = toff[s];
  const int eo = ell[s], en = ell2[s];
  const int m = eo > en ? eo : en;
  for (int l = 0; l < m; l++) {
    const uint32_t y = tp[o + l];
    const int d = (l < en ? 1 : 0) - (l < eo ? 1 : 0);
    if (d) atomicAdd(&delta[y], d);
    evdirty[y] = 1;  // a kept point may change label or time stamp only
  }
}

__global__ void k_newcnt(uint32_t N, const uint64_t *__restrict__ eoff, int32_t *__restrict__ delta,
                         uint32_t *__restrict__ cnt) {
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y > N) return;
  if (y == N) {
    cnt[N] = 0;
    return;
  }
  cnt[y] = (uint32_t)((int64_t)(eoff[y + 1] - eoff[y]) + delta[y]);
  delta[y] = 0;
}

// Moves every list to its new place, dropping the events of changed starts.
__global__ void k_ev_keep(BParams P, uint32_t N, const uint64_t *__restrict__ eoff,
                          const uint64_t *__restrict__ ev, const uint64_t *__restrict__ eoff2,
                          uint64_t *__restrict__ ev2, const uint8_t *__restrict__ evdirty,
                          const uint8_t *__restrict__ chg, uint32_t *__restrict__ cur, int *err) {
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= N) return;
  const uint64_t b = eoff[y], e = eoff[y + 1], b2 = eoff2[y], n2 = eoff2[y + 1] - b2;
  if (!evdirty[y]) {
    if (e - b != n2) {
      atomicOr(err, 4);  // an untouched list must keep its size
      return;
    }
    for (uint64_t x = b; x < e; x++) ev2[b2 + (x - b)] = ev[x];
    return;
  }
  uint32_t kept = 0;
  for (uint64_t x = b; x < e; x++) {
    const uint64_t v = ev[x];
    if (!chg[event_start(P, v)]) {
      if (kept < n2)
        ev2[b2 + kept] = v;
      else
        atomicOr(err, 4);
      kept++;
    }
  }
  cur[y] = kept;
}

// Starts to decide again: those with an event at a point whose list or
// known==2 time changed, and those points themselves.
__global__ void k_mark_decide(BParams P, uint32_t N, const uint64_t *__restrict__ eoff,
                              const uint64_t *__restrict__ ev, const uint8_t *__restrict__ evdirty,
                              const uint8_t *__restrict__ ktchg, uint8_t *__restrict__ dflag) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N || !(evdirty[x] || ktchg[x])) return;
  dflag[x] = 1;
  for (uint64_t q = eoff[x]; q < eoff[x + 1]; q++) dflag[event_start(P, ev[q])] = 1;
}

__global__ void k_apply(uint32_t nc, const uint32_t *__restrict__ chlist, int32_t *__restrict__ ell,
                        int32_t *__restrict__ lam, const int32_t *__restrict__ ell2,
                        const int32_t *__restrict__ lam2, uint8_t *__restrict__ chg) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nc) return;
  const uint32_t s = chlist[t];
  ell[s] = ell2[s];
  lam[s] = lam2[s];
  chg[s] = 0;
}

__global__ void k_final_state(uint32_t N, const uint64_t *__restrict__ eoff, const uint64_t *__restrict__ ev,
                              const uint64_t *__restrict__ Kt, int32_t *__restrict__ V,
                              int8_t *__restrict__ known, unsigned long long *nstarts,
                              const int32_t *__restrict__ ell) {
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t st = 0;
  if (y < N) {
    const uint64_t b = eoff[y], e = eoff[y + 1];
    V[y] = (e > b) ? ev_lab(ev[e - 1]) : 0;
    known[y] = (Kt[y] != KT_INF) ? 2 : 0;
    st = ell[y] > 0;
  }
  wave_add(nstarts, st);
}

// ---------------------------------------------------------------- host driver

struct SweepBufs {
  uint64_t *eoff[2];
  uint64_t *ev[2];
  int cur = 0;
};

static void sort_lists(c2gpu_ctx *ctx, uint32_t N, const uint8_t *flag, const uint64_t *eoff, uint64_t *ev,
                       uint64_t nev, uint8_t *uni, uint32_t *longlist, uint32_t *hugelist) {
  hipStream_t st = ctx->stream;
  uint32_t *d_n = buf<uint32_t>(ctx, "sweep_nlist", 2);
  C2_HIP(hipMemsetAsync(d_n, 0, 2 * sizeof(uint32_t), st));
  C2_LAUNCH(ctx, k_ev_sort, dim3(grid_for(N, 256)), dim3(256), 0, N, flag, eoff, ev, uni, longlist, d_n,
            hugelist, d_n + 1);
  uint32_t nlh[2];
  C2_HIP(hipMemcpyAsync(nlh, d_n, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  C2_HIP(hipStreamSynchronize(st));
  if (nlh[0]) C2_LAUNCH(ctx, k_ev_sort_long, dim3(nlh[0]), dim3(256), 0, longlist, eoff, ev, uni);
  if (nlh[1]) {
    const uint32_t nh = nlh[1];
    uint64_t *hb = buf<uint64_t>(ctx, "sweep_hugeb", nh), *he = buf<uint64_t>(ctx, "sweep_hugee", nh);
    C2_LAUNCH(ctx, k_huge_bounds, dim3(grid_for(nh, 64)), dim3(64), 0, nh, hugelist, eoff, hb, he);
    uint64_t *alt = buf<uint64_t>(ctx, "sweep_ev_alt", nev);
    size_t tmp = 0;
    C2_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tmp, ev, alt, (int)nev, (int)nh, hb, he, 0, 64, st));
    void *t = buffer(ctx, "sweep_segsort_tmp", tmp);
    C2_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(t, tmp, ev, alt, (int)nev, (int)nh, hb, he, 0, 64, st));
    C2_LAUNCH(ctx, k_huge_copyback, dim3(nh), dim3(256), 0, hugelist, eoff, alt, ev, uni);
  }
}

struct SweepResult {
  int iters;
  int64_t events_total;
  int64_t nstarts;
};

// Runs the fixpoint; on return V holds bas%idg after the sweep (volnum) and
// known the known==2 marks (2 or 0).
SweepResult sweep_run(c2gpu_ctx *ctx, const BParams &P, uint32_t N, const uint64_t *toff, const uint32_t *tp,
                      const int32_t *maxlab, int32_t *V, int8_t *known) {
  hipStream_t st = ctx->stream;
  const int B = 256;
  const unsigned gN = grid_for(N, B);
  const bool full_every_step = std::getenv("C2GPU_SWEEP_FULL") != nullptr;
  const bool trace = std::getenv("C2GPU_SWEEP_TRACE") != nullptr;
  SweepResult R{0, 0, 0};
  int32_t *ell = buf<int32_t>(ctx, "sweep_ell", N), *lam = buf<int32_t>(ctx, "sweep_lam", N);
  int32_t *ell2 = buf<int32_t>(ctx, "sweep_ell2", N), *lam2 = buf<int32_t>(ctx, "sweep_lam2", N);
  uint32_t *cnt = buf<uint32_t>(ctx, "sweep_cnt", (size_t)N + 1);
  uint32_t *cur = buf<uint32_t>(ctx, "sweep_cur", N);
  int32_t *delta = buf<int32_t>(ctx, "sweep_delta", N);
  uint64_t *Kt = buf<uint64_t>(ctx, "sweep_kt", N);
  uint8_t *uni = buf<uint8_t>(ctx, "sweep_uni", N);
  uint8_t *chg = buf<uint8_t>(ctx, "sweep_chg", N);
  uint8_t *evdirty = buf<uint8_t>(ctx, "sweep_evdirty", N);
  uint8_t *ktchg = buf<uint8_t>(ctx, "sweep_ktchg", N);
  uint8_t *dflag = buf<uint8_t>(ctx, "sweep_dflag", N);
  uint32_t *longlist = buf<uint32_t>(ctx, "sweep_long", N), *hugelist = buf<uint32_t>(ctx, "sweep_huge", N);
  uint32_t *chlist = buf<uint32_t>(ctx, "sweep_chlist", N), *dlist = buf<uint32_t>(ctx, "sweep_dlist", N);
  uint32_t *d_nch = buf<uint32_t>(ctx, "sweep_nch", 1);
  int *d_err = buf<int>(ctx, "sweep_err", 1);
  SweepBufs E;
  E.eoff[0] = buf<uint64_t>(ctx, "sweep_eoff0", (size_t)N + 1);
  E.eoff[1] = buf<uint64_t>(ctx, "sweep_eoff1", (size_t)N + 1);
  C2_HIP(hipMemsetAsync(d_err, 0, sizeof(int), st));
  C2_HIP(hipMemsetAsync(delta, 0, sizeof(int32_t) * N, st));
  C2_HIP(hipMemsetAsync(chg, 0, N, st));

  // full build of the event state of the current story
  auto full_build = [&](int slot) -> uint64_t {
    C2_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t) * ((size_t)N + 1), st));
    C2_LAUNCH(ctx, k_ev_count, dim3(gN), dim3(B), 0, N, toff, tp, ell, cnt);
    exclusive_scan(ctx, "ev", cnt, E.eoff[slot], (int64_t)N + 1);
    const uint64_t nev = read_scalar(ctx, E.eoff[slot] + N);
    E.ev[slot] = buf<uint64_t>(ctx, slot ? "sweep_ev1" : "sweep_ev0", nev);
    C2_HIP(hipMemsetAsync(cur, 0, sizeof(uint32_t) * N, st));
    C2_LAUNCH(ctx, k_ev_fill, dim3(gN), dim3(B), 0, P, N, (const uint32_t *)nullptr, toff, tp, ell, lam,
              E.eoff[slot], cur, E.ev[slot], d_err);
    sort_lists(ctx, N, nullptr, E.eoff[slot], E.ev[slot], nev, uni, longlist, hugelist);
    C2_LAUNCH(ctx, k_kentry_all, dim3(gN), dim3(B), 0, P, N, E.eoff[slot], E.ev[slot], uni, Kt);
    return nev;
  };

  C2_LAUNCH(ctx, k_story_init, dim3(gN), dim3(B), 0, N, toff, tp, maxlab, ell, lam);
  R.events_total += (int64_t)full_build(E.cur);
  // first step: every start
  C2_HIP(hipMemsetAsync(d_nch, 0, sizeof(uint32_t), st));
  C2_LAUNCH(ctx, k_decide, dim3(gN), dim3(B), 0, P, N, (const uint32_t *)nullptr, toff, tp, E.eoff[E.cur],
            E.ev[E.cur], Kt, maxlab, ell, lam, ell2, lam2, chg, chlist, d_nch, d_err);
  uint32_t nc = read_scalar(ctx, d_nch);
  int it = 1;
  const int max_iters = 20000;
  while (nc > 0) {
    if (it >= max_iters) fail(C2GPU_ERR_NOCONVERGE, "c2gpu_bader: first-sweep fixpoint did not converge");
    it++;
    const unsigned gC = grid_for(nc, B);
    if (full_every_step) {
      C2_LAUNCH(ctx, k_apply, dim3(gC), dim3(B), 0, nc, chlist, ell, lam, ell2, lam2, chg);
      R.events_total += (int64_t)full_build(E.cur);
      C2_HIP(hipMemsetAsync(d_nch, 0, sizeof(uint32_t), st));
      C2_LAUNCH(ctx, k_decide, dim3(gN), dim3(B), 0, P, N, (const uint32_t *)nullptr, toff, tp, E.eoff[E.cur],
                E.ev[E.cur], Kt, maxlab, ell, lam, ell2, lam2, chg, chlist, d_nch, d_err);
      nc = read_scalar(ctx, d_nch);
      continue;
    }
    // 1. per-point count changes and the lists to rebuild
    C2_HIP(hipMemsetAsync(evdirty, 0, N, st));
    C2_HIP(hipMemsetAsync(ktchg, 0, N, st));
    C2_HIP(hipMemsetAsync(dflag, 0, N, st));
    C2_LAUNCH(ctx, k_delta, dim3(gC), dim3(B), 0, nc, chlist, toff, tp, ell, ell2, delta, evdirty);
    const int nx = 1 - E.cur;
    C2_LAUNCH(ctx, k_newcnt, dim3(grid_for((int64_t)N + 1, B)), dim3(B), 0, N, E.eoff[E.cur], delta, cnt);
    exclusive_scan(ctx, "ev", cnt, E.eoff[nx], (int64_t)N + 1);
    const uint64_t nev = read_scalar(ctx, E.eoff[nx] + N);
    R.events_total += (int64_t)nev;
    E.ev[nx] = buf<uint64_t>(ctx, nx ? "sweep_ev1" : "sweep_ev0", nev);
    // 2. move the lists, drop the changed starts' events, add their new ones
    C2_LAUNCH(ctx, k_ev_keep, dim3(gN), dim3(B), 0, P, N, E.eoff[E.cur], E.ev[E.cur], E.eoff[nx], E.ev[nx],
              evdirty, chg, cur, d_err);
    C2_LAUNCH(ctx, k_ev_fill, dim3(gC), dim3(B), 0, P, nc, chlist, toff, tp, ell2, lam2, E.eoff[nx], cur,
              E.ev[nx], d_err);
    C2_LAUNCH(ctx, k_apply, dim3(gC), dim3(B), 0, nc, chlist, ell, lam, ell2, lam2, chg);
    E.cur = nx;
    sort_lists(ctx, N, evdirty, E.eoff[nx], E.ev[nx], nev, uni, longlist, hugelist);
    // 3. known == 2 times around the rebuilt lists
    C2_LAUNCH(ctx, k_kentry_dirty, dim3(gN), dim3(B), 0, P, N, E.eoff[nx], E.ev[nx], uni, evdirty, Kt, ktchg);
    // 4. starts whose inputs changed decide again
    C2_LAUNCH(ctx, k_mark_decide, dim3(gN), dim3(B), 0, P, N, E.eoff[nx], E.ev[nx], evdirty, ktchg, dflag);
    const uint32_t nd = compact_flags(ctx, "decide", dflag, N, dlist);
    C2_HIP(hipMemsetAsync(d_nch, 0, sizeof(uint32_t), st));
    if (nd)
      C2_LAUNCH(ctx, k_decide, dim3(grid_for(nd, B)), dim3(B), 0, P, nd, dlist, toff, tp, E.eoff[nx], E.ev[nx],
                Kt, maxlab, ell, lam, ell2, lam2, chg, chlist, d_nch, d_err);
    nc = read_scalar(ctx, d_nch);
    if (read_scalar(ctx, d_err) & 4) fail(C2GPU_ERR_ARG, "c2gpu_bader: internal inconsistency in the sweep event lists");
    if (trace) {
      const uint32_t ndirty = compact_flags(ctx, "trace", evdirty, N, longlist);
      const uint32_t nkt = compact_flags(ctx, "trace", ktchg, N, longlist);
      std::fprintf(stderr, "[sweep] it %d changed %u decided %u dirty %u ktchg %u events %llu\n", it, nc, nd,
                   ndirty, nkt, (unsigned long long)nev);
    }
  }
  const int serr = read_scalar(ctx, d_err);
  if (serr & 2) fail(C2GPU_ERR_ARG, "c2gpu_bader: internal inconsistency in the sweep state");
  if (serr & 4) fail(C2GPU_ERR_ARG, "c2gpu_bader: internal inconsistency in the sweep event lists");
  if (std::getenv("C2GPU_SWEEP_VERIFY")) {
    // rebuild from the converged story and confirm it is a fixpoint
    uint64_t *Kt_inc = buf<uint64_t>(ctx, "sweep_kt_check", N);
    C2_HIP(hipMemcpyAsync(Kt_inc, Kt, 8 * (size_t)N, hipMemcpyDeviceToDevice, st));
    full_build(E.cur);
    C2_HIP(hipMemsetAsync(d_nch, 0, sizeof(uint32_t), st));
    C2_LAUNCH(ctx, k_decide, dim3(gN), dim3(B), 0, P, N, (const uint32_t *)nullptr, toff, tp, E.eoff[E.cur],
              E.ev[E.cur], Kt, maxlab, ell, lam, ell2, lam2, chg, chlist, d_nch, d_err);
    if (read_scalar(ctx, d_nch) != 0) fail(C2GPU_ERR_NOCONVERGE, "sweep verify: converged story is not a fixpoint");
    std::vector<uint64_t> a(N), b(N);
    C2_HIP(hipMemcpy(a.data(), Kt, 8 * (size_t)N, hipMemcpyDeviceToHost));
    C2_HIP(hipMemcpy(b.data(), Kt_inc, 8 * (size_t)N, hipMemcpyDeviceToHost));
    if (a != b) fail(C2GPU_ERR_NOCONVERGE, "sweep verify: incremental known==2 times differ from a rebuild");
  }
  unsigned long long *d_ns = buf<unsigned long long>(ctx, "sweep_nstarts", 1);
  C2_HIP(hipMemsetAsync(d_ns, 0, 8, st));
  C2_LAUNCH(ctx, k_final_state, dim3(gN), dim3(B), 0, N, E.eoff[E.cur], E.ev[E.cur], Kt, V, known, d_ns, ell);
  R.nstarts = (int64_t)read_scalar(ctx, d_ns);
  R.iters = it;
  return R;
}

}  // namespace c2
