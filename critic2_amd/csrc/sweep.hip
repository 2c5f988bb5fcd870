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
// Event lists live in an append-only pool: point y's list is
// pool[off[y] .. off[y]+cnt[y]).  The first step builds every list; later steps
// are incremental and proportional to what changed: the lists of points
// covered by a changed path are rewritten at the pool's end, the known==2
// times of their neighbourhoods recomputed, and only starts whose inputs
// changed are decided again.  A start's inputs are the events at the points
// of its committed path, so the starts to redo are read off those points'
// lists.  When the pool fills up, it is rebuilt from the story.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdio>
#include <cstdlib>
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

// The per-point event lists.
struct Lists {
  uint64_t *off;
  uint32_t *cnt;
  uint64_t *ev;
  __device__ __forceinline__ uint64_t b(uint32_t y) const { return off[y]; }
  __device__ __forceinline__ uint64_t e(uint32_t y) const { return off[y] + cnt[y]; }
};

// Label of y just after the events with time <= tau (inclusive) or < tau.
__device__ __forceinline__ int32_t v_at(const Lists &L, uint32_t y, uint64_t tau, bool inclusive) {
  const uint64_t b = L.b(y);
  uint64_t lo = b, hi = L.e(y);
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    const uint64_t t = ev_time(L.ev[mid]);
    if (inclusive ? (t <= tau) : (t < tau))
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo > b) ? ev_lab(L.ev[lo - 1]) : 0;
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

// Set flag[x]; the first setter appends x to list (each x at most once).
__device__ __forceinline__ void flag_append(uint32_t *flag, uint32_t x, uint32_t *list, uint32_t *n) {
  if (flag[x] == 0 && atomicExch(&flag[x], 1u) == 0) list[atomicAdd(n, 1u)] = x;
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

// Writes the events of the starts in `list` (all starts when list == null)
// into their points' lists at cursor cur[y].
__global__ void k_ev_fill(BParams P, uint32_t n, const uint32_t *__restrict__ list,
                          const uint64_t *__restrict__ toff, const uint32_t *__restrict__ tp,
                          const int32_t *__restrict__ ell, const int32_t *__restrict__ lam, Lists L,
                          uint32_t *__restrict__ cur, int *err) {
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
    if (slot < L.cnt[y])
      L.ev[L.off[y] + slot] = make_event(t0, l, lab);
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

// Sorts the short lists of the points in `list` (every point when null) in
// place and records whether each carries a single label; longer lists are
// queued for the workgroup and radix sorts.
__global__ void k_ev_sort(uint32_t n, const uint32_t *__restrict__ list, Lists L, uint8_t *__restrict__ uni,
                          uint32_t *__restrict__ longlist, uint32_t *nlong, uint32_t *__restrict__ hugelist,
                          uint32_t *nhuge) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t y = list ? list[t] : t;
  const uint64_t b = L.b(y), e = L.e(y);
  const uint64_t m = e - b;
  if (m > SORT_LDS) {
    hugelist[atomicAdd(nhuge, 1u)] = y;
    return;
  }
  if (m > SORT_INLINE) {
    longlist[atomicAdd(nlong, 1u)] = y;
    return;
  }
  uint64_t *ev = L.ev;
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
__global__ void __launch_bounds__(256) k_ev_sort_long(const uint32_t *__restrict__ longlist, Lists L,
                                                      uint8_t *__restrict__ uni) {
  __shared__ uint64_t sh[SORT_LDS];
  __shared__ int notuni;
  const uint32_t y = longlist[blockIdx.x];
  const uint64_t b = L.b(y), n = L.cnt[y];
  uint32_t m = 1;
  while (m < n) m <<= 1;
  for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) sh[i] = (i < n) ? L.ev[b + i] : ~0ull;
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
    L.ev[b + i] = sh[i];
    if (ev_lab(sh[i]) != ev_lab(sh[0])) notuni = 1;
  }
  __syncthreads();
  if (threadIdx.x == 0) uni[y] = notuni ? 0 : 1;
}

__global__ void k_huge_len(uint32_t nhuge, const uint32_t *__restrict__ hugelist, Lists L,
                           uint32_t *__restrict__ len) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > nhuge) return;
  len[q] = (q < nhuge) ? L.cnt[hugelist[q]] : 0u;
}

// Very long lists are gathered into one compact buffer for a segmented sort.
__global__ void k_huge_gather(const uint32_t *__restrict__ hugelist, Lists L, const uint64_t *__restrict__ hoff,
                              uint64_t *__restrict__ dst) {
  const uint32_t y = hugelist[blockIdx.x];
  const uint64_t b = L.b(y), n = L.cnt[y], o = hoff[blockIdx.x];
  for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) dst[o + i] = L.ev[b + i];
}

// Copies a sorted very long list back and records its single-label flag.
__global__ void k_huge_copyback(const uint32_t *__restrict__ hugelist, Lists L, const uint64_t *__restrict__ hoff,
                                const uint64_t *__restrict__ src, uint8_t *__restrict__ uni) {
  __shared__ int notuni;
  const uint32_t y = hugelist[blockIdx.x];
  const uint64_t b = L.b(y), n = L.cnt[y], o = hoff[blockIdx.x];
  if (threadIdx.x == 0) notuni = 0;
  __syncthreads();
  for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) {
    L.ev[b + i] = src[o + i];
    if (ev_lab(src[o + i]) != ev_lab(src[o])) notuni = 1;
  }
  __syncthreads();
  if (threadIdx.x == 0) uni[y] = notuni ? 0 : 1;
}

// ---------------------------------------------------------------- known == 2

// Time at which q enters known == 2 (assign_surrounding_pts and
// known_volnum_ongrid, bader@proc.f90:665-726).
__device__ uint64_t kentry_point(const BParams &P, uint32_t q, const Lists &L, const uint8_t *__restrict__ uni) {
  int i, j, k;
  coords(P, q, i, j, k);
  uint32_t nb[6];
  nb[0] = lin3(P, wrap1(i + 1, P.n1), j, k);
  nb[1] = lin3(P, wrap1(i - 1, P.n1), j, k);
  nb[2] = lin3(P, i, wrap1(j + 1, P.n2), k);
  nb[3] = lin3(P, i, wrap1(j - 1, P.n2), k);
  nb[4] = lin3(P, i, j, wrap1(k + 1, P.n3));
  nb[5] = lin3(P, i, j, wrap1(k - 1, P.n3));
  const uint64_t *ev = L.ev;
  // all seven points must carry a label
  if (L.cnt[q] == 0) return KT_INF;
  uint64_t tmin = ev_time(ev[L.b(q)]);
  bool alluni = uni[q] != 0;
  for (int t = 0; t < 6; t++) {
    if (L.cnt[nb[t]] == 0) return KT_INF;
    const uint64_t ft = ev_time(ev[L.b(nb[t])]);
    tmin = ft > tmin ? ft : tmin;
    alluni = alluni && uni[nb[t]];
  }
  if (alluni) {
    // labels are constant after tmin: the first candidate at or after it
    const int32_t l0 = ev_lab(ev[L.b(q)]);
    if (l0 <= 0) return KT_INF;
    for (int t = 0; t < 6; t++)
      if (ev_lab(ev[L.b(nb[t])]) != l0) return KT_INF;
    uint64_t kt = KT_INF;
    for (int t = 0; t < 6; t++) {
      const uint64_t e = L.e(nb[t]);
      const uint64_t x = first_at_or_after(ev, L.b(nb[t]), e, tmin);
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
    cend[t] = L.e(nb[t]);
    cpos[t] = first_at_or_after(ev, L.b(nb[t]), cend[t], tmin);
  }
  const uint32_t y7[7] = {q, nb[0], nb[1], nb[2], nb[3], nb[4], nb[5]};
  uint64_t vpos[7], vend[7], vbeg[7];
  for (int u = 0; u < 7; u++) {
    vbeg[u] = L.b(y7[u]);
    vend[u] = L.e(y7[u]);
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

// Kt for the points in `list` (every point when null); changed entries are
// flagged in ktchg when ktchg != null.
__global__ void k_kentry(BParams P, uint32_t n, const uint32_t *__restrict__ list, Lists L,
                         const uint8_t *__restrict__ uni, uint64_t *__restrict__ Kt, uint8_t *__restrict__ ktchg) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t q = list ? list[t] : t;
  const uint64_t kt = kentry_point(P, q, L, uni);
  if (ktchg && kt != Kt[q]) ktchg[q] = 1;
  Kt[q] = kt;
}

// ---------------------------------------------------------------- decisions

// One start's decision against the state of the current story
// (bader@proc.f90:154-211).
__device__ void decide_start(const BParams &P, uint32_t s, const uint64_t *__restrict__ toff,
                             const uint32_t *__restrict__ tp, const Lists &L, const uint64_t *__restrict__ Kt,
                             const int32_t *__restrict__ maxlab, int &e, int32_t &la, int *err) {
  const uint64_t ts = sweep_index(P, s) << POS_BITS;
  e = 0;
  la = 0;
  if (v_at(L, s, ts, false) != 0) return;  // volnum /= 0: not a start
  const uint64_t o = toff[s];
  const int len = (int)(toff[s + 1] - o);
  e = len;
  for (int q = 1; q < len; q++)
    if (Kt[tp[o + q]] < ts) {  // known == 2 when the path reaches it
      e = q + 1;
      break;
    }
  const uint32_t end = tp[o + e - 1];
  la = v_at(L, end, ts, false);
  if (la == 0) {
    if (e != len) atomicOr(err, 2);  // an unlabeled known==2 point cannot exist
    la = maxlab[end];                // new maximum -> its attractor
  }
}

// Decides the starts in `list` (every point when list == null); changed
// decisions are flagged and appended to `chlist`.
__global__ void k_decide(BParams P, uint32_t n, const uint32_t *__restrict__ list,
                         const uint64_t *__restrict__ toff, const uint32_t *__restrict__ tp, Lists L,
                         const uint64_t *__restrict__ Kt, const int32_t *__restrict__ maxlab,
                         const int32_t *__restrict__ ell, const int32_t *__restrict__ lam,
                         int32_t *__restrict__ ell2, int32_t *__restrict__ lam2, uint8_t *__restrict__ chg,
                         uint32_t *__restrict__ chlist, uint32_t *nch, int *err) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t s = list ? list[t] : t;
  int e;
  int32_t la;
  decide_start(P, s, toff, tp, L, Kt, maxlab, e, la, err);
  ell2[s] = e;
  lam2[s] = la;
  if (e != ell[s] || la != lam[s]) {
    chg[s] = 1;
    chlist[atomicAdd(nch, 1u)] = s;
  }
}

// ---------------------------------------------------------------- incremental step

// Count changes per point from the changed starts' old and new paths; every
// point on either path gets its list rewritten (listed once in dl).
__global__ void k_delta(uint32_t nc, const uint32_t *__restrict__ chlist, const uint64_t *__restrict__ toff,
                        const uint32_t *__restrict__ tp, const int32_t *__restrict__ ell,
                        const int32_t *__restrict__ ell2, int32_t *__restrict__ delta, uint32_t *__restrict__ dflag,
                        uint32_t *__restrict__ dl, uint32_t *ndl) {
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
    flag_append(dflag, y, dl, ndl);
  }
}

__global__ void k_patch_count(uint32_t nd, const uint32_t *__restrict__ dl, const uint32_t *__restrict__ cnt,
                              int32_t *__restrict__ delta, uint32_t *__restrict__ ncnt) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > nd) return;
  if (t == nd) {
    ncnt[nd] = 0;
    return;
  }
  const uint32_t y = dl[t];
  ncnt[t] = (uint32_t)((int64_t)cnt[y] + delta[y]);
  delta[y] = 0;
}

// New list of each rewritten point at the pool's end: its old events minus
// those of changed starts; the changed starts' new events follow (k_ev_fill).
__global__ void k_patch_keep(BParams P, uint32_t nd, const uint32_t *__restrict__ dl,
                             const uint64_t *__restrict__ noff, uint64_t base, Lists L,
                             const uint8_t *__restrict__ chg, uint32_t *__restrict__ cur, int *err) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nd) return;
  const uint32_t y = dl[t];
  const uint64_t ob = L.off[y], oe = ob + L.cnt[y];
  const uint64_t nb = base + noff[t];
  const uint64_t nn = noff[t + 1] - noff[t];
  uint32_t kept = 0;
  for (uint64_t x = ob; x < oe; x++) {
    const uint64_t v = L.ev[x];
    if (!chg[event_start(P, v)]) {
      if (kept < nn)
        L.ev[nb + kept] = v;
      else
        atomicOr(err, 4);
      kept++;
    }
  }
  cur[y] = kept;
  L.off[y] = nb;
  L.cnt[y] = (uint32_t)nn;
}

// Points whose known==2 time may change: the rewritten points and their face
// neighbours (listed once in ql).
__global__ void k_mark_kt(BParams P, uint32_t nd, const uint32_t *__restrict__ dl, uint32_t *__restrict__ qflag,
                          uint32_t *__restrict__ ql, uint32_t *nql) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nd) return;
  const uint32_t y = dl[t];
  int i, j, k;
  coords(P, y, i, j, k);
  flag_append(qflag, y, ql, nql);
  flag_append(qflag, lin3(P, wrap1(i + 1, P.n1), j, k), ql, nql);
  flag_append(qflag, lin3(P, wrap1(i - 1, P.n1), j, k), ql, nql);
  flag_append(qflag, lin3(P, i, wrap1(j + 1, P.n2), k), ql, nql);
  flag_append(qflag, lin3(P, i, wrap1(j - 1, P.n2), k), ql, nql);
  flag_append(qflag, lin3(P, i, j, wrap1(k + 1, P.n3)), ql, nql);
  flag_append(qflag, lin3(P, i, j, wrap1(k - 1, P.n3)), ql, nql);
}

// Starts to decide again: those with an event at a point whose list or
// known==2 time changed, and those points themselves.
__global__ void k_mark_decide(BParams P, uint32_t nq, const uint32_t *__restrict__ ql, Lists L,
                              const uint32_t *__restrict__ dflag, const uint8_t *__restrict__ ktchg,
                              uint32_t *__restrict__ sflag, uint32_t *__restrict__ sl, uint32_t *nsl) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nq) return;
  const uint32_t x = ql[t];
  if (!(dflag[x] || ktchg[x])) return;
  flag_append(sflag, x, sl, nsl);
  for (uint64_t q = L.b(x); q < L.e(x); q++) flag_append(sflag, event_start(P, L.ev[q]), sl, nsl);
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

template <class F>
__global__ void k_clear_list(uint32_t n, const uint32_t *__restrict__ list, F *__restrict__ flag) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) flag[list[t]] = 0;
}

__global__ void k_final_state(uint32_t N, Lists L, const uint64_t *__restrict__ Kt, int32_t *__restrict__ V,
                              int8_t *__restrict__ known, unsigned long long *nstarts,
                              const int32_t *__restrict__ ell) {
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t st = 0;
  if (y < N) {
    V[y] = L.cnt[y] ? ev_lab(L.ev[L.e(y) - 1]) : 0;
    known[y] = (Kt[y] != KT_INF) ? 2 : 0;
    st = ell[y] > 0;
  }
  wave_add(nstarts, st);
}

// ---------------------------------------------------------------- host driver

// Sorts the lists of the points in `list` (all points when null).
static void sort_lists(c2gpu_ctx *ctx, uint32_t n, const uint32_t *list, const Lists &L, uint8_t *uni,
                       uint32_t *longlist, uint32_t *hugelist) {
  hipStream_t st = ctx->stream;
  if (n == 0) return;
  uint32_t *d_n = buf<uint32_t>(ctx, "sweep_nlist", 2);
  C2_HIP(hipMemsetAsync(d_n, 0, 2 * sizeof(uint32_t), st));
  C2_LAUNCH(ctx, k_ev_sort, dim3(grid_for(n, 256)), dim3(256), 0, n, list, L, uni, longlist, d_n, hugelist,
            d_n + 1);
  uint32_t nlh[2];
  C2_HIP(hipMemcpyAsync(nlh, d_n, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  C2_HIP(hipStreamSynchronize(st));
  if (nlh[0]) C2_LAUNCH(ctx, k_ev_sort_long, dim3(nlh[0]), dim3(256), 0, longlist, L, uni);
  if (nlh[1]) {
    const uint32_t nh = nlh[1];
    uint32_t *hlen = buf<uint32_t>(ctx, "sweep_hugelen", (size_t)nh + 1);
    uint64_t *hoff = buf<uint64_t>(ctx, "sweep_hugeoff", (size_t)nh + 1);
    C2_LAUNCH(ctx, k_huge_len, dim3(grid_for((int64_t)nh + 1, 64)), dim3(64), 0, nh, hugelist, L, hlen);
    exclusive_scan(ctx, "huge", hlen, hoff, (int64_t)nh + 1);
    const uint64_t tot = read_scalar(ctx, hoff + nh);
    if (tot >= (1ull << 31)) fail(C2GPU_ERR_ARG, "c2gpu_bader: sweep event lists too long to sort");
    uint64_t *keys = buf<uint64_t>(ctx, "sweep_hugekeys", tot), *alt = buf<uint64_t>(ctx, "sweep_hugealt", tot);
    C2_LAUNCH(ctx, k_huge_gather, dim3(nh), dim3(256), 0, hugelist, L, hoff, keys);
    size_t tmp = 0;
    C2_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tmp, keys, alt, (int)tot, (int)nh, hoff, hoff + 1, 0,
                                                      64, st));
    void *t = buffer(ctx, "sweep_segsort_tmp", tmp);
    C2_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(t, tmp, keys, alt, (int)tot, (int)nh, hoff, hoff + 1, 0, 64,
                                                      st));
    C2_LAUNCH(ctx, k_huge_copyback, dim3(nh), dim3(256), 0, hugelist, L, hoff, alt, uni);
  }
}

struct SweepResult {
  int iters;
  int64_t events_total;
  int64_t nstarts;
};

// Runs the fixpoint; on return V holds volnum after the sweep and known the
// known==2 marks (2 or 0).
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
  uint32_t *cntN = buf<uint32_t>(ctx, "sweep_cnt", (size_t)N + 1);
  uint64_t *offN = buf<uint64_t>(ctx, "sweep_off", (size_t)N + 1);
  uint32_t *cur = buf<uint32_t>(ctx, "sweep_cur", N);
  int32_t *delta = buf<int32_t>(ctx, "sweep_delta", N);
  uint64_t *Kt = buf<uint64_t>(ctx, "sweep_kt", N);
  uint8_t *uni = buf<uint8_t>(ctx, "sweep_uni", N);
  uint8_t *chg = buf<uint8_t>(ctx, "sweep_chg", N);
  uint8_t *ktchg = buf<uint8_t>(ctx, "sweep_ktchg", N);
  uint32_t *dflag = buf<uint32_t>(ctx, "sweep_dflag", N);  // list rewritten this step
  uint32_t *qflag = buf<uint32_t>(ctx, "sweep_qflag", N);  // known==2 time to recompute
  uint32_t *sflag = buf<uint32_t>(ctx, "sweep_sflag", N);  // start to decide again
  uint32_t *dl = buf<uint32_t>(ctx, "sweep_dl", N), *ql = buf<uint32_t>(ctx, "sweep_ql", N);
  uint32_t *sl = buf<uint32_t>(ctx, "sweep_sl", N);
  uint32_t *longlist = buf<uint32_t>(ctx, "sweep_long", N), *hugelist = buf<uint32_t>(ctx, "sweep_huge", N);
  uint32_t *chlist = buf<uint32_t>(ctx, "sweep_chlist", N);
  uint32_t *ncnt = buf<uint32_t>(ctx, "sweep_ncnt", (size_t)N + 1);
  uint64_t *noff = buf<uint64_t>(ctx, "sweep_noff", (size_t)N + 1);
  uint32_t *d_n = buf<uint32_t>(ctx, "sweep_counts", 4);  // nch, ndl, nql, nsl
  int *d_err = buf<int>(ctx, "sweep_err", 1);
  C2_HIP(hipMemsetAsync(d_err, 0, sizeof(int), st));
  C2_HIP(hipMemsetAsync(delta, 0, sizeof(int32_t) * N, st));
  C2_HIP(hipMemsetAsync(chg, 0, N, st));
  C2_HIP(hipMemsetAsync(ktchg, 0, N, st));
  C2_HIP(hipMemsetAsync(dflag, 0, sizeof(uint32_t) * N, st));
  C2_HIP(hipMemsetAsync(qflag, 0, sizeof(uint32_t) * N, st));
  C2_HIP(hipMemsetAsync(sflag, 0, sizeof(uint32_t) * N, st));

  Lists L;
  L.off = offN;
  L.cnt = cntN;
  L.ev = nullptr;
  uint64_t pool_cap = 0, pool_top = 0;

  // every list from the story: counts, offsets, events, order, known==2 times
  auto full_build = [&]() {
    C2_HIP(hipMemsetAsync(cntN, 0, sizeof(uint32_t) * ((size_t)N + 1), st));
    C2_LAUNCH(ctx, k_ev_count, dim3(gN), dim3(B), 0, N, toff, tp, ell, cntN);
    exclusive_scan(ctx, "ev", cntN, offN, (int64_t)N + 1);
    const uint64_t nev = read_scalar(ctx, offN + N);
    if (nev + (uint64_t)N * 4 > pool_cap) {  // room for later rewrites too
      pool_cap = nev * 3 + (uint64_t)N * 4;
      L.ev = buf<uint64_t>(ctx, "sweep_pool", pool_cap);
      pool_cap = ctx->bufs["sweep_pool"].bytes / 8;
    }
    pool_top = nev;
    C2_HIP(hipMemsetAsync(cur, 0, sizeof(uint32_t) * N, st));
    C2_LAUNCH(ctx, k_ev_fill, dim3(gN), dim3(B), 0, P, N, (const uint32_t *)nullptr, toff, tp, ell, lam, L, cur,
              d_err);
    sort_lists(ctx, N, nullptr, L, uni, longlist, hugelist);
    C2_LAUNCH(ctx, k_kentry, dim3(gN), dim3(B), 0, P, N, (const uint32_t *)nullptr, L, uni, Kt,
              (uint8_t *)nullptr);
    R.events_total += (int64_t)nev;
  };
  auto decide_all = [&]() -> uint32_t {
    C2_HIP(hipMemsetAsync(d_n, 0, sizeof(uint32_t), st));
    C2_LAUNCH(ctx, k_decide, dim3(gN), dim3(B), 0, P, N, (const uint32_t *)nullptr, toff, tp, L, Kt, maxlab, ell,
              lam, ell2, lam2, chg, chlist, d_n, d_err);
    return read_scalar(ctx, d_n);
  };

  C2_LAUNCH(ctx, k_story_init, dim3(gN), dim3(B), 0, N, toff, tp, maxlab, ell, lam);
  full_build();
  uint32_t nc = decide_all();
  int it = 1;
  const int max_iters = 20000;
  while (nc > 0) {
    if (it >= max_iters) fail(C2GPU_ERR_NOCONVERGE, "c2gpu_bader: first-sweep fixpoint did not converge");
    it++;
    const unsigned gC = grid_for(nc, B);
    // 1. points whose lists change, and their new sizes
    uint32_t counts[4] = {0, 0, 0, 0};
    C2_HIP(hipMemsetAsync(d_n + 1, 0, 3 * sizeof(uint32_t), st));
    C2_LAUNCH(ctx, k_delta, dim3(gC), dim3(B), 0, nc, chlist, toff, tp, ell, ell2, delta, dflag, dl, d_n + 1);
    const uint32_t nd = read_scalar(ctx, d_n + 1);
    C2_LAUNCH(ctx, k_patch_count, dim3(grid_for((int64_t)nd + 1, B)), dim3(B), 0, nd, dl, cntN, delta, ncnt);
    exclusive_scan(ctx, "patch", ncnt, noff, (int64_t)nd + 1);
    const uint64_t added = read_scalar(ctx, noff + nd);
    if (full_every_step || pool_top + added > pool_cap) {
      // rebuild every list from the updated story (compacts the pool)
      C2_LAUNCH(ctx, k_clear_list<uint32_t>, dim3(grid_for(nd, B)), dim3(B), 0, nd, dl, dflag);
      C2_LAUNCH(ctx, k_apply, dim3(gC), dim3(B), 0, nc, chlist, ell, lam, ell2, lam2, chg);
      full_build();
      nc = decide_all();
      if (trace) std::fprintf(stderr, "[sweep] it %d full rebuild, changed %u\n", it, nc);
      continue;
    }
    // 2. rewrite those lists at the pool's end; add the changed starts' events
    C2_LAUNCH(ctx, k_patch_keep, dim3(grid_for(nd, B)), dim3(B), 0, P, nd, dl, noff, pool_top, L, chg, cur, d_err);
    C2_LAUNCH(ctx, k_ev_fill, dim3(gC), dim3(B), 0, P, nc, chlist, toff, tp, ell2, lam2, L, cur, d_err);
    C2_LAUNCH(ctx, k_apply, dim3(gC), dim3(B), 0, nc, chlist, ell, lam, ell2, lam2, chg);
    pool_top += added;
    R.events_total += (int64_t)added;
    sort_lists(ctx, nd, dl, L, uni, longlist, hugelist);
    // 3. known==2 times around the rewritten lists
    C2_LAUNCH(ctx, k_mark_kt, dim3(grid_for(nd, B)), dim3(B), 0, P, nd, dl, qflag, ql, d_n + 2);
    const uint32_t nq = read_scalar(ctx, d_n + 2);
    C2_LAUNCH(ctx, k_kentry, dim3(grid_for(nq, B)), dim3(B), 0, P, nq, ql, L, uni, Kt, ktchg);
    // 4. starts whose inputs changed decide again
    C2_LAUNCH(ctx, k_mark_decide, dim3(grid_for(nq, B)), dim3(B), 0, P, nq, ql, L, dflag, ktchg, sflag, sl,
              d_n + 3);
    C2_HIP(hipMemcpyAsync(counts, d_n, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    C2_HIP(hipStreamSynchronize(st));
    const uint32_t ns = counts[3];
    C2_HIP(hipMemsetAsync(d_n, 0, sizeof(uint32_t), st));
    if (ns)
      C2_LAUNCH(ctx, k_decide, dim3(grid_for(ns, B)), dim3(B), 0, P, ns, sl, toff, tp, L, Kt, maxlab, ell, lam,
                ell2, lam2, chg, chlist, d_n, d_err);
    // 5. clear the step's flags through their lists
    C2_LAUNCH(ctx, k_clear_list<uint32_t>, dim3(grid_for(nd, B)), dim3(B), 0, nd, dl, dflag);
    if (nq) {
      C2_LAUNCH(ctx, k_clear_list<uint32_t>, dim3(grid_for(nq, B)), dim3(B), 0, nq, ql, qflag);
      C2_LAUNCH(ctx, k_clear_list<uint8_t>, dim3(grid_for(nq, B)), dim3(B), 0, nq, ql, ktchg);
    }
    if (ns) C2_LAUNCH(ctx, k_clear_list<uint32_t>, dim3(grid_for(ns, B)), dim3(B), 0, ns, sl, sflag);
    nc = read_scalar(ctx, d_n);
    if (read_scalar(ctx, d_err) & 4) fail(C2GPU_ERR_ARG, "c2gpu_bader: internal inconsistency in the sweep event lists");
    if (trace)
      std::fprintf(stderr, "[sweep] it %d changed %u rewritten %u kt %u decided %u added %llu pool %llu/%llu\n", it,
                   nc, nd, nq, ns, (unsigned long long)added, (unsigned long long)pool_top,
                   (unsigned long long)pool_cap);
  }
  const int serr = read_scalar(ctx, d_err);
  if (serr & 2) fail(C2GPU_ERR_ARG, "c2gpu_bader: internal inconsistency in the sweep state");
  if (serr & 4) fail(C2GPU_ERR_ARG, "c2gpu_bader: internal inconsistency in the sweep event lists");
  if (std::getenv("C2GPU_SWEEP_VERIFY")) {
    // rebuild from the converged story and confirm it is a fixpoint
    uint64_t *Kt_inc = buf<uint64_t>(ctx, "sweep_kt_check", N);
    C2_HIP(hipMemcpyAsync(Kt_inc, Kt, 8 * (size_t)N, hipMemcpyDeviceToDevice, st));
    full_build();
    if (decide_all() != 0) fail(C2GPU_ERR_NOCONVERGE, "sweep verify: converged story is not a fixpoint");
    std::vector<uint64_t> a(N), b(N);
    C2_HIP(hipMemcpy(a.data(), Kt, 8 * (size_t)N, hipMemcpyDeviceToHost));
    C2_HIP(hipMemcpy(b.data(), Kt_inc, 8 * (size_t)N, hipMemcpyDeviceToHost));
    if (a != b) fail(C2GPU_ERR_NOCONVERGE, "sweep verify: incremental known==2 times differ from a rebuild");
  }
  unsigned long long *d_ns = buf<unsigned long long>(ctx, "sweep_nstarts", 1);
  C2_HIP(hipMemsetAsync(d_ns, 0, 8, st));
  C2_LAUNCH(ctx, k_final_state, dim3(gN), dim3(B), 0, N, L, Kt, V, known, d_ns, ell);
  R.nstarts = (int64_t)read_scalar(ctx, d_ns);
  R.iters = it;
  return R;
}

}  // namespace c2
