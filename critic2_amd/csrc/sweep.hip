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
// Every commit is an event stamped time = (sweep index << 11 | position on
// the path); per point, the events form a time-sorted list from which the
// state of the serial sweep at any moment is read back: volnum(y) at time t is
// the label of y's last event before t, and the time y enters known == 2 is
// the first commit at a face neighbour of y after which y and its six face
// neighbours carry one common positive label.  Recomputing every decision
// against the state its story implies is one Jacobi step.  The serial
// recurrence has exactly one solution and every fixpoint satisfies it, so the
// converged story is the serial sweep.
//
// Event lists live in one pool: point y's list is pool[off[y] .. +cnt[y])
// inside a reserved block of cap[y] slots.  The first step builds every list;
// later steps are incremental and proportional to what changed: the lists of
// points covered by a changed path are rewritten (in place when they fit,
// else moved to the pool's end with room to grow), the known==2 times of their
// neighbourhoods recomputed, and only starts whose inputs changed are decided
// again.  A start's inputs are the events at the points of its committed path,
// so the starts to redo are read off those points' lists.  Working sets are
// byte flags gathered into index lists by one streaming pass each, with one
// atomic per workgroup.  When the pool fills up, it is rebuilt from the story.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

#include "bader_dev.hpp"
#include "common.hpp"
#include "devutil.hpp"

namespace c2 {

static constexpr uint64_t KT_INF = ~0ull;
static constexpr int POS_BITS = 11;     // position on a path: < 2048
static constexpr int LAB_BITS = 20;     // label: < 2^20
static constexpr int SORT_INLINE = 8;   // longer event lists are sorted by wavefronts
static constexpr int SORT_WAVE = 64;    // longer still: by workgroups
static constexpr int SORT_LDS = 4096;   // longer still: segmented radix sort

__device__ __forceinline__ uint64_t ev_time(uint64_t e) { return e >> LAB_BITS; }
__device__ __forceinline__ int32_t ev_lab(uint64_t e) { return (int32_t)(e & ((1u << LAB_BITS) - 1)); }
__device__ __forceinline__ uint32_t ev_pos(uint64_t e) { return (uint32_t)(ev_time(e) & ((1u << POS_BITS) - 1)); }
__device__ __forceinline__ uint64_t make_event(uint64_t t0, int l, int32_t lab) {
  return ((t0 | (uint64_t)l) << LAB_BITS) | (uint64_t)(uint32_t)lab;
}
// the start point (linear index) whose path produced an event; the sweep
// index of a grid of fewer than 2^32 points fits 32 bits
__device__ __forceinline__ uint32_t event_start(const BParams &P, uint64_t e) {
  const uint32_t sg = (uint32_t)(ev_time(e) >> POS_BITS);
  uint32_t k, j;
  const uint32_t r = udiv(sg, (uint32_t)P.n3, P.inv3, k);
  const uint32_t i = udiv(r, (uint32_t)P.n2, P.inv2, j);
  return lin3(P, (int)i, (int)j, (int)k);
}

// The per-point event lists.
// Lists longer than this are filtered / scanned by a whole workgroup.
constexpr uint32_t LIST_LONG = 32;

struct Lists {
  uint64_t *off;
  uint32_t *cnt;
  uint32_t *cap;
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

// first_at_or_after over K ranges at once ([lo[k], hi[k]) -> pos[k]): the
// searches advance in lockstep, so their loads overlap and the dependent
// chain is that of the longest range alone.
template <int K>
__device__ __forceinline__ void first_at_or_after_k(const uint64_t *__restrict__ ev, const uint64_t *lo,
                                                    const uint64_t *hi, uint64_t t, uint64_t *pos) {
  uint64_t n[K];
  for (int k = 0; k < K; k++) {
    pos[k] = lo[k];
    n[k] = hi[k] - lo[k];
  }
  for (;;) {
    bool any = false;
    for (int k = 0; k < K; k++)
      if (n[k] > 0) {
        const uint64_t half = n[k] >> 1;
        if (ev_time(ev[pos[k] + half]) < t) {
          pos[k] += half + 1;
          n[k] -= half + 1;
        } else {
          n[k] = half;
        }
        any = true;
      }
    if (!any) break;
  }
}

// Counters that many wavefronts add to are spread over CSLOTS slots (slot
// q of a counter at dst[q * CSTRIDE]; the host sums them): one address
// taking an atomic from every wavefront of a full-grid launch serialises
// them (k_final_state took 25 ms for its count of starts, 1.4 ms without).
constexpr int CSLOTS = 64, CSTRIDE = 16;
__device__ __forceinline__ void wave_add(unsigned long long *dst, uint64_t v) {
  if (!dst) return;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v)
    atomicAdd(dst + CSTRIDE * ((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (CSLOTS - 1)),
              (unsigned long long)v);
}

// Appends val to list where pred holds: one atomic per wavefront.  Every lane
// of the wavefront must call it.
__device__ __forceinline__ void wave_append(bool pred, uint32_t val, uint32_t *list, uint32_t *n) {
  const uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(n, (uint32_t)__popcll(mask));
  base = __shfl(base, leader, 64);
  if (pred) list[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = val;
}

// Gathers the indices of set byte flags (flags are 0 or 1; the flag array is
// padded with zeros to a multiple of 16 bytes) into list, in no particular
// order, clearing the flags if clear != 0.  Each thread takes 16 flags; one
// atomic per workgroup and pass.
__global__ void __launch_bounds__(256) k_collect(uint32_t N, uint8_t *__restrict__ flag, int clear,
                                                 uint32_t *__restrict__ list, uint32_t *n) {
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t bbase;
  const uint32_t nq = (N + 15) >> 4;
  uint4 *fq = reinterpret_cast<uint4 *>(flag);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // COLLECT_Q groups of 16 flags per thread and pass: the loads of a pass
  // are in flight together, and the scan and barriers are paid once per
  // COLLECT_Q groups
  constexpr int COLLECT_Q = 4;
  const uint32_t span = blockDim.x * COLLECT_Q;
  for (uint32_t q0 = blockIdx.x * span; q0 < nq; q0 += gridDim.x * span) {
    uint4 vv[COLLECT_Q];
    int cc[COLLECT_Q];
    int c = 0;
#pragma unroll
    for (int g = 0; g < COLLECT_Q; g++) {
      const uint32_t q = q0 + g * blockDim.x + threadIdx.x;
      vv[g] = (q < nq) ? fq[q] : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int g = 0; g < COLLECT_Q; g++) {
      const uint4 v = vv[g];
      if (clear && (v.x | v.y | v.z | v.w) != 0u) fq[q0 + g * blockDim.x + threadIdx.x] = make_uint4(0u, 0u, 0u, 0u);
      cc[g] = __popc(v.x & 0x01010101u) + __popc(v.y & 0x01010101u) + __popc(v.z & 0x01010101u) +
              __popc(v.w & 0x01010101u);
      c += cc[g];
    }
    int incl = c;
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[wv] = (uint32_t)incl;
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
      bbase = tot ? atomicAdd(n, tot) : 0u;
    }
    __syncthreads();
    uint32_t o = bbase + (uint32_t)(incl - c);
    for (int w = 0; w < wv; w++) o += wsum[w];
    if (c)
      for (int g = 0; g < COLLECT_Q; g++) {
        if (!cc[g]) continue;
        const uint32_t q = q0 + g * blockDim.x + threadIdx.x;
        const uint32_t wd[4] = {vv[g].x, vv[g].y, vv[g].z, vv[g].w};
        for (int a = 0; a < 4; a++)
          for (int b = 0; b < 4; b++)
            if ((wd[a] >> (8 * b)) & 1u) list[o++] = 16 * q + 4 * a + b;
      }
    __syncthreads();  // wsum and bbase are rewritten by the next pass
  }
}

template <class F>
__global__ void k_clear_list(uint32_t n, const uint32_t *__restrict__ list, F *__restrict__ flag) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) flag[list[t]] = 0;
}

// ---------------------------------------------------------------- event lists

// The plain first story (C2GPU_SWEEP_PRE=0): every point a start whose path
// holds its first INIT_PATH points, with the label of its free trajectory's
// maximum.  Any first story converges to the same fixpoint (each decision
// depends only on earlier events); of these plain ones this is closest in
// work: on the 512^3 step the fixpoint took 3.85 s from paths of 1 point,
// 2.91 s from 2, 3.01 s from 3, 3.12 s from 4, 3.22 s from 8, 3.41 s from 16
// and 4.19 s from whole paths.  The default first story is the approximate
// one of k_pre_* below.
constexpr int32_t INIT_PATH = 2;

__global__ void k_story_init(uint32_t N, const uint64_t *__restrict__ toff, const uint32_t *__restrict__ tp,
                             const int32_t *__restrict__ maxlab, int32_t *__restrict__ ell,
                             int32_t *__restrict__ lam) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const int32_t len = (int32_t)(toff[s + 1] - toff[s]);
  ell[s] = len < INIT_PATH ? len : INIT_PATH;
  lam[s] = maxlab[tp[toff[s + 1] - 1]];
}

// Path positions examined per batch of independent loads.
constexpr int PRE_CHUNK = 4;

// Length of a committed path: the trajectory tp[o .. o+len) up to and
// including its first point (from position 1) whose known == 2 time is
// before ts, or all of it.  PRE_CHUNK positions are loaded at a time, so the
// loads of a batch are in flight together instead of one after another.
// Position q (>= 1) of start s's trajectory: from the position-major
// prefix table when it holds it, else from the CSR copy.
__device__ __forceinline__ uint32_t traj_at(const BParams &P, uint32_t s, const uint32_t *__restrict__ tp, uint64_t o,
                                            int q) {
  return (P.tpk && q <= TPK) ? P.tpk[(size_t)(q - 1) * P.npts + s] : tp[o + q];
}

__device__ __forceinline__ int stop_position(const BParams &P, uint32_t s, const uint32_t *__restrict__ tp,
                                             uint64_t o, int len, const uint64_t *__restrict__ Kt, uint64_t ts) {
  constexpr int SC = 8;  // path points whose indices and Kt are loaded together
  for (int q0 = 1; q0 < len; q0 += SC) {
    uint32_t x[SC];
#pragma unroll
    for (int u = 0; u < SC; u++) x[u] = q0 + u < len ? traj_at(P, s, tp, o, q0 + u) : 0u;
    uint64_t kt[SC];
#pragma unroll
    for (int u = 0; u < SC; u++) kt[u] = q0 + u < len ? Kt[x[u]] : KT_INF;
#pragma unroll
    for (int u = 0; u < SC; u++)
      if (kt[u] < ts) return q0 + u + 1;
  }
  return len;
}

// A closer first story from a few cheap rounds without event lists.  Every
// point's label is taken as its free trajectory's attractor (flab), so a
// point enters known == 2 once it and its six face neighbours are covered
// and their labels agree: at the latest of their covering times (Kt).  The
// covering times C follow from the paths those Kt allow (a start is a point
// not covered before its own turn; its path stops at the first point whose
// Kt is earlier); rounds alternate C and Kt from C = each point's own time.
// Labels that later paths change are what the fixpoint then corrects.
//
// The rounds compare covering times only with a start's own time, whose
// path position is 0, so they keep sweep indices alone (32 bits; the
// position on the covering path never decides a comparison).
constexpr uint32_t PRE_INF = 0xffffffffu;

__device__ __forceinline__ uint32_t sweep_index32(const BParams &P, uint32_t l) {
  return (uint32_t)sweep_index(P, l);
}

// The rounds keep each point's known == 2 time and covering time in one
// 64-bit word (time low, covering time high), so a path point costs one
// gather; the covering time is lowered with a 32-bit atomic on its half.
__device__ __forceinline__ uint64_t pre_word(uint32_t kt, uint32_t c) { return ((uint64_t)c << 32) | kt; }
__device__ __forceinline__ uint32_t pre_kt(uint64_t w) { return (uint32_t)w; }
__device__ __forceinline__ uint32_t pre_c(uint64_t w) { return (uint32_t)(w >> 32); }
__device__ __forceinline__ uint32_t *pre_c_ptr(uint64_t *W, uint32_t x) {
  return reinterpret_cast<uint32_t *>(W + x) + 1;
}

__global__ void k_pre_flab(uint32_t N, const uint64_t *__restrict__ toff, const uint32_t *__restrict__ tp,
                           const int32_t *__restrict__ maxlab, int32_t *__restrict__ flab,
                           uint64_t *__restrict__ W, BParams P) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  flab[s] = maxlab[tp[toff[s + 1] - 1]];
  W[s] = pre_word(PRE_INF, sweep_index32(P, s));
}

// Covering by every point's first step (the plain first story's paths of
// two points), a middle ground to start the rounds from.
__global__ void k_pre_seed(BParams P, uint32_t N, const uint64_t *__restrict__ toff, const uint32_t *__restrict__ tp,
                           uint64_t *__restrict__ W) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const uint64_t o = toff[s];
  if (toff[s + 1] - o < 2) return;
  atomicMin(pre_c_ptr(W, traj_at(P, s, tp, o, 1)), sweep_index32(P, s));
}

// Whether a point and its six face neighbours carry one positive free
// label: fixed for the rounds, computed once.
__global__ void k_pre_agree(BParams P, uint32_t N, const int32_t *__restrict__ flab, uint8_t *__restrict__ agree) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  int i, j, k;
  coords(P, s, i, j, k);
  const int32_t la = flab[s];
  bool same = la > 0;
  same = same && flab[lin3(P, wrap1(i + 1, P.n1), j, k)] == la && flab[lin3(P, wrap1(i - 1, P.n1), j, k)] == la;
  same = same && flab[lin3(P, i, wrap1(j + 1, P.n2), k)] == la && flab[lin3(P, i, wrap1(j - 1, P.n2), k)] == la;
  same = same && flab[lin3(P, i, j, wrap1(k + 1, P.n3))] == la && flab[lin3(P, i, j, wrap1(k - 1, P.n3))] == la;
  agree[s] = same;
}

// Win: this round's covering times (high halves).  Wout[s] = the next
// round's known == 2 time from them, and its covering time starting at the
// point's own time (a point no path covers is a start and covers itself).
// nchanged (diagnostics) compares with Wout's previous covering times.
__global__ void k_pre_kt(BParams P, uint32_t N, const uint64_t *__restrict__ Win, const uint8_t *__restrict__ agree,
                         uint64_t *__restrict__ Wout, unsigned long long *nchanged) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (nchanged) wave_add(nchanged, s < N && pre_c(Win[s]) != pre_c(Wout[s]) ? 1 : 0);
  if (s >= N) return;
  uint32_t m = PRE_INF;
  if (agree[s]) {
    int i, j, k;
    coords(P, s, i, j, k);
    m = pre_c(Win[s]);
    const uint32_t z[6] = {lin3(P, wrap1(i + 1, P.n1), j, k), lin3(P, wrap1(i - 1, P.n1), j, k),
                           lin3(P, i, wrap1(j + 1, P.n2), k), lin3(P, i, wrap1(j - 1, P.n2), k),
                           lin3(P, i, j, wrap1(k + 1, P.n3)), lin3(P, i, j, wrap1(k - 1, P.n3))};
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const uint32_t c = pre_c(Win[z[q]]);
      m = c > m ? c : m;
    }
  }
  Wout[s] = pre_word(m, sweep_index32(P, s));
}

__global__ void k_pre_cover(BParams P, uint32_t N, const uint64_t *__restrict__ toff, const uint32_t *__restrict__ tp,
                            const uint64_t *__restrict__ Wold, uint64_t *__restrict__ W, int cap) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const uint32_t ts = sweep_index32(P, s);
  if (pre_c(Wold[s]) < ts) return;  // covered before its turn: not a start
  // Cnew starts at each point's own time, so the start itself needs no
  // update; a covering time only decreases, so a value already at or below
  // this path's needs no atomic either (most points lie on many paths)
  const uint64_t o = toff[s];
  int len = (int)(toff[s + 1] - o);
  if (len > cap) len = cap;
  // PRE_CHUNK path points at a time: their indices and Kt are loaded
  // together (independent loads), then the stop is found among them
  for (int q0 = 1; q0 < len; q0 += PRE_CHUNK) {
    uint32_t x[PRE_CHUNK];
    uint64_t w[PRE_CHUNK];
#pragma unroll
    for (int u = 0; u < PRE_CHUNK; u++) x[u] = q0 + u < len ? traj_at(P, s, tp, o, q0 + u) : 0u;
    // each path point's Kt and current covering time in one load; a stale
    // (larger) covering time only costs an atomic that changes nothing
#pragma unroll
    for (int u = 0; u < PRE_CHUNK; u++)
      w[u] = q0 + u < len ? __hip_atomic_load(&W[x[u]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0ull;
    bool stop = false;
#pragma unroll
    for (int u = 0; u < PRE_CHUNK; u++) {
      if (stop || q0 + u >= len) break;
      if (ts < pre_c(w[u])) atomicMin(pre_c_ptr(W, x[u]), ts);
      stop = pre_kt(w[u]) < ts;
    }
    if (stop) break;
  }
}

// WC: the final covering times (high halves); WK: the known == 2 times
// from them (low halves).
__global__ void k_pre_story(BParams P, uint32_t N, const uint64_t *__restrict__ toff, const uint32_t *__restrict__ tp,
                            const uint64_t *__restrict__ WC, const uint64_t *__restrict__ WK,
                            const int32_t *__restrict__ flab, int32_t *__restrict__ ell, int32_t *__restrict__ lam) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const uint32_t ts = sweep_index32(P, s);
  if (pre_c(WC[s]) < ts) {
    ell[s] = 0;
    lam[s] = 0;
    return;
  }
  const uint64_t o = toff[s];
  const int len = (int)(toff[s + 1] - o);
  int e = len;
  for (int q0 = 1; q0 < len && e == len; q0 += PRE_CHUNK) {
    uint32_t x[PRE_CHUNK], kt[PRE_CHUNK];
#pragma unroll
    for (int u = 0; u < PRE_CHUNK; u++) x[u] = q0 + u < len ? traj_at(P, s, tp, o, q0 + u) : 0u;
#pragma unroll
    for (int u = 0; u < PRE_CHUNK; u++) kt[u] = q0 + u < len ? pre_kt(WK[x[u]]) : PRE_INF;
#pragma unroll
    for (int u = 0; u < PRE_CHUNK; u++)
      if (e == len && kt[u] < ts) e = q0 + u + 1;
  }
  ell[s] = e;
  lam[s] = flab[e > 1 ? traj_at(P, s, tp, o, e - 1) : s];
}

__global__ void k_ev_count(uint32_t N, const uint64_t *__restrict__ toff, const uint32_t *__restrict__ tp,
                           const int32_t *__restrict__ ell, uint32_t *__restrict__ cnt) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const int e = ell[s];
  const uint64_t o = toff[s];
  for (int l = 0; l < e; l++) atomicAdd(&cnt[tp[o + l]], 1u);
}

// Reserved block per list: the count plus a quarter, plus one.  capv has a
// zero sentinel at N for the offset scan.
__global__ void k_reserve(uint32_t N, const uint32_t *__restrict__ cnt, uint32_t *__restrict__ cap,
                          uint32_t *__restrict__ capv) {
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y > N) return;
  if (y == N) {
    capv[N] = 0;
    return;
  }
  const uint32_t c = cnt[y];
  const uint32_t r = c + (c >> 2) + 1;
  cap[y] = r;
  capv[y] = r;
}

// Writes the events of the starts in `list` (all starts when list == null)
// into their points' lists at cursor cur[y].
__global__ void __launch_bounds__(256) k_ev_fill(BParams P, uint32_t n, const uint32_t *__restrict__ list,
                                                 const uint64_t *__restrict__ toff, const uint32_t *__restrict__ tp,
                                                 const int32_t *__restrict__ ell, const int32_t *__restrict__ lam,
                                                 const uint16_t *__restrict__ from, Lists L,
                                                 uint32_t *__restrict__ cur, int *err, unsigned long long *work,
                                                 int32_t *__restrict__ ell_to, int32_t *__restrict__ lam_to,
                                                 uint8_t *__restrict__ chg) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t s = t < n ? (list ? list[t] : t) : 0u;
  const int e = t < n ? ell[s] : 0;
  const int f = (t < n && from) ? (int)from[s] : 0;  // events before f are already in place
  if (t < n && ell_to) {
    // k_apply's commit of the new decision, for the starts of this launch
    ell_to[s] = e;
    lam_to[s] = lam[s];
    chg[s] = 0;
  }
  if (e > f) {
    const uint64_t t0 = sweep_index(P, s) << POS_BITS;
    const int32_t lab = lam[s];
    const uint64_t o = toff[s];
    constexpr int FB = 16;  // slot reservations in flight at once
    for (int l0 = f; l0 < e; l0 += FB) {
      const int m = (e - l0) < FB ? e - l0 : FB;
      uint32_t y[FB], slot[FB], cnt[FB];
      uint64_t off[FB];
#pragma unroll
      for (int u = 0; u < FB; u++) y[u] = u < m ? tp[o + l0 + u] : 0u;
#pragma unroll
      for (int u = 0; u < FB; u++) {
        slot[u] = u < m ? atomicAdd(&cur[y[u]], 1u) : 0u;
        cnt[u] = u < m ? L.cnt[y[u]] : 0u;
        off[u] = u < m ? L.off[y[u]] : 0u;
      }
#pragma unroll
      for (int u = 0; u < FB; u++)
        if (u < m) {
          if (slot[u] < cnt[u])
            L.ev[off[u] + slot[u]] = make_event(t0, l0 + u, lab);
          else
            atomicOr(err, 4);  // count/fill mismatch (never expected): refuse to write
        }
    }
  }
  wave_add(work, (uint64_t)(e > f ? e - f : 0));
}

__device__ __forceinline__ bool list_uniform(const uint64_t *__restrict__ ev, uint64_t b, uint64_t e) {
  if (e <= b) return true;
  const int32_t l0 = ev_lab(ev[b]);
  for (uint64_t x = b + 1; x < e; x++)
    if (ev_lab(ev[x]) != l0) return false;
  return true;
}

// Sorts the short lists (up to SORT_INLINE events) of the points in `list`
// (every point when null) in registers and records whether each carries a
// single label; longer lists are queued for the wavefront sort (from the
// back of longlist, n - 1 - index), the workgroup sort (from its front) and
// the radix sort.  A list is in one queue only, so the two ends never meet.
__global__ void k_ev_sort(uint32_t n, const uint32_t *__restrict__ list, Lists L, uint8_t *__restrict__ uni,
                          uint32_t *__restrict__ longlist, uint32_t *nlong, uint32_t *nwave,
                          uint32_t *__restrict__ hugelist, uint32_t *nhuge, unsigned long long *work) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t sorted = 0;
  const uint32_t y = t < n ? (list ? list[t] : t) : 0u;
  const uint64_t m = t < n ? (uint64_t)L.cnt[y] : 0;
  // the queues take one atomic per wavefront each
  wave_append(m > SORT_LDS, y, hugelist, nhuge);
  wave_append(m > SORT_WAVE && m <= SORT_LDS, y, longlist, nlong);
  {
    const bool w = m > SORT_INLINE && m <= SORT_WAVE;
    const uint64_t mask = __ballot(w);
    if (mask) {
      const int lane = threadIdx.x & 63;
      const int leader = __ffsll((unsigned long long)mask) - 1;
      uint32_t base = 0;
      if (lane == leader) base = atomicAdd(nwave, (uint32_t)__popcll(mask));
      base = __shfl(base, leader, 64);
      if (w) longlist[n - 1 - (base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull)))] = y;
    }
  }
  if (t < n && m <= SORT_INLINE) {
    const uint64_t b = L.b(y), e = b + m;
    {
      uint64_t *ev = L.ev;
      uint64_t r[SORT_INLINE];
#pragma unroll
      for (int u = 0; u < SORT_INLINE; u++) r[u] = u < (int)m ? ev[b + u] : ~0ull;
#pragma unroll
      for (int i = 1; i < SORT_INLINE; i++)
#pragma unroll
        for (int j = i; j > 0; j--)
          if (r[j - 1] > r[j]) {
            const uint64_t x = r[j - 1];
            r[j - 1] = r[j];
            r[j] = x;
          }
      bool same = true;
#pragma unroll
      for (int u = 0; u < SORT_INLINE; u++)
        if (u < (int)m) {
          ev[b + u] = r[u];
          same = same && ev_lab(r[u]) == ev_lab(r[0]);
        }
      uni[y] = same;
      sorted = m;
    }
  }
  wave_add(work, sorted);
}

// One wavefront per list of SORT_INLINE+1 .. SORT_WAVE events: a bitonic
// sort across the lanes (one event each, padding sorts last).
__global__ void __launch_bounds__(256) k_ev_sort_wave(const uint32_t *__restrict__ wavelist, uint32_t nw, Lists L,
                                                      uint8_t *__restrict__ uni) {
  const uint32_t w = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  if (w >= nw) return;  // the whole wavefront
  const int lane = threadIdx.x & 63;
  const uint32_t y = wavelist[w];
  const uint64_t b = L.b(y);
  const int m = (int)L.cnt[y];
  uint64_t v = lane < m ? L.ev[b + lane] : ~0ull;
  for (int k = 2; k <= 64; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      const uint64_t o = __shfl_xor(v, j, 64);
      const bool up = (lane & k) == 0, low = (lane & j) == 0;
      v = (low == up) ? (o < v ? o : v) : (o > v ? o : v);
    }
  if (lane < m) L.ev[b + lane] = v;
  const int32_t l0 = __shfl(ev_lab(v), 0, 64);
  const uint64_t diff = __ballot(lane < m && ev_lab(v) != l0);
  if (lane == 0) uni[y] = diff == 0;
}

// One workgroup per long list, in LDS.  A rewritten list is its kept
// events (still in order) followed by the new ones: up to its first descent
// k it is sorted.  When at most 256 events follow, they alone are sorted
// (bitonic) and merged with the first k (each event's place is its index
// plus the count of the other run's smaller events, a binary search);
// otherwise the whole list is sorted (bitonic).
__global__ void __launch_bounds__(256) k_ev_sort_long(const uint32_t *__restrict__ longlist, Lists L,
                                                      uint8_t *__restrict__ uni) {
  __shared__ uint64_t sh[SORT_LDS];
  __shared__ uint64_t tl[256];
  __shared__ int notuni;
  __shared__ uint32_t kd;
  const uint32_t y = longlist[blockIdx.x];
  const uint64_t b = L.b(y), n = L.cnt[y];
  uint32_t m = 1;
  while (m < n) m <<= 1;
  for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) sh[i] = (i < n) ? L.ev[b + i] : ~0ull;
  if (threadIdx.x == 0) {
    notuni = 0;
    kd = (uint32_t)n;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x + 1; i < n; i += blockDim.x)
    if (sh[i] < sh[i - 1]) atomicMin(&kd, i);
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
    if (ev_lab(sh[i]) != ev_lab(sh[0])) notuni = 1;
  __syncthreads();
  const uint32_t k = kd, t = (uint32_t)n - k;
  if (t == 0) {
    // already in order
  } else if (t <= 256) {
    uint32_t tp = 1;
    while (tp < t) tp <<= 1;
    if (threadIdx.x < tp) tl[threadIdx.x] = threadIdx.x < t ? sh[k + threadIdx.x] : ~0ull;
    __syncthreads();
    for (uint32_t kk = 2; kk <= tp; kk <<= 1)
      for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
        const uint32_t i = threadIdx.x;
        if (i < tp) {
          const uint32_t ixj = i ^ j;
          if (ixj > i) {
            const bool up = ((i & kk) == 0);
            const uint64_t a = tl[i], c = tl[ixj];
            if ((a > c) == up) {
              tl[i] = c;
              tl[ixj] = a;
            }
          }
        }
        __syncthreads();
      }
    // the kept run: index + tail events below it
    for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
      const uint64_t v = sh[i];
      uint32_t lo = 0, hi = t;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tl[mid] < v)
          lo = mid + 1;
        else
          hi = mid;
      }
      L.ev[b + i + lo] = v;
    }
    // the tail: index + kept events below it
    if (threadIdx.x < t) {
      const uint64_t v = tl[threadIdx.x];
      uint32_t lo = 0, hi = k;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sh[mid] < v)
          lo = mid + 1;
        else
          hi = mid;
      }
      L.ev[b + threadIdx.x + lo] = v;
    }
  } else {
    for (uint32_t kk = 2; kk <= m; kk <<= 1)
      for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
        for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
          const uint32_t ixj = i ^ j;
          if (ixj > i) {
            const bool up = ((i & kk) == 0);
            const uint64_t a = sh[i], c = sh[ixj];
            if ((a > c) == up) {
              sh[i] = c;
              sh[ixj] = a;
            }
          }
        }
        __syncthreads();
      }
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) L.ev[b + i] = sh[i];
  }
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

// A rewritten very long list is its kept events (still in order) followed
// by the new ones, few next to millions.  Such a list is merged instead of
// radix-sorted: its first descent k ends the sorted run, the tail after it
// (at most HUGE_TAIL events, else the batch takes the radix sort) is sorted
// in LDS, and every event goes to its index in its run plus the count of
// the other run's smaller events (distinct times: a permutation).
constexpr uint32_t HUGE_TAIL = 4096;

__global__ void __launch_bounds__(256) k_huge_desc(const uint64_t *__restrict__ boff,
                                                   const uint64_t *__restrict__ keys, uint32_t *__restrict__ kd,
                                                   uint32_t *bigtail) {
  __shared__ uint32_t m;
  const uint64_t o = boff[blockIdx.x], n = boff[blockIdx.x + 1] - o;
  if (threadIdx.x == 0) m = (uint32_t)n;
  __syncthreads();
  for (uint64_t i = threadIdx.x + 1; i < n; i += blockDim.x)
    if (keys[o + i] < keys[o + i - 1]) atomicMin(&m, (uint32_t)i);
  __syncthreads();
  if (threadIdx.x == 0) {
    kd[blockIdx.x] = m;
    if (n - m > HUGE_TAIL) atomicOr(bigtail, 1u);
  }
}

__global__ void __launch_bounds__(256) k_huge_tail_sort(const uint64_t *__restrict__ boff,
                                                        const uint64_t *__restrict__ keys,
                                                        const uint32_t *__restrict__ kd, uint64_t *__restrict__ tails) {
  __shared__ uint64_t sh[HUGE_TAIL];
  const uint64_t o = boff[blockIdx.x], n = boff[blockIdx.x + 1] - o;
  const uint32_t k = kd[blockIdx.x], t = (uint32_t)(n - k);
  if (t == 0) return;  // the whole workgroup
  uint32_t m = 1;
  while (m < t) m <<= 1;
  for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) sh[i] = i < t ? keys[o + k + i] : ~0ull;
  __syncthreads();
  for (uint32_t kk = 2; kk <= m; kk <<= 1)
    for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        const uint32_t ixj = i ^ j;
        if (ixj > i) {
          const bool up = ((i & kk) == 0);
          const uint64_t a = sh[i], c = sh[ixj];
          if ((a > c) == up) {
            sh[i] = c;
            sh[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  for (uint32_t i = threadIdx.x; i < t; i += blockDim.x) tails[o + k + i] = sh[i];
}

// Every event of the batch to its merged place in its list (L.ev; the runs
// are read from the gathered copies); notuni[q] is set when list q holds
// more than one label.
__global__ void k_huge_merge(uint32_t nb, const uint32_t *__restrict__ hugelist, Lists L,
                             const uint64_t *__restrict__ boff, const uint64_t *__restrict__ keys,
                             const uint64_t *__restrict__ tails, const uint32_t *__restrict__ kd,
                             uint32_t *__restrict__ notuni) {
  const uint64_t total = boff[nb];
  for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < total;
       x += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = nb;  // the list: the last q with boff[q] <= x
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (boff[mid] <= x)
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t q = lo;
    const uint64_t o = boff[q], r = x - o, n = boff[q + 1] - o;
    const uint32_t k = kd[q];
    const uint64_t t = n - k;
    uint64_t v, pos;
    if (r < k) {
      v = keys[x];
      uint64_t a = 0, b = t;  // tail events below v
      while (a < b) {
        const uint64_t mid = (a + b) >> 1;
        if (tails[o + k + mid] < v)
          a = mid + 1;
        else
          b = mid;
      }
      pos = r + a;
    } else {
      v = tails[x];
      uint64_t a = 0, b = k;  // kept events below v
      while (a < b) {
        const uint64_t mid = (a + b) >> 1;
        if (keys[o + mid] < v)
          a = mid + 1;
        else
          b = mid;
      }
      pos = (r - k) + a;
    }
    if (t) L.ev[L.b(hugelist[q]) + pos] = v;  // (no tail: the list is in order already)
    if (ev_lab(v) != ev_lab(keys[o])) notuni[q] = 1;
  }
}

__global__ void k_huge_uni(uint32_t nb, const uint32_t *__restrict__ hugelist, const uint32_t *__restrict__ notuni,
                           uint8_t *__restrict__ uni) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nb) uni[hugelist[q]] = notuni[q] ? 0 : 1;
}

// ---------------------------------------------------------------- known == 2

// The six face neighbours of q (periodic).
__device__ __forceinline__ void face_neighbours(const BParams &P, uint32_t q, uint32_t nb[6]) {
  int i, j, k;
  coords(P, q, i, j, k);
  nb[0] = lin3(P, wrap1(i + 1, P.n1), j, k);
  nb[1] = lin3(P, wrap1(i - 1, P.n1), j, k);
  nb[2] = lin3(P, i, wrap1(j + 1, P.n2), k);
  nb[3] = lin3(P, i, wrap1(j - 1, P.n2), k);
  nb[4] = lin3(P, i, j, wrap1(k + 1, P.n3));
  nb[5] = lin3(P, i, j, wrap1(k - 1, P.n3));
}

// Candidate events above which a point's merge is cut into work items
// (k_kt_items) instead of running on one thread.
constexpr uint64_t KT_LONG = 4;

// First index at or after pos (in [pos, end)) whose event time exceeds c:
// a gallop from pos, then a binary search (cursors that jump far over a
// long list cost a logarithm, not the jump).
__device__ __forceinline__ uint64_t advance_past(const uint64_t *__restrict__ ev, uint64_t pos, uint64_t end,
                                                 uint64_t c) {
  if (pos >= end || ev_time(ev[pos]) > c) return pos;
  uint64_t lo = pos, step = 1;  // ev[lo] <= c
  while (lo + step < end && ev_time(ev[lo + step]) <= c) {
    lo += step;
    step <<= 1;
  }
  const uint64_t hi = lo + step < end ? lo + step : end;
  return first_at_or_after(ev, lo + 1, hi, c + 1);
}

// Whether the labels of the seven points agree (and are > 0) at time c:
// each cursor moves to the first event after c, whose predecessor carries
// the label at c (none: 0).
__device__ __forceinline__ bool labels_agree(const uint64_t *__restrict__ ev, uint64_t c, uint64_t vpos[7],
                                             const uint64_t vbeg[7], const uint64_t vend[7]) {
  int32_t lab[7];
  for (int u = 0; u < 7; u++) {
    vpos[u] = advance_past(ev, vpos[u], vend[u], c);
    lab[u] = (vpos[u] > vbeg[u]) ? ev_lab(ev[vpos[u] - 1]) : 0;
  }
  bool ok = lab[0] > 0;
  for (int u = 1; u < 7 && ok; u++) ok = (lab[u] == lab[0]);
  return ok;
}

// The merge of the general case over the candidates with times in
// [tlo, thi): the six neighbour lists' events in time order, the labels of
// the seven points followed with forward-only cursors; the first time at
// which all seven labels agree (and are > 0), or KT_INF.  At most `steps`
// candidates are taken per call; cpos/vpos carry the state between calls
// and *more says whether candidates remain.
__device__ __forceinline__ uint64_t kt_merge(const uint64_t *__restrict__ ev, uint64_t cpos[6],
                                             const uint64_t cend[6], uint64_t vpos[7], const uint64_t vbeg[7],
                                             const uint64_t vend[7], int steps, bool *more) {
  for (int it = 0; it < steps; it++) {
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
    if (tc < 0) {
      *more = false;
      return KT_INF;
    }
    cpos[tc]++;
    if (labels_agree(ev, c, vpos, vbeg, vend)) {
      *more = false;
      return c;
    }
  }
  *more = true;
  return KT_INF;
}

// Time at which q enters known == 2 (assign_surrounding_pts and
// known_volnum_ongrid, bader@proc.f90:665-726).  A general-case merge of
// more than KT_LONG candidates is not done here: *defer is set instead.
__device__ uint64_t kentry_point(const BParams &P, uint32_t q, const Lists &L, const uint8_t *__restrict__ uni,
                                 bool *defer) {
  *defer = false;
  uint32_t nb[6];
  face_neighbours(P, q, nb);
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
    uint64_t b6[6], e6[6], x6[6];
    for (int t = 0; t < 6; t++) {
      b6[t] = L.b(nb[t]);
      e6[t] = L.e(nb[t]);
    }
    first_at_or_after_k<6>(ev, b6, e6, tmin, x6);
    uint64_t kt = KT_INF;
    for (int t = 0; t < 6; t++)
      if (x6[t] < e6[t]) {
        const uint64_t tt = ev_time(ev[x6[t]]);
        kt = tt < kt ? tt : kt;
      }
    return kt;
  }
  // general case: merge the six neighbour lists in time order from tmin;
  // the label cursors of the seven points start where the candidates do
  const uint32_t y7[7] = {q, nb[0], nb[1], nb[2], nb[3], nb[4], nb[5]};
  uint64_t vpos[7], vend[7], vbeg[7];
  for (int u = 0; u < 7; u++) {
    vbeg[u] = L.b(y7[u]);
    vend[u] = L.e(y7[u]);
  }
  first_at_or_after_k<7>(ev, vbeg, vend, tmin, vpos);
  uint64_t cpos[6], cend[6], total = 0;
  for (int t = 0; t < 6; t++) {
    cpos[t] = vpos[t + 1];
    cend[t] = vend[t + 1];
    total += cend[t] - cpos[t];
  }
  if (total > KT_LONG) {
    *defer = true;
    return KT_INF;
  }
  bool more;
  return kt_merge(ev, cpos, cend, vpos, vbeg, vend, (int)KT_LONG + 1, &more);
}

// Kt for the points in `list` (every point when null); changed entries are
// flagged in ktchg when ktchg != null.  Points with long merges are queued
// in longq for k_kt_plan / k_kt_items / k_kt_commit.
__global__ void k_kentry(BParams P, uint32_t n, const uint32_t *__restrict__ list, Lists L,
                         const uint8_t *__restrict__ uni, uint64_t *__restrict__ Kt, uint8_t *__restrict__ ktchg,
                         uint32_t *__restrict__ longq, uint32_t *nlong) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t q = list ? list[t] : t;
  bool defer;
  const uint64_t kt = kentry_point(P, q, L, uni, &defer);
  if (!defer) {
    if (ktchg && kt != Kt[q]) ktchg[q] = 1;
    Kt[q] = kt;
  }
  wave_append(defer, q, longq, nlong);
}

// The long merges.  Every candidate of such a point can be checked on its
// own (labels_agree from cursors placed by binary search), so the
// candidates are cut into work items of KT_CHUNK consecutive events of one
// neighbour list; an item checks its candidates in time order and lowers
// the point's time to its first agreement.  The minimum over the items is
// the first agreement in the merged order, the serial merge's answer.

constexpr uint64_t KT_CHUNK = 4;

// Per deferred point w: its lists' first candidates (lb) and the prefix
// of its item counts over the six lists (cpre, 7 entries); cnt[w] = items.
// nbound bounds the deferred count *nlp (read on the device); entries from
// *nlp to nbound get no items.
__global__ void k_kt_plan(BParams P, const uint32_t *nlp, uint32_t nbound, const uint32_t *__restrict__ longq,
                          Lists L, uint64_t *__restrict__ lbw, uint32_t *__restrict__ cpre,
                          uint32_t *__restrict__ cnt, uint64_t *__restrict__ ktmp) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w > nbound) return;
  if (w >= *nlp) {  // past the deferred points, and the scan sentinel
    cnt[w] = 0;
    return;
  }
  const uint32_t q = longq[w];
  uint32_t nb[6];
  face_neighbours(P, q, nb);
  const uint64_t *ev = L.ev;
  // deferred points have events at all seven points
  uint64_t tmin = ev_time(ev[L.b(q)]);
  for (int t = 0; t < 6; t++) {
    const uint64_t ft = ev_time(ev[L.b(nb[t])]);
    tmin = ft > tmin ? ft : tmin;
  }
  uint64_t b6[6], e6[6], lb[6];
  for (int t = 0; t < 6; t++) {
    b6[t] = L.b(nb[t]);
    e6[t] = L.e(nb[t]);
  }
  first_at_or_after_k<6>(ev, b6, e6, tmin, lb);
  uint32_t acc = 0;
  for (int t = 0; t < 6; t++) {
    lbw[6 * (size_t)w + t] = lb[t];
    cpre[7 * (size_t)w + t] = acc;
    acc += (uint32_t)((e6[t] - lb[t] + KT_CHUNK - 1) / KT_CHUNK);
  }
  cpre[7 * (size_t)w + 6] = acc;
  cnt[w] = acc;
  ktmp[w] = KT_INF;
}

__global__ void __launch_bounds__(256) k_kt_items(BParams P, uint32_t nl, const uint32_t *__restrict__ longq,
                                                  Lists L, const uint64_t *__restrict__ lbw,
                                                  const uint32_t *__restrict__ cpre,
                                                  const uint64_t *__restrict__ off, uint64_t *__restrict__ ktmp) {
  // nl: the bound of k_kt_plan (points past the deferred ones own no items)
  const uint64_t total = off[nl];
  const uint64_t *ev = L.ev;
  for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < total;
       x += (uint64_t)gridDim.x * blockDim.x) {
    // the point: the last w with off[w] <= x
    uint32_t lo = 0, hi = nl;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (off[mid] <= x)
        lo = mid;
      else
        hi = mid;
    }
    const uint32_t w = lo;
    const uint32_t r = (uint32_t)(x - off[w]);
    int t = 0;
    while (t < 5 && cpre[7 * (size_t)w + t + 1] <= r) t++;
    const uint32_t q = longq[w];
    uint32_t nb[6];
    face_neighbours(P, q, nb);
    const uint64_t c0 = lbw[6 * (size_t)w + t] + (uint64_t)(r - cpre[7 * (size_t)w + t]) * KT_CHUNK;
    uint32_t nbt = nb[0];  // nb[t] by selects (a dynamic index would put nb in scratch)
#pragma unroll
    for (int u = 1; u < 6; u++) nbt = u == t ? nb[u] : nbt;
    const uint64_t le = L.e(nbt);
    const uint64_t c1 = c0 + KT_CHUNK < le ? c0 + KT_CHUNK : le;
    const uint64_t tfirst = ev_time(ev[c0]);
    if (__hip_atomic_load(&ktmp[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < tfirst) continue;
    const uint32_t y7[7] = {q, nb[0], nb[1], nb[2], nb[3], nb[4], nb[5]};
    uint64_t vpos[7], vend[7], vbeg[7];
    for (int u = 0; u < 7; u++) {
      vbeg[u] = L.b(y7[u]);
      vend[u] = L.e(y7[u]);
    }
    first_at_or_after_k<7>(ev, vbeg, vend, tfirst, vpos);
    for (uint64_t i = c0; i < c1; i++) {
      const uint64_t c = ev_time(ev[i]);
      if (labels_agree(ev, c, vpos, vbeg, vend)) {
        atomicMin((unsigned long long *)&ktmp[w], (unsigned long long)c);
        break;
      }
    }
  }
}

__global__ void k_kt_commit(const uint32_t *nlp, uint32_t nb, const uint32_t *__restrict__ longq,
                            const uint64_t *__restrict__ ktmp, uint64_t *__restrict__ Kt, uint8_t *__restrict__ ktchg) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nb || w >= *nlp) return;
  const uint32_t q = longq[w];
  const uint64_t kt = ktmp[w];
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
  e = stop_position(P, s, tp, o, len, Kt, ts);  // stops at a point known == 2 when the path reaches it
  // (e positions of the path were read: tp and Kt)
  const uint32_t end = e > 1 ? traj_at(P, s, tp, o, e - 1) : s;
  la = v_at(L, end, ts, false);
  if (la == 0) {
    if (e != len) atomicOr(err, 2);  // an unlabeled known==2 point cannot exist
    la = maxlab[end];                // new maximum -> its attractor
  }
}

// Decides the starts in `list` (every point when list == null); changed
// decisions are flagged and appended to `chlist`.
__global__ void __launch_bounds__(256) k_decide(BParams P, uint32_t n, const uint32_t *__restrict__ list,
                                                const uint64_t *__restrict__ toff, const uint32_t *__restrict__ tp,
                                                Lists L, const uint64_t *__restrict__ Kt,
                                                const int32_t *__restrict__ maxlab, const int32_t *__restrict__ ell,
                                                const int32_t *__restrict__ lam, int32_t *__restrict__ ell2,
                                                int32_t *__restrict__ lam2, uint8_t *__restrict__ chg,
                                                uint32_t *__restrict__ chlist, uint32_t *nch, int *err,
                                                unsigned long long *first_change, unsigned long long *work) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  bool changed = false;
  uint32_t s = 0;
  int e = 0;
  if (t < n) {
    s = list ? list[t] : t;
    int32_t la;
    decide_start(P, s, toff, tp, L, Kt, maxlab, e, la, err);
    ell2[s] = e;
    lam2[s] = la;
    changed = (e != ell[s] || la != lam[s]);
    if (changed) chg[s] = 1;
  }
  wave_add(work, (uint64_t)e);
  if (first_change) {  // earliest changed start in sweep order (diagnostics)
    uint64_t v = changed ? sweep_index(P, s) : ~0ull;
    for (int o = 32; o > 0; o >>= 1) {
      const uint64_t w = __shfl_down(v, o, 64);
      v = w < v ? w : v;
    }
    if ((threadIdx.x & 63) == 0 && v != ~0ull) atomicMin(first_change, (unsigned long long)v);
  }
  wave_append(changed, s, chlist, nch);
}

// ---------------------------------------------------------------- incremental step

// Count changes per point from the changed starts' old and new paths.  Both
// paths are prefixes of the start's trajectory, so when the label stays the
// same the events of the common prefix are identical and only the positions
// from min(old, new) length on change (from[s]); every point holding such a
// position gets its list rewritten (dflag).
__global__ void __launch_bounds__(256) k_delta(uint32_t nc, const uint32_t *__restrict__ chlist,
                                               const uint64_t *__restrict__ toff, const uint32_t *__restrict__ tp,
                                               const int32_t *__restrict__ ell, const int32_t *__restrict__ ell2,
                                               const int32_t *__restrict__ lam, const int32_t *__restrict__ lam2,
                                               int32_t *__restrict__ delta, uint8_t *__restrict__ dflag,
                                               uint16_t *__restrict__ from, unsigned long long *stats,
                                               unsigned long long *work) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t st[4] = {0, 0, 0, 0};
  uint64_t upd = 0;
  if (t < nc) {
    const uint32_t s = chlist[t];
    const uint64_t o = toff[s];
    const int eo = ell[s], en = ell2[s];
    const int m = eo > en ? eo : en, mn = eo < en ? eo : en;
    const bool same = lam[s] == lam2[s];
    const int f = same ? mn : 0;
    from[s] = (uint16_t)f;
    for (int l0 = f; l0 < m; l0 += 8) {  // 8 path points read ahead of their updates
      uint32_t y[8];
#pragma unroll
      for (int u = 0; u < 8; u++) y[u] = l0 + u < m ? tp[o + l0 + u] : 0u;
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int l = l0 + u;
        if (l < m) {
          const int d = (l < en ? 1 : 0) - (l < eo ? 1 : 0);
          if (d) atomicAdd(&delta[y[u]], d);
          dflag[y[u]] = 1;
        }
      }
    }
    st[0] = same && mn > 0;           // changes that keep the label
    st[1] = same ? (uint64_t)mn : 0;  // unchanged prefix events
    st[2] = (uint64_t)(m - f);        // positions touched
    st[3] = same ? 0 : (uint64_t)mn;  // prefix events relabelled
    upd = (uint64_t)(m - mn);
  }
  if (stats)
    for (int q = 0; q < 4; q++) wave_add(stats + q, st[q]);
  wave_add(work, st[2]);
  wave_add(work + 1, upd);
}

// New size of each rewritten list; a list that outgrows its block gets a new
// block of 1.5x its size (+2) at the pool's end (alloc > 0, scanned).
__global__ void k_patch_count(uint32_t nd, const uint32_t *__restrict__ dl, const uint32_t *__restrict__ cnt,
                              const uint32_t *__restrict__ cap, int32_t *__restrict__ delta,
                              uint32_t *__restrict__ ncnt, uint32_t *__restrict__ alloc, int *err) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > nd) return;
  if (t == nd) {
    alloc[nd] = 0;
    return;
  }
  const uint32_t y = dl[t];
  const int64_t c = (int64_t)cnt[y] + delta[y];
  delta[y] = 0;
  if (c < 0) atomicOr(err, 4);  // more events removed than present (never expected)
  const uint32_t nc = c < 0 ? 0u : (uint32_t)c;
  ncnt[t] = nc;
  alloc[t] = nc > cap[y] ? nc + (nc >> 1) + 2 : 0u;
}

// Rewrites each list: its old events minus those of changed starts, in place
// (the write index never passes the read index) or in its new block; the
// changed starts' new events follow (k_ev_fill).
__global__ void __launch_bounds__(256) k_patch_keep(BParams P, uint32_t nd, const uint32_t *__restrict__ dl,
                                                    const uint32_t *__restrict__ ncnt,
                                                    const uint32_t *__restrict__ alloc,
                                                    const uint64_t *__restrict__ aoff, uint64_t base, Lists L,
                                                    const uint8_t *__restrict__ chg,
                                                    const uint16_t *__restrict__ from, uint32_t *__restrict__ cur,
                                                    uint32_t *__restrict__ longq, uint32_t *nlong, int *err,
                                                    unsigned long long *work) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t nold = 0, nkept = 0;
  const bool lng = t < nd && L.cnt[dl[t]] > LIST_LONG;
  wave_append(lng, t, longq, nlong);  // a workgroup each (k_patch_keep_long)
  if (t < nd && !lng) {
    const uint32_t y = dl[t];
    const uint64_t ob = L.off[y], oe = ob + L.cnt[y];
    const uint32_t nn = ncnt[t];
    const bool move = alloc[t] != 0;
    const uint64_t nb = move ? base + aoff[t] : ob;
    const uint32_t room = move ? alloc[t] : L.cap[y];
    const uint32_t lim = nn <= room ? nn : room;  // room >= nn by construction
    if (nn > room) atomicOr(err, 4);
    uint32_t kept = 0;
    // 8 events and their starts' flags are read before any of them is
    // written: the writes (in place when the list does not move) never
    // reach past the events already read
    for (uint64_t x0 = ob; x0 < oe; x0 += 8) {
      uint64_t v[8];
      bool keep[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = x0 + u < oe ? L.ev[x0 + u] : 0;
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const uint32_t st = event_start(P, v[u]);
        keep[u] = x0 + u < oe && (!chg[st] || ev_pos(v[u]) < from[st]);
      }
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (keep[u]) {
          if (kept < lim)
            L.ev[nb + kept] = v[u];
          else
            atomicOr(err, 4);
          kept++;
        }
    }
    cur[y] = kept;
    L.off[y] = nb;
    L.cnt[y] = lim;
    if (move) L.cap[y] = room;
    nold = oe - ob;
    nkept = kept;
  }
  wave_add(work, nold);
  wave_add(work + 1, nkept);
}

// Exclusive prefix of pred over the workgroup (256 threads) and its total;
// every thread must call it.
__device__ __forceinline__ uint32_t block_prefix(bool pred, uint32_t *lds4, uint32_t &total) {
  const uint64_t m = __ballot(pred);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) lds4[w] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t before = 0;
  total = 0;
  for (int i = 0; i < 4; i++) {
    if (i < w) before += lds4[i];
    total += lds4[i];
  }
  __syncthreads();
  return before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

// k_patch_keep for the long lists, one workgroup per list at a time: the
// list is filtered in chunks of 256 events (coalesced reads, one flag gather
// each, a workgroup prefix for the write positions).  In place the writes of
// a chunk land at or below its own positions, which every lane has read
// before the prefix's barrier, and below the next chunk's.
__global__ void __launch_bounds__(256) k_patch_keep_long(BParams P, const uint32_t *__restrict__ longq,
                                                         const uint32_t *nlong, const uint32_t *__restrict__ dl,
                                                         const uint32_t *__restrict__ ncnt,
                                                         const uint32_t *__restrict__ alloc,
                                                         const uint64_t *__restrict__ aoff, uint64_t base, Lists L,
                                                         const uint8_t *__restrict__ chg,
                                                         const uint16_t *__restrict__ from,
                                                         uint32_t *__restrict__ cur, int *err,
                                                         unsigned long long *work) {
  __shared__ uint32_t lds4[4];
  const uint32_t nl = *nlong;
  uint64_t wold = 0, wkept = 0;  // this workgroup's work, added once at the end
  for (uint32_t q = blockIdx.x; q < nl; q += gridDim.x) {
    const uint32_t t = longq[q];
    const uint32_t y = dl[t];
    const uint64_t ob = L.off[y], oe = ob + L.cnt[y];
    const uint32_t nn = ncnt[t];
    const bool move = alloc[t] != 0;
    const uint64_t nb = move ? base + aoff[t] : ob;
    const uint32_t room = move ? alloc[t] : L.cap[y];
    const uint32_t lim = nn <= room ? nn : room;
    uint32_t kept = 0;
    for (uint64_t c = ob; c < oe; c += 256) {
      const uint64_t x = c + threadIdx.x;
      uint64_t v = 0;
      bool keep = false;
      if (x < oe) {
        v = L.ev[x];
        const uint32_t st = event_start(P, v);
        keep = !chg[st] || ev_pos(v) < from[st];
      }
      uint32_t tot;
      const uint32_t pos = kept + block_prefix(keep, lds4, tot);
      if (keep) {
        if (pos < lim)
          L.ev[nb + pos] = v;
        else
          atomicOr(err, 4);
      }
      kept += tot;
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      if (nn > room) atomicOr(err, 4);
      cur[y] = kept;
      L.off[y] = nb;
      L.cnt[y] = lim;
      if (move) L.cap[y] = room;
      wold += oe - ob;
      wkept += kept;
    }
    __syncthreads();
  }
  // one slotted add per workgroup (a shared address per list serialises)
  if (work && threadIdx.x == 0 && wold) {
    unsigned long long *w = work + CSTRIDE * (blockIdx.x & (CSLOTS - 1));
    atomicAdd(w, (unsigned long long)wold);
    atomicAdd(w + 1, (unsigned long long)wkept);
  }
}

// Points whose known==2 time may change: the rewritten points and their face
// neighbours.
__global__ void k_mark_kt(BParams P, uint32_t nd, const uint32_t *__restrict__ dl, uint8_t *__restrict__ qflag) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nd) return;
  const uint32_t y = dl[t];
  int i, j, k;
  coords(P, y, i, j, k);
  qflag[y] = 1;
  qflag[lin3(P, wrap1(i + 1, P.n1), j, k)] = 1;
  qflag[lin3(P, wrap1(i - 1, P.n1), j, k)] = 1;
  qflag[lin3(P, i, wrap1(j + 1, P.n2), k)] = 1;
  qflag[lin3(P, i, wrap1(j - 1, P.n2), k)] = 1;
  qflag[lin3(P, i, j, wrap1(k + 1, P.n3))] = 1;
  qflag[lin3(P, i, j, wrap1(k - 1, P.n3))] = 1;
}

// The step's flags cleared through their lists: dflag over dl, ktchg over
// ql.
__global__ void k_clear_step(uint32_t nd, const uint32_t *__restrict__ dl, uint8_t *__restrict__ dflag, uint32_t nq,
                             const uint32_t *__restrict__ ql, uint8_t *__restrict__ ktchg) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nd) dflag[dl[t]] = 0;
  if (t < nq) ktchg[ql[t]] = 0;
}

// Starts to decide again: those with an event at a point whose list or
// known==2 time changed, and those points themselves.  (A start's decision
// reads volnum at the start and at its path's end and the known==2 marks
// along its committed path, bader@proc.f90:154-211; every such point carries
// one of its events.)
__global__ void __launch_bounds__(256) k_mark_decide(BParams P, uint32_t nq, const uint32_t *__restrict__ ql,
                                                     Lists L, const uint8_t *__restrict__ dflag,
                                                     const uint8_t *__restrict__ ktchg, uint8_t *__restrict__ sflag,
                                                     uint32_t *__restrict__ longq, uint32_t *nlong,
                                                     unsigned long long *work) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t scanned = 0, marked = 0;
  const uint32_t x = t < nq ? ql[t] : 0u;
  const bool act = t < nq && (dflag[x] || ktchg[x]);
  const bool lng = act && L.cnt[x] > LIST_LONG;
  wave_append(lng, x, longq, nlong);  // a workgroup each (k_mark_decide_long)
  if (act) {
    sflag[x] = 1;
    marked = 1;
    if (!lng) {
      // events loaded 8 at a time ahead of their flag stores (which the
      // compiler must otherwise order after each load)
      const uint64_t b = L.b(x), e = L.e(x);
      for (uint64_t q0 = b; q0 < e; q0 += 8) {
        uint64_t v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = q0 + u < e ? L.ev[q0 + u] : 0;
#pragma unroll
        for (int u = 0; u < 8; u++)
          if (q0 + u < e) sflag[event_start(P, v[u])] = 1;
      }
      scanned = e - b;
    }
  }
  wave_add(work, scanned);
  wave_add(work + 1, marked);
}

__global__ void __launch_bounds__(256) k_mark_decide_long(BParams P, const uint32_t *__restrict__ longq,
                                                          const uint32_t *nlong, Lists L,
                                                          uint8_t *__restrict__ sflag, unsigned long long *work) {
  const uint32_t nl = *nlong;
  uint64_t scanned = 0;
  for (uint32_t q = blockIdx.x; q < nl; q += gridDim.x) {
    const uint32_t x = longq[q];
    const uint64_t b = L.b(x), e = L.e(x);
    for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x) sflag[event_start(P, L.ev[i])] = 1;
    scanned += (e - b + blockDim.x - 1 - threadIdx.x) / blockDim.x;
  }
  wave_add(work, scanned);
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

// Counter q of a slotted counter array (wave_add), summed over its slots.
static unsigned long long sum_slots(c2gpu_ctx *ctx, const unsigned long long *d, int q) {
  unsigned long long h[CSLOTS * CSTRIDE];
  copy_to_host(ctx, h, d, sizeof(h));
  C2_HIP(hipStreamSynchronize(ctx->stream));
  unsigned long long t = 0;
  for (int s = 0; s < CSLOTS; s++) t += h[s * CSTRIDE + q];
  return t;
}

constexpr uint64_t HUGE_BATCH = 1ull << 27;

// Sorts the lists of the points in `list` (all points when null).
// d_n: three counters (long, huge, wavefront lists), zero on entry.
// end_bit: the events' time bits end there (the radix sort of the very long
// lists orders the time bits only: a list's times are distinct).
static void sort_lists(c2gpu_ctx *ctx, uint32_t n, const uint32_t *list, const Lists &L, uint8_t *uni,
                       uint32_t *longlist, uint32_t *hugelist, unsigned long long *work, uint32_t *d_n,
                       int end_bit) {
  hipStream_t st = ctx->stream;
  if (n == 0) return;
  C2_LAUNCH(ctx, k_ev_sort, dim3(grid_for(n, 256)), dim3(256), 0, n, list, L, uni, longlist, d_n, d_n + 2,
            hugelist, d_n + 1, work);
  uint32_t nlh[3];
  copy_to_host(ctx, nlh, d_n, 3 * sizeof(uint32_t));
  C2_HIP(hipStreamSynchronize(st));
  if (nlh[2])
    C2_LAUNCH(ctx, k_ev_sort_wave, dim3((nlh[2] + 3) / 4), dim3(256), 0, longlist + (n - nlh[2]), nlh[2], L, uni);
  if (nlh[0]) C2_LAUNCH(ctx, k_ev_sort_long, dim3(nlh[0]), dim3(256), 0, longlist, L, uni);
  if (nlh[1]) {
    // very long lists: gathered in batches of about HUGE_BATCH events (a
    // single longer list is a batch of its own), each batch sorted by one
    // segmented radix sort in a compact buffer and copied back
    const uint32_t nh = nlh[1];
    uint32_t *hlen = buf<uint32_t>(ctx, "sweep_hugelen", (size_t)nh + 1);
    uint64_t *hoff = buf<uint64_t>(ctx, "sweep_hugeoff", (size_t)nh + 1);
    C2_LAUNCH(ctx, k_huge_len, dim3(grid_for((int64_t)nh + 1, 64)), dim3(64), 0, nh, hugelist, L, hlen);
    exclusive_scan(ctx, "huge", hlen, hoff, (int64_t)nh + 1);
    std::vector<uint64_t> ho((size_t)nh + 1);
    copy_to_host(ctx, ho.data(), hoff, 8 * ((size_t)nh + 1));
    C2_HIP(hipStreamSynchronize(st));
    std::vector<uint64_t> bo;  // batch-relative segment offsets
    for (uint32_t q0 = 0; q0 < nh;) {
      uint32_t q1 = q0 + 1;
      while (q1 < nh && ho[q1 + 1] - ho[q0] <= HUGE_BATCH) q1++;
      const uint64_t base = ho[q0], tot = ho[q1] - base;
      if (tot >= (1ull << 31)) fail(C2GPU_ERR_ARG, "c2gpu_bader: sweep event list too long to sort");
      const uint32_t nb = q1 - q0;
      bo.resize((size_t)nb + 1);
      for (uint32_t q = 0; q <= nb; q++) bo[q] = ho[q0 + q] - base;
      uint64_t *boff = buf<uint64_t>(ctx, "sweep_hugeboff", (size_t)nb + 1);
      C2_HIP(hipMemcpyAsync(boff, bo.data(), 8 * ((size_t)nb + 1), hipMemcpyHostToDevice, st));
      uint64_t *keys = buf<uint64_t>(ctx, "sweep_hugekeys", tot), *alt = buf<uint64_t>(ctx, "sweep_hugealt", tot);
      C2_LAUNCH(ctx, k_huge_gather, dim3(nb), dim3(256), 0, hugelist + q0, L, boff, keys);
      // a sorted run and a short tail each: merge (see k_huge_desc)
      uint32_t *kd = buf<uint32_t>(ctx, "sweep_hugekd", nb), *hf = buf<uint32_t>(ctx, "sweep_hugeflag", nb + 1);
      C2_HIP(hipMemsetAsync(hf, 0, 4 * ((size_t)nb + 1), st));
      C2_LAUNCH(ctx, k_huge_desc, dim3(nb), dim3(256), 0, boff, keys, kd, hf);
      if (read_scalar(ctx, hf) == 0) {
        C2_LAUNCH(ctx, k_huge_tail_sort, dim3(nb), dim3(256), 0, boff, keys, kd, alt);
        C2_LAUNCH(ctx, k_huge_merge, dim3(std::min<uint64_t>((tot + 255) / 256, 8192)), dim3(256), 0, nb,
                  hugelist + q0, L, boff, keys, alt, kd, hf + 1);
        C2_LAUNCH(ctx, k_huge_uni, dim3(grid_for(nb, 256)), dim3(256), 0, nb, hugelist + q0, hf + 1, uni);
        C2_HIP(hipStreamSynchronize(st));
        q0 = q1;
        continue;
      }
      size_t tmp = 0;
      C2_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tmp, keys, alt, (int)tot, (int)nb, boff, boff + 1,
                                                        LAB_BITS, end_bit, st));
      void *t = buffer(ctx, "sweep_segsort_tmp", tmp);
      C2_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(t, tmp, keys, alt, (int)tot, (int)nb, boff, boff + 1,
                                                        LAB_BITS, end_bit, st));
      C2_LAUNCH(ctx, k_huge_copyback, dim3(nb), dim3(256), 0, hugelist + q0, L, boff, alt, uni);
      // the host offsets are rewritten by the next batch: wait for this one's copy
      C2_HIP(hipStreamSynchronize(st));
      q0 = q1;
    }
  }
}

struct SweepResult {
  int iters;
  int64_t events_total;
  int64_t nstarts;
  const int32_t *ell;  // the converged story's committed path lengths (device)
};

// Runs the fixpoint; on return V holds volnum after the sweep and known the
// known==2 marks (2 or 0).
SweepResult sweep_run(c2gpu_ctx *ctx, const BParams &P, uint32_t N, const uint64_t *toff, const uint32_t *tp,
                      const int32_t *maxlab, int32_t *V, int8_t *known) {
  hipStream_t st = ctx->stream;
  const int B = 256;
  const unsigned gN = grid_for(N, B);
  const unsigned gW = grid_for(((int64_t)N + 15) / 16, B);
  const unsigned gCollect = gW < 2048 ? gW : 2048;
  const bool full_every_step = std::getenv("C2GPU_SWEEP_FULL") != nullptr;
  // while a step changes more than this many decisions (a quarter of the
  // points), the lists are rebuilt from the story instead of patched: an
  // early step rewrites most of the story, and a full build writes each
  // event once where the patch path moves the lists that overflow their
  // blocks (at 512^3: 3.14 s against 3.40 s for patching every step)
  uint64_t full_above = N / 4;
  if (const char *fa = std::getenv("C2GPU_SWEEP_FULL_ABOVE")) full_above = std::strtoull(fa, nullptr, 10);
  const bool trace = std::getenv("C2GPU_SWEEP_TRACE") != nullptr;
  SweepResult R{0, 0, 0, nullptr};
  const size_t NF = ((size_t)N + 16) & ~(size_t)15;  // byte flags padded to whole 16-byte groups
  int32_t *ell = buf<int32_t>(ctx, "sweep_ell", N), *lam = buf<int32_t>(ctx, "sweep_lam", N);
  int32_t *ell2 = buf<int32_t>(ctx, "sweep_ell2", N), *lam2 = buf<int32_t>(ctx, "sweep_lam2", N);
  uint32_t *cntN = buf<uint32_t>(ctx, "sweep_cnt", (size_t)N + 1);
  uint32_t *capN = buf<uint32_t>(ctx, "sweep_cap", N);
  uint32_t *capv = buf<uint32_t>(ctx, "sweep_capv", (size_t)N + 1);
  uint64_t *offN = buf<uint64_t>(ctx, "sweep_off", (size_t)N + 1);
  uint32_t *cur = buf<uint32_t>(ctx, "sweep_cur", N);
  int32_t *delta = buf<int32_t>(ctx, "sweep_delta", N);
  uint16_t *cfrom = buf<uint16_t>(ctx, "sweep_from", N);  // first changed position of each changed start
  uint64_t *Kt = buf<uint64_t>(ctx, "sweep_kt", N);
  uint8_t *uni = buf<uint8_t>(ctx, "sweep_uni", N);
  uint8_t *chg = buf<uint8_t>(ctx, "sweep_chg", NF);
  uint8_t *ktchg = buf<uint8_t>(ctx, "sweep_ktchg", NF);
  uint8_t *dflag = buf<uint8_t>(ctx, "sweep_dflag", NF);  // list rewritten this step
  uint8_t *qflag = buf<uint8_t>(ctx, "sweep_qflag", NF);  // known==2 time to recompute
  uint8_t *sflag = buf<uint8_t>(ctx, "sweep_sflag", NF);  // start to decide again
  uint32_t *dl = buf<uint32_t>(ctx, "sweep_dl", N), *ql = buf<uint32_t>(ctx, "sweep_ql", N);
  uint32_t *sl = buf<uint32_t>(ctx, "sweep_sl", N);
  uint32_t *longlist = buf<uint32_t>(ctx, "sweep_long", N), *hugelist = buf<uint32_t>(ctx, "sweep_huge", N);
  uint32_t *chlist = buf<uint32_t>(ctx, "sweep_chlist", N);
  uint32_t *ncnt = buf<uint32_t>(ctx, "sweep_ncnt", N);
  uint32_t *alloc = buf<uint32_t>(ctx, "sweep_alloc", (size_t)N + 1);
  uint64_t *aoff = buf<uint64_t>(ctx, "sweep_aoff", (size_t)N + 1);
  // the step's counters: nch, ndl, nql, nsl, long patch / mark / known==2
  // lists, sorted long / huge / wavefront lists (all zeroed by one memset
  // per step), then the error bits, kept over the whole sweep
  constexpr int NCOUNT = 10, CERR = 11;
  uint32_t *d_n = buf<uint32_t>(ctx, "sweep_counts", 12);
  uint32_t *longq = buf<uint32_t>(ctx, "sweep_longq", N);
  const unsigned gLong = (unsigned)persistent_threads(ctx->device, 256) / 256 / 2;  // 4 workgroups per CU
  int *d_err = reinterpret_cast<int *>(d_n + CERR);
  unsigned long long *d_stats = buf<unsigned long long>(ctx, "sweep_stats", CSLOTS * CSTRIDE);
  unsigned long long *d_work = buf<unsigned long long>(ctx, "sweep_work", CSLOTS * CSTRIDE);  // algorithmic work counters
  C2_HIP(hipMemsetAsync(d_n, 0, 12 * sizeof(uint32_t), st));
  C2_HIP(hipMemsetAsync(d_work, 0, CSLOTS * CSTRIDE * sizeof(unsigned long long), st));
  // host-side totals for the algorithmic-bytes report
  double s_nc = 0, s_nd = 0, s_nq = 0, s_ns = 0, s_coll = 0, n_coll = 0, n_full = 0, n_decall = 0, s_nev = 0;
  C2_HIP(hipMemsetAsync(delta, 0, sizeof(int32_t) * N, st));
  for (uint8_t *f : {chg, ktchg, dflag, qflag, sflag}) C2_HIP(hipMemsetAsync(f, 0, NF, st));

  // events carry their time (sweep index, then position) above the label
  int sort_end_bit = LAB_BITS + POS_BITS;
  while (sort_end_bit < 64 && (1ull << (sort_end_bit - LAB_BITS - POS_BITS)) < (uint64_t)N) sort_end_bit++;
  Lists L;
  L.off = offN;
  L.cnt = cntN;
  L.cap = capN;
  L.ev = nullptr;
  uint64_t pool_cap = 0, pool_top = 0;

  // d_cnt is zero on entry (the step's memset)
  auto collect = [&](uint8_t *flag, int clear, uint32_t *list, uint32_t *d_cnt) -> uint32_t {
    C2_LAUNCH(ctx, k_collect, dim3(gCollect), dim3(B), 0, N, flag, clear, list, d_cnt);
    const uint32_t c = read_scalar(ctx, d_cnt);
    s_coll += c;
    n_coll += 1;
    return c;
  };

  // known==2 times of the points in list (every point when null), the long
  // merges as work items.  The plan's buffers (96 bytes a point) are sized
  // by the deferred count read back, or, within a step (list != null,
  // counter zeroed by the step's memset) of at most 1 M points, by n with
  // no read-back (the plan reads the count on the device).
  auto kentry = [&](uint32_t n, const uint32_t *list, uint8_t *flags) {
    uint32_t *nlp = d_n + 6;
    if (!list) C2_HIP(hipMemsetAsync(nlp, 0, sizeof(uint32_t), st));
    C2_LAUNCH(ctx, k_kentry, dim3(grid_for(n, B)), dim3(B), 0, P, n, list, L, uni, Kt, flags, longq, nlp);
    const uint32_t nb = list && n <= (1u << 20) ? n : read_scalar(ctx, nlp);
    if (nb == 0) return;
    uint64_t *lbw = buf<uint64_t>(ctx, "kt_lb", 6 * (size_t)nb);
    uint32_t *cpre = buf<uint32_t>(ctx, "kt_cpre", 7 * (size_t)nb);
    uint32_t *kcnt = buf<uint32_t>(ctx, "kt_cnt", (size_t)nb + 1);
    uint64_t *koff = buf<uint64_t>(ctx, "kt_off", (size_t)nb + 1);
    uint64_t *ktmp = buf<uint64_t>(ctx, "kt_tmp", nb);
    C2_LAUNCH(ctx, k_kt_plan, dim3(grid_for((int64_t)nb + 1, B)), dim3(B), 0, P, (const uint32_t *)nlp, nb, longq,
              L, lbw, cpre, kcnt, ktmp);
    exclusive_scan(ctx, "kt", kcnt, koff, (int64_t)nb + 1);
    C2_LAUNCH(ctx, k_kt_items, dim3(2 * gLong), dim3(B), 0, P, nb, longq, L, lbw, cpre, koff, ktmp);
    C2_LAUNCH(ctx, k_kt_commit, dim3(grid_for(nb, B)), dim3(B), 0, (const uint32_t *)nlp, nb, longq, ktmp, Kt,
              flags);
  };

  // every list from the story: counts, reserved blocks, events, order,
  // known==2 times
  auto full_build = [&]() {
    C2_HIP(hipMemsetAsync(cntN, 0, sizeof(uint32_t) * ((size_t)N + 1), st));
    C2_LAUNCH(ctx, k_ev_count, dim3(gN), dim3(B), 0, N, toff, tp, ell, cntN);
    const uint64_t nev = sum_u32(ctx, "ev", cntN, N);
    C2_LAUNCH(ctx, k_reserve, dim3(grid_for((int64_t)N + 1, B)), dim3(B), 0, N, cntN, capN, capv);
    exclusive_scan(ctx, "ev", capv, offN, (int64_t)N + 1);
    const uint64_t reserved = read_scalar(ctx, offN + N);
    // room for moved lists: a quarter more, or what the device has left
    // (a fuller pool only rebuilds more often)
    const uint64_t want = reserved + (reserved >> 2) + (uint64_t)N;
    if (want > pool_cap) {
      const uint64_t least = reserved + (reserved >> 5) + (uint64_t)N / 8;
      // the pool is the arena that held the walk (no longer needed)
      L.ev = static_cast<uint64_t *>(buffer_within(ctx, "bader_arena", 8 * least, 8 * want, 6ull << 30));
      pool_cap = ctx->bufs["bader_arena"].bytes / 8;
    }
    pool_top = reserved;
    C2_HIP(hipMemsetAsync(cur, 0, sizeof(uint32_t) * N, st));
    C2_LAUNCH(ctx, k_ev_fill, dim3(gN), dim3(B), 0, P, N, (const uint32_t *)nullptr, toff, tp, ell, lam,
              (const uint16_t *)nullptr, L, cur, d_err, d_work + 4, (int32_t *)nullptr, (int32_t *)nullptr,
              (uint8_t *)nullptr);
    C2_HIP(hipMemsetAsync(d_n + 7, 0, 3 * sizeof(uint32_t), st));
    sort_lists(ctx, N, nullptr, L, uni, longlist, hugelist, d_work + 8, d_n + 7, sort_end_bit);
    kentry(N, nullptr, nullptr);
    R.events_total += (int64_t)nev;
    n_full += 1;
    s_nev += (double)nev;
  };
  auto decide_all = [&]() -> uint32_t {
    C2_HIP(hipMemsetAsync(d_n, 0, sizeof(uint32_t), st));
    C2_LAUNCH(ctx, k_decide, dim3(gN), dim3(B), 0, P, N, (const uint32_t *)nullptr, toff, tp, L, Kt, maxlab, ell,
              lam, ell2, lam2, chg, chlist, d_n, d_err, (unsigned long long *)nullptr, d_work + 7);
    n_decall += 1;
    return read_scalar(ctx, d_n);
  };

  // rounds of the approximate first story: the bulk of what the early
  // fixpoint steps would correct, at a fraction of their cost (512^3: the
  // fixpoint from 2.64 s to 1.03 s with 32 rounds, 73 steps instead of 104;
  // 768^3: 10.4 s to 3.7 s).  32 is best at 512^3 (20: +11%, 40: +3%) and
  // on cells of 48-160 points an edge too, where a round costs far less
  // than a fixpoint step (8-10 rounds: +27%, 24: +2%, 48: +3%)
  int pre_rounds = 32;
  if (const char *pr = std::getenv("C2GPU_SWEEP_PRE")) pre_rounds = std::atoi(pr);
  // the rounds follow each start's path for at most 32 points: longer ones
  // (rare, and mostly in the early rounds when the known == 2 times are
  // late) cost more than their coverage helps (512^3: 925 ms against 950 ms
  // uncapped; 8 points: 1289 ms)
  const int pre_cap = 32;
  if (pre_rounds > 0) {
    // buffers the first full build overwrites: flab in cur, the rounds'
    // (known == 2 time, covering time) words in offN / aoff
    int32_t *flab = reinterpret_cast<int32_t *>(cur);
    uint64_t *Wa = offN, *Wb = aoff;
    C2_LAUNCH(ctx, k_pre_flab, dim3(gN), dim3(B), 0, N, toff, tp, maxlab, flab, Wb, P);
    // start from every point's first step covered (512^3: 900 ms of
    // fixpoint and rounds against 923 ms from own times only)
    C2_LAUNCH(ctx, k_pre_seed, dim3(gN), dim3(B), 0, P, N, toff, tp, Wb);
    uint8_t *agree = uni;  // rewritten by the first full build
    C2_LAUNCH(ctx, k_pre_agree, dim3(gN), dim3(B), 0, P, N, flab, agree);
    // Wb: the covering times a round starts from; Wa: Kt and the covering
    // times it builds
    C2_LAUNCH(ctx, k_pre_kt, dim3(gN), dim3(B), 0, P, N, (const uint64_t *)Wb, agree, Wa,
              (unsigned long long *)nullptr);
    unsigned long long *d_pc = buf<unsigned long long>(ctx, "sweep_precount", CSLOTS * CSTRIDE);
    const bool count = trace;  // covering times changed per round (diagnostics)
    for (int r = 1; r < pre_rounds; r++) {
      if (count) C2_HIP(hipMemsetAsync(d_pc, 0, CSLOTS * CSTRIDE * sizeof(unsigned long long), st));
      // the first rounds, whose known == 2 times are still late, follow 8
      // points only (512^3: 773 ms of rounds and fixpoint against 813 ms)
      const int cap_r = r < 8 ? 8 : pre_cap;
      C2_LAUNCH(ctx, k_pre_cover, dim3(gN), dim3(B), 0, P, N, toff, tp, (const uint64_t *)Wb, Wa, cap_r);
      // Wa now holds this round's covering times; Wb gets the next round's
      // Kt and starting covering times
      C2_LAUNCH(ctx, k_pre_kt, dim3(gN), dim3(B), 0, P, N, (const uint64_t *)Wa, agree, Wb,
                count ? d_pc : (unsigned long long *)nullptr);
      std::swap(Wa, Wb);
      if (count)
        std::fprintf(stderr, "[sweep] pre round %d: %llu covering times changed\n", r,
                     (unsigned long long)sum_slots(ctx, d_pc, 0));
    }
    C2_LAUNCH(ctx, k_pre_story, dim3(gN), dim3(B), 0, P, N, toff, tp, (const uint64_t *)Wb, (const uint64_t *)Wa,
              flab, ell, lam);
  } else {
    C2_LAUNCH(ctx, k_story_init, dim3(gN), dim3(B), 0, N, toff, tp, maxlab, ell, lam);
  }
  full_build();
  uint32_t nc = decide_all();
  int it = 1;
  const int max_iters = 20000;
  auto t_last = std::chrono::steady_clock::now();  // per-step wall time in the trace
  while (nc > 0) {
    if (it >= max_iters) fail(C2GPU_ERR_NOCONVERGE, "c2gpu_bader: first-sweep fixpoint did not converge");
    it++;
    const unsigned gC = grid_for(nc, B);
    if ((uint64_t)nc > full_above) {
      // a large step: take the new decisions and rebuild every list
      C2_LAUNCH(ctx, k_apply, dim3(gC), dim3(B), 0, nc, chlist, ell, lam, ell2, lam2, chg);
      full_build();
      nc = decide_all();
      if (trace) {
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[sweep] it %d %.3f ms\n", it,
                     std::chrono::duration<double, std::milli>(now - t_last).count());
        t_last = now;
        std::fprintf(stderr, "[sweep] it %d full build, changed %u\n", it, nc);
      }
      continue;
    }
    // 1. points whose lists change, and their new sizes
    C2_HIP(hipMemsetAsync(d_n, 0, NCOUNT * sizeof(uint32_t), st));
    if (trace) C2_HIP(hipMemsetAsync(d_stats, 0, CSLOTS * CSTRIDE * sizeof(unsigned long long), st));
    C2_LAUNCH(ctx, k_delta, dim3(gC), dim3(B), 0, nc, chlist, toff, tp, ell, ell2, lam, lam2, delta, dflag, cfrom,
              trace ? d_stats : nullptr, d_work + 5);
    s_nc += nc;
    const uint32_t nd = collect(dflag, 0, dl, d_n + 1);
    const unsigned gD = grid_for(nd, B);
    C2_LAUNCH(ctx, k_patch_count, dim3(grid_for((int64_t)nd + 1, B)), dim3(B), 0, nd, dl, cntN, capN, delta, ncnt,
              alloc, d_err);
    exclusive_scan(ctx, "patch", alloc, aoff, (int64_t)nd + 1);
    const uint64_t added = read_scalar(ctx, aoff + nd);
    if (full_every_step || pool_top + added > pool_cap) {
      // rebuild every list from the updated story (compacts the pool)
      C2_LAUNCH(ctx, k_clear_list<uint8_t>, dim3(gD), dim3(B), 0, nd, dl, dflag);
      C2_LAUNCH(ctx, k_apply, dim3(gC), dim3(B), 0, nc, chlist, ell, lam, ell2, lam2, chg);
      full_build();
      nc = decide_all();
      if (trace) std::fprintf(stderr, "[sweep] it %d full rebuild, changed %u\n", it, nc);
      continue;
    }
    // 2. rewrite those lists; add the changed starts' events (and commit
    // the new decisions, k_apply's work)
    C2_LAUNCH(ctx, k_patch_keep, dim3(gD), dim3(B), 0, P, nd, dl, ncnt, alloc, aoff, pool_top, L, chg, cfrom, cur,
              longq, d_n + 4, d_err, d_work + 0);
    C2_LAUNCH(ctx, k_patch_keep_long, dim3(gLong), dim3(B), 0, P, longq, d_n + 4, dl, ncnt, alloc, aoff, pool_top,
              L, chg, cfrom, cur, d_err, d_work + 0);
    C2_LAUNCH(ctx, k_ev_fill, dim3(gC), dim3(B), 0, P, nc, chlist, toff, tp, ell2, lam2, cfrom, L, cur, d_err,
              d_work + 4, ell, lam, chg);
    s_nd += nd;
    pool_top += added;
    sort_lists(ctx, nd, dl, L, uni, longlist, hugelist, d_work + 8, d_n + 7, sort_end_bit);
    // 3. known==2 times around the rewritten lists
    C2_LAUNCH(ctx, k_mark_kt, dim3(gD), dim3(B), 0, P, nd, dl, qflag);
    const uint32_t nq = collect(qflag, 1, ql, d_n + 2);
    const unsigned gQ = grid_for(nq, B);
    if (nq) kentry(nq, ql, ktchg);
    // 4. starts whose inputs changed decide again
    C2_LAUNCH(ctx, k_mark_decide, dim3(gQ), dim3(B), 0, P, nq, ql, L, dflag, ktchg, sflag, longq, d_n + 5,
              d_work + 2);
    C2_LAUNCH(ctx, k_mark_decide_long, dim3(gLong), dim3(B), 0, P, longq, d_n + 5, L, sflag, d_work + 2);
    s_nq += nq;
    const uint32_t ns = collect(sflag, 1, sl, d_n + 3);
    if (trace) C2_HIP(hipMemsetAsync(d_stats + 4, 0xff, sizeof(unsigned long long), st));
    if (ns)
      C2_LAUNCH(ctx, k_decide, dim3(grid_for(ns, B)), dim3(B), 0, P, ns, sl, toff, tp, L, Kt, maxlab, ell, lam,
                ell2, lam2, chg, chlist, d_n, d_err, trace ? d_stats + 4 : nullptr, d_work + 7);
    s_ns += ns;
    // 5. clear the step's flags through their lists; read the changes and
    // the error bits together
    C2_LAUNCH(ctx, k_clear_step, dim3(grid_for(nd > nq ? nd : nq, B)), dim3(B), 0, nd, dl, dflag, nq, ql, ktchg);
    uint32_t hc[CERR + 1];
    copy_to_host(ctx, hc, d_n, sizeof(hc));
    nc = hc[0];
    if ((int)hc[CERR] & 4) fail(C2GPU_ERR_ARG, "c2gpu_bader: internal inconsistency in the sweep event lists");
    R.events_total += (int64_t)added;
    if (trace) {
      unsigned long long h[5];
      for (int q = 0; q < 4; q++) h[q] = sum_slots(ctx, d_stats, q);
      h[4] = read_scalar(ctx, d_stats + 4);  // slot 0 only: an atomic minimum
      const auto now = std::chrono::steady_clock::now();
      std::fprintf(stderr, "[sweep] it %d %.3f ms\n", it,
                   std::chrono::duration<double, std::milli>(now - t_last).count());
      t_last = now;
      std::fprintf(stderr,
                   "[sweep] it %d changed %u rewritten %u kt %u decided %u moved %llu pool %llu/%llu | "
                   "same-label %llu kept-prefix %llu touched %llu relabel %llu first-change %lld\n",
                   it, nc, nd, nq, ns, (unsigned long long)added, (unsigned long long)pool_top,
                   (unsigned long long)pool_cap, h[0], h[1], h[2], h[3], (long long)h[4]);
    }
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
  if (ctx->timing) {
    // algorithmic bytes per kernel: each element a launch must read or write
    // once (read-modify-write twice), from the host-side launch sizes and the
    // device work counters
    unsigned long long w[16];
    for (int q = 0; q < 16; q++) w[q] = sum_slots(ctx, d_work, q);
    const double dN = (double)N;
    add_kernel_bytes(ctx, "k_patch_keep", 56.0 * s_nd + 9.0 * w[0] + 8.0 * w[1]);
    add_kernel_bytes(ctx, "k_mark_decide", 6.0 * s_nq + 13.0 * w[3] + 9.0 * w[2]);
    add_kernel_bytes(ctx, "k_ev_fill", 20.0 * (s_nc + dN * n_full) + 32.0 * w[4]);
    add_kernel_bytes(ctx, "k_delta", 28.0 * s_nc + 5.0 * w[5] + 8.0 * w[6]);
    add_kernel_bytes(ctx, "k_decide", 77.0 * (s_ns + dN * n_decall) + 12.0 * w[7]);
    add_kernel_bytes(ctx, "k_kentry", 168.0 * (s_nq + dN * n_full));
    add_kernel_bytes(ctx, "k_collect", n_coll * (double)NF + 5.0 * s_coll);
    add_kernel_bytes(ctx, "k_ev_sort", 17.0 * (s_nd + dN * n_full) + 16.0 * w[8]);
    add_kernel_bytes(ctx, "k_ev_count", n_full * 12.0 * dN + 12.0 * s_nev);
  }
  unsigned long long *d_ns = buf<unsigned long long>(ctx, "sweep_nstarts", CSLOTS * CSTRIDE);
  C2_HIP(hipMemsetAsync(d_ns, 0, CSLOTS * CSTRIDE * sizeof(unsigned long long), st));
  C2_LAUNCH(ctx, k_final_state, dim3(gN), dim3(B), 0, N, L, Kt, V, known, d_ns, ell);
  R.nstarts = (int64_t)sum_slots(ctx, d_ns, 0);
  R.iters = it;
  R.ell = ell;
  return R;
}

}  // namespace c2
