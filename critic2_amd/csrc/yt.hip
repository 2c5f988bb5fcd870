// Yu-Trinkle basins and weights on the GPU (yt_integrate, yt@proc.f90:38-224;
// yt_weights, :399-531; the imtype_yt sums of intgrid_fields,
// integration@proc.f90:1208-1218, :1289-1300).
//
// The reference ranks the points by rho and runs down the ranking: a point
// with no higher neighbour (over the Voronoi-relevant neighbours of the grid
// lattice) is a maximum; a point whose higher neighbours share one basin
// joins it; the others lie on an interatomic surface (basin 0) and pass on
// weight to the basins of their higher neighbours, fractions chi_k / sum chi
// with chi_k = area_k * (rho_j - rho_i).  None of this needs the ranking
// itself, only the comparisons between neighbours; where equal values meet,
// the comparisons take the ranks of the reference's quicksort instead
// (qcksort.cpp, on the host):
//   k_yt_class    higher-neighbour mask per point, maxima listed;
//   host          maxima in decreasing rho, attractor rules (as for Bader);
//   k_yt_lab_step per point the least and greatest basin reachable upwards,
//                 computed once, in topological order: a point is done when
//                 its last higher neighbour is (a count of pending higher
//                 neighbours per point, frontier after frontier from the
//                 maxima); equal means the point's basin (a point joins a
//                 basin exactly when every upward path ends in it),
//                 different means a surface point;
//   k_yt_w_narrow / k_yt_w_wide_step
//                 the surface points' weights, a sparse vector of basins
//                 each (16 lanes per point up to 32 basins, a workgroup
//                 beyond), in the same topological order over the surface
//                 points; each basin's weight sums the contributions in
//                 decreasing rho of the higher neighbour, the order in which
//                 yt_weights adds them, so the weights are the reference's
//                 bit for bit;
//   sums          basin interiors through the per-basin sums kernel, surface
//                 points through k_yt_surface_sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "devutil.hpp"

namespace c2 {

int attractor_rule(const c2gpu_attractor_opts *o, uint32_t nmax, const uint32_t *order, const double *dv,
                   int32_t *lab, std::vector<double> &xattr, std::vector<int8_t> &discarded);
int ws_facets(const double m[9], int vec[3 * 14], double area[14]);
void qcksort_order(const double *arr, uint32_t *iord, int64_t first, int64_t last);
void basin_sums_run(c2gpu_ctx *, const int32_t *, int64_t, int, int, const double *const *, int, double *,
                    int64_t *);

namespace {

constexpr int YB = 256;
constexpr int KW = 32;  // basins a surface point keeps in registers; wider points use a workgroup
constexpr uint32_t LO_NONE = 0xffffffffu;
constexpr double VSMALL = 1e-40;  // param.F90 vsmall

struct YtGeo {
  int n1, n2, n3, nvec;
  int di[14], dj[14], dk[14];
  double al[14];
  // the reference sort's rank of every point (qcksort_order), set when the
  // grid has equal values where the order matters; else null and values
  // decide (they are distinct wherever compared)
  const uint32_t *rk;
};

// Whether point a is ranked above point b (yt@proc.f90:117: jj > ii).
__device__ __forceinline__ bool yt_above(const YtGeo &G, const double *__restrict__ g, uint32_t a, uint32_t b) {
  return G.rk ? G.rk[a] > G.rk[b] : g[a] > g[b];
}
// A key that orders points as their ranks do.
__device__ __forceinline__ double yt_key(const YtGeo &G, const double *__restrict__ g, uint32_t a) {
  return G.rk ? (double)G.rk[a] : g[a];
}

__device__ __forceinline__ int wrapn(int x, int n) {
  x %= n;
  return x < 0 ? x + n : x;
}

__device__ __forceinline__ uint32_t nbr(const YtGeo &G, int i, int j, int k, int q) {
  return (uint32_t)wrapn(i + G.di[q], G.n1) +
         (uint32_t)G.n1 * ((uint32_t)wrapn(j + G.dj[q], G.n2) + (uint32_t)G.n2 * (uint32_t)wrapn(k + G.dk[q], G.n3));
}

__device__ __forceinline__ void coords3(const YtGeo &G, uint32_t p, int &i, int &j, int &k) {
  i = (int)(p % (uint32_t)G.n1);
  const uint32_t r = p / (uint32_t)G.n1;
  j = (int)(r % (uint32_t)G.n2);
  k = (int)(r / (uint32_t)G.n2);
}

__device__ __forceinline__ void yt_append(bool pred, uint32_t val, uint32_t *list, uint32_t *n) {
  const uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(n, (uint32_t)__popcll(mask));
  base = __shfl(base, leader, 64);
  if (pred) list[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = val;
}

__global__ void __launch_bounds__(YB) k_yt_class(YtGeo G, uint32_t N, const double *__restrict__ g,
                                                  uint16_t *__restrict__ hmask, uint32_t *__restrict__ pend,
                                                  uint32_t *__restrict__ maxl, uint32_t *nmax, int *err) {
  const uint32_t p = blockIdx.x * YB + threadIdx.x;
  uint32_t mask = 0;
  if (p < N) {
    int i, j, k;
    coords3(G, p, i, j, k);
    const double gp = g[p];
    bool tie = false;
    double hv[14];
    int nh = 0;
    for (int q = 0; q < G.nvec; q++) {
      const uint32_t jn = nbr(G, i, j, k, q);
      if (yt_above(G, g, jn, p)) {
        mask |= 1u << q;
        hv[nh++] = g[jn];
      }
      if (!G.rk && g[jn] == gp) tie = true;
    }
    // without ranks, equal values among the higher neighbours would also be
    // ordered by the reference sort (the order their weights add in)
    if (!G.rk)
      for (int a = 1; a < nh; a++)
        for (int b = 0; b < a; b++) tie = tie || hv[a] == hv[b];
    hmask[p] = (uint16_t)mask;
    pend[p] = (uint32_t)__popc(mask);
    if (tie) atomicOr(err, 1);
  }
  yt_append(p < N && mask == 0, p, maxl, nmax);
}

__global__ void k_yt_gather(uint32_t n, const uint32_t *__restrict__ idx, const double *__restrict__ g,
                            double *__restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = g[idx[t]];
}

__global__ void k_yt_init(uint32_t N, uint64_t *__restrict__ lohi) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N) lohi[p] = (uint64_t)LO_NONE << 32;
}

__global__ void k_yt_seed(uint32_t n, const uint32_t *__restrict__ maxl, const int32_t *__restrict__ lab,
                          uint64_t *__restrict__ lohi) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) lohi[maxl[t]] = ((uint64_t)(uint32_t)lab[t] << 32) | (uint32_t)lab[t];
}

__device__ __forceinline__ uint64_t yt_lohi_of(const YtGeo &G, uint32_t p, const uint16_t *__restrict__ hmask,
                                                const uint64_t *lohi) {
  int i, j, k;
  coords3(G, p, i, j, k);
  uint32_t mask = hmask[p];
  uint32_t lo = LO_NONE, hi = 0;
  while (mask) {
    const int q = __ffs(mask) - 1;
    mask &= mask - 1;
    const uint64_t v = lohi[nbr(G, i, j, k, q)];
    const uint32_t l = (uint32_t)(v >> 32), h = (uint32_t)v;
    lo = l < lo ? l : lo;
    hi = h > hi ? h : hi;
  }
  return ((uint64_t)lo << 32) | hi;
}

// One frontier of the labels: every point of `front` is final; each lower
// neighbour loses one pending higher neighbour, and the thread that takes
// the last computes it and puts it on the next frontier.
// (The frontier's size is read from nfp, written by the step before, so the
// host need not read it back between frontiers; a grid-stride loop covers
// it, and an empty frontier makes the launch a no-op.)
__global__ void __launch_bounds__(YB) k_yt_lab_step(YtGeo G, const uint32_t *nfp, const uint32_t *__restrict__ front,
                                                     const double *__restrict__ g,
                                                     const uint16_t *__restrict__ hmask, uint32_t *__restrict__ pend,
                                                     uint64_t *lohi, uint32_t *__restrict__ next, uint32_t *nnext) {
  const uint32_t nf = *nfp;
  for (uint32_t b0 = blockIdx.x * YB; b0 < nf; b0 += gridDim.x * YB) {  // uniform per workgroup
    const uint32_t t = b0 + threadIdx.x;
    uint32_t done[14];
    int nd = 0;
    if (t < nf) {
      const uint32_t p = front[t];
      int i, j, k;
      coords3(G, p, i, j, k);
      for (int q = 0; q < G.nvec; q++) {
        const uint32_t jn = nbr(G, i, j, k, q);
        if (yt_above(G, g, p, jn) && atomicSub(&pend[jn], 1u) == 1u) {
          lohi[jn] = yt_lohi_of(G, jn, hmask, lohi);
          done[nd++] = jn;
        }
      }
    }
    for (int q = 0; q < 14; q++) yt_append(q < nd, q < nd ? done[q] : 0u, next, nnext);
  }
}

__global__ void k_yt_label(uint32_t N, const uint64_t *__restrict__ lohi, int32_t *__restrict__ lab,
                           uint8_t *__restrict__ ias) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const uint64_t v = lohi[p];
  const uint32_t lo = (uint32_t)(v >> 32), hi = (uint32_t)v;
  const bool one = lo == hi;
  lab[p] = one ? (int32_t)lo : 0;
  ias[p] = one ? 0 : 1;
}

__global__ void k_yt_slots(uint32_t n, const uint32_t *__restrict__ ias, int32_t *__restrict__ slot) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) slot[ias[t]] = (int32_t)t;
}

// Sparse weight vectors of the surface points: slot t holds c[t] entries
// (basin b, weight w) from off[t] in a pool filled by a bump allocator.
struct WBuf {
  uint32_t *b;
  double *w;
  uint64_t *off;
  uint16_t *c;
  unsigned long long *top;  // entries handed out
  uint64_t cap;             // pool entries
};

// The higher neighbours of p in decreasing rho (hq) with their fractions
// chi_k / max(sum chi, vsmall), chi_k = max(area_k (rho_j - rho_p), vsmall)
// summed in facet order (yt@proc.f90:108-119, :175).
__device__ __forceinline__ int yt_fracs(const YtGeo &G, uint32_t p, const double *__restrict__ g,
                                        const uint16_t *__restrict__ hmask, uint32_t hq[14], double fr[14]) {
  int i, j, k;
  coords3(G, p, i, j, k);
  const double gp = g[p];
  uint32_t q0[14];
  double hg[14], chi[14];
  int nh = 0;
  double csum = 0.0;
  const uint32_t mask = hmask[p];
  for (int q = 0; q < G.nvec; q++) {
    if (!(mask & (1u << q))) continue;
    const uint32_t jn = nbr(G, i, j, k, q);
    const double gj = g[jn];
    const double c = fmax(G.al[q] * (gj - gp), VSMALL);
    q0[nh] = jn;
    hg[nh] = yt_key(G, g, jn);
    chi[nh] = c;
    csum = csum + c;
    nh++;
  }
  const double den = fmax(csum, VSMALL);
  int ord[14];
  for (int a = 0; a < nh; a++) {
    int b = a;
    while (b > 0 && hg[ord[b - 1]] < hg[a]) {
      ord[b] = ord[b - 1];
      b--;
    }
    ord[b] = a;
  }
  for (int a = 0; a < nh; a++) {
    hq[a] = q0[ord[a]];
    fr[a] = chi[ord[a]] / den;
  }
  return nh;
}

// The frontier's weights, GL lanes per surface point p.  Lane a takes p's
// a-th higher neighbour in facet order: its fraction chi_a / max(sum chi,
// vsmall) (the sum formed in facet order, as yt_fracs) and its rank in
// decreasing rho (stable, as yt_fracs orders them).  The neighbours' weight
// lists (one basin for an interior neighbour, else its sorted entries) are
// merged basin by basin in increasing order; each basin's weight adds the
// neighbours' contributions in rank order, a neighbour without the basin
// adding +0.0, which leaves the sum's bits as yt_weights forms them.  Entry
// e lives in lane e % GL until written; points where more than KW basins
// meet go to the wide list.  No per-lane arrays: nothing spills to scratch.
constexpr int GL = 16;
__device__ __forceinline__ void yt_narrow_point(const YtGeo &G, uint32_t p, int r, const double *__restrict__ g,
                                                const uint16_t *__restrict__ hmask, const int32_t *__restrict__ lab,
                                                const int32_t *__restrict__ slot, const WBuf &W,
                                                uint32_t *__restrict__ wide, uint32_t *nwide,
                                                unsigned long long *widetot, int *err) {
  int i, j, k;
  coords3(G, p, i, j, k);
  const double gp = g[p];
  const uint32_t mask = hmask[p];
  const int nh = __popc(mask);
  // this lane's neighbour: the r-th set bit of the mask
  int q = -1;
  for (int u = 0, c = 0; u < G.nvec; u++)
    if (mask & (1u << u)) {
      if (c == r) q = u;
      c++;
    }
  uint32_t jn = 0;
  double gj = 0.0, chi = 0.0, kj = 0.0;
  if (q >= 0) {
    jn = nbr(G, i, j, k, q);
    gj = g[jn];
    kj = yt_key(G, g, jn);
    chi = fmax(G.al[q] * (gj - gp), VSMALL);
  }
  double csum = 0.0;
  for (int u = 0; u < nh; u++) csum = csum + __shfl(chi, u, GL);
  const double fr = chi / fmax(csum, VSMALL);
  int rank = 0;
  for (int u = 0; u < nh; u++) {
    const double ku = __shfl(kj, u, GL);
    rank += (ku > kj || (ku == kj && u < r)) ? 1 : 0;
  }
  // the lane of each rank, four bits each
  uint64_t src = 0;
  for (int kk = 0; kk < nh; kk++) {
    int who = (q >= 0 && rank == kk) ? r : 0;
    for (int o = GL / 2; o > 0; o >>= 1) who = max(who, __shfl_xor(who, o, GL));
    src |= (uint64_t)who << (4 * kk);
  }
  // this lane's list: one basin (interior neighbour) or its weight entries
  const int32_t lj = q >= 0 ? lab[jn] : 0;
  int ce = 0;
  uint64_t so = 0;
  if (q >= 0 && lj == 0) {
    const int32_t s = slot[jn];
    ce = W.c[s];
    so = W.off[s];
  }
  const int len = q < 0 ? 0 : (lj != 0 ? 1 : ce);
  int head = 0;
  uint32_t hb = head < len ? (lj != 0 ? (uint32_t)lj : W.b[so]) : 0xffffffffu;
  int nb = 0;
  bool over = false;
  uint32_t eb0 = 0, eb1 = 0;
  double ew0 = 0.0, ew1 = 0.0;
  for (;;) {
    uint32_t m = hb;
    for (int o = GL / 2; o > 0; o >>= 1) m = min(m, __shfl_xor(m, o, GL));
    if (m == 0xffffffffu) break;  // every list consumed
    if (nb == KW) {
      over = true;
      break;
    }
    const bool has = hb == m;
    const double x = has ? fr * (lj != 0 ? 1.0 : W.w[so + head]) : 0.0;
    double acc = 0.0;
    for (int kk = 0; kk < nh; kk++) acc = acc + __shfl(x, (int)((src >> (4 * kk)) & 15u), GL);
    if (r == nb % GL) {
      if (nb < GL) {
        eb0 = m;
        ew0 = acc;
      } else {
        eb1 = m;
        ew1 = acc;
      }
    }
    nb++;
    if (has) {
      head++;
      hb = head < len ? W.b[so + head] : 0xffffffffu;  // an interior neighbour has one entry
    }
  }
  const uint32_t t = (uint32_t)slot[p];
  if (over) {
    if (r == 0) {
      wide[atomicAdd(nwide, 1u)] = p;
      atomicAdd(widetot, 1ull);
    }
    return;
  }
  unsigned long long o = 0;
  if (r == 0) o = atomicAdd(W.top, (unsigned long long)nb);
  o = __shfl(o, 0, GL);
  if (o + (unsigned long long)nb > W.cap) {  // pool full: the host enlarges it and runs the weights again
    if (r == 0) {
      atomicOr(err, 4);
      W.c[t] = 0;
      W.off[t] = 0;
    }
    return;
  }
  if (r == 0) {
    W.c[t] = (uint16_t)nb;
    W.off[t] = o;
  }
  if (r < nb) {
    W.b[o + r] = eb0;
    W.w[o + r] = ew0;
  }
  if (r + GL < nb) {
    W.b[o + r + GL] = eb1;
    W.w[o + r + GL] = ew1;
  }
}

// The frontier (its size read from nfp) in GL-lane groups, grid-stride.
__global__ void __launch_bounds__(YB) k_yt_w_narrow(YtGeo G, const uint32_t *nfp, const uint32_t *__restrict__ front,
                                                     const double *__restrict__ g, const uint16_t *__restrict__ hmask,
                                                     const int32_t *__restrict__ lab, const int32_t *__restrict__ slot,
                                                     WBuf W, uint32_t *__restrict__ wide, uint32_t *nwide,
                                                     unsigned long long *widetot, int *err) {
  const uint64_t nl = (uint64_t)*nfp * GL;
  for (uint64_t b0 = (uint64_t)blockIdx.x * YB; b0 < nl; b0 += (uint64_t)gridDim.x * YB) {
    const uint64_t l = b0 + threadIdx.x;
    if (l < nl)  // whole groups: GL divides YB
      yt_narrow_point(G, front[l / GL], (int)(l % GL), g, hmask, lab, slot, W, wide, nwide, widetot, err);
  }
}

// One wide point p by the whole workgroup: a dense accumulator per basin in
// LDS (acc, pres); the neighbours are added one after the other in
// decreasing rho (a neighbour's entries have distinct basins, so its
// additions do not collide), then the basins present are written in
// increasing order.  Every thread calls it.
__device__ void yt_wide_point(const YtGeo &G, uint32_t p, int nattr, const double *__restrict__ g,
                              const uint16_t *__restrict__ hmask, const int32_t *__restrict__ lab,
                              const int32_t *__restrict__ slot, const WBuf &W, int *err, double *acc, uint8_t *pres,
                              uint32_t *hq, double *fr, int &nh, uint32_t *wsum, unsigned long long &base) {
  for (int a = threadIdx.x; a < nattr; a += YB) {
    acc[a] = 0.0;
    pres[a] = 0;
  }
  if (threadIdx.x == 0) nh = yt_fracs(G, p, g, hmask, hq, fr);
  __syncthreads();
  for (int a = 0; a < nh; a++) {
    const uint32_t jn = hq[a];
    const int32_t lj = lab[jn];
    if (lj != 0) {
      if (threadIdx.x == 0) {
        acc[lj - 1] = acc[lj - 1] + fr[a] * 1.0;
        pres[lj - 1] = 1;
      }
    } else {
      const int32_t s = slot[jn];
      const int ce = W.c[s];
      const uint64_t so = W.off[s];
      for (int e = threadIdx.x; e < ce; e += YB) {
        const int b = (int)W.b[so + e] - 1;
        acc[b] = acc[b] + fr[a] * W.w[so + e];
        pres[b] = 1;
      }
    }
    __syncthreads();
  }
  // the basins present, in order: block-wide count, then positions
  uint32_t mine = 0;
  for (int a = threadIdx.x; a < nattr; a += YB) mine += pres[a];
  uint32_t v = mine;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
  __syncthreads();
  const uint32_t na = wsum[0] + wsum[1] + wsum[2] + wsum[3];
  const uint32_t t = (uint32_t)slot[p];
  if (threadIdx.x == 0) {
    base = atomicAdd(W.top, (unsigned long long)na);
    if (base + na > W.cap) {
      atomicOr(err, 4);
      W.c[t] = 0;
      W.off[t] = 0;
    } else {
      W.c[t] = (uint16_t)na;
      W.off[t] = base;
    }
  }
  __syncthreads();
  if (base + na > W.cap) return;
  if (threadIdx.x == 0) {  // serial compaction keeps increasing basin order
    uint64_t o = base;
    for (int a = 0; a < nattr; a++)
      if (pres[a]) {
        W.b[o] = (uint32_t)(a + 1);
        W.w[o] = acc[a];
        o++;
      }
  }
}

// The frontier's wide points, each by one workgroup (yt_wide_point), the
// workgroups striding over their list; then the next frontier: each surface
// point below a frontier point loses one pending neighbour, and the thread
// that takes the last lists it.  One launch: the two parts touch different
// data (weight entries; pending counts and the next list), and it runs every
// frontier, where a wide launch of its own would most often find nothing.
__global__ void __launch_bounds__(YB) k_yt_w_wide_step(YtGeo G, const uint32_t *__restrict__ wide,
                                                        const uint32_t *nwp, int nattr, const double *__restrict__ g,
                                                        const uint16_t *__restrict__ hmask,
                                                        const int32_t *__restrict__ lab,
                                                        const int32_t *__restrict__ slot, WBuf W, int *err,
                                                        const uint32_t *nfp, const uint32_t *__restrict__ front,
                                                        uint32_t *__restrict__ pend, uint32_t *__restrict__ next,
                                                        uint32_t *nnext) {
  extern __shared__ double acc[];  // nattr doubles, then nattr flags
  uint8_t *pres = reinterpret_cast<uint8_t *>(acc + nattr);
  __shared__ uint32_t hq[14];
  __shared__ double fr[14];
  __shared__ int nh;
  __shared__ uint32_t wsum[YB / 64];
  __shared__ unsigned long long base;
  const uint32_t nw = *nwp;
  for (uint32_t qi = blockIdx.x; qi < nw; qi += gridDim.x) {
    yt_wide_point(G, wide[qi], nattr, g, hmask, lab, slot, W, err, acc, pres, hq, fr, nh, wsum, base);
    __syncthreads();
  }
  const uint32_t nf = *nfp;
  for (uint32_t b0 = blockIdx.x * YB; b0 < nf; b0 += gridDim.x * YB) {  // uniform per workgroup
    const uint32_t t = b0 + threadIdx.x;
    uint32_t done[14];
    int nd = 0;
    if (t < nf) {
      const uint32_t p = front[t];
      int i, j, k;
      coords3(G, p, i, j, k);
      for (int q = 0; q < G.nvec; q++) {
        const uint32_t jn = nbr(G, i, j, k, q);
        if (yt_above(G, g, p, jn) && lab[jn] == 0 && atomicSub(&pend[jn], 1u) == 1u) done[nd++] = jn;
      }
    }
    for (int q = 0; q < 14; q++) yt_append(q < nd, q < nd ? done[q] : 0u, next, nnext);
  }
}

// Pending surface neighbours above each surface point; those with none
// form the first frontier.
__global__ void __launch_bounds__(YB) k_yt_w_init(YtGeo G, uint32_t nias, const uint32_t *__restrict__ ias,
                                                   const uint16_t *__restrict__ hmask,
                                                   const int32_t *__restrict__ lab, uint32_t *__restrict__ pend,
                                                   uint32_t *__restrict__ front, uint32_t *nfront) {
  const uint32_t t = blockIdx.x * YB + threadIdx.x;
  bool ready = false;
  uint32_t p = 0;
  if (t < nias) {
    p = ias[t];
    int i, j, k;
    coords3(G, p, i, j, k);
    uint32_t mask = hmask[p], c = 0;
    while (mask) {
      const int q = __ffs(mask) - 1;
      mask &= mask - 1;
      if (lab[nbr(G, i, j, k, q)] == 0) c++;
    }
    pend[p] = c;
    ready = c == 0;
  }
  yt_append(ready, p, front, nfront);
}

struct YtFields {
  const double *p[4];
};

// Surface points' weighted sums: acc[q*nattr + a], q = 0 the weights;
// binned per workgroup in LDS, one global addition per bin and workgroup.
// With direct (more bins than the LDS holds) every entry adds straight
// into acc; there are a few entries per surface point, spread over nattr
// basins.
__global__ void __launch_bounds__(YB) k_yt_surface_sums(uint32_t nias, const uint32_t *__restrict__ ias, WBuf W,
                                                         int nattr, int nf, YtFields F, double *__restrict__ acc,
                                                         bool direct) {
  extern __shared__ double lbins[];  // (nf + 1) * nattr
  const int nb = (nf + 1) * nattr;
  double *bins = direct ? acc : lbins;
  if (!direct) {
    for (int q = threadIdx.x; q < nb; q += YB) bins[q] = 0.0;
    __syncthreads();
  }
  for (uint32_t t = blockIdx.x * YB + threadIdx.x; t < nias; t += gridDim.x * YB) {
    const uint32_t p = ias[t];
    const int ce = W.c[t];
    const uint64_t so = W.off[t];
    double fv[4];
    for (int q = 0; q < nf; q++) fv[q] = F.p[q][p];
    for (int e = 0; e < ce; e++) {
      const int a = (int)W.b[so + e] - 1;
      const double w = W.w[so + e];
      if (a < 0 || a >= nattr) continue;
      atomicAdd(&bins[a], w);
      for (int q = 0; q < nf; q++) atomicAdd(&bins[(q + 1) * nattr + a], w * fv[q]);
    }
  }
  if (direct) return;
  __syncthreads();
  for (int q = threadIdx.x; q < nb; q += YB)
    if (bins[q] != 0.0) atomicAdd(&acc[q], bins[q]);
}

// Frontier sizes kept on the device: at(p) holds the size of frontier p
// (the step over frontier p counts frontier p + 1 there), so the host reads
// a size only every FRONT_BATCH frontiers; steps past the last frontier see
// size 0 and do nothing.  reserve(n) makes indices below n exist (zeroed);
// finish() reads the sizes back and returns the number of non-empty
// frontiers and their sum.
constexpr int FRONT_BATCH = 32;
struct FrontCounts {
  c2gpu_ctx *ctx;
  std::string name;
  uint32_t *d = nullptr;
  size_t cap = 0;
  FrontCounts(c2gpu_ctx *c, const char *nm) : ctx(c), name(nm) { reserve(4096); }
  void reserve(size_t n) {
    if (n <= cap) return;
    size_t nc = cap ? cap : 4096;
    while (nc < n) nc *= 2;
    uint32_t *nd = buf<uint32_t>(ctx, name + "_" + std::to_string(nc), nc);
    C2_HIP(hipMemsetAsync(nd, 0, 4 * nc, ctx->stream));
    if (cap) C2_HIP(hipMemcpyAsync(nd, d, 4 * cap, hipMemcpyDeviceToDevice, ctx->stream));
    d = nd;
    cap = nc;
  }
  uint32_t *at(size_t p) const { return d + p; }
  void set0(uint32_t v) {
    C2_HIP(hipMemcpyAsync(d, &v, 4, hipMemcpyHostToDevice, ctx->stream));
    C2_HIP(hipStreamSynchronize(ctx->stream));
  }
  int finish(int upto, uint64_t &sum) {
    std::vector<uint32_t> h((size_t)upto + 1);
    copy_to_host(ctx, h.data(), d, 4 * h.size());
    C2_HIP(hipStreamSynchronize(ctx->stream));
    int z = 0;
    sum = 0;
    while (z <= upto && h[z]) sum += h[z++];
    return z;
  }
};

// The basin ids of the stored surface weights through map (1-based ids).
__global__ void __launch_bounds__(YB) k_yt_remap(uint32_t nias, const uint64_t *__restrict__ woff,
                                                  const uint16_t *__restrict__ wc, uint32_t *__restrict__ wb,
                                                  const int32_t *__restrict__ map, uint32_t nmap) {
  const uint32_t t = blockIdx.x * YB + threadIdx.x;
  if (t >= nias) return;
  const uint64_t so = woff[t];
  const int ce = wc[t];
  for (int e = 0; e < ce; e++) {
    const uint32_t b = wb[so + e];
    if (b >= 1 && b <= nmap) wb[so + e] = (uint32_t)map[b - 1];
  }
}

}  // namespace

void yt_run(c2gpu_ctx *ctx, const double *rho, const int n[3], const double x2c[9], const c2gpu_attractor_opts *opts,
            int nfields, const double *const *fields, int flags, int32_t *label_out, int *nattr_out,
            double *xattr_out, int xattr_cap, double *vol_out, double *sums_out, c2gpu_yt_stats *stats) {
  if (!rho || !n || !x2c || !opts || !nattr_out || !xattr_out || !vol_out) fail(C2GPU_ERR_ARG, "c2gpu_yt: NULL argument");
  // DISCARD in yt_integrate (yt@proc.f90:153-167) turns the points below a
  // discarded maximum into surface points; not built here
  if (opts->discard) fail(C2GPU_ERR_UNSUPPORTED, "c2gpu_yt: the DISCARD callback is not supported");
  ctx->yt.valid = false;
  if (nfields < 0 || nfields > 4 || (nfields > 0 && (!fields || !sums_out)))
    fail(C2GPU_ERR_ARG, "c2gpu_yt: 0 <= nfields <= 4, with fields and sums_out");
  for (int a = 0; a < 3; a++)
    if (n[a] < 3) fail(C2GPU_ERR_ARG, "c2gpu_yt: grids need n >= 3 per axis");
  const uint64_t N64 = (uint64_t)n[0] * n[1] * n[2];
  if (N64 >= (1ull << 32)) fail(C2GPU_ERR_ARG, "c2gpu_yt: grids need fewer than 2^32 points");
  const uint32_t N = (uint32_t)N64;
  hipStream_t st = ctx->stream;
  Timer ttot(st);
  c2gpu_yt_stats S{};
  const bool dev = flags & C2GPU_PTR_DEVICE;
  // geometry: the Voronoi-relevant neighbours of the grid lattice
  YtGeo G{};
  G.n1 = n[0];
  G.n2 = n[1];
  G.n3 = n[2];
  double m[9];
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) m[r + 3 * c] = x2c[r + 3 * c] / (double)n[c];
  int vec[42];
  G.nvec = ws_facets(m, vec, G.al);
  for (int q = 0; q < G.nvec; q++) {
    G.di[q] = vec[3 * q];
    G.dj[q] = vec[3 * q + 1];
    G.dk[q] = vec[3 * q + 2];
  }
  // inputs on the device
  const double *g = rho;
  YtFields F{};
  if (dev) {
    require_device_ptr(rho, "rho");
    for (int q = 0; q < nfields; q++) {
      require_device_ptr(fields[q], "fields[q]");
      F.p[q] = fields[q];
    }
    if (label_out) require_device_ptr(label_out, "label_out");
  } else {
    double *d = buf<double>(ctx, "yt_rho", N);
    C2_HIP(hipMemcpyAsync(d, rho, 8 * (size_t)N, hipMemcpyHostToDevice, st));
    g = d;
    for (int q = 0; q < nfields; q++) {
      double *df = buf<double>(ctx, "yt_f" + std::to_string(q), N);
      C2_HIP(hipMemcpyAsync(df, fields[q], 8 * (size_t)N, hipMemcpyHostToDevice, st));
      F.p[q] = df;
    }
  }
  uint16_t *hmask = buf<uint16_t>(ctx, "yt_hmask", N);
  uint32_t *maxl = buf<uint32_t>(ctx, "yt_maxl", N);
  uint32_t *d_cnt = buf<uint32_t>(ctx, "yt_cnt", 2);
  int *d_err = buf<int>(ctx, "yt_err", 2);  // err, changed
  C2_HIP(hipMemsetAsync(d_cnt, 0, 8, st));
  C2_HIP(hipMemsetAsync(d_err, 0, 8, st));
  const unsigned gN = grid_for(N, YB);
  const unsigned gF = (unsigned)persistent_threads(ctx->device, YB) / YB;  // frontier kernels, grid-stride
  uint32_t *pend = buf<uint32_t>(ctx, "yt_pend", N);
  C2_LAUNCH(ctx, k_yt_class, dim3(gN), dim3(YB), 0, G, N, g, hmask, pend, maxl, d_cnt, d_err);
  add_kernel_bytes(ctx, "k_yt_class", (double)N * (8.0 * (G.nvec + 1) + 2.0));
  uint32_t nmax = read_scalar(ctx, d_cnt);
  // Equal values where the order matters (neighbours, the higher neighbours
  // of one point, or maxima): every point is ranked as the reference's
  // quicksort ranks it (yt@proc.f90:85-93), on the host, and the ranks
  // decide every order from here on.
  std::vector<uint32_t> h_rk;
  auto rank_points = [&]() {
    std::vector<double> h_g(N);
    copy_to_host(ctx, h_g.data(), g, 8 * (size_t)N);
    C2_HIP(hipStreamSynchronize(st));
    std::vector<uint32_t> io(N);
    for (uint32_t q = 0; q < N; q++) io[q] = q;
    qcksort_order(h_g.data(), io.data(), 0, (int64_t)N - 1);
    h_rk.assign(N, 0);
    for (uint32_t q = 0; q < N; q++) h_rk[io[q]] = q;
    uint32_t *d_rk = buf<uint32_t>(ctx, "yt_rank", N);
    C2_HIP(hipMemcpyAsync(d_rk, h_rk.data(), 4 * (size_t)N, hipMemcpyHostToDevice, st));
    C2_HIP(hipStreamSynchronize(st));
    G.rk = d_rk;
    S.ranked = 1;
  };
  if (read_scalar(ctx, d_err) & 1) {
    rank_points();
    C2_HIP(hipMemsetAsync(d_cnt, 0, 8, st));
    C2_HIP(hipMemsetAsync(d_err, 0, 8, st));
    C2_LAUNCH(ctx, k_yt_class, dim3(gN), dim3(YB), 0, G, N, g, hmask, pend, maxl, d_cnt, d_err);
    nmax = read_scalar(ctx, d_cnt);
  }
  // maxima in decreasing rho (rank); the attractor rules
  std::vector<uint32_t> h_max(nmax);
  std::vector<double> h_gmax(nmax);
  if (nmax) {
    double *d_gm = buf<double>(ctx, "yt_gmax", nmax);
    C2_LAUNCH(ctx, k_yt_gather, dim3(grid_for(nmax, YB)), dim3(YB), 0, nmax, maxl, g, d_gm);
    copy_to_host(ctx, h_max.data(), maxl, 4 * (size_t)nmax);
    copy_to_host(ctx, h_gmax.data(), d_gm, 8 * (size_t)nmax);
    C2_HIP(hipStreamSynchronize(st));
  }
  std::vector<uint32_t> order(nmax);
  for (uint32_t q = 0; q < nmax; q++) order[q] = q;
  auto by_value = [&](uint32_t a, uint32_t b) {
    return h_gmax[a] != h_gmax[b] ? h_gmax[a] > h_gmax[b] : h_max[a] < h_max[b];
  };
  if (h_rk.empty()) {
    std::sort(order.begin(), order.end(), by_value);
    bool tied = false;
    for (uint32_t q = 1; q < nmax; q++) tied = tied || h_gmax[order[q]] == h_gmax[order[q - 1]];
    if (tied) rank_points();  // maxima of equal value: their order is the reference sort's
  }
  if (!h_rk.empty())
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return h_rk[h_max[a]] > h_rk[h_max[b]]; });
  std::vector<double> dv(3 * (size_t)nmax);
  for (uint32_t q = 0; q < nmax; q++) {
    const uint32_t l = h_max[q];
    dv[3 * q] = (double)(l % (uint32_t)n[0]) / n[0];
    dv[3 * q + 1] = (double)((l / (uint32_t)n[0]) % (uint32_t)n[1]) / n[1];
    dv[3 * q + 2] = (double)(l / ((uint32_t)n[0] * (uint32_t)n[1])) / n[2];
  }
  std::vector<int32_t> mlab(nmax);
  std::vector<double> xattr;
  std::vector<int8_t> discarded;
  const int nattr = attractor_rule(opts, nmax, order.data(), dv.data(), mlab.data(), xattr, discarded);
  if (nattr > xattr_cap) fail(C2GPU_ERR_ARG, "c2gpu_yt: xattr_cap too small for " + std::to_string(nattr));
  if (nattr < 1) fail(C2GPU_ERR_ARG, "c2gpu_yt: no attractors");
  S.nmaxima = (int)nmax;
  S.nattr = nattr;
  // labels: least and greatest basin reachable upwards
  uint64_t *lohi = buf<uint64_t>(ctx, "yt_lohi", N);
  int32_t *d_mlab = buf<int32_t>(ctx, "yt_mlab", nmax > 0 ? nmax : 1);
  C2_HIP(hipMemcpyAsync(d_mlab, mlab.data(), 4 * (size_t)nmax, hipMemcpyHostToDevice, st));
  C2_LAUNCH(ctx, k_yt_init, dim3(gN), dim3(YB), 0, N, lohi);
  if (nmax) C2_LAUNCH(ctx, k_yt_seed, dim3(grid_for(nmax, YB)), dim3(YB), 0, nmax, maxl, d_mlab, lohi);
  // frontier after frontier from the maxima
  uint32_t *fr[2] = {buf<uint32_t>(ctx, "yt_front0", N), buf<uint32_t>(ctx, "yt_front1", N)};
  C2_HIP(hipMemcpyAsync(fr[0], maxl, 4 * (size_t)nmax, hipMemcpyDeviceToDevice, st));
  // the frontier sizes stay on the device (fc[p] = size of frontier p); the
  // host looks at them every FRONT_BATCH frontiers only
  FrontCounts fc(ctx, "yt_fcnt");
  fc.set0(nmax);
  int passes = 0;
  for (;;) {
    fc.reserve(passes + FRONT_BATCH + 1);
    for (int u = 0; u < FRONT_BATCH; u++, passes++)
      C2_LAUNCH(ctx, k_yt_lab_step, dim3(gF), dim3(YB), 0, G, fc.at(passes), fr[passes & 1], g, hmask, pend, lohi,
                fr[(passes & 1) ^ 1], fc.at(passes + 1));
    if (read_scalar(ctx, fc.at(passes)) == 0) break;
  }
  uint64_t ndone = 0;
  passes = fc.finish(passes, ndone);
  add_kernel_bytes(ctx, "k_yt_lab_step", (double)ndone * (4.0 + 8.0 * G.nvec * 2.0));
  if (ndone != N) fail(C2GPU_ERR_ARG, "c2gpu_yt: internal inconsistency in the label order");
  S.label_passes = passes;
  int32_t *lab = dev && label_out ? label_out : buf<int32_t>(ctx, "yt_lab", N);
  uint8_t *iasf = buf<uint8_t>(ctx, "yt_iasf", N);
  C2_LAUNCH(ctx, k_yt_label, dim3(gN), dim3(YB), 0, N, lohi, lab, iasf);
  uint32_t *ias = buf<uint32_t>(ctx, "yt_ias", N);
  const uint32_t nias = compact_flags(ctx, "yt", iasf, N, ias);
  S.nias = nias;
  // surface weights, frontier after frontier over the surface points; the
  // pool of weight entries starts at 8 per point and grows if it fills
  WBuf W{};
  int wpasses = 0;
  if (nias) {
    int32_t *slot = buf<int32_t>(ctx, "yt_slot", N);
    C2_LAUNCH(ctx, k_yt_slots, dim3(grid_for(nias, YB)), dim3(YB), 0, nias, ias, slot);
    W.off = buf<uint64_t>(ctx, "yt_woff", nias);
    W.c = buf<uint16_t>(ctx, "yt_wc", nias);
    W.top = buf<unsigned long long>(ctx, "yt_wtop", 1);
    uint32_t *wide = buf<uint32_t>(ctx, "yt_wide", nias);
    unsigned long long *widetot = buf<unsigned long long>(ctx, "yt_widetot", 1);
    const size_t wide_lds = 9 * (size_t)nattr;
    if (wide_lds > 120 * 1024) fail(C2GPU_ERR_UNSUPPORTED, "c2gpu_yt: too many attractors for the LDS accumulator");
    uint64_t cap = 8ull * nias;
    for (;;) {
      S.wide = 0;
      W.cap = cap;
      W.b = buf<uint32_t>(ctx, "yt_wb", cap);
      W.w = buf<double>(ctx, "yt_ww", cap);
      C2_HIP(hipMemsetAsync(W.top, 0, 8, st));
      C2_HIP(hipMemsetAsync(d_err, 0, 4, st));
      C2_HIP(hipMemsetAsync(widetot, 0, 8, st));
      FrontCounts wc(ctx, "yt_wcnt");
      C2_LAUNCH(ctx, k_yt_w_init, dim3(grid_for(nias, YB)), dim3(YB), 0, G, nias, ias, hmask, lab, pend, fr[0],
                wc.at(0));
      wpasses = 0;
      for (;;) {
        wc.reserve(wpasses + FRONT_BATCH + 1);
        for (int u = 0; u < FRONT_BATCH; u++, wpasses++) {
          uint32_t *cf = fr[wpasses & 1], *nx = fr[(wpasses & 1) ^ 1];
          C2_HIP(hipMemsetAsync(d_cnt, 0, 4, st));  // the frontier's wide points
          C2_LAUNCH(ctx, k_yt_w_narrow, dim3(gF), dim3(YB), 0, G, wc.at(wpasses), cf, g, hmask, lab, slot, W, wide,
                    d_cnt, widetot, d_err);
          C2_LAUNCH(ctx, k_yt_w_wide_step, dim3(gF), dim3(YB), wide_lds, G, wide, (const uint32_t *)d_cnt, nattr,
                    g, hmask, lab, slot, W, d_err, (const uint32_t *)wc.at(wpasses), cf, pend, nx,
                    wc.at(wpasses + 1));
        }
        if (read_scalar(ctx, wc.at(wpasses)) == 0) break;
      }
      uint64_t wdone = 0;
      wpasses = wc.finish(wpasses, wdone);
      S.wide = (int)read_scalar(ctx, widetot);
      // each point: its and its higher neighbours' rho, masks, labels and slots
      add_kernel_bytes(ctx, "k_yt_w_narrow", (double)wdone * (4.0 + 2.0 + 8.0 * (G.nvec + 1) + 8.0 * G.nvec));
      if (wdone != nias) fail(C2GPU_ERR_ARG, "c2gpu_yt: internal inconsistency in the weight order");
      // weight entries: each written once and read back by the points below
      add_kernel_bytes(ctx, "k_yt_w_narrow", 24.0 * (double)read_scalar(ctx, W.top));
      const int e = read_scalar(ctx, d_err);
      if (!(e & 4)) break;
      cap = read_scalar(ctx, W.top) + (cap >> 1);  // every entry asked for, and room to spare
    }
  }
  S.weight_passes = wpasses;
  // sums: interiors, then surface points
  std::vector<int64_t> counts(nattr);
  std::vector<double> isums((size_t)nfields * nattr);
  basin_sums_run(ctx, lab, N, nattr, nfields, F.p, C2GPU_PTR_DEVICE, nfields ? isums.data() : nullptr,
                 counts.data());
  const size_t nb = (size_t)(nfields + 1) * nattr;
  const bool direct = 8 * nb > 60000;  // bins past the LDS budget: add into acc itself
  double *acc = buf<double>(ctx, "yt_acc", nb);
  C2_HIP(hipMemsetAsync(acc, 0, 8 * nb, st));
  if (nias) {
    const unsigned gs = std::min<unsigned>(grid_for(nias, YB), (unsigned)persistent_threads(ctx->device, YB) / YB);
    C2_LAUNCH(ctx, k_yt_surface_sums, dim3(gs), dim3(YB), direct ? 0 : 8 * nb, nias, ias, W, nattr, nfields, F,
              acc, direct);
    add_kernel_bytes(ctx, "k_yt_surface_sums", (double)nias * (14.0 + 8.0 * nfields) +
                                                   12.0 * (double)read_scalar(ctx, W.top));
  }
  std::vector<double> h(nb);
  copy_to_host(ctx, h.data(), acc, 8 * nb);
  if (label_out && !dev) C2_HIP(hipMemcpyAsync(label_out, lab, 4 * (size_t)N, hipMemcpyDeviceToHost, st));
  C2_HIP(hipStreamSynchronize(st));
  for (int a = 0; a < nattr; a++) vol_out[a] = (double)counts[a] + h[a];
  for (int q = 0; q < nfields; q++)
    for (int a = 0; a < nattr; a++)
      sums_out[(size_t)q * nattr + a] = isums[(size_t)q * nattr + a] + h[(size_t)(q + 1) * nattr + a];
  for (size_t c = 0; c < xattr.size(); c++) xattr_out[c] = xattr[c];
  *nattr_out = nattr;
  // the surface weights stay in the context for c2gpu_yt_multipoles
  ctx->yt.valid = true;
  for (int c = 0; c < 3; c++) ctx->yt.n[c] = n[c];
  ctx->yt.nattr = nattr;
  ctx->yt.nias = nias;
  ctx->yt.nent = nias ? read_scalar(ctx, W.top) : 0;
  ctx->yt.ias = ias;
  ctx->yt.woff = W.off;
  ctx->yt.wc = W.c;
  ctx->yt.wb = W.b;
  ctx->yt.ww = W.w;
  S.ms_total = ttot.stop();
  if (stats) *stats = S;
}

// The YT checkpoint update of int_reorder_gridout (integration@proc.f90:
// 1125-1158): ibasin2 = assigned(ibasin) on the surface weights kept by the
// last c2gpu_yt call; afterwards they belong to nattr_new attractors.
void yt_remap_run(c2gpu_ctx *ctx, const int32_t *map, int nmap, int nattr_new) {
  auto &Y = ctx->yt;
  if (!map) fail(C2GPU_ERR_ARG, "c2gpu_yt_remap: NULL map");
  if (!Y.valid) fail(C2GPU_ERR_ARG, "c2gpu_yt_remap: no c2gpu_yt call on this context has left its weights");
  if (nmap != Y.nattr)
    fail(C2GPU_ERR_ARG, "c2gpu_yt_remap: inconsistent number of attractors in the yt weights (" +
                            std::to_string(nmap) + " vs " + std::to_string(Y.nattr) + ")");
  if (nattr_new < 1) fail(C2GPU_ERR_ARG, "c2gpu_yt_remap: nattr_new >= 1");
  for (int i = 0; i < nmap; i++)
    if (map[i] < 1 || map[i] > nattr_new) fail(C2GPU_ERR_ARG, "c2gpu_yt_remap: map entries must lie in 1..nattr_new");
  hipStream_t st = ctx->stream;
  if (Y.nias) {
    int32_t *d_map = buf<int32_t>(ctx, "yt_map", (size_t)nmap);
    C2_HIP(hipMemcpyAsync(d_map, map, 4 * (size_t)nmap, hipMemcpyHostToDevice, st));
    C2_LAUNCH(ctx, k_yt_remap, dim3(grid_for(Y.nias, YB)), dim3(YB), 0, Y.nias, Y.woff, Y.wc,
              Y.wb, d_map, (uint32_t)nmap);
    C2_HIP(hipStreamSynchronize(st));
  }
  Y.nattr = nattr_new;
}

}  // namespace c2
