// Device-side near-grid ascent primitives (gfx950).  Every function follows
// the FP64 operation order of critic2's bader@proc.f90 so that paths are
// bit-identical to the serial reference; this file must be compiled with
// -ffp-contract=off (no FMA contraction: the reference build emits none).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace c2 {

struct BParams {
  const double *__restrict__ f;
  int n1, n2, n3;
  double inv1, inv2, inv3;  // 1/n1, 1/n2, 1/n3 (index arithmetic)
  double c[9];    // car2lat, column-major
  double lid[27]; // lat_i_dist
  // when set: each point's normalised step direction (x, y, z; all 0 where
  // max|g| < 1e-30) and value (w), so a step reads one record instead of
  // seven values
  const double4 *__restrict__ G;
  // when set: positions 1..TPK of every free trajectory, position-major
  // (tpk[(q-1)*npts + s]), so the lanes of a wavefront (neighbouring starts)
  // read a path position from one contiguous run; valid for q < its length
  const uint32_t *__restrict__ tpk;
  uint32_t npts;
};
constexpr int TPK = 8;

// x / d and x % d for x < 2^32 from the FP64 reciprocal: the product is
// within 2^-20 of x/d, so truncation is off by at most one, which the
// remainder test corrects.
__device__ __forceinline__ uint32_t udiv(uint32_t x, uint32_t d, double inv, uint32_t &rem) {
  uint32_t q = (uint32_t)((double)x * inv);
  int64_t r = (int64_t)x - (int64_t)q * (int64_t)d;
  if (r < 0) {
    q--;
    r += d;
  } else if (r >= (int64_t)d) {
    q++;
    r -= d;
  }
  rem = (uint32_t)r;
  return q;
}

__device__ __forceinline__ int wrap1(int p, int n) {
  return p < 0 ? p + n : (p >= n ? p - n : p);
}
__device__ __forceinline__ uint32_t lin3(const BParams &P, int i, int j, int k) {
  return (uint32_t)i + (uint32_t)P.n1 * ((uint32_t)j + (uint32_t)P.n2 * (uint32_t)k);
}
__device__ __forceinline__ void coords(const BParams &P, uint32_t l, int &i, int &j, int &k) {
  uint32_t ri, rj;
  const uint32_t r = udiv(l, (uint32_t)P.n1, P.inv1, ri);
  k = (int)udiv(r, (uint32_t)P.n2, P.inv2, rj);
  i = (int)ri;
  j = (int)rj;
}
// sweep-order index: do i; do j; do k (k innermost), bader@proc.f90:151-153
__device__ __forceinline__ uint64_t sweep_index(const BParams &P, uint32_t l) {
  int i, j, k;
  coords(P, l, i, j, k);
  return ((uint64_t)i * (uint64_t)P.n2 + (uint64_t)j) * (uint64_t)P.n3 + (uint64_t)k;
}
// rho_val with periodic wrap (offsets of at most one cell), :619-639
__device__ __forceinline__ double rv(const BParams &P, int i, int j, int k) {
  return P.f[lin3(P, wrap1(i, P.n1), wrap1(j, P.n2), wrap1(k, P.n3))];
}

// is_max, bader@proc.f90:571-597 (strict >, plateaus count as maxima)
__device__ inline bool is_max(const BParams &P, int i, int j, int k) {
  const double r = rv(P, i, j, k);
  bool m = true;
#pragma unroll
  for (int d1 = -1; d1 <= 1; d1++)
#pragma unroll
    for (int d2 = -1; d2 <= 1; d2++)
#pragma unroll
      for (int d3 = -1; d3 <= 1; d3++)
        if (rv(P, i + d1, j + d2, k + d3) > r) m = false;
  return m;
}

// step_ongrid, bader@proc.f90:500-527: first strict maximum in (d1,d2,d3)
// order of rho_ctr + (rho - rho_ctr)*lat_i_dist
__device__ inline void step_ongrid(const BParams &P, int p[3]) {
  int pm0 = p[0], pm1 = p[1], pm2 = p[2];
  const double rc = rv(P, p[0], p[1], p[2]);
  double rmax = rc;
#pragma unroll
  for (int d1 = -1; d1 <= 1; d1++)
#pragma unroll
    for (int d2 = -1; d2 <= 1; d2++)
#pragma unroll
      for (int d3 = -1; d3 <= 1; d3++) {
        double t = rv(P, p[0] + d1, p[1] + d2, p[2] + d3);
        t = rc + (t - rc) * P.lid[(d1 + 1) + 3 * (d2 + 1) + 9 * (d3 + 1)];
        if (t > rmax) {
          rmax = t;
          pm0 = p[0] + d1;
          pm1 = p[1] + d2;
          pm2 = p[2] + d3;
        }
      }
  p[0] = wrap1(pm0, P.n1);
  p[1] = wrap1(pm1, P.n2);
  p[2] = wrap1(pm2, P.n3);
}

// rho_grad_dir, bader@proc.f90:532-567
__device__ inline void grad_dir(const BParams &P, const int p[3], double res[3]) {
  const int a = p[0], b = p[1], c = p[2];
  const double r000 = rv(P, a, b, c);
  const double r001 = rv(P, a, b, c + 1);
  const double r010 = rv(P, a, b + 1, c);
  const double r100 = rv(P, a + 1, b, c);
  const double r00m = rv(P, a, b, c - 1);
  const double rm00 = rv(P, a - 1, b, c);
  const double r0m0 = rv(P, a, b - 1, c);
  double gl[3];
  gl[0] = (r100 - rm00) / 2.0;
  gl[1] = (r010 - r0m0) / 2.0;
  gl[2] = (r001 - r00m) / 2.0;
  if (r100 < r000 && rm00 < r000) gl[0] = 0.0;
  if (r010 < r000 && r0m0 < r000) gl[1] = 0.0;
  if (r001 < r000 && r00m < r000) gl[2] = 0.0;
  double gc[3];
#pragma unroll
  for (int j = 0; j < 3; j++) {  // matmul(rho_grad_lat, car2lat)
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) acc = acc + gl[i] * P.c[i + 3 * j];
    gc[j] = acc;
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {  // matmul(car2lat, rho_grad_car)
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) acc = acc + P.c[i + 3 * j] * gc[j];
    res[i] = acc;
  }
}

// Fortran nint for the small values met here: round half away from zero.
__device__ __forceinline__ int fnint(double x) { return (int)round(x); }

// One step_neargrid call (bader@proc.f90:455-494) up to, but excluding, the
// known(pm)==1 test, which needs the caller's path.  Returns false when the
// is_max early return fires (no move, known(p) untouched); otherwise pm
// holds the wrapped candidate target.
// g of step_neargrid after g = g / max|g| (bader@proc.f90:466-478); false
// when max|g| < 1e-30.
__device__ __forceinline__ bool step_dir(const BParams &P, const int p[3], double g[3]) {
  grad_dir(P, p, g);
  double gm = fabs(g[0]);
  if (fabs(g[1]) > gm) gm = fabs(g[1]);
  if (fabs(g[2]) > gm) gm = fabs(g[2]);
  if (gm < 1e-30) return false;
  const double coeff = 1.0 / gm;
#pragma unroll
  for (int c = 0; c < 3; c++) g[c] = coeff * g[c];
  return true;
}

// t: p's record of the step direction table when the table is set (a
// point with no step direction has g = 0; a moving point has a component
// of magnitude ~1 after the normalisation).
__device__ inline bool step_near_candidate(const BParams &P, const int p[3], const double4 &t, double dr[3],
                                           int pm[3]) {
  double g[3];
  bool moving;
  if (P.G) {
    g[0] = t.x;
    g[1] = t.y;
    g[2] = t.z;
    moving = g[0] != 0.0 || g[1] != 0.0 || g[2] != 0.0;
  } else {
    moving = step_dir(P, p, g);
  }
  if (!moving) {
    dr[0] = dr[1] = dr[2] = 0.0;
    if (is_max(P, p[0], p[1], p[2])) return false;
    pm[0] = p[0];
    pm[1] = p[1];
    pm[2] = p[2];
    step_ongrid(P, pm);
    return true;
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const int ng = fnint(g[c]);
    int q = p[c] + ng;
    dr[c] = (dr[c] + g[c]) - (double)ng;
    const int nd = fnint(dr[c]);
    q = q + nd;
    dr[c] = dr[c] - (double)nd;
    pm[c] = q;
  }
  pm[0] = wrap1(pm[0], P.n1);
  pm[1] = wrap1(pm[1], P.n2);
  pm[2] = wrap1(pm[2], P.n3);
  return true;
}

// max_neargrid (bader@proc.f90:427-450) from `start` with dr = 0.
// path[] (capacity maxlen) receives the visited points; known(pm)==1 is
// membership in path[].  stop(x) is the known(x)==2 test (false for free
// trajectories).  Returns the path length, or -1 if it exceeds maxlen.
// 256-bit membership filter of the points on the current path: a clear bit
// proves absence, so the exact scan of the path runs only on filter hits.
// The four words are separate members picked by selects: an array indexed
// by the hash would live in scratch memory, one more memory access in the
// dependent chain of every step.
struct PathFilter {
  uint64_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
  __device__ __forceinline__ static uint32_t bit(uint32_t l) { return (l * 0x9E3779B1u) >> 24; }
  __device__ __forceinline__ void add(uint32_t l) {
    const uint32_t b = bit(l), k = b >> 6;
    const uint64_t m = 1ull << (b & 63);
    w0 |= k == 0 ? m : 0ull;
    w1 |= k == 1 ? m : 0ull;
    w2 |= k == 2 ? m : 0ull;
    w3 |= k == 3 ? m : 0ull;
  }
  __device__ __forceinline__ bool maybe(uint32_t l) const {
    const uint32_t b = bit(l), k = b >> 6;
    const uint64_t x = k == 0 ? w0 : k == 1 ? w1 : k == 2 ? w2 : w3;
    return (x >> (b & 63)) & 1ull;
  }
};

// Path storage of walk_path: set(q, v) appends entry q, get(q, len) reads
// entry q < len.  FlatPath is a plain array.
struct FlatPath {
  uint32_t *p;
  __device__ __forceinline__ void set(int q, uint32_t v) const { p[q] = v; }
  __device__ __forceinline__ uint32_t get(int q, int) const { return p[q]; }
};

// The newest entries in an LDS ring of RING words per thread (interleaved
// over the workgroup's threads), written to the thread's scratch a whole
// ring (64 bytes) at a time: the scratch lines are written once and whole
// instead of a word per step.  Entries below len rounded down to RING are in
// the scratch, the others in the ring.
constexpr int RING = 16;
struct RingPath {
  uint32_t *ring;   // this thread's slot 0; slot j at ring[j * stride]
  uint32_t *spill;  // this thread's scratch (16-word aligned)
  uint32_t stride;
  __device__ __forceinline__ void set(int q, uint32_t v) const {
    ring[(q & (RING - 1)) * stride] = v;
    if ((q & (RING - 1)) == RING - 1) {
      uint4 *d = reinterpret_cast<uint4 *>(spill + (q - (RING - 1)));
#pragma unroll
      for (int u = 0; u < RING / 4; u++)
        d[u] = make_uint4(ring[(4 * u) * stride], ring[(4 * u + 1) * stride], ring[(4 * u + 2) * stride],
                          ring[(4 * u + 3) * stride]);
    }
  }
  __device__ __forceinline__ uint32_t get(int q, int len) const {
    return q < (len & ~(RING - 1)) ? spill[q] : ring[(q & (RING - 1)) * stride];
  }
};

template <class Stop, class Visit, class Path = FlatPath>
// A candidate whose value exceeds every value on the path so far cannot be
// on it: the membership filter and the exact scan run only otherwise (a
// path mostly climbs, and a long path saturates the filter).  The table
// record of the candidate (its step direction and value) is loaded once
// and carried to the next step.
__device__ inline int walk_path(const BParams &P, uint32_t start, Path path, int maxlen, Stop stop,
                                Visit visit) {
  int p[3];
  coords(P, start, p[0], p[1], p[2]);
  double dr[3] = {0.0, 0.0, 0.0};
  int len = 1;
  path.set(0, start);
  PathFilter filt;
  filt.add(start);
  visit(start);
  double4 rec = P.G ? P.G[start] : make_double4(0.0, 0.0, 0.0, 0.0);
  double fmax = P.G ? rec.w : P.f[start];
  for (;;) {
    int pm[3];
    if (!step_near_candidate(P, p, rec, dr, pm)) break;  // maximum: no move
    uint32_t lm = lin3(P, pm[0], pm[1], pm[2]);
    double4 recm = P.G ? P.G[lm] : make_double4(0.0, 0.0, 0.0, 0.0);
    double fm = P.G ? recm.w : P.f[lm];
    bool seen = false;
    if (!(fm > fmax) && filt.maybe(lm))
      for (int q = len - 1; q >= 0; q--)
        if (path.get(q, len) == lm) {
          seen = true;
          break;
        }
    if (seen) {  // known(pm) == 1: on-grid step, dr = 0
      pm[0] = p[0];
      pm[1] = p[1];
      pm[2] = p[2];
      step_ongrid(P, pm);
      dr[0] = dr[1] = dr[2] = 0.0;
      lm = lin3(P, pm[0], pm[1], pm[2]);
      if (P.G) recm = P.G[lm];
      fm = P.G ? recm.w : P.f[lm];
    }
    if (pm[0] == p[0] && pm[1] == p[1] && pm[2] == p[2]) break;  // no move
    if (len >= maxlen) return -1;
    path.set(len++, lm);
    filt.add(lm);
    visit(lm);
    p[0] = pm[0];
    p[1] = pm[1];
    p[2] = pm[2];
    rec = recm;
    fmax = fm > fmax ? fm : fmax;
    if (stop(lm)) break;
  }
  return len;
}

}  // namespace c2
