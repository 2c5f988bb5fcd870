// Gaussian cube input: the list-directed text read of grid3mod@proc.f90:512-572
// and the unformatted binary read of grid3mod@proc.f90:575-635.
//
// Text cubes.  The header (two title records, natoms and origin, three axis
// records, the atom records and, for orbital cubes, one more record) is read
// on the host with the same conversion code the GPU uses.  The grid values
// are tokenised and converted on the GPU:
//   k_tok_count   one 16-byte segment per thread, token starts (a non-blank
//                 byte after a blank) counted per workgroup;
//   exclusive scan of the counts -> index of each workgroup's first token;
//   k_tok_parse   the workgroup's 4 KiB span (plus a 256-byte tail) staged
//                 in LDS, each token converted to binary64 (correctly
//                 rounded, number.hpp) and stored at its index, file order;
//                 the rare tokens that need the exact big-integer tier (or
//                 do not fit the staged window) go to k_tok_slow;
//   k_cube_transpose  file order (z fastest, critic2's read loop order
//                 (((f(i,j,k),k=1,n3),j=1,n2),i=1,n1)) -> f(n1,n2,n3), x
//                 fastest, through 64x64 LDS tiles.
// Values past the n1*n2*n3-th are ignored, as the Fortran read ignores them;
// fewer values, non-numeric tokens and list-directed forms other than plain
// blank-separated constants (r*c repeats, commas, slashes) are errors.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "devutil.hpp"
#include "format.hpp"
#include "number.hpp"
#include "pow5tab.hpp"

namespace c2 {

// ---- the 5^q table

const Pow5Host &pow5_host() {
  static Pow5Host h;
  static std::once_flag once;
  std::call_once(once, [] {
    h.t.resize(2 * (POW5_QMAX - POW5_QMIN + 1));
    h.s.resize(POW5_QMAX - POW5_QMIN + 1);
    pow5_table(h.t.data(), h.s.data());
  });
  return h;
}

void pow5_device(c2gpu_ctx *ctx, const uint64_t *&t, const int16_t *&s) {
  const Pow5Host &h = pow5_host();
  const bool fresh = ctx->bufs.find("cube_pow5") == ctx->bufs.end();
  uint64_t *dt = buf<uint64_t>(ctx, "cube_pow5", h.t.size());
  int16_t *ds = buf<int16_t>(ctx, "cube_pow5s", h.s.size());
  if (fresh) {
    C2_HIP(hipMemcpyAsync(dt, h.t.data(), h.t.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    C2_HIP(hipMemcpyAsync(ds, h.s.data(), h.s.size() * 2, hipMemcpyHostToDevice, ctx->stream));
    C2_HIP(hipStreamSynchronize(ctx->stream));
  }
  t = dt;
  s = ds;
}

namespace {

constexpr int TOK_BLOCK = 256;
constexpr int TOK_SEG = 16;                    // bytes per thread
constexpr int TOK_SPAN = TOK_BLOCK * TOK_SEG;  // bytes per workgroup
constexpr int TOK_TAIL = 256;                  // staged bytes past the span
constexpr int TOK_MAXG = 4096;                 // longest token read from global memory

// ---- host: list-directed header records

// Over a prefix of the file (whole == false) any record or token that
// reaches the end of the bytes reports "need more bytes" (false, why empty).
struct HdrReader {
  const char *b;
  uint64_t n;
  bool whole;
  uint64_t p = 0;  // start of the next record
  bool at_end() const { return p >= n; }
  // Skips one record (read (lu,*) with an empty list).
  bool skip_record() {
    if (p >= n) return false;
    while (p < n && b[p] != '\n') p++;
    if (p >= n && !whole) return false;
    if (p < n) p++;
    return true;
  }
  // Reads `count` items starting at the next record, continuing over
  // records as list-directed input does, then moves past the record where
  // the list ended.  kinds[i]: 0 integer, 1 real.
  bool items(int count, const int *kinds, int64_t *iv, double *rv, std::string &why) {
    if (p >= n) {
      if (whole) why = "end of file";
      return false;
    }
    const Pow5Host &h = pow5_host();
    uint64_t q = p;
    for (int it = 0; it < count; it++) {
      while (q < n && num_sep(b[q])) q++;
      if (q >= n) {
        if (whole) why = "end of file";
        return false;
      }
      uint64_t e = q;
      while (e < n && !num_sep(b[e])) e++;
      if (e >= n && !whole) return false;
      if (e - q > 4096) {
        why = "token too long";
        return false;
      }
      const int len = (int)(e - q);
      if (kinds[it] == 0) {
        int64_t v = 0;
        int k = 0, neg = 0;
        if (b[q] == '+' || b[q] == '-') {
          neg = b[q] == '-';
          k = 1;
        }
        if (k >= len) {
          why = "bad integer";
          return false;
        }
        for (; k < len; k++) {
          if (b[q + k] < '0' || b[q + k] > '9' || v > (int64_t)1e15) {
            why = "bad integer '" + std::string(b + q, len) + "'";
            return false;
          }
          v = v * 10 + (b[q + k] - '0');
        }
        iv[it] = neg ? -v : v;
      } else {
        DecNum d;
        int rc = dec_parse(b + q, len, d);
        double v = 0;
        if (rc == 0) rc = dec_to_double(b + q, d, h.t.data(), h.s.data(), v);
        if (rc) {
          why = "bad real '" + std::string(b + q, len) + "'";
          return false;
        }
        rv[it] = v;
      }
      q = e;
    }
    p = q;
    return skip_record() || p >= n;
  }
};

// Parses the header of a text cube (grid3mod@proc.f90:533-551).  Returns
// false if the bytes end before the header does (need more text).
bool parse_text_header(const char *b, uint64_t n, c2gpu_cube_header &h, bool whole) {
  HdrReader r{b, n, whole};
  std::string why;
  auto need = [&](bool ok) {
    if (ok) return;
    if (!why.empty() || whole)
      fail(C2GPU_ERR_FORMAT, "cube header: " + (why.empty() ? std::string("end of file") : why));
    throw 0;
  };
  try {
    need(r.skip_record());
    need(r.skip_record());
    int64_t iv[4];
    double rv[4];
    const int k_nat[4] = {0, 1, 1, 1};
    need(r.items(4, k_nat, iv, rv, why));
    h.ismo = iv[0] < 0;
    const int64_t nat = iv[0] < 0 ? -iv[0] : iv[0];
    if (nat > 100000000) fail(C2GPU_ERR_FORMAT, "cube header: atom count out of range");
    h.nat = (int)nat;
    for (int a = 0; a < 3; a++) h.x0[a] = rv[1 + a];
    for (int i = 0; i < 3; i++) {
      need(r.items(4, k_nat, iv, rv, why));
      if (iv[0] < 1 || iv[0] > 100000) fail(C2GPU_ERR_FORMAT, "cube header: grid dimension out of range");
      h.n[i] = (int)iv[0];
      for (int a = 0; a < 3; a++) h.step[a + 3 * i] = rv[1 + a];
    }
    for (int64_t i = 0; i < nat; i++) need(r.skip_record());
    if (h.ismo) need(r.skip_record());
    need(!r.at_end() || whole);  // a header-only file is complete at its end
    h.data_offset = r.p;
    return true;
  } catch (int) {
    return false;
  }
}

// ---- device: tokenisation and conversion

__device__ __forceinline__ uint8_t byte_at(const uint8_t *base, uint64_t lo, uint64_t hi, uint64_t g) {
  return (g >= lo && g < hi) ? base[g] : (uint8_t)' ';
}

// The 16 bytes at g (16-aligned, relative to base); bytes outside [lo, hi)
// read as blanks.  Whole-vector loads only where the segment lies inside.
__device__ __forceinline__ uint4 load_seg(const uint8_t *base, uint64_t lo, uint64_t hi, uint64_t g) {
  if (g >= lo && g + 16 <= hi) return *reinterpret_cast<const uint4 *>(base + g);
  uint32_t w[4];
  for (int k = 0; k < 4; k++) {
    uint32_t v = 0;
    for (int c = 0; c < 4; c++) v |= (uint32_t)byte_at(base, lo, hi, g + 4 * k + c) << (8 * c);
    w[k] = v;
  }
  return make_uint4(w[0], w[1], w[2], w[3]);
}

__device__ __forceinline__ uint32_t start_mask(const uint8_t *c, uint8_t prev) {
  uint32_t m = 0;
  bool ps = num_sep((char)prev);
  for (int k = 0; k < 16; k++) {
    const bool s = num_sep((char)c[k]);
    if (!s && ps) m |= 1u << k;
    ps = s;
  }
  return m;
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t *lds) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) lds[w] = v;
  __syncthreads();
  uint32_t t = 0;
  for (int i = 0; i < TOK_BLOCK / 64; i++) t += lds[i];
  return t;
}

__global__ void __launch_bounds__(TOK_BLOCK) k_tok_count(const uint8_t *base, uint64_t lo, uint64_t hi,
                                                         uint32_t *blockcnt) {
  __shared__ uint32_t lds[TOK_BLOCK / 64];
  const uint64_t g = ((uint64_t)blockIdx.x * TOK_BLOCK + threadIdx.x) * TOK_SEG;
  uint4 v = load_seg(base, lo, hi, g);
  const uint8_t *c = reinterpret_cast<const uint8_t *>(&v);
  const uint8_t prev = g == 0 ? (uint8_t)' ' : byte_at(base, lo, hi, g - 1);
  const uint32_t t = block_sum(__popc(start_mask(c, prev)), lds);
  if (threadIdx.x == 0) blockcnt[blockIdx.x] = t;
}

// Tokens that end inside the staged window and need no exact tier are
// converted from LDS here; the others (the exact tier, or a token longer
// than the window) are appended to `slow` as (index, byte position) for
// k_tok_slow.  stat[0]: smallest (index << 2 | error kind); stat[1]: the
// number of slow tokens (entries past cap are dropped and the launch
// repeated with room for all).
__global__ void __launch_bounds__(TOK_BLOCK) k_tok_parse(const uint8_t *base, uint64_t lo, uint64_t hi,
                                                         const uint64_t *blockoff, uint64_t npts,
                                                         const uint64_t *pow5, const int16_t *pow5s,
                                                         double *vals, unsigned long long *stat, uint64_t *slow,
                                                         uint64_t cap) {
  __shared__ uint4 stage[(TOK_SPAN + TOK_TAIL) / 16];
  __shared__ uint16_t tokpos[TOK_SPAN / 2];  // a token and its blank take >= 2 bytes
  __shared__ uint32_t wtot[TOK_BLOCK / 64];
  const uint64_t gb = (uint64_t)blockIdx.x * TOK_SPAN;
  for (int k = threadIdx.x; k < (TOK_SPAN + TOK_TAIL) / 16; k += TOK_BLOCK)
    stage[k] = load_seg(base, lo, hi, gb + 16 * (uint64_t)k);
  __syncthreads();
  const uint8_t *s = reinterpret_cast<const uint8_t *>(stage);
  const int l0 = threadIdx.x * TOK_SEG;
  const uint8_t prev = threadIdx.x == 0 ? (gb == 0 ? (uint8_t)' ' : byte_at(base, lo, hi, gb - 1)) : s[l0 - 1];
  uint32_t mask = start_mask(s + l0, prev);
  // exclusive prefix of the token counts over the workgroup; the token
  // starts are listed in LDS so that lane j converts the j-th token (even
  // work per lane, and neighbouring lanes read bytes about a token apart)
  const uint32_t cnt = __popc(mask);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = cnt;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(inc, o);
    if (lane >= o) inc += u;
  }
  if (lane == 63) wtot[w] = inc;
  __syncthreads();
  uint32_t before = 0, total = 0;
  for (int i = 0; i < TOK_BLOCK / 64; i++) {
    if (i < w) before += wtot[i];
    total += wtot[i];
  }
  uint32_t j = before + inc - cnt;
  while (mask) {
    tokpos[j++] = (uint16_t)(l0 + __ffs(mask) - 1);
    mask &= mask - 1;
  }
  __syncthreads();
  const uint64_t t0 = blockoff[blockIdx.x];
  for (uint32_t q = threadIdx.x; q < total; q += TOK_BLOCK) {
    const uint64_t t = t0 + q;
    if (t >= npts) break;
    const int loc = tokpos[q];
    const int avail = TOK_SPAN + TOK_TAIL - loc;
    int L = 0;
    while (L < avail && !num_sep((char)s[loc + L])) L++;
    bool done = false;
    if (L < avail) {  // the token and its terminator are staged (bytes past hi read as blanks)
      DecNum d;
      const int rc = dec_parse(reinterpret_cast<const char *>(s + loc), avail, d);
      if (rc) {
        atomicMin(stat, (unsigned long long)((t << 2) | (uint64_t)rc));
        done = true;
      } else {
        double v;
        uint64_t m;
        int eb;
        if (dec_fast(d, pow5, pow5s, v, m, eb) == 0) {
          vals[t] = v;
          done = true;
        }
      }
    }
    if (!done) {
      const unsigned long long qq = atomicAdd(stat + 1, 1ull);
      if (qq < cap) {
        slow[2 * qq] = t;
        slow[2 * qq + 1] = gb + loc;
      }
    }
  }
}

// The slow tokens, read from global memory with all three tiers.
__global__ void __launch_bounds__(256) k_tok_slow(const uint8_t *base, uint64_t hi, const uint64_t *slow,
                                                  uint64_t count, const uint64_t *pow5, const int16_t *pow5s,
                                                  double *vals, unsigned long long *stat) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint64_t t = slow[2 * i], gpos = slow[2 * i + 1];
  const char *p = reinterpret_cast<const char *>(base + gpos);
  const uint64_t rem = hi - gpos;
  const int n = rem < (uint64_t)TOK_MAXG ? (int)rem : TOK_MAXG;
  DecNum d;
  int rc = dec_parse(p, n, d);
  if (rc == 0 && d.len >= n && n == TOK_MAXG) rc = 2;  // token longer than the read window
  double v = 0;
  if (rc == 0) rc = dec_to_double(p, d, pow5, pow5s, v) ? 2 : 0;
  if (rc) atomicMin(stat, (unsigned long long)((t << 2) | (uint64_t)rc));
  vals[t] = v;
}

constexpr int TR = 64;  // transpose tile

// vals: file order, index k + n3 (j + n2 i); f: index i + n1 (j + n2 k).
// Grid: x = (k tile) + (k tiles) * j, y = i tile.
__global__ void __launch_bounds__(256) k_cube_transpose(const double *vals, double *f, int n1, int n2, int n3) {
  __shared__ double tile[TR][TR + 1];
  const int nkt = (n3 + TR - 1) / TR;
  const int j = blockIdx.x / nkt;
  const int k0 = (blockIdx.x % nkt) * TR, i0 = blockIdx.y * TR;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < TR; r += 4) {
    const int i = i0 + r, k = k0 + tx;
    if (i < n1 && k < n3) tile[r][tx] = vals[(uint64_t)k + (uint64_t)n3 * (j + (uint64_t)n2 * i)];
  }
  __syncthreads();
  for (int r = ty; r < TR; r += 4) {
    const int k = k0 + r, i = i0 + tx;
    if (i < n1 && k < n3) f[(uint64_t)i + (uint64_t)n1 * (j + (uint64_t)n2 * k)] = tile[tx][r];
  }
}

// ---- device: cube body text (writegrid_cube, crystalmod@write.f90:3557-3568;
// nciplot's write_cube_body, nci@proc.f90:916-932)
//
// Per (i,j) row, the n3 values in records of six, each " " + the edit
// descriptor's field; a record that ends before the sixth value also holds
// the next " " (the format's character edit descriptor is processed before
// the data list is found exhausted).  Every field has the style's fixed
// width, so each value's byte offset follows from its index.

struct BodyLayout {
  int W;          // field width
  int n3;
  uint64_t row;   // bytes per (i,j) row
};

__host__ __device__ inline BodyLayout body_layout(int n3, int style) {
  BodyLayout L;
  L.W = fmt_width(style);
  L.n3 = n3;
  const uint64_t nrec = (uint64_t)(n3 + 5) / 6;
  L.row = (uint64_t)n3 * (L.W + 1) + nrec + (n3 % 6 ? 1 : 0);
  return L;
}

__device__ __forceinline__ void put_field(char *text, const BodyLayout &L, uint64_t t, const char *fld) {
  const uint64_t r = t / (uint64_t)L.n3;
  const int k = (int)(t - r * (uint64_t)L.n3);
  char *o = text + r * L.row + (uint64_t)k * (L.W + 1) + (uint64_t)(k / 6);
  o[0] = ' ';
  for (int c = 0; c < L.W; c++) o[1 + c] = fld[c];
  if (k % 6 == 5) {
    o[L.W + 1] = '\n';
  } else if (k == L.n3 - 1) {
    o[L.W + 1] = ' ';
    o[L.W + 2] = '\n';
  }
}

// zf: values in file order (k fastest); one thread per value.  Values whose
// rounding needs the exact tier go to `slow` (count in stat[1], entries past
// cap dropped and the launch repeated) for k_cube_format_slow.
__global__ void __launch_bounds__(256) k_cube_format(const double *zf, uint64_t npts, BodyLayout L, int style,
                                                     const uint64_t *pow5, const int16_t *pow5s, char *text,
                                                     unsigned long long *stat, uint64_t *slow, uint64_t cap) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npts) return;
  char fld[24];
  if (fmt_value<false>(zf[t], style, pow5, pow5s, fld) == 0) {
    put_field(text, L, t, fld);
  } else {
    const unsigned long long q = atomicAdd(stat + 1, 1ull);
    if (q < cap) slow[q] = t;
  }
}

__global__ void __launch_bounds__(256) k_cube_format_slow(const double *zf, const uint64_t *slow, uint64_t count,
                                                          BodyLayout L, int style, const uint64_t *pow5,
                                                          const int16_t *pow5s, char *text,
                                                          unsigned long long *stat) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint64_t t = slow[i];
  char fld[24];
  if (fmt_value<true>(zf[t], style, pow5, pow5s, fld)) atomicOr(stat, 1ull);
  put_field(text, L, t, fld);
}

// Converts the grid values in [lo, hi) of base (device memory, base 16-byte
// aligned) into f (device, x fastest).
void parse_values(c2gpu_ctx *ctx, const uint8_t *base, uint64_t lo, uint64_t hi, const int n[3], double *f) {
  const uint64_t npts = (uint64_t)n[0] * n[1] * n[2];
  const uint64_t nblk = (hi + TOK_SPAN - 1) / TOK_SPAN;
  if (hi <= lo) fail(C2GPU_ERR_FORMAT, "cube: file holds no grid values");
  if (nblk > 0x7fffffffull) fail(C2GPU_ERR_ARG, "cube text too large");
  uint32_t *cnt = buf<uint32_t>(ctx, "cube_blockcnt", nblk + 1);
  uint64_t *off = buf<uint64_t>(ctx, "cube_blockoff", nblk + 1);
  unsigned long long *stat = buf<unsigned long long>(ctx, "cube_stat", 2);
  double *vals = buf<double>(ctx, "cube_vals", npts);
  const uint64_t *p5;
  const int16_t *p5s;
  pow5_device(ctx, p5, p5s);
  C2_HIP(hipMemsetAsync(cnt + nblk, 0, 4, ctx->stream));
  C2_LAUNCH(ctx, k_tok_count, dim3((unsigned)nblk), dim3(TOK_BLOCK), 0, base, lo, hi, cnt);
  add_kernel_bytes(ctx, "k_tok_count", (double)(hi - lo) + 4.0 * nblk);
  exclusive_scan(ctx, "cube", cnt, off, (int64_t)nblk + 1);
  const uint64_t ntok = read_scalar(ctx, off + nblk);
  if (ntok < npts)
    fail(C2GPU_ERR_FORMAT, "cube: file holds " + std::to_string(ntok) + " grid values, expected " +
                               std::to_string(npts));
  uint64_t cap = 1ull << 16;
  unsigned long long st[2];
  for (;;) {
    uint64_t *slow = buf<uint64_t>(ctx, "cube_slow", 2 * cap);
    const unsigned long long init[2] = {~0ull, 0ull};
    C2_HIP(hipMemcpyAsync(stat, init, sizeof init, hipMemcpyHostToDevice, ctx->stream));
    C2_LAUNCH(ctx, k_tok_parse, dim3((unsigned)nblk), dim3(TOK_BLOCK), 0, base, lo, hi, off, npts, p5, p5s,
              vals, stat, slow, cap);
    add_kernel_bytes(ctx, "k_tok_parse", (double)(hi - lo) + 8.0 * nblk + 8.0 * (double)npts);
    copy_to_host(ctx, st, stat, sizeof st);
    C2_HIP(hipStreamSynchronize(ctx->stream));
    if (st[1] <= cap) break;
    cap = st[1];  // room for every slow token, then once more
  }
  if (st[1]) {
    C2_LAUNCH(ctx, k_tok_slow, dim3(grid_for((int64_t)st[1], 256)), dim3(256), 0, base, hi,
              buf<uint64_t>(ctx, "cube_slow", 2 * cap), (uint64_t)st[1], p5, p5s, vals, stat);
  }
  const unsigned long long e = st[1] ? read_scalar(ctx, stat) : st[0];
  if (e != ~0ull) {
    const uint64_t t = e >> 2;
    const int kk = (int)(t % n[2]), jj = (int)((t / n[2]) % n[1]), ii = (int)(t / ((uint64_t)n[2] * n[1]));
    fail(C2GPU_ERR_FORMAT, std::string("cube: grid value ") + std::to_string(t + 1) + " (i,j,k = " +
                               std::to_string(ii + 1) + "," + std::to_string(jj + 1) + "," +
                               std::to_string(kk + 1) + ") " +
                               ((e & 3) == 1 ? "is not a number" : "uses an unsupported list-directed form"));
  }
  dim3 grid((unsigned)(((n[2] + TR - 1) / TR) * (uint64_t)n[1]), (n[0] + TR - 1) / TR);
  C2_LAUNCH(ctx, k_cube_transpose, grid, dim3(256), 0, vals, f, n[0], n[1], n[2]);
  add_kernel_bytes(ctx, "k_cube_transpose", 16.0 * (double)npts);
}

uint64_t npts_of(const c2gpu_cube_header &h) { return (uint64_t)h.n[0] * h.n[1] * h.n[2]; }

uint64_t body_bytes(const int n[3], int style) {
  return (uint64_t)n[0] * n[1] * body_layout(n[2], style).row;
}

// Body text of f (device; x fastest unless zfast) into text (device).
void format_body(c2gpu_ctx *ctx, const double *f, const int n[3], int style, int zfast, char *text) {
  const uint64_t npts = (uint64_t)n[0] * n[1] * n[2];
  const double *zf = f;
  if (!zfast) {
    double *t = buf<double>(ctx, "cube_zf", npts);
    // the transpose kernel with the axes swapped: x fastest -> z fastest
    dim3 grid((unsigned)(((n[0] + TR - 1) / TR) * (uint64_t)n[1]), (n[2] + TR - 1) / TR);
    C2_LAUNCH(ctx, k_cube_transpose, grid, dim3(256), 0, f, t, n[2], n[1], n[0]);
    add_kernel_bytes(ctx, "k_cube_transpose", 16.0 * (double)npts);
    zf = t;
  }
  const uint64_t *p5;
  const int16_t *p5s;
  pow5_device(ctx, p5, p5s);
  unsigned long long *stat = buf<unsigned long long>(ctx, "cube_fmt_stat", 2);
  const BodyLayout L = body_layout(n[2], style);
  uint64_t cap = 1ull << 16;
  unsigned long long st[2];
  for (;;) {
    uint64_t *slow = buf<uint64_t>(ctx, "cube_fmt_slow", cap);
    const unsigned long long init[2] = {0ull, 0ull};
    C2_HIP(hipMemcpyAsync(stat, init, sizeof init, hipMemcpyHostToDevice, ctx->stream));
    C2_LAUNCH(ctx, k_cube_format, dim3(grid_for((int64_t)npts, 256)), dim3(256), 0, zf, npts, L, style, p5, p5s,
              text, stat, slow, cap);
    add_kernel_bytes(ctx, "k_cube_format", 8.0 * (double)npts + (double)body_bytes(n, style));
    copy_to_host(ctx, st, stat, sizeof st);
    C2_HIP(hipStreamSynchronize(ctx->stream));
    if (st[1] <= cap) break;
    cap = st[1];
  }
  if (st[1]) {
    C2_LAUNCH(ctx, k_cube_format_slow, dim3(grid_for((int64_t)st[1], 256)), dim3(256), 0, zf,
              buf<uint64_t>(ctx, "cube_fmt_slow", cap), (uint64_t)st[1], L, style, p5, p5s, text, stat);
    if (read_scalar(ctx, stat)) fail(C2GPU_ERR_HIP, "cube: value formatting overflowed its big integers");
  }
}

// Header of writegrid_cube (crystalmod@write.f90:3537-3556): precise
// (E22.14E3 fields) for FMT_E22_14, else the F12.6 / F11.6 form.
std::string grid_header(const c2gpu_cube_header &h, const int *zat, const double *xat, int style) {
  const Pow5Host &p = pow5_host();
  std::string s = "critic2-cube\ncritic2-cube\n";
  char b[64];
  auto e22 = [&](double v) {
    char f[24];
    fmt_value(v, FMT_E22_14, p.t.data(), p.s.data(), f);
    s.append(f, 22);
  };
  const bool precise = style == FMT_E22_14;
  auto line = [&](int iv, const double *v3) {
    snprintf(b, sizeof b, "%5d", iv);
    s += b;
    for (int a = 0; a < 3; a++) {
      if (precise) {
        s += ' ';
        e22(v3[a]);
      } else {
        snprintf(b, sizeof b, "%12.6f", v3[a]);
        s += b;
      }
    }
    s += precise ? " \n" : "\n";
  };
  line(h.nat, h.x0);
  for (int i = 0; i < 3; i++) line(h.n[i], h.step + 3 * i);
  for (int a = 0; a < h.nat; a++) {
    if (precise) {
      snprintf(b, sizeof b, "%5d %5.1f", zat[a], 0.0);
      s += b;
      for (int c = 0; c < 3; c++) {
        s += ' ';
        e22(xat[3 * a + c]);
      }
      s += " \n";
    } else {
      snprintf(b, sizeof b, "%4d%5.1f%11.6f%11.6f%11.6f\n", zat[a], 0.0, xat[3 * a], xat[3 * a + 1], xat[3 * a + 2]);
      s += b;
    }
  }
  return s;
}

// Device output: the caller's (PTR_DEVICE) or a context buffer.
double *out_grid(c2gpu_ctx *ctx, int flags, double *f, uint64_t npts) {
  if (flags & C2GPU_PTR_DEVICE) {
    require_device_ptr(f, "f");
    return f;
  }
  return buf<double>(ctx, "cube_out", npts);
}

// Pinned staging for host->device streaming of files.
struct Pinned {
  void *p[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  size_t bytes = 0;
  explicit Pinned(size_t b) : bytes(b) {
    for (int i = 0; i < 2; i++) {
      C2_HIP(hipHostMalloc(&p[i], b, hipHostMallocDefault));
      C2_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
    }
  }
  ~Pinned() {
    for (int i = 0; i < 2; i++) {
      if (ev[i]) {
        (void)hipEventSynchronize(ev[i]);
        (void)hipEventDestroy(ev[i]);
      }
      if (p[i]) (void)hipHostFree(p[i]);
    }
  }
};

constexpr size_t CHUNK = 64ull << 20;

struct File {
  FILE *fp = nullptr;
  explicit File(const char *path) {
    if (!path) fail(C2GPU_ERR_ARG, "file name is NULL");
    fp = fopen(path, "rb");
    if (!fp) fail(C2GPU_ERR_FORMAT, std::string("Error reading file: ") + path);
  }
  ~File() {
    if (fp) fclose(fp);
  }
  uint64_t size() {
    fseeko(fp, 0, SEEK_END);
    const uint64_t s = (uint64_t)ftello(fp);
    fseeko(fp, 0, SEEK_SET);
    return s;
  }
};

// Reads bytes [from, from + len) of fp into device memory dst, through
// double-buffered pinned chunks.
void stream_to_device(c2gpu_ctx *ctx, FILE *fp, uint64_t from, uint64_t len, uint8_t *dst) {
  if (len == 0) return;
  Pinned pin(len < CHUNK ? len : CHUNK);
  fseeko(fp, (off_t)from, SEEK_SET);
  uint64_t done = 0;
  int which = 0;
  bool used[2] = {false, false};
  while (done < len) {
    const size_t want = (size_t)((len - done) < pin.bytes ? (len - done) : pin.bytes);
    if (used[which]) C2_HIP(hipEventSynchronize(pin.ev[which]));
    if (fread(pin.p[which], 1, want, fp) != want) fail(C2GPU_ERR_FORMAT, "short read");
    C2_HIP(hipMemcpyAsync(dst + done, pin.p[which], want, hipMemcpyHostToDevice, ctx->stream));
    C2_HIP(hipEventRecord(pin.ev[which], ctx->stream));
    used[which] = true;
    done += want;
    which ^= 1;
  }
  C2_HIP(hipStreamSynchronize(ctx->stream));
}

// Reads the text header of a file: the first MiB, growing until the header
// is complete.
void file_text_header(FILE *fp, uint64_t fsize, c2gpu_cube_header &h) {
  std::vector<char> b;
  uint64_t want = 1ull << 20;
  for (;;) {
    const uint64_t n = want < fsize ? want : fsize;
    b.resize(n);
    fseeko(fp, 0, SEEK_SET);
    if (n && fread(b.data(), 1, n, fp) != n) fail(C2GPU_ERR_FORMAT, "short read");
    if (parse_text_header(b.data(), n, h, n == fsize)) return;
    want *= 4;
  }
}

// ---- binary cubes (gfortran unformatted sequential records)

struct RecordReader {
  FILE *fp;
  // Length of the next record (sum of its subrecords) without consuming it
  // is not needed: records are read straight into their destination.
  // Reads one logical record; `sink(ptr, bytes)` receives its data in pieces.
  template <class Sink>
  uint64_t read(Sink &&sink) {
    uint64_t total = 0;
    for (;;) {
      int32_t lead = 0, trail = 0;
      if (fread(&lead, 4, 1, fp) != 1) fail(C2GPU_ERR_FORMAT, "binary cube: end of file");
      const uint64_t len = (uint64_t)(lead < 0 ? -(int64_t)lead : lead);
      sink(len);
      if (fread(&trail, 4, 1, fp) != 1) fail(C2GPU_ERR_FORMAT, "binary cube: end of file");
      const uint64_t tl = (uint64_t)(trail < 0 ? -(int64_t)trail : trail);
      if (tl != len) fail(C2GPU_ERR_FORMAT, "binary cube: inconsistent record markers");
      total += len;
      if (lead >= 0) return total;
    }
  }
};

// Small record into a host buffer: at least `need` bytes.
void small_record(RecordReader &r, void *dst, uint64_t need, const char *what) {
  std::vector<char> tmp;
  r.read([&](uint64_t len) {
    const size_t old = tmp.size();
    tmp.resize(old + len);
    if (len && fread(tmp.data() + old, 1, len, r.fp) != len) fail(C2GPU_ERR_FORMAT, "binary cube: end of file");
  });
  if (tmp.size() < need) fail(C2GPU_ERR_FORMAT, std::string("binary cube: short record (") + what + ")");
  memcpy(dst, tmp.data(), need);
}

void bin_header(RecordReader &r, c2gpu_cube_header &h) {
  char rec1[28];
  small_record(r, rec1, 28, "natoms, origin");
  int32_t nat;
  memcpy(&nat, rec1, 4);
  memcpy(h.x0, rec1 + 4, 24);
  h.ismo = nat < 0;
  h.nat = nat < 0 ? -nat : nat;
  char rec2[84];
  small_record(r, rec2, 84, "grid dimensions, axes");
  memcpy(h.n, rec2, 12);
  memcpy(h.step, rec2 + 12, 72);
  for (int i = 0; i < 3; i++)
    if (h.n[i] < 1 || h.n[i] > 100000) fail(C2GPU_ERR_FORMAT, "binary cube: grid dimension out of range");
  for (int i = 0; i < h.nat; i++) {
    char rec[36];
    small_record(r, rec, 36, "atom");
  }
  h.data_offset = (uint64_t)ftello(r.fp);
}

}  // namespace

void cube_header(const char *path, c2gpu_cube_header *h, int binary) {
  File f(path);
  if (binary) {
    RecordReader r{f.fp};
    bin_header(r, *h);
  } else {
    file_text_header(f.fp, f.size(), *h);
  }
}

void cube_parse_run(c2gpu_ctx *ctx, const char *text, uint64_t len, int flags, c2gpu_cube_header *h, double *f) {
  if (!text || !h) fail(C2GPU_ERR_ARG, "cube: NULL argument");
  memset(h, 0, sizeof(*h));
  const bool dev = flags & C2GPU_PTR_DEVICE;
  if (dev) {
    require_device_ptr(text, "text");
    // header bytes to the host, growing until the header is complete
    std::vector<char> hb;
    uint64_t want = 1ull << 20;
    for (;;) {
      const uint64_t n = want < len ? want : len;
      hb.resize(n);
      copy_to_host(ctx, hb.data(), text, n);
      C2_HIP(hipStreamSynchronize(ctx->stream));
      if (parse_text_header(hb.data(), n, *h, n == len)) break;
      want *= 4;
    }
  } else {
    parse_text_header(text, len, *h, true);
  }
  if (!f) return;  // header only
  const uint64_t npts = npts_of(*h);
  double *out = out_grid(ctx, flags, f, npts);
  const uint64_t off = h->data_offset;
  if (dev) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(text) + off;
    const uint8_t *base = reinterpret_cast<const uint8_t *>(a & ~(uintptr_t)15);
    const uint64_t lo = a & 15;
    parse_values(ctx, base, lo, lo + (len - off), h->n, out);
  } else {
    uint8_t *d = buf<uint8_t>(ctx, "cube_text", len - off + 16);
    C2_HIP(hipMemcpyAsync(d, text + off, len - off, hipMemcpyHostToDevice, ctx->stream));
    parse_values(ctx, d, 0, len - off, h->n, out);
    C2_HIP(hipMemcpyAsync(f, out, npts * 8, hipMemcpyDeviceToHost, ctx->stream));
  }
  C2_HIP(hipStreamSynchronize(ctx->stream));
}

void cube_read_run(c2gpu_ctx *ctx, const char *path, int flags, c2gpu_cube_header *h, double *f) {
  if (!h || !f) fail(C2GPU_ERR_ARG, "cube: NULL argument");
  memset(h, 0, sizeof(*h));
  File file(path);
  const uint64_t fsize = file.size();
  file_text_header(file.fp, fsize, *h);
  const uint64_t npts = npts_of(*h);
  double *out = out_grid(ctx, flags, f, npts);
  const uint64_t len = fsize - h->data_offset;
  uint8_t *d = buf<uint8_t>(ctx, "cube_text", len + 16);
  stream_to_device(ctx, file.fp, h->data_offset, len, d);
  parse_values(ctx, d, 0, len, h->n, out);
  if (!(flags & C2GPU_PTR_DEVICE))
    C2_HIP(hipMemcpyAsync(f, out, npts * 8, hipMemcpyDeviceToHost, ctx->stream));
  C2_HIP(hipStreamSynchronize(ctx->stream));
}

void bincube_read_run(c2gpu_ctx *ctx, const char *path, int flags, c2gpu_cube_header *h, double *f) {
  if (!h || !f) fail(C2GPU_ERR_ARG, "bincube: NULL argument");
  memset(h, 0, sizeof(*h));
  File file(path);
  RecordReader r{file.fp};
  bin_header(r, *h);
  const uint64_t npts = npts_of(*h);
  const uint64_t need = npts * 8;
  const bool dev = flags & C2GPU_PTR_DEVICE;
  if (dev) require_device_ptr(f, "f");
  uint64_t got = 0;
  std::unique_ptr<Pinned> pin;
  int which = 0;
  bool used[2] = {false, false};
  r.read([&](uint64_t len) {
    // the grid record: the first npts doubles go to f, the rest is skipped
    uint64_t take = got < need ? (need - got < len ? need - got : len) : 0;
    if (!dev) {
      if (take && fread(reinterpret_cast<char *>(f) + got, 1, take, file.fp) != take)
        fail(C2GPU_ERR_FORMAT, "binary cube: end of file");
    } else {
      if (!pin && take) pin.reset(new Pinned(need < CHUNK ? need : CHUNK));
      uint64_t done = 0;
      while (done < take) {
        const size_t want = (size_t)((take - done) < pin->bytes ? (take - done) : pin->bytes);
        if (used[which]) C2_HIP(hipEventSynchronize(pin->ev[which]));
        if (fread(pin->p[which], 1, want, file.fp) != want) fail(C2GPU_ERR_FORMAT, "binary cube: end of file");
        C2_HIP(hipMemcpyAsync(reinterpret_cast<char *>(f) + got + done, pin->p[which], want,
                              hipMemcpyHostToDevice, ctx->stream));
        C2_HIP(hipEventRecord(pin->ev[which], ctx->stream));
        used[which] = true;
        done += want;
        which ^= 1;
      }
    }
    got += take;
    if (len > take) fseeko(file.fp, (off_t)(len - take), SEEK_CUR);
  });
  if (got < need) fail(C2GPU_ERR_FORMAT, "binary cube: grid record holds fewer values than n1*n2*n3");
  C2_HIP(hipStreamSynchronize(ctx->stream));
}

void cube_format_run(c2gpu_ctx *ctx, const double *f, const int n[3], int style, int zfast, int flags, char *text,
                     uint64_t cap, uint64_t *len) {
  if (!n || !len) fail(C2GPU_ERR_ARG, "cube: NULL argument");
  if (style < FMT_E22_14 || style > FMT_1P_E13_5) fail(C2GPU_ERR_ARG, "cube: unknown body style");
  for (int a = 0; a < 3; a++)
    if (n[a] < 1) fail(C2GPU_ERR_ARG, "cube: grid dimensions must be positive");
  const uint64_t nb = body_bytes(n, style);
  *len = nb;
  if (!text) return;  // size query
  if (!f) fail(C2GPU_ERR_ARG, "cube: f is NULL");
  if (cap < nb) fail(C2GPU_ERR_ARG, "cube: text buffer holds " + std::to_string(cap) + " bytes, need " +
                                        std::to_string(nb));
  const uint64_t npts = (uint64_t)n[0] * n[1] * n[2];
  if (flags & C2GPU_PTR_DEVICE) {
    require_device_ptr(f, "f");
    require_device_ptr(text, "text");
    format_body(ctx, f, n, style, zfast, text);
  } else {
    double *df = buf<double>(ctx, "cube_fin", npts);
    C2_HIP(hipMemcpyAsync(df, f, npts * 8, hipMemcpyHostToDevice, ctx->stream));
    char *dt = buf<char>(ctx, "cube_text_out", nb);
    format_body(ctx, df, n, style, zfast, dt);
    C2_HIP(hipMemcpyAsync(text, dt, nb, hipMemcpyDeviceToHost, ctx->stream));
  }
  C2_HIP(hipStreamSynchronize(ctx->stream));
}

void cube_write_run(c2gpu_ctx *ctx, const char *path, const c2gpu_cube_header *h, const int *zat,
                    const double *xat, const double *f, int style, int flags) {
  if (!path || !h) fail(C2GPU_ERR_ARG, "cube: NULL argument");
  if (style != FMT_E22_14 && style != FMT_1P_E12_5) fail(C2GPU_ERR_ARG, "cube: grid files use style 0 or 1");
  if (h->nat < 0 || (h->nat > 0 && (!zat || !xat))) fail(C2GPU_ERR_ARG, "cube: atoms missing");
  for (int a = 0; a < 3; a++)
    if (h->n[a] < 1) fail(C2GPU_ERR_ARG, "cube: grid dimensions must be positive");
  const std::string head = grid_header(*h, zat, xat, style);
  char *dt = nullptr;
  uint64_t nb = 0;
  if (f) {  // f NULL: header only (writegrid_cube's onlyheader)
    const uint64_t npts = npts_of(*h);
    nb = body_bytes(h->n, style);
    dt = buf<char>(ctx, "cube_text_out", nb);
    if (flags & C2GPU_PTR_DEVICE) {
      require_device_ptr(f, "f");
      format_body(ctx, f, h->n, style, 0, dt);
    } else {
      double *df = buf<double>(ctx, "cube_fin", npts);
      C2_HIP(hipMemcpyAsync(df, f, npts * 8, hipMemcpyHostToDevice, ctx->stream));
      format_body(ctx, df, h->n, style, 0, dt);
    }
  }
  FILE *fp = fopen(path, "wb");
  if (!fp) fail(C2GPU_ERR_FORMAT, std::string("Error writing file: ") + path);
  struct Closer {
    FILE *fp;
    ~Closer() {
      if (fp) fclose(fp);
    }
  } closer{fp};
  if (fwrite(head.data(), 1, head.size(), fp) != head.size()) fail(C2GPU_ERR_FORMAT, "cube: write failed");
  if (nb) {
    Pinned pin(nb < CHUNK ? nb : CHUNK);
    uint64_t done = 0;
    int which = 0;
    // copy chunk c+1 while chunk c is written
    size_t pending = 0;
    int pend_buf = -1;
    while (done < nb || pend_buf >= 0) {
      size_t want = 0;
      if (done < nb) {
        want = (size_t)((nb - done) < pin.bytes ? (nb - done) : pin.bytes);
        C2_HIP(hipMemcpyAsync(pin.p[which], dt + done, want, hipMemcpyDeviceToHost, ctx->stream));
        C2_HIP(hipEventRecord(pin.ev[which], ctx->stream));
      }
      if (pend_buf >= 0) {
        if (fwrite(pin.p[pend_buf], 1, pending, fp) != pending) fail(C2GPU_ERR_FORMAT, "cube: write failed");
        pend_buf = -1;
      }
      if (want) {
        C2_HIP(hipEventSynchronize(pin.ev[which]));
        pending = want;
        pend_buf = which;
        done += want;
        which ^= 1;
      }
    }
  }
  if (fclose(fp) != 0) {
    closer.fp = nullptr;
    fail(C2GPU_ERR_FORMAT, "cube: write failed");
  }
  closer.fp = nullptr;
}

}  // namespace c2
