// Decimal text -> binary64, correctly rounded (nearest, ties to even), as
// Fortran list-directed input delivers it: the reading side of the cube
// parser (grid3mod@proc.f90:512-572).
//
// A token is [sign] digits [. digits] [exponent], the exponent being a letter
// E/e/D/d/Q/q with an optional signed integer, or a bare signed integer.
// Up to 19 significant digits form an exact 64-bit integer w and the value
// is w * 10^q.  Three tiers, each exact where it answers:
//  1. w < 2^53 and |q| <= 22: one correctly rounded multiply or divide.
//  2. the 192-bit product of w (normalised) and a 128-bit truncation T(q) of
//     5^q gives the binary mantissa; the truncation can only raise the
//     product by less than 2^64, so when the bits below the rounding position
//     are not within that margin of a rounding boundary the result is final.
//  3. otherwise an exact comparison with big integers of the decimal value
//     against the candidate's bounds and midpoint.
// Tokens with more than 40 significant digits, repeat counts (r*c), value
// separators other than blanks (commas, slashes) and non-numeric values are
// reported as unsupported.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace c2 {

static constexpr int POW5_QMIN = -400, POW5_QMAX = 340;  // reading needs up to 310, writing 338
static constexpr int NUM_MAXDIG = 40;

struct DecNum {
  uint64_t w;    // first up-to-19 significant digits
  int q;         // value = w * 10^q  (exact when nsig <= 19)
  int nsig;      // significant digits seen (leading zeros excluded)
  int neg;
  int first;     // offset of the first significant digit (for tier 3)
  int dot;       // offset of the decimal point, or -1
  int mant_end;  // offset just past the mantissa digits
  int exp10;     // explicit exponent
  int len;       // characters consumed
};

__host__ __device__ __forceinline__ bool num_sep(char c) {
  return c == ' ' || c == '\n' || c == '\r' || c == '\t' || c == '\f' || c == '\v';
}
__host__ __device__ __forceinline__ bool num_special(char c) { return c == ',' || c == '/' || c == '*'; }

// Parses one token at s (at most n characters).  Returns 0 on success,
// 1 if not a number, 2 if unsupported (see above).
__host__ __device__ inline int dec_parse(const char *s, int n, DecNum &d) {
  int p = 0;
  d.neg = 0;
  if (p < n && (s[p] == '+' || s[p] == '-')) {
    d.neg = s[p] == '-';
    p++;
  }
  uint64_t w = 0;
  int nsig = 0, ndig = 0, dropped = 0, after_dot = 0;
  d.first = -1;
  d.dot = -1;
  for (; p < n; p++) {
    const char c = s[p];
    if (c >= '0' && c <= '9') {
      ndig++;
      if (d.dot >= 0) after_dot++;
      if (nsig == 0 && c == '0') continue;
      if (nsig == 0) d.first = p;
      nsig++;
      if (nsig <= 19)
        w = w * 10 + (uint64_t)(c - '0');
      else
        dropped++;
    } else if (c == '.' && d.dot < 0) {
      d.dot = p;
    } else {
      break;
    }
  }
  if (ndig == 0) return 1;
  d.mant_end = p;
  int e = 0;
  if (p < n && !num_sep(s[p])) {
    const char c = s[p];
    if (num_special(c)) return 2;  // r*c repeat counts, null values, slash
    int q = p;
    if (c == 'E' || c == 'e' || c == 'D' || c == 'd' || c == 'Q' || c == 'q') q++;
    else if (c != '+' && c != '-') return 1;
    int es = 1;
    if (q < n && (s[q] == '+' || s[q] == '-')) {
      es = s[q] == '-' ? -1 : 1;
      q++;
    }
    int ne = 0;
    while (q < n && s[q] >= '0' && s[q] <= '9') {
      if (e < 100000) e = e * 10 + (s[q] - '0');
      q++;
      ne++;
    }
    if (ne == 0) return 1;
    e *= es;
    p = q;
  }
  if (p < n && !num_sep(s[p])) return num_special(s[p]) ? 2 : 1;
  if (nsig > NUM_MAXDIG) return 2;
  d.w = w;
  d.nsig = nsig;
  d.exp10 = e;
  d.q = e - after_dot + dropped;
  d.len = p;
  return 0;
}

// ---- 64x64 and 64x128 products

__host__ __device__ __forceinline__ void mul64(uint64_t a, uint64_t b, uint64_t &hi, uint64_t &lo) {
  const uint64_t a0 = a & 0xffffffffu, a1 = a >> 32, b0 = b & 0xffffffffu, b1 = b >> 32;
  const uint64_t p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
  const uint64_t mid = (p00 >> 32) + (p01 & 0xffffffffu) + (p10 & 0xffffffffu);
  lo = (p00 & 0xffffffffu) | (mid << 32);
  hi = p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}

__host__ __device__ __forceinline__ int clz64(uint64_t x) {
  int n = 0;
  if (x == 0) return 64;
  while (!(x & (1ull << 63))) {
    x <<= 1;
    n++;
  }
  return n;
}

// ---- tier 3: exact big-integer comparison

static constexpr int BIG_LIMBS = 40;  // 1280 bits: the largest operand is below 1000 bits

struct Big {
  uint32_t l[BIG_LIMBS];
  int n;  // limbs in use
};

__host__ __device__ inline void big_set(Big &b, uint64_t v) {
  b.n = 0;
  while (v) {
    b.l[b.n++] = (uint32_t)v;
    v >>= 32;
  }
}
__host__ __device__ inline bool big_muladd(Big &b, uint32_t m, uint32_t a) {
  uint64_t carry = a;
  for (int i = 0; i < b.n; i++) {
    const uint64_t t = (uint64_t)b.l[i] * m + carry;
    b.l[i] = (uint32_t)t;
    carry = t >> 32;
  }
  if (carry) {
    if (b.n >= BIG_LIMBS) return false;
    b.l[b.n++] = (uint32_t)carry;
  }
  return true;
}
__host__ __device__ inline bool big_pow5(Big &b, int k) {
  while (k >= 13) {  // 5^13 < 2^32
    if (!big_muladd(b, 1220703125u, 0)) return false;
    k -= 13;
  }
  uint32_t m = 1;
  while (k-- > 0) m *= 5;
  return big_muladd(b, m, 0);
}
__host__ __device__ inline bool big_shl(Big &b, int k) {
  if (b.n == 0) return true;
  const int words = k >> 5, bits = k & 31;
  if (b.n + words + 1 > BIG_LIMBS) return false;
  int nn = b.n + words;
  b.l[nn] = 0;
  for (int i = b.n - 1; i >= 0; i--) {
    const uint32_t v = b.l[i];
    if (bits) {
      b.l[i + words + 1] |= v >> (32 - bits);
      b.l[i + words] = v << bits;
    } else {
      b.l[i + words] = v;
    }
  }
  for (int i = 0; i < words; i++) b.l[i] = 0;
  b.n = nn + 1;
  while (b.n > 0 && b.l[b.n - 1] == 0) b.n--;
  return true;
}
__host__ __device__ inline int big_cmp(const Big &a, const Big &b) {
  if (a.n != b.n) return a.n < b.n ? -1 : 1;
  for (int i = a.n - 1; i >= 0; i--)
    if (a.l[i] != b.l[i]) return a.l[i] < b.l[i] ? -1 : 1;
  return 0;
}

// sign(D * 10^q10 - K * 2^eb), D the significant digits of the token as an
// integer (all of them), q10 their decimal exponent.  ok = false if the big
// integers would overflow.
__host__ __device__ inline int cmp_dec_bin(const char *s, const DecNum &d, int q10, uint64_t K, int eb, bool &ok) {
  Big a, b;
  a.n = 0;
  ok = true;
  int taken = 0;
  for (int p = d.first; p < d.mant_end && taken < d.nsig; p++) {
    const char c = s[p];
    if (c < '0' || c > '9') continue;
    if (a.n == 0 && c == '0' && taken == 0) continue;
    if (a.n == 0) big_set(a, (uint64_t)(c - '0'));
    else ok = ok && big_muladd(a, 10u, (uint32_t)(c - '0'));
    taken++;
  }
  big_set(b, K);
  if (q10 >= 0) ok = ok && big_pow5(a, q10);
  else ok = ok && big_pow5(b, -q10);
  const int dsh = q10 - eb;  // powers of two: a carries 2^q10, b carries 2^eb
  if (dsh >= 0) ok = ok && big_shl(a, dsh);
  else ok = ok && big_shl(b, -dsh);
  return ok ? big_cmp(a, b) : 0;
}

// ---- conversion

__host__ __device__ __forceinline__ double bits_to_double(uint64_t v) {
  union {
    uint64_t u;
    double d;
  } x;
  x.u = v;
  return x.d;
}

// Assembles the double from a 53-bit significand m (with the implicit bit
// for normal numbers) whose least significant bit weighs 2^eb.
__host__ __device__ __forceinline__ uint64_t pack_double(uint64_t m, int eb, int neg) {
  uint64_t bits;
  if (m == 0) {
    bits = 0;
  } else if (m < (1ull << 52)) {  // subnormal (eb == -1074)
    bits = m;
  } else {
    const int be = eb + 52 + 1023;  // biased exponent of the leading bit
    if (be >= 2047) bits = 0x7ff0000000000000ull;
    else bits = ((uint64_t)be << 52) | (m & ((1ull << 52) - 1));
  }
  return bits | (neg ? (1ull << 63) : 0);
}

// T(q): 128-bit truncation of 5^q scaled into [2^127, 2^128), and its
// binary scale s: 5^q ~= T * 2^s, T <= 5^q * 2^-s < T + 1.  pow5: (hi, lo)
// per q from POW5_QMIN; pow5s: s per q.
//
// Tiers 1 and 2 (no access to the text): returns 0 with out set, or 3 when
// the exact tier must decide, with the candidate significand m (floor of the
// value at scale 2^eb, up to +-1) and eb.
__host__ __device__ inline int dec_fast(const DecNum &d, const uint64_t *pow5, const int16_t *pow5s, double &out,
                                        uint64_t &m_out, int &eb_out) {
  const int neg = d.neg;
  const uint64_t sign = neg ? (1ull << 63) : 0;
  if (d.w == 0) {
    out = bits_to_double(sign);
    return 0;
  }
  const bool exact_w = d.nsig <= 19;
  const int q = d.q;
  // tier 1
  if (exact_w && d.w < (1ull << 53) && q >= -22 && q <= 22) {
    double p10 = 1.0;  // 10^|q|: every partial product is an exact double up to 10^22
    for (int i = q < 0 ? -q : q; i > 0; i--) p10 *= 10.0;
    const double v = q >= 0 ? (double)d.w * p10 : (double)d.w / p10;
    out = neg ? -v : v;
    return 0;
  }
  if (q > 309) {  // w >= 1: at least 10^310
    out = bits_to_double(0x7ff0000000000000ull | sign);
    return 0;
  }
  if (q < -364) {  // w < 10^19: below 10^-345, under half the least subnormal
    out = bits_to_double(sign);
    return 0;
  }
  // tier 2: P = w' T (192 bits), value ~= P * 2^(s5 + q - lz)
  const int lz = clz64(d.w);
  const uint64_t wn = d.w << lz;
  const uint64_t th = pow5[2 * (q - POW5_QMIN)], tl = pow5[2 * (q - POW5_QMIN) + 1];
  const int s5 = pow5s[q - POW5_QMIN];
  uint64_t h1, l1, h2, l2;
  mul64(wn, th, h1, l1);
  mul64(wn, tl, h2, l2);
  const uint64_t mid = l1 + h2;
  const uint64_t hi = h1 + (mid < l1 ? 1 : 0);
  const uint64_t lo = l2;
  const int top = (hi >> 63) ? 191 : 190;  // leading bit of P
  const int ev = top + s5 + q - lz;        // its binary exponent
  int width = 53;                          // significand bits (fewer when subnormal)
  if (ev < -1022) width = 53 - (-1022 - ev);
  const int cut = top - width + 1;  // kept bits of P: [cut, top]; cut >= 138
  const int eb = cut + s5 + q - lz; // weight of the last kept bit
  const int nb = cut - 1 - 64;      // the round bit as a bit of U = P >> 64; nb >= 73
  const uint64_t Uh = hi, Ul = mid;
  auto u_shr = [&](int b) -> uint64_t {  // low 64 bits of U >> b
    if (b >= 128) return 0;
    if (b >= 64) return Uh >> (b - 64);
    if (b == 0) return Ul;
    return (Ul >> b) | (Uh << (64 - b));
  };
  uint64_t m = u_shr(nb + 1);
  const uint64_t rbit = u_shr(nb) & 1;
  // bits [0, nb) of U: the sticky bits above P's uncertain low word
  uint64_t fl = Ul, fh = Uh, ml = ~0ull, mh = ~0ull;
  if (nb < 128) {
    mh = nb > 64 ? ((1ull << (nb - 64)) - 1) : 0;
    if (nb < 64) ml = (1ull << nb) - 1;
    fl &= ml;
    fh &= mh;
  }
  const bool f_zero = (fl | fh) == 0, f_ones = (fl == ml) && (fh == mh);
  const bool texact = q >= 0 && q <= 55;  // 5^q < 2^128: T and P are exact
  int decided = 0;                       // 1 down, 2 up, 3 exact tie
  if (exact_w && texact) {
    decided = !rbit ? 1 : ((f_zero && lo == 0) ? 3 : 2);
  } else if (exact_w) {
    // true P lies in [P, P + 2^64): a carry reaches bit 64 at most once
    if (!rbit) decided = f_ones ? 0 : 1;
    else decided = f_zero ? 0 : 2;
  }
  if (decided) {
    if (decided == 2 || (decided == 3 && (m & 1))) m++;
    int e2 = eb;
    if (width == 53 && m == (1ull << 53)) {  // carry into the next binade
      m >>= 1;
      e2++;
    }
    out = bits_to_double(pack_double(m, e2, neg) | 0);
    return 0;
  }
  m_out = m;
  eb_out = eb;
  return 3;
}

// Tier 3: exact big-integer comparisons of the token's digits at s against
// the candidate's neighbours and midpoint.  Returns 0, or 2 if the big
// integers ran out of room.
__host__ __device__ inline int dec_exact(const char *s, const DecNum &d, uint64_t m, int eb, double &out) {
  const int neg = d.neg;
  const int q10 = d.exp10 - (d.dot >= 0 ? (d.mant_end - d.dot - 1) : 0);
  bool ok = true;
  uint64_t mm = m;
  int ebx = eb;
  for (int pass = 0; pass < 4; pass++) {
    for (int it = 0; it < 6; it++) {
      if (cmp_dec_bin(s, d, q10, mm, ebx, ok) < 0) {
        if (!ok) return 2;
        if (mm == 0) break;
        mm--;
        continue;
      }
      if (!ok) return 2;
      if (cmp_dec_bin(s, d, q10, mm + 1, ebx, ok) >= 0) {
        if (!ok) return 2;
        mm++;
        continue;
      }
      if (!ok) return 2;
      break;
    }
    if (mm >= (1ull << 53)) {  // one more significant bit than assumed
      mm >>= 1;
      ebx++;
      continue;
    }
    if (mm < (1ull << 52) && ebx > -1074) {  // one fewer
      mm <<= 1;
      ebx--;
      continue;
    }
    break;
  }
  const int c = cmp_dec_bin(s, d, q10, 2 * mm + 1, ebx - 1, ok);
  if (!ok) return 2;
  if (c > 0 || (c == 0 && (mm & 1))) mm++;
  if (mm == (1ull << 53)) {
    mm >>= 1;
    ebx++;
  }
  out = bits_to_double(pack_double(mm, ebx, neg));
  return 0;
}

// Tiers 1-3.  Returns 0 (out set), or 2 if the exact tier ran out of room.
__host__ __device__ inline int dec_to_double(const char *s, const DecNum &d, const uint64_t *pow5,
                                             const int16_t *pow5s, double &out) {
  uint64_t m;
  int eb;
  if (dec_fast(d, pow5, pow5s, out, m, eb) == 0) return 0;
  return dec_exact(s, d, m, eb, out);
}

// ---- host: the 5^q table used by tier 2

inline int big_bitlen(const Big &b) {
  if (b.n == 0) return 0;
  uint32_t top = b.l[b.n - 1];
  int k = 0;
  while (top) {
    top >>= 1;
    k++;
  }
  return 32 * (b.n - 1) + k;
}
inline int big_bit(const Big &b, int pos) {
  if (pos < 0 || pos >= 32 * b.n) return 0;
  return (b.l[pos >> 5] >> (pos & 31)) & 1;
}
inline void big_div5(Big &b) {
  uint64_t rem = 0;
  for (int i = b.n - 1; i >= 0; i--) {
    const uint64_t cur = (rem << 32) | b.l[i];
    b.l[i] = (uint32_t)(cur / 5);
    rem = cur % 5;
  }
  while (b.n > 0 && b.l[b.n - 1] == 0) b.n--;
}
// Bits [pos, pos + 128) of b as (hi, lo); bits below zero read as zero.
inline void big_window128(const Big &b, int pos, uint64_t &hi, uint64_t &lo) {
  hi = lo = 0;
  for (int i = 127; i >= 0; i--) {
    const uint64_t bit = (uint64_t)big_bit(b, pos + i);
    if (i >= 64) hi |= bit << (i - 64);
    else lo |= bit << i;
  }
}

// pow5[2 (q - POW5_QMIN)] = high word of T(q), [+1] = low word;
// pow5s[q - POW5_QMIN] = s(q), for POW5_QMIN <= q <= POW5_QMAX, where
// 5^q = (T + delta) 2^s with 2^127 <= T < 2^128, 0 <= delta < 1.
inline void pow5_table(uint64_t *pow5, int16_t *pow5s) {
  for (int q = POW5_QMIN; q <= POW5_QMAX; q++) {
    const int n = q < 0 ? -q : q;
    Big b;
    big_set(b, 1);
    big_pow5(b, n);
    const int k = big_bitlen(b) - 1;  // 2^k <= 5^n < 2^(k+1)
    uint64_t hi, lo;
    int sc;
    if (q >= 0) {  // T = floor(5^q / 2^(k-127))
      big_window128(b, k - 127, hi, lo);
      sc = k - 127;
    } else {  // T = floor(2^(k+128) / 5^n)
      Big x;
      big_set(x, 1);
      big_shl(x, k + 128);
      for (int i = 0; i < n; i++) big_div5(x);
      big_window128(x, 0, hi, lo);
      sc = -k - 128;
    }
    pow5[2 * (q - POW5_QMIN)] = hi;
    pow5[2 * (q - POW5_QMIN) + 1] = lo;
    pow5s[q - POW5_QMIN] = (int16_t)sc;
  }
}

}  // namespace c2
