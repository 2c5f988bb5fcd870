// binary64 -> Fortran E-edited text, correctly rounded (nearest, ties to
// even on the exact binary value), as gfortran's formatted WRITE gives it
// for the three cube body formats of critic2:
//   FMT_E22_14  E22.14E3      writegrid_cube with precisecube
//                             (crystalmod@write.f90:3563)
//   FMT_1P_E12_5  1p,E12.5E3  writegrid_cube without (:3565)
//   FMT_1P_E13_5  1p,e13.5e3  nciplot's write_cube_body (nci@proc.f90:927)
// and the two-digit-exponent field of nciplot's .dat rows:
//   FMT_1P_E15_7  1p,E15.7    (nci@proc.f90:638); an exponent beyond 99
//                             drops the letter: d.dddddddE-05, d.ddddddd-100
// P significant digits (14 or 6) of x: n = round(x 10^k), k = P-1-E with
// E = floor(log10 x), from the 192-bit product of the normalised
// significand and the same 128-bit truncation T(k) of 5^k the decimal
// reader uses; the truncation lowers the product by less than 2^64, so the
// rounding is final unless the discarded bits lie within that margin of
// one half, where an exact big-integer comparison decides.
#pragma once
#include "number.hpp"

namespace c2 {

enum { FMT_E22_14 = 0, FMT_1P_E12_5 = 1, FMT_1P_E13_5 = 2, FMT_1P_E15_7 = 3 };

__host__ __device__ __forceinline__ constexpr int fmt_width(int style) {
  return style == FMT_E22_14 ? 22 : (style == FMT_1P_E12_5 ? 12 : (style == FMT_1P_E13_5 ? 13 : 15));
}
__host__ __device__ __forceinline__ constexpr int fmt_digits(int style) {
  return style == FMT_E22_14 ? 14 : (style == FMT_1P_E15_7 ? 8 : 6);
}

__host__ __device__ __forceinline__ uint64_t pow10_u64(int p) {
  uint64_t v = 1;
  for (int i = 0; i < p; i++) v *= 10;
  return v;
}

// sign(2 m 2^e 10^k - (2 I + 1)): x = m 2^e against the midpoint above I
// at scale 10^-k.  ok = false on big-integer overflow.
__host__ __device__ inline int cmp_mid(uint64_t m, int e, int k, uint64_t I, bool &ok) {
  Big a, b;
  big_set(a, m);
  big_set(b, 2 * I + 1);
  ok = true;
  if (k >= 0) ok = ok && big_pow5(a, k);
  else ok = ok && big_pow5(b, -k);
  const int sh = e + k + 1;  // a carries 2^(e+k+1)
  if (sh >= 0) ok = ok && big_shl(a, sh);
  else ok = ok && big_shl(b, -sh);
  return ok ? big_cmp(a, b) : 0;
}

// P significant digits of x > 0 (finite): n in [10^(P-1), 10^P) and the
// decimal exponent E with x ~= n 10^(E-P+1).  Returns 0; 1 when the exact
// tier is needed and EXACT is false; 2 on an internal overflow of the exact
// tier (not reachable for binary64).
template <bool EXACT = true>
__host__ __device__ inline int sig_digits(double x, int P, const uint64_t *pow5, const int16_t *pow5s,
                                          uint64_t &n, int &E) {
  union {
    double d;
    uint64_t u;
  } cv;
  cv.d = x;
  const int be = (int)((cv.u >> 52) & 0x7ff);
  uint64_t m = cv.u & ((1ull << 52) - 1);
  int e;
  if (be) {
    m |= 1ull << 52;
    e = be - 1075;
  } else {
    e = -1074;
  }
  const int lz = clz64(m);
  const uint64_t mn = m << lz;
  const int en = e - lz;  // x = mn 2^en, mn in [2^63, 2^64)
  int Eg = (int)floor((double)(en + 63) * 0.30102999566398120);  // E or E - 1
  const uint64_t lo10 = pow10_u64(P - 1), hi10 = pow10_u64(P);
  for (int pass = 0; pass < 3; pass++) {
    const int k = P - 1 - Eg;
    const uint64_t th = pow5[2 * (k - POW5_QMIN)], tl = pow5[2 * (k - POW5_QMIN) + 1];
    const int s5 = pow5s[k - POW5_QMIN];
    // Q = mn * T, 192 bits as (q2, q1, q0)
    uint64_t h1, l1, h2, l2;
    mul64(mn, th, h1, l1);
    mul64(mn, tl, h2, l2);
    const uint64_t q0 = l2;
    const uint64_t q1 = l1 + h2;
    const uint64_t q2 = h1 + (q1 < l1 ? 1 : 0);
    const int fb = -(en + k + s5);  // fraction bits: y = Q 2^-fb; fb in (64, 192)
    // integer part I = Q >> fb and fraction F = Q mod 2^fb against half
    uint64_t I;
    if (fb >= 128) I = q2 >> (fb - 128);
    else I = (q2 << (128 - fb)) | (q1 >> (fb - 64));
    if (I >= hi10 && pass < 2) {  // E was underestimated
      Eg++;
      continue;
    }
    if (I < lo10 && pass < 2) {
      Eg--;
      continue;
    }
    // F's bits: fb-1 is the half bit
    uint64_t f2, f1;  // F >> 64 split: bits [64, fb) as (f2:f1) with f1 = bits 64..127
    f1 = q1;
    f2 = q2;
    if (fb < 128) {
      f1 &= (1ull << (fb - 64)) - 1;
      f2 = 0;
    } else if (fb < 192) {
      f2 &= fb == 128 ? 0 : ((1ull << (fb - 128)) - 1);
    }
    // half = 2^(fb-1); compare F (f2:f1:q0) with it, margin 2^64 on the upper side
    const int hb = fb - 1;
    uint64_t h2w = 0, h1w = 0;  // half as (word2, word1), word0 = 0
    if (hb >= 128) h2w = 1ull << (hb - 128);
    else h1w = 1ull << (hb - 64);
    // F_hi = F >> 64 = (f2:f1); half_hi = (h2w:h1w)
    int c;  // sign(F_hi - half_hi)
    if (f2 != h2w) c = f2 < h2w ? -1 : 1;
    else if (f1 != h1w) c = f1 < h1w ? -1 : 1;
    else c = 0;
    const bool texact = k >= 0 && k <= 55;
    int up;  // 1 up, 0 down, -1 undecided
    if (c > 0) {
      up = 1;  // F > half (and the true product only larger)
    } else if (c < 0) {
      // F_hi < half_hi: F + 2^64 <= half unless F_hi == half_hi - 1
      uint64_t d1 = h1w - f1, d2 = h2w - f2 - (h1w < f1 ? 1 : 0);
      up = (d2 != 0 || d1 > 1) ? 0 : (texact ? 0 : -1);
    } else {
      if (q0 != 0) up = 1;
      else up = texact ? 2 : -1;  // 2: exact tie
    }
    if (up == 2) up = (I & 1) ? 1 : 0;
    if (up < 0) {
      if (!EXACT) return 1;
      bool ok;
      const int cm = cmp_mid(mn, en, k, I, ok);
      if (!ok) return 2;
      up = cm > 0 ? 1 : (cm < 0 ? 0 : (int)(I & 1));
    }
    n = I + (uint64_t)up;
    E = Eg;
    if (n >= hi10) {
      n /= 10;  // 10^P -> 10^(P-1)
      E++;
    }
    return 0;
  }
  return 2;
}

// Writes exactly fmt_width(style) characters of x at out.  Returns 0; 1
// (nothing written) when the exact tier is needed and EXACT is false; 2 on
// an internal overflow.  The fields have fixed layouts (sign slot, digits,
// exponent), so every character goes to a position known at compile time.
template <bool EXACT, int STYLE>
__host__ __device__ inline int fmt_value_s(double x, const uint64_t *pow5, const int16_t *pow5s, char *out) {
  constexpr int W = fmt_width(STYLE);
  constexpr int P = fmt_digits(STYLE);
  constexpr bool p1 = STYLE != FMT_E22_14;
  union {
    double d;
    uint64_t u;
  } cv;
  cv.d = x;
  const bool neg = cv.u >> 63;
  const uint64_t mag = cv.u & 0x7fffffffffffffffull;
  if (mag >= 0x7ff0000000000000ull) {  // Infinity / NaN, right-justified
    const bool nan = mag > 0x7ff0000000000000ull;
    const char *w = nan ? "NaN" : (neg ? "-Infinity" : "Infinity");
    const int len = nan ? 3 : (neg ? 9 : 8);
    for (int i = 0; i < W; i++) out[i] = i < W - len ? ' ' : w[i - (W - len)];
    return 0;
  }
  if (STYLE == FMT_1P_E12_5 && neg) {  // no room for the sign: W asterisks
    for (int i = 0; i < W; i++) out[i] = '*';
    return 0;
  }
  uint64_t n = 0;
  int E = 0;
  if (mag != 0) {
    cv.u = mag;
    const int rc = sig_digits<EXACT>(cv.d, P, pow5, pow5s, n, E);
    if (rc) return rc;
  }
  const int ex = mag ? (p1 ? E : E + 1) : 0;
  // layout: [blank] [sign slot] lead digits, exponent
  constexpr int B = STYLE == FMT_1P_E15_7 ? 1 : 0;  // leading blank (field wider than needed)
  constexpr int S = STYLE == FMT_1P_E12_5 ? 0 : 1;  // sign slot width
  if (B) out[0] = ' ';
  if (S) out[B] = neg ? '-' : ' ';
  // digits from two 32-bit halves (n < 10^14): 32-bit divisions by the
  // constant 10 are a multiply-high and a shift, 64-bit ones are not
  uint32_t lo7 = (uint32_t)(n % 10000000ull), hi7 = (uint32_t)(n / 10000000ull);
  char dg[P];
#pragma unroll
  for (int i = P - 1; i >= 0; i--) {
    if (i == P - 8) lo7 = hi7;
    dg[i] = (char)('0' + lo7 % 10u);
    lo7 /= 10u;
  }
  int o = B + S;
  if (p1) {  // d.ddddd
    out[o++] = dg[0];
    out[o++] = '.';
#pragma unroll
    for (int i = 1; i < P; i++) out[o++] = dg[i];
  } else {  // 0.dddd...
    out[o++] = '0';
    out[o++] = '.';
#pragma unroll
    for (int i = 0; i < P; i++) out[o++] = dg[i];
  }
  const uint32_t a = (uint32_t)(ex < 0 ? -ex : ex);
  if (STYLE == FMT_1P_E15_7) {  // Ew.d: E+dd, or +ddd past 99
    if (a <= 99u) {
      out[o++] = 'E';
      out[o++] = ex < 0 ? '-' : '+';
    } else {
      out[o++] = ex < 0 ? '-' : '+';
      out[o++] = (char)('0' + a / 100u);
    }
    out[o++] = (char)('0' + (a / 10u) % 10u);
    out[o++] = (char)('0' + a % 10u);
    return 0;
  }
  out[o++] = 'E';
  out[o++] = ex < 0 ? '-' : '+';
  out[o++] = (char)('0' + a / 100u);
  out[o++] = (char)('0' + (a / 10u) % 10u);
  out[o++] = (char)('0' + a % 10u);
  return 0;
}

template <bool EXACT = true>
__host__ __device__ inline int fmt_value(double x, int style, const uint64_t *pow5, const int16_t *pow5s,
                                         char *out) {
  if (style == FMT_E22_14) return fmt_value_s<EXACT, FMT_E22_14>(x, pow5, pow5s, out);
  if (style == FMT_1P_E12_5) return fmt_value_s<EXACT, FMT_1P_E12_5>(x, pow5, pow5s, out);
  if (style == FMT_1P_E13_5) return fmt_value_s<EXACT, FMT_1P_E13_5>(x, pow5, pow5s, out);
  return fmt_value_s<EXACT, FMT_1P_E15_7>(x, pow5, pow5s, out);
}

}  // namespace c2
