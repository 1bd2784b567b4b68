// Decimal text to IEEE binary32 / binary64, as glibc's strtof / strtod read
// it in the C locale, for the depth-file parser (host and device).
//
// The depth file's columns are read with strtof (means) and strtod
// (contigLen) (src/rabbitbin.cpp:2890-2925).  Both round correctly (to
// nearest, ties to even).  This conversion is exact by construction for the
// inputs it accepts, and reports the rest as unsupported instead of guessing:
//   * accepted: leading white space (' ', \t, \n, \v, \f, \r, as strtof
//     skips it), an optional sign, then decimal digits with an optional '.'
//     and an optional exponent (e|E, optional sign, digits; an 'e' without
//     digits is not consumed), or "inf" / "infinity" / "nan" in any case;
//     text that is none of these converts nothing (value 0, end = start);
//   * unsupported: hexadecimal floats, "nan(...)" payloads, more than 19
//     significant digits, and magnitudes whose exact conversion needs more
//     than 128-bit integers (below about 1e-30 for binary32, 1e-21 for
//     binary64 given 19 digits, or above 1e38 for binary64); none of these
//     occurs in a depth file, and the parser fails loudly on them.
// Method: the digits give an integer M (< 10^19) and a power of ten 10^e.
// For e >= 0 the value is the integer M * 10^e, rounded at the format's
// precision with its dropped bits as guard and sticky; for e < 0 it is
// floor(M * 2^s / 10^-e) with s chosen so that the quotient holds two more
// bits than the format's significand, the remainder joining the sticky bit.
// Both are exact integer arithmetic, so the rounding is the correct one.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define RBG_DEC_FN __host__ __device__ __forceinline__
#else
#define RBG_DEC_FN inline
#endif

namespace rbg {
namespace dec {

typedef unsigned __int128 u128;

enum Status : int { kOk = 0, kUnsupported = 1 };

RBG_DEC_FN int bitlen128(u128 x) {
  const uint64_t hi = (uint64_t)(x >> 64), lo = (uint64_t)x;
  if (hi) return 128 - __builtin_clzll(hi);
  if (lo) return 64 - __builtin_clzll(lo);
  return 0;
}

RBG_DEC_FN u128 pow10_128(int k) {  // k in [0, 38]
  u128 p = 1;
  for (int i = 0; i < k; ++i) p *= 10u;
  return p;
}

// a * b for a < 2^128, b < 2^64; false when the product needs more than 128 bits.
RBG_DEC_FN bool mul_checked(u128 a, uint64_t b, u128& out) {
  const u128 lo = (u128)(uint64_t)a * b;
  const u128 hi = (u128)(uint64_t)(a >> 64) * b;
  if ((uint64_t)(hi >> 64)) return false;
  out = (hi << 64) + lo;
  return out >= lo;
}

// floor(num / den) and num % den, den > 0, by shift-and-subtract over the
// quotient's bits only.
RBG_DEC_FN void divmod128(u128 num, u128 den, u128& q, u128& r) {
  q = 0;
  r = num;
  const int ln = bitlen128(num), ld = bitlen128(den);
  if (ln < ld) return;
  int sh = ln - ld;
  u128 d = den << sh;
  for (; sh >= 0; --sh, d >>= 1) {
    q <<= 1;
    if (r >= d) {
      r -= d;
      q |= 1u;
    }
  }
}

// Round value = (x + frac) * 2^e2, frac in [0, 1) with frac > 0 iff sticky, to
// a p-bit significand (ties to even); writes the IEEE bits (exponent bias
// `bias`, emax the largest unbiased exponent).  Returns kUnsupported below
// the normal range (never reached by the accepted inputs).
RBG_DEC_FN int round_pack(u128 x, bool sticky, int e2, int p, int bias, int emax, uint64_t& bits) {
  const int L = bitlen128(x);  // x > 0
  int exp2 = L - 1 + e2;       // unbiased exponent of the leading bit
  uint64_t mant;
  if (L <= p) {
    if (sticky) return kUnsupported;  // not reached: a quotient always has p + 2 bits
    mant = (uint64_t)x << (p - L);
  } else {
    const int drop = L - p;
    mant = (uint64_t)(x >> drop);
    const u128 rest = x & ((((u128)1) << drop) - 1u);
    const u128 half = ((u128)1) << (drop - 1);
    const bool up = rest > half || (rest == half && (sticky || (mant & 1u)));
    if (up && (++mant >> p)) {  // carried into a new leading bit
      mant >>= 1;
      ++exp2;
    }
  }
  if (exp2 > emax) {  // overflow: infinity, as strtof / strtod return HUGE_VAL
    bits = (uint64_t)(2 * bias + 1) << (p - 1);
    return kOk;
  }
  if (exp2 < 1 - bias) return kUnsupported;  // subnormal range
  bits = ((uint64_t)(exp2 + bias) << (p - 1)) | (mant & ((1ull << (p - 1)) - 1u));
  return kOk;
}

RBG_DEC_FN bool is_space(uint8_t c) { return c == ' ' || (c >= '\t' && c <= '\r'); }
RBG_DEC_FN uint8_t lower(uint8_t c) { return (c >= 'A' && c <= 'Z') ? (uint8_t)(c + 32) : c; }

// Parse one number at s (not past end).  p = 24 (binary32) or 53 (binary64).
// On kOk: bits = the IEEE bits, *consumed = bytes read from s (0 when nothing
// converts, as strtof then sets its end pointer to s).
RBG_DEC_FN int parse_decimal(const uint8_t* s, const uint8_t* end, int p, uint64_t& bits, uint64_t* consumed) {
  const int bias = p == 24 ? 127 : 1023;
  const int emax = bias;
  const uint64_t sign_bit = p == 24 ? (1ull << 31) : (1ull << 63);
  const uint64_t inf_bits = (uint64_t)(2 * bias + 1) << (p - 1);
  const uint64_t qnan_bits = inf_bits | (1ull << (p - 2));
  const uint8_t* c = s;
  bits = 0;
  *consumed = 0;
  while (c < end && is_space(*c)) ++c;
  bool neg = false;
  if (c < end && (*c == '+' || *c == '-')) {
    neg = *c == '-';
    ++c;
  }
  // inf / infinity / nan
  if (c < end && (lower(*c) == 'i' || lower(*c) == 'n')) {
    const char* word = lower(*c) == 'i' ? "inf" : "nan";
    int k = 0;
    while (k < 3 && c + k < end && lower(c[k]) == (uint8_t)word[k]) ++k;
    if (k < 3) return kOk;  // no conversion
    if (word[0] == 'i') {
      const char* rest = "inity";
      int t = 0;
      while (t < 5 && c + 3 + t < end && lower(c[3 + t]) == (uint8_t)rest[t]) ++t;
      c += t == 5 ? 8 : 3;
      bits = inf_bits | (neg ? sign_bit : 0);
    } else {
      if (c + 3 < end && c[3] == '(') return kUnsupported;  // NaN payloads
      c += 3;
      bits = qnan_bits | (neg ? sign_bit : 0);
    }
    *consumed = (uint64_t)(c - s);
    return kOk;
  }
  if (c + 1 < end && c[0] == '0' && (c[1] == 'x' || c[1] == 'X')) return kUnsupported;  // hex floats
  uint64_t M = 0;
  int nd = 0;         // significant digits kept in M
  int exp10 = 0;      // value = M * 10^exp10 (before the explicit exponent)
  bool any = false;   // at least one digit
  bool dropped = false;
  bool point = false;
  for (; c < end; ++c) {
    const uint8_t ch = *c;
    if (ch == '.') {
      if (point) break;
      point = true;
      continue;
    }
    if (ch < '0' || ch > '9') break;
    any = true;
    const uint32_t d = ch - '0';
    if (M == 0 && d == 0) {  // leading zeros
      if (point) --exp10;
      continue;
    }
    if (nd < 19) {
      M = M * 10u + d;
      ++nd;
      if (point) --exp10;
    } else {
      if (d) dropped = true;
      if (!point) ++exp10;
    }
  }
  if (!any) return kOk;  // "." or sign alone: no conversion
  if (c < end && (*c == 'e' || *c == 'E')) {
    const uint8_t* e = c + 1;
    bool eneg = false;
    if (e < end && (*e == '+' || *e == '-')) {
      eneg = *e == '-';
      ++e;
    }
    if (e < end && *e >= '0' && *e <= '9') {
      int ev = 0;
      for (; e < end && *e >= '0' && *e <= '9'; ++e)
        if (ev < 100000) ev = ev * 10 + (*e - '0');
      exp10 += eneg ? -ev : ev;
      c = e;
    }
  }
  *consumed = (uint64_t)(c - s);
  if (dropped) return kUnsupported;
  if (M == 0) {
    bits = neg ? sign_bit : 0;
    return kOk;
  }
  int st;
  if (exp10 >= 0) {
    if (exp10 > 38) {
      if (p == 24) {  // M * 10^39 > FLT_MAX: infinity
        bits = inf_bits | (neg ? sign_bit : 0);
        return kOk;
      }
      return kUnsupported;
    }
    u128 x;
    if (!mul_checked(pow10_128(exp10), M, x)) {  // at least 2^128 > FLT_MAX
      if (p == 24) {
        bits = inf_bits | (neg ? sign_bit : 0);
        return kOk;
      }
      return kUnsupported;
    }
    st = round_pack(x, false, 0, p, bias, emax, bits);
  } else {
    const int k = -exp10;
    if (k > 38) return kUnsupported;
    const u128 D = pow10_128(k);
    const int s_shift = p + 2 + bitlen128(D) - bitlen128((u128)M) + 1;
    if (s_shift < 0 || bitlen128((u128)M) + s_shift > 128) return kUnsupported;
    u128 q, r;
    divmod128(((u128)M) << s_shift, D, q, r);
    st = round_pack(q, r != 0, -s_shift, p, bias, emax, bits);
  }
  if (st != kOk) return st;
  if (neg) bits |= sign_bit;
  return kOk;
}

}  // namespace dec
}  // namespace rbg
