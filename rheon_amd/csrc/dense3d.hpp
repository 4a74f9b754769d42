// Exact scientific formatting of finite binary floating-point values, as
// Rust's `{:.Pe}` prints them (the dense3d frame writer's number format,
// src/dense3d_sequence.rs:512-548: `{v:.9e}` for f32 fields, `{v:.17e}` for
// f64 fields and diagnostics).  N = P + 1 significant digits:
//
//   [-]d.ddd...e[-]X
//
// with no '+' and no padding in the exponent, a '-' for every negative
// value including -0.0, and the digits of the exact binary value rounded to
// N places, ties to even (core::fmt's exact mode).  Host and device share
// this code.  The value v = m 2^e is scaled exactly to D = round(v 10^(N-1-E))
// with E = floor(log10 v):
//
//   * 0 <= N-1-E <= 27 (|v| >= 1e-10 for N = 18, >= 1e-18 for N = 10): one
//     64 x 64 -> 128-bit product m 5^(N-1-E) and a shift by -(e + N-1-E);
//   * otherwise (tiny or huge magnitudes): the same quotient in multi-word
//     integers (shift, or long division by 5^c 2^d when N-1-E < 0).
//
// E starts from an estimate at most one below and is corrected until D has
// exactly N digits; a round-up to 10^N becomes 10^(N-1) with E + 1.
#pragma once
#include <stdint.h>

#include <cmath>

#include <hip/hip_runtime.h>

namespace rh {
namespace d3 {

#define RH_HD __host__ __device__ __forceinline__

constexpr int kMaxF32Chars = 16;  // -1.234567890e-45
constexpr int kMaxF64Chars = 25;  // -1.23456789012345678e-308

RH_HD uint64_t pow10u(int k) {
  uint64_t r = 1;
  for (int i = 0; i < k; ++i) r *= 10u;
  return r;
}
RH_HD uint64_t pow5u(int k) {  // k <= 27
  uint64_t r = 1, b = 5;
  while (k) {
    if (k & 1) r *= b;
    b *= b;
    k >>= 1;
  }
  return r;
}

// Multi-word unsigned integer, little-endian 32-bit words (the rare
// magnitudes outside the 128-bit fast path: at most 2^1100 here).
struct Big {
  static constexpr int kWords = 40;
  uint32_t w[kWords];
  int n;  // words in use
  RH_HD void set(uint64_t v) {
    n = 0;
    while (v) {
      w[n++] = (uint32_t)v;
      v >>= 32;
    }
  }
  RH_HD void mul_small(uint32_t f) {
    uint64_t carry = 0;
    for (int i = 0; i < n; ++i) {
      const uint64_t t = (uint64_t)w[i] * f + carry;
      w[i] = (uint32_t)t;
      carry = t >> 32;
    }
    if (carry) w[n++] = (uint32_t)carry;
  }
  RH_HD void mul_pow5(int k) {
    while (k >= 13) {  // 5^13 = 1220703125 < 2^32
      mul_small(1220703125u);
      k -= 13;
    }
    if (k) mul_small((uint32_t)pow5u(k));
  }
  RH_HD void shl(int bits) {
    if (!n || !bits) return;
    const int ws = bits >> 5, bs = bits & 31;
    int m = n + ws + 1;
    for (int i = m - 1; i >= 0; --i) {
      const int src = i - ws;
      uint32_t hi = (src >= 0 && src < n) ? w[src] : 0u;
      uint32_t lo = (src - 1 >= 0 && src - 1 < n) ? w[src - 1] : 0u;
      w[i] = bs ? (hi << bs) | (lo >> (32 - bs)) : hi;
    }
    n = m;
    while (n && !w[n - 1]) --n;
  }
  // bit b (0 = least significant)
  RH_HD int bit(int b) const {
    const int i = b >> 5;
    return i < n ? (int)((w[i] >> (b & 31)) & 1u) : 0;
  }
  // (this >> s) as 64 bits (the caller knows it fits), the low s bits
  // compared with 2^(s-1): -1 below, 0 equal, 1 above.
  RH_HD uint64_t shr_round_info(int s, int* cmp_half) const {
    uint64_t q = 0;
    for (int b = 63; b >= 0; --b) q = (q << 1) | (uint64_t)bit(s + b);
    if (s == 0) {
      *cmp_half = -1;  // no remainder
      return q;
    }
    int c = bit(s - 1) ? 0 : -1;  // the half bit
    if (c == 0) {
      for (int b = 0; b < s - 1; ++b)
        if (bit(b)) {
          c = 1;
          break;
        }
    }
    *cmp_half = c;
    return q;
  }
  RH_HD int cmp(const Big& o) const {
    if (n != o.n) return n < o.n ? -1 : 1;
    for (int i = n - 1; i >= 0; --i)
      if (w[i] != o.w[i]) return w[i] < o.w[i] ? -1 : 1;
    return 0;
  }
  RH_HD void sub(const Big& o) {  // this >= o
    int64_t borrow = 0;
    for (int i = 0; i < n; ++i) {
      int64_t t = (int64_t)w[i] - (i < o.n ? (int64_t)o.w[i] : 0) - borrow;
      borrow = t < 0;
      w[i] = (uint32_t)(t + (borrow ? ((int64_t)1 << 32) : 0));
    }
    while (n && !w[n - 1]) --n;
  }
};

// floor(num / den) (known to be < 2^64) by restoring division; num becomes
// the remainder.
RH_HD uint64_t big_divmod(Big& num, const Big& den) {
  uint64_t q = 0;
  for (int b = 63; b >= 0; --b) {
    Big t = den;
    t.shl(b);
    if (num.cmp(t) >= 0) {
      num.sub(t);
      q |= (uint64_t)1 << b;
    }
  }
  return q;
}

// The multi-word form of scale_digits (magnitudes outside the fast path),
// kept out of line so its registers do not weigh on the fast path.
RH_HD uint64_t scale_digits_big(uint64_t m, int e, int s, int* round_up);

// D_trunc = floor(|v| 10^s) and its rounding: 1 when it rounds up, ties to
// even.  m, e: |v| = m 2^e.
RH_HD uint64_t scale_digits(uint64_t m, int e, int s, int* round_up) {
  const int g = e + s;
  if (s >= 0 && s <= 16 && m < (1ull << 26)) {
    // f32 mantissas at moderate magnitudes: the product fits 64 bits
    const uint64_t p = m * pow5u(s);
    if (g >= 0) {
      *round_up = 0;
      return p << g;
    }
    const int d = -g;
    if (d >= 64) {
      *round_up = 0;
      return 0;
    }
    const uint64_t q = p >> d;
    const uint64_t rem = p & ((1ull << d) - 1);
    const uint64_t half = d == 0 ? 0 : 1ull << (d - 1);
    *round_up = d > 0 && (rem > half || (rem == half && (q & 1u)));
    return q;
  }
  if (s >= 0 && s <= 27) {
    const unsigned __int128 p = (unsigned __int128)m * pow5u(s);
    if (g >= 0) {
      *round_up = 0;
      return (uint64_t)(p << g);
    }
    const int d = -g;
    if (d >= 128) {
      *round_up = 0;
      return 0;
    }
    const uint64_t q = (uint64_t)(p >> d);
    const unsigned __int128 rem = d == 0 ? 0 : p & ((((unsigned __int128)1) << d) - 1);
    const unsigned __int128 half = d == 0 ? 0 : ((unsigned __int128)1) << (d - 1);
    *round_up = d > 0 && (rem > half || (rem == half && (q & 1u)));
    return q;
  }
  return scale_digits_big(m, e, s, round_up);
}

RH_HD uint64_t scale_digits_big(uint64_t m, int e, int s, int* round_up) {
  const int g = e + s;
  Big num;
  num.set(m);
  if (s >= 0) {
    num.mul_pow5(s);
    if (g >= 0) {
      num.shl(g);
      int c;
      const uint64_t q = num.shr_round_info(0, &c);
      *round_up = 0;
      return q;
    }
    int c;
    const uint64_t q = num.shr_round_info(-g, &c);
    *round_up = c > 0 || (c == 0 && (q & 1u));
    return q;
  }
  // |v| >= 10^N: divide by 5^-s 2^max(-g, 0)
  if (g > 0) num.shl(g);
  Big den;
  den.set(1);
  den.mul_pow5(-s);
  if (g < 0) den.shl(-g);
  const uint64_t q = big_divmod(num, den);
  num.shl(1);  // 2 R vs den
  const int c = num.cmp(den);
  *round_up = c > 0 || (c == 0 && (q & 1u));
  return q;
}

// |v| = D 10^(E - n + 1) rounded as above (D = 0 for zeros), and the sign.
struct Sci {
  uint64_t D;
  int E;
  bool neg;
};
template <int n>
RH_HD Sci decompose(double v) {
  uint64_t bits;
  __builtin_memcpy(&bits, &v, 8);
  Sci r{0, 0, (bits >> 63) != 0};
  const uint64_t frac = bits & ((1ull << 52) - 1);
  const int bexp = (int)((bits >> 52) & 0x7ff);
  if (bexp == 0 && frac == 0) return r;
  const uint64_t m = bexp ? (frac | (1ull << 52)) : frac;
  const int e = bexp ? bexp - 1075 : -1074;
  // E from below: log10 v = (b + log2(1 + x)) log10(2) with x the mantissa
  // fraction and log2(1 + x) >= x on [0, 1], so the estimate is E or E - 1
  // (the loop below adds the one); subnormals take log10.
  int E = bexp ? (int)floor(((double)(bexp - 1023) + (double)frac * 0x1p-52) * 0.30102999566398120)
               : (int)floor(log10(r.neg ? -v : v));
  const uint64_t lo = pow10u(n - 1), hi = lo * 10u;
  int up = 0;
  uint64_t D = 0;
  for (int guard = 0; guard < 4; ++guard) {
    D = scale_digits(m, e, n - 1 - E, &up);
    if (D >= hi) {
      ++E;
    } else if (D < lo) {
      --E;
    } else {
      break;
    }
  }
  D += (uint64_t)up;
  if (D == hi) {
    D = lo;
    ++E;
  }
  r.D = D;
  r.E = E;
  return r;
}

RH_HD int sci_length(const Sci& s, int n) {
  const int x = s.E < 0 ? -s.E : s.E;
  return (s.neg ? 1 : 0) + n + 1 + 1 + (s.E < 0 ? 1 : 0) + (x >= 100 ? 3 : (x >= 10 ? 2 : 1));
}

// Writes s (n <= 18 significant digits) at out; returns the length.
RH_HD int sci_write(const Sci& s, int n, char* out) {
  int pos = 0;
  if (s.neg) out[pos++] = '-';
  // digit i of D at out[pos + i + (i > 0)], the point after digit 0
  const uint64_t hi9 = s.D / 1000000000u;
  uint32_t lo9 = (uint32_t)(s.D - hi9 * 1000000000u);
  uint32_t h = (uint32_t)hi9;
  for (int i = n - 1; i >= 0; --i) {
    uint32_t d;
    if (i >= n - 9) {
      d = lo9 % 10u;
      lo9 /= 10u;
    } else {
      d = h % 10u;
      h /= 10u;
    }
    out[pos + i + (i > 0 ? 1 : 0)] = (char)('0' + d);
  }
  out[pos + 1] = '.';
  pos += n + 1;
  out[pos++] = 'e';
  int x = s.E;
  if (x < 0) {
    out[pos++] = '-';
    x = -x;
  }
  const int nd = x >= 100 ? 3 : (x >= 10 ? 2 : 1);
  for (int i = nd - 1; i >= 0; --i) {
    out[pos + i] = (char)('0' + x % 10);
    x /= 10;
  }
  return pos + nd;
}

// Writes v with n significant digits to out; returns the length.
template <int n>
RH_HD int format_sci(double v, char* out) { return sci_write(decompose<n>(v), n, out); }

}  // namespace d3
}  // namespace rh
