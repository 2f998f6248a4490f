// ring_div.h -- the reference's signed division of ring elements through IEEE
// double, (S)((double)(S)x / (double)(S)y) (functors.cuh:33-45, 94-105),
// restated in 64-bit integer arithmetic: no floating-point instruction is
// involved, and the result is the one an IEEE-754 binary64 unit with
// round-to-nearest-even gives, followed by truncation toward zero.
//
// Self-contained (only <cstdint>) so that a host-only compiler can build it.
#pragma once

#include <cstdint>

#ifndef PIR_HD
#if defined(__HIPCC__)
#define PIR_HD __host__ __device__
#else
#define PIR_HD
#endif
#endif

namespace pir {

PIR_HD inline int bit_length_u64(uint64_t v) { return v ? 64 - __builtin_clzll(v) : 0; }

// |v| as binary64 would hold it: m * 2^e with m < 2^53 (round-to-nearest-even
// to 53 significant bits; magnitudes up to 2^64 - 1).
struct RoundedMag {
    uint64_t m;
    int e;
};

PIR_HD inline RoundedMag round_to_53_bits(uint64_t mag) {
    int len = bit_length_u64(mag);
    if (len <= 53) return {mag, 0};
    int sh = len - 53;
    uint64_t m = mag >> sh, rem = mag & ((1ull << sh) - 1), half = 1ull << (sh - 1);
    if (rem > half || (rem == half && (m & 1))) m++;
    if (m >> 53) {  // rounded up to the next power of two
        m >>= 1;
        sh++;
    }
    return {m, sh};
}

// (S)((double)(S)t / 2^k) for 1 <= k <= 62: the int64 -> double conversion
// rounds |t| to 53 significant bits, the division by 2^k is exact, and the
// conversion back truncates toward zero (sign symmetric).
PIR_HD inline uint64_t trunc_div_pow2(uint64_t t, int k) {
    bool neg = (int64_t)t < 0;
    RoundedMag r = round_to_53_bits(neg ? (uint64_t)0 - t : t);
    int e = r.e - k;
    uint64_t q = e >= 0 ? r.m << e : (e <= -64 ? 0 : r.m >> -e);
    return neg ? (uint64_t)0 - q : q;
}

// (S)((double)(S)x / (double)(S)y) for any x, y.  The reference evaluates the
// functor on the GPU, whose double -> int64 conversion (cvt.rzi.s64.f64)
// clamps to the int64 range and maps NaN to 0: a quotient >= 2^63 (or x / 0
// with x > 0) gives INT64_MAX, one <= -2^63 (or x / 0 with x < 0) INT64_MIN,
// and 0 / 0 gives 0.
//
// Method (schoolbook long division with IEEE 754 round-to-nearest-even):
// both operands are rounded to 53 bits as binary64 would and scaled
// to mantissas in [2^52, 2^53); their ratio (in (1/2, 2)) is formed to 55
// fractional bits by long division in 11-bit steps (remainders stay below
// 2^53, so every step fits in 64 bits) with a sticky bit for the rest; that
// 55- or 56-bit quotient is rounded to 53 bits (round-to-nearest-even), and
// the value m * 2^E is truncated toward zero.
PIR_HD inline uint64_t signed_div_u64(uint64_t x, uint64_t y) {
    const uint64_t kMin = 1ull << 63, kMax = kMin - 1;
    if (y >= 2 && y < kMin && (y & (y - 1)) == 0) return trunc_div_pow2(x, bit_length_u64(y) - 1);
    if (x == 0) return 0;  // 0 / y is +-0, and 0 / 0 a NaN
    bool neg = (int64_t)(x ^ y) < 0;
    if (y == 0) return (int64_t)x < 0 ? kMin : kMax;  // +-infinity
    RoundedMag a = round_to_53_bits((int64_t)x < 0 ? (uint64_t)0 - x : x);
    RoundedMag b = round_to_53_bits((int64_t)y < 0 ? (uint64_t)0 - y : y);
    int sa = 53 - bit_length_u64(a.m), sb = 53 - bit_length_u64(b.m);
    uint64_t ma = a.m << sa, mb = b.m << sb;  // both in [2^52, 2^53)
    int ea = a.e - sa, eb = b.e - sb;
    uint64_t q = ma >= mb ? 1 : 0, r = ma - q * mb;  // integer part
#pragma unroll
    for (int i = 0; i < 5; i++) {
        r <<= 11;
        uint64_t d = r / mb;
        r -= d * mb;
        q = (q << 11) | d;
    }
    // q = floor(ma * 2^55 / mb) in [2^54, 2^56)
    int sh = bit_length_u64(q) - 53;  // 2 or 3
    uint64_t m = q >> sh, rem = q & ((1ull << sh) - 1), half = 1ull << (sh - 1);
    if (rem > half || (rem == half && (r != 0 || (m & 1)))) m++;
    int E = ea - eb - 55 + sh;
    if (m >> 53) {
        m >>= 1;
        E++;
    }
    uint64_t t;
    if (E >= 0) {
        if (E > 10) return neg ? kMin : kMax;  // m >= 2^52, so |q| >= 2^63
        t = m << E;
    } else {
        t = E <= -64 ? 0 : m >> -E;
    }
    return neg ? (uint64_t)0 - t : t;
}

}  // namespace pir
