// Host check of the integer restatement of the reference's division
// (S)((double)(S)x / (double)(S)y) (functors.cuh:33-45) against the host's own
// IEEE-754 binary64 division and conversions.  Prints the number of
// mismatches and exits non-zero when there is one.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "ring_div.h"

// binary64 division, then the clamping double -> int64 conversion of the
// reference's GPU (NaN -> 0, out-of-range -> INT64_MIN / INT64_MAX)
static uint64_t native_div(uint64_t x, uint64_t y) {
    volatile double a = (double)(int64_t)x, b = (double)(int64_t)y;
    double q = a / b;
    if (q != q) return 0;
    if (q >= 9223372036854775808.0) return (1ull << 63) - 1;
    if (q < -9223372036854775808.0) return 1ull << 63;
    return (uint64_t)(int64_t)q;
}

int main(int argc, char **argv) {
    long count = argc > 1 ? atol(argv[1]) : 1000000;
    std::mt19937_64 g(argc > 2 ? atoll(argv[2]) : 1);
    auto of_length = [&](int len) -> uint64_t {  // random value of bit length len (0..64)
        if (len == 0) return 0;
        uint64_t v = g();
        if (len < 64) v &= (1ull << len) - 1;
        return v | (1ull << (len - 1));
    };
    auto signed_of_length = [&](int len) -> uint64_t {
        uint64_t v = of_length(len);
        return (g() & 1) ? (uint64_t)0 - v : v;
    };
    std::vector<uint64_t> edge = {0, 1, 2, 3, 9, 10, 1000, (1ull << 53) - 1, 1ull << 53, (1ull << 53) + 1,
                                  (1ull << 53) + 2, (1ull << 53) + 3, (1ull << 54) + 2, (1ull << 54) + 6,
                                  (1ull << 62) - 1, (1ull << 62) + 1, (1ull << 63) - 1, 1ull << 63,
                                  (1ull << 63) + 1, ~0ull, ~0ull - 1, 0x0123456789abcdefull};
    long bad = 0, total = 0;
    auto check = [&](uint64_t x, uint64_t y) {
        uint64_t want = native_div(x, y), got = pir::signed_div_u64(x, y);
        total++;
        if (want != got) {
            if (bad < 10)
                printf("mismatch x=%016llx y=%016llx want=%016llx got=%016llx\n", (unsigned long long)x,
                       (unsigned long long)y, (unsigned long long)want, (unsigned long long)got);
            bad++;
        }
    };
    for (uint64_t x : edge)
        for (uint64_t y : edge) {
            check(x, y);
            check((uint64_t)0 - x, y);
            check(x, (uint64_t)0 - y);
        }
    for (int k = 0; k < 64; k++)  // every power of two, both signs, both operands
        for (int i = 0; i < 2000; i++) {
            uint64_t x = signed_of_length((int)(g() % 65));
            check(x, 1ull << k);
            check(x, (uint64_t)0 - (1ull << k));
            check(1ull << k, signed_of_length((int)(g() % 65)));
        }
    for (long i = 0; i < count; i++) {  // bit lengths drawn uniformly, so every magnitude pairing occurs
        uint64_t x = signed_of_length((int)(g() % 65)), y = signed_of_length((int)(g() % 65));
        check(x, y);
        check(x, (uint64_t)(int64_t)(1 + (int64_t)(g() % 100)));  // small divisors (pool sizes, counts)
    }
    printf("%ld checked, %ld mismatches\n", total, bad);
    return bad ? 1 : 0;
}
