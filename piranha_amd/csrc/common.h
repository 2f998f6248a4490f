// common.h -- shared types, error handling and small device helpers for the
// MI355X (gfx950) secret-share compute path.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>

#include "ring_div.h"

namespace pir {

typedef uint64_t u64;
typedef int64_t s64;
typedef uint32_t u32;
typedef uint8_t u8;

// Party indices of the 3-party replicated sharing (RSS.h:24, enum Party).
enum { PARTY_A = 0, PARTY_B = 1, PARTY_C = 2, NUM_PARTIES = 3 };

__host__ __device__ inline int next_party(int p) { return p == 2 ? 0 : p + 1; }
__host__ __device__ inline int prev_party(int p) { return p == 0 ? 2 : p - 1; }

// Status codes returned through the C ABI.
enum Status { OK = 0, ERR_ARG = 1, ERR_HIP = 2, ERR_UNSUPPORTED = 3, ERR_STATE = 4 };

void set_error(const std::string &msg);
const char *last_error();

#define PIR_HIP(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) {                                                                \
            ::pir::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));               \
            return ::pir::ERR_HIP;                                                             \
        }                                                                                      \
    } while (0)

#define PIR_REQUIRE(cond, msg)                                                                 \
    do {                                                                                       \
        if (!(cond)) {                                                                         \
            ::pir::set_error(msg);                                                             \
            return ::pir::ERR_ARG;                                                             \
        }                                                                                      \
    } while (0)

#define PIR_TRY(expr)                                                                          \
    do {                                                                                       \
        int _s = (expr);                                                                       \
        if (_s != 0) return _s;                                                                \
    } while (0)

// Launch check: kernel launches do not return a status, query it right away.
#define PIR_LAUNCH_CHECK() PIR_HIP(hipGetLastError())

inline unsigned grid_for(size_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

// Unsigned division by a runtime constant d >= 1 with one multiply-high:
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1, n / d = (mulhi(n, m) + n) >> l,
// exact for every 32-bit n (the sum is formed in 64 bits).
struct FastDiv {
    u32 d = 1, mul = 1, shr = 0;
    FastDiv() = default;
    explicit FastDiv(u32 dv) : d(dv) {
        shr = 0;
        while (shr < 32 && (1ull << shr) < dv) shr++;
        mul = (u32)((((u64)1 << 32) * (((u64)1 << shr) - dv)) / dv + 1);
    }
    __host__ __device__ __forceinline__ u32 div(u32 n) const {
        return (u32)(((((u64)n * mul) >> 32) + n) >> shr);
    }
};

// 64-bit left shift with the amount clamped as the reference's GPU shifts
// clamp it: x << y for y >= 64 is 0.
__host__ __device__ __forceinline__ u64 shl_clamped(u64 x, u64 y) { return y >= 64 ? 0 : x << y; }

// The coefficient functors of taylorSeries (RSS.inl:1237-1320) tabulated over
// their argument x, the opened power plus one.  Both functors scale by the
// 32-bit `1 << x`, which is 0 for every x >= 32, so entry TAYLOR_TAB - 1 holds
// the value of every x >= 32 and a lookup clamps its index there.
constexpr int TAYLOR_TAB = 33;
struct TaylorTab {
    u64 v[TAYLOR_TAB];
    __host__ __device__ __forceinline__ u64 at(u64 x) const { return v[x < TAYLOR_TAB - 1 ? x : TAYLOR_TAB - 1]; }
};
struct TaylorTables {
    TaylorTab c2, c1, c0;  // tofixed_variable_precision_functor(a2 / a1 / a0)
    TaylorTab f;           // calc_fn(fn)
    int fp;                // FLOAT_PRECISION
    int use_a2, use_a1;    // a2 != 0, a1 != 0
};

}  // namespace pir
