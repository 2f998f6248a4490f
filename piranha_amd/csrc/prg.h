// prg.h -- the seeded counter-based PRG of the masked precompute mode.
//
// Generator: Philox4x32-10 (J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
// "Parallel Random Numbers: As Easy as 1, 2, 3", SC'11): a 4 x 32-bit counter
// and a 2 x 32-bit key, ten rounds of two 32 x 32 -> 64-bit multiplications
// (multipliers 0xD2511F53, 0xCD9E8D57) with the key bumped by the Weyl
// constants 0x9E3779B9, 0xBB67AE85 between rounds.  It stands in for the
// reference's AES-256 counter-mode PRG (src/mpc/AESObject.cpp:11-90), which the
// reference never instantiates (its correlated randomness is all placeholder
// constants, src/mpc/Precompute.h:73-134).
//
// Word stream of the masked protocols: a protocol call owns a 32-bit call
// number; random word j of element e of that call is
//     prg_word(seed, call, e, j) = 64-bit half (j & 1) of
//     Philox4x32-10(key = (seed_lo, seed_hi), ctr = (e_lo, e_hi, call, j >> 1)),
// half 0 = out0 | out1 << 32, half 1 = out2 | out3 << 32.  Every element of a
// call therefore draws from its own counter range and any thread can produce
// any word without sequential state.
#pragma once

#include "common.h"

namespace pir {

__host__ __device__ __forceinline__ void philox4x32_10(const u32 ctr_in[4], const u32 key_in[2], u32 out[4]) {
    u32 c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    u32 k0 = key_in[0], k1 = key_in[1];
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const u64 p0 = (u64)0xD2511F53u * c0, p1 = (u64)0xCD9E8D57u * c2;
        const u32 hi0 = (u32)(p0 >> 32), lo0 = (u32)p0, hi1 = (u32)(p1 >> 32), lo1 = (u32)p1;
        const u32 n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// The random stream of one protocol call (on = 0: placeholder precompute, no
// randomness; the kernels then follow the reference's constants).
struct PrgStream {
    u64 seed;
    u32 call;
    int on;
};

// Words 2m and 2m+1 of element e: one Philox block.
__host__ __device__ __forceinline__ void prg_pair(const PrgStream &g, u64 e, u32 m, u64 &w0, u64 &w1) {
    const u32 ctr[4] = {(u32)e, (u32)(e >> 32), g.call, m};
    const u32 key[2] = {(u32)g.seed, (u32)(g.seed >> 32)};
    u32 o[4];
    philox4x32_10(ctr, key, o);
    w0 = (u64)o[0] | ((u64)o[1] << 32);
    w1 = (u64)o[2] | ((u64)o[3] << 32);
}

__host__ __device__ __forceinline__ u64 prg_word(const PrgStream &g, u64 e, u32 j) {
    u64 w0, w1;
    prg_pair(g, e, j >> 1, w0, w1);
    return (j & 1) ? w1 : w0;
}

// Word indices of the masked protocols for one element (all relative to the
// element's base index within its call).  Blocks of the dReLU / selectShare /
// ReLU / maxpool layouts start at even words, so each Philox block drawn
// yields two used words:
constexpr u32 PRG_RESHARE = 0;       // 2 words: zero-sharing alpha_0, alpha_1 (alpha_2 closes the sum)
constexpr u32 PRG_DIV = 0;           // 5 words: R, r_1, r_2, r'_a, r'_b of the truncation pair
constexpr u32 PRG_DRELU = 0;         // 10 words: 2 + 2 carry-tree masks, R, 2 arithmetic, 2 binary share words, 1 spare
constexpr u32 PRG_DRELU_WORDS = 10;
constexpr u32 PRG_SELECT_WORDS = 8;  // 2 reshare masks, bit, 2 binary and 2 arithmetic share words, 1 spare
constexpr u32 PRG_RELU_WORDS = PRG_DRELU_WORDS + PRG_SELECT_WORDS;

// Host launchers (prg.hip).  prg_philox: out[4i..4i+3] = Philox4x32-10 of
// ctr[4i..4i+3] under key[2i..2i+1].  prg_fill: out[2b + h] = word h of
// element b of the call.  rss_deal: a random consistent RSS sharing
// [3][2][n] of values[i] (or, with values null, of the value party A opens
// from shares [3][2][n]) from words 0 and 1 of element i of the call.
int prg_philox(const u32 *ctr, const u32 *key, size_t n, u32 *out, hipStream_t s);
int prg_fill(u64 seed, u32 call, size_t n, u64 *out, hipStream_t s);
int rss_deal(const u64 *values, const u64 *shares, size_t n, u64 seed, u32 call, u64 *out, hipStream_t s);

}  // namespace pir
