// rss_masked.h -- per-element device form of the 3-party RSS protocol in the
// masked precompute mode: the reference's protocol steps with the correlated
// randomness that its placeholder precompute leaves constant (src/mpc/
// Precompute.h:73-134, the unmasked reshare of RSS.inl:1052-1074, r = 0 and
// rbits = 1 in dReLU RSS.inl:806-852, c = cbits = 0 in selectShare
// RSS.inl:542-569) drawn from the Philox stream of prg.h, as a trusted dealer
// would deal them.  Two steps that only work while every share but x0 is zero
// are replaced by their general forms:
//   * dReLU's tail `result.fill(1); result -= preResult` (RSS.inl:850-851)
//     becomes result = preResult XOR 1 (1 enters the x0 slot: A.sA, C.sB);
//   * selectShare's opening of the XOR-shared b ^ cbits by addition
//     (RSS.inl:553-558) becomes an opening by XOR.
// Truncation follows the reference's own plan for getDividedShares (r' = r * d,
// the comment of Precompute.h:107): r is random with |r * d| < 2^61 and the
// opened t - r*d is divided exactly, so the result is floor(t/d) or ceil(t/d).
//
// Word layout per element (prg.h): reshare / mult / AND: words 0, 1 (zero
// sharings alpha_0, alpha_1, alpha_2 = -alpha_0 - alpha_1 or alpha_0 ^ alpha_1);
// truncation: 0 = R (sign-extended to |R * 2^k| < 2^61), 1, 2 = r_1, r_2 with
// r_0 = R - r_1 - r_2, 3, 4 = r'_1, r'_2 of the op-by-op form; dReLU: 0, 1 =
// the zero sharing of the carry tree's g-ANDs, 2, 3 = that of its p-ANDs (the
// bit positions each AND writes are its own: level L's g-AND takes the words'
// bits at its result positions, which no other level uses, and level L's
// p-AND the bits 2^L - 1 above its result positions, again disjoint across
// levels), 4 = R, 5, 6 = arithmetic shares of R - 1, 7, 8 = binary shares of
// -R; selectShare: 0, 1 = the product's reshare, 2 = the random bit c, 3, 4 =
// its binary shares, 5, 6 = its arithmetic shares.
#pragma once

#include "prg.h"
#include "rss_device.h"

namespace pir {

__device__ __forceinline__ void prg_two(const PrgStream &g, u64 e, u32 j, u64 &w0, u64 &w1) {
    if ((j & 1) == 0) {
        prg_pair(g, e, j >> 1, w0, w1);
    } else {
        w0 = prg_word(g, e, j);
        w1 = prg_word(g, e, j + 1);
    }
}

// reshare (RSS.inl:1052-1074) of the 3-of-3 values z with an arithmetic
// (XOR = false) or binary (XOR = true) zero sharing from words j, j + 1;
// `width` keeps the masks inside the values' range (all ones, or 1 for bits).
template <typename T, bool XOR>
__device__ __forceinline__ Sh3<T> reshare_m(const T z[3], const PrgStream &g, u64 e, u32 j, T width) {
    u64 w0, w1;
    prg_two(g, e, j, w0, w1);
    T a[3];
    a[0] = (T)w0 & width;
    a[1] = (T)w1 & width;
    a[2] = XOR ? (T)(a[0] ^ a[1]) : (T)((T)0 - a[0] - a[1]);
    Sh3<T> o;
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const int q = next_party(p);
        o.s[p][0] = XOR ? (T)(z[p] ^ a[p]) : (T)(z[p] + a[p]);
        o.s[p][1] = XOR ? (T)(z[q] ^ a[q]) : (T)(z[q] + a[q]);
    }
    return o;
}

// RSS *= RSS (RSS.inl:283-320) with a masked reshare.
template <typename T>
__device__ __forceinline__ void rss_mult_m(Sh3<T> &x, const Sh3<T> &y, const PrgStream &g, u64 e, u32 j) {
    T z[3];
#pragma unroll
    for (int p = 0; p < 3; p++) z[p] = (T)(x.s[p][0] * (T)(y.s[p][0] + y.s[p][1]) + x.s[p][1] * y.s[p][0]);
    x = reshare_m<T, false>(z, g, e, j, (T)~(T)0);
}

// RSS &= RSS (RSS.inl:330-345) with a masked binary reshare.
template <typename T>
__device__ __forceinline__ void rss_and_m(Sh3<T> &x, const Sh3<T> &y, const PrgStream &g, u64 e, u32 j, T width) {
    T z[3];
#pragma unroll
    for (int p = 0; p < 3; p++) z[p] = (T)((x.s[p][0] & (T)(y.s[p][0] ^ y.s[p][1])) ^ (x.s[p][1] & y.s[p][0]));
    x = reshare_m<T, true>(z, g, e, j, width);
}

// One AND of the carry tree on bit-packed words restricted to `mask` (see
// rss_and_masked), with the binary zero sharing (w0, w1, w0 ^ w1) on the
// positions the view covers.
__device__ __forceinline__ void rss_and_masked_m(u64 xa[3], u64 xb[3], const u64 ya[3], const u64 yb[3], u64 mask,
                                                 u64 w0, u64 w1) {
    const u64 a[3] = {w0, w1, w0 ^ w1};
    u64 z[3];
#pragma unroll
    for (int p = 0; p < 3; p++) z[p] = ((xa[p] & (ya[p] ^ yb[p])) ^ (xb[p] & ya[p])) ^ a[p];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        xa[p] = (xa[p] & ~mask) | (z[p] & mask);
        xb[p] = (xb[p] & ~mask) | (z[next_party(p)] & mask);
    }
}

// dReLU (RSS.inl:806-852) with a random mask: r = arithmetic sharing of R - 1,
// rbits = binary sharing of -R, so a = open(input + r) + 1 = x + R and the
// carry circuit adds a + (-R) = x; words j .. j + 9 (j even).
template <typename U>
__device__ __forceinline__ Sh3<U> drelu_m(const Sh3<u64> &in, const PrgStream &g, u64 e, u32 j) {
    u64 g0, g1, p0, p1, R, s1, s2, b1, b2;
    prg_two(g, e, j + 0, g0, g1);
    prg_two(g, e, j + 2, p0, p1);
    prg_two(g, e, j + 4, R, s1);
    prg_two(g, e, j + 6, s2, b1);
    b2 = prg_word(g, e, j + 8);
    const u64 s[3] = {R - 1 - s1 - s2, s1, s2};
    u64 a[3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const int q = next_party(p), pr = prev_party(p);
        a[p] = (in.s[p][0] + s[p]) + (in.s[p][1] + s[q]) + (in.s[pr][0] + s[pr]) + 1;
    }
    const u64 rb3[3] = {((u64)0 - R) ^ b1 ^ b2, b1, b2};
    u64 ra[3], rb[3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        ra[p] = rb3[p];
        rb[p] = rb3[next_party(p)];
    }
    const u64 top = 1ull << 63;
    U msba[3], msbb[3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        // setCarryOutMSB (bitwise.cuh:41-64)
        u64 a63 = a[p] >> 63;
        msba[p] = (U)(((ra[p] >> 63) ^ (p == PARTY_A ? a63 : 0ull)) & 1ull);
        msbb[p] = (U)(((rb[p] >> 63) ^ (p == PARTY_C ? a63 : 0ull)) & 1ull);
        a[p] |= top;
        ra[p] &= ~top;
        rb[p] &= ~top;
    }
    u64 ga[3], gb[3], pa[3], pb[3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        ga[p] = ra[p] & a[p];
        gb[p] = rb[p] & a[p];
        pa[p] = ra[p] ^ (p == PARTY_A ? a[p] : 0ull);
        pb[p] = rb[p] ^ (p == PARTY_C ? a[p] : 0ull);
    }
#pragma unroll
    for (int L = 0; L < 6; L++) {
        const int h = 1 << L;
        const u64 mge = tree_mask_geven(L), mpe = tree_mask_peven(L);
        u64 ya[3], yb[3];
#pragma unroll
        for (int p = 0; p < 3; p++) {
            ya[p] = pa[p] >> 1;
            yb[p] = pb[p] >> 1;
        }
        rss_and_masked_m(ga, gb, ya, yb, mge, g0, g1);
#pragma unroll
        for (int p = 0; p < 3; p++) {
            ya[p] = pa[p] >> h;
            yb[p] = pb[p] >> h;
        }
        rss_and_masked_m(pa, pb, ya, yb, mpe, p0 >> (h - 1), p1 >> (h - 1));
#pragma unroll
        for (int p = 0; p < 3; p++) {
            ga[p] ^= (ga[p] & mge) << h;
            gb[p] ^= (gb[p] & mge) << h;
        }
    }
    Sh3<U> res;
#pragma unroll
    for (int p = 0; p < 3; p++) {
        U prea = (U)((ga[p] >> 63) & 1ull) ^ msba[p];
        U preb = (U)((gb[p] >> 63) & 1ull) ^ msbb[p];
        res.s[p][0] = (U)(prea ^ (U)(p == PARTY_A ? 1 : 0));
        res.s[p][1] = (U)(preb ^ (U)(p == PARTY_C ? 1 : 0));
    }
    return res;
}

// selectShare (RSS.inl:542-569) with a random bit c (binary shares cbits,
// arithmetic shares c) and b ^ cbits opened by XOR; words j .. j + 7 (j even).
template <typename U>
__device__ __forceinline__ Sh3<u64> select_m(const Sh3<u64> &x, const Sh3<u64> &y, const Sh3<U> &b,
                                             const PrgStream &g, u64 e, u32 j) {
    u64 wc, wb1, wb2, wc1, wc2;
    prg_two(g, e, j + 2, wc, wb1);
    prg_two(g, e, j + 4, wb2, wc1);
    wc2 = prg_word(g, e, j + 6);
    const U cb = (U)(wc & 1), cb1 = (U)(wb1 & 1), cb2 = (U)(wb2 & 1);
    const U cbits[3] = {(U)(cb ^ cb1 ^ cb2), cb1, cb2};
    const u64 c[3] = {(u64)cb - wc1 - wc2, wc1, wc2};
    Sh3<u64> d, r;
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const int q = next_party(p), pr = prev_party(p);
        const U ev = (U)((b.s[p][0] ^ cbits[p]) ^ (b.s[p][1] ^ cbits[q]) ^ (b.s[pr][0] ^ cbits[pr]));
        // convex_comb (RSS.inl:33-61): d = e ? 1 - c : c
        if (ev == (U)1) {
            d.s[p][0] = (p == PARTY_A ? 1ull : 0ull) - c[p];
            d.s[p][1] = (p == PARTY_C ? 1ull : 0ull) - c[q];
        } else {
            d.s[p][0] = c[p];
            d.s[p][1] = c[q];
        }
        r.s[p][0] = y.s[p][0] - x.s[p][0];
        r.s[p][1] = y.s[p][1] - x.s[p][1];
    }
    rss_mult_m(r, d, g, e, j + 0);
#pragma unroll
    for (int p = 0; p < 3; p++) {
        r.s[p][0] += x.s[p][0];
        r.s[p][1] += x.s[p][1];
    }
    return r;
}

// ReLU (RSS.inl:854-865): dReLU on words 0..9, selectShare(0, x, b) on 10..17.
template <typename U>
__device__ __forceinline__ Sh3<u64> relu_m(const Sh3<u64> &in, Sh3<U> &b, const PrgStream &g, u64 e) {
    b = drelu_m<U>(in, g, e, 0);
    Sh3<u64> zero;
#pragma unroll
    for (int p = 0; p < 3; p++) zero.s[p][0] = zero.s[p][1] = 0;
    return select_m<U>(zero, in, b, g, e, PRG_DRELU_WORDS);
}

// Signed division of the opened value by 2^k, truncated toward zero, exactly.
__device__ __forceinline__ u64 trunc_exact_pow2(u64 u, int k) {
    const s64 v = (s64)u;
    const u64 mag = v < 0 ? (u64)0 - u : u;
    const u64 q = mag >> k;
    return v < 0 ? (u64)0 - q : q;
}

// The random truncation pair of element e: R with |R * 2^k| < 2^61, and the
// arithmetic shares r_0, r_1, r_2 of R.
__device__ __forceinline__ void trunc_pair_m(const PrgStream &g, u64 e, int k, u64 &R, u64 r[3]) {
    u64 w0, w1;
    prg_pair(g, e, 0, w0, w1);
    const int sh = k + 2 < 63 ? k + 2 : 63;
    R = (u64)((s64)w0 >> sh);
    r[1] = w1;
    r[2] = prg_word(g, e, 2);
    r[0] = R - r[1] - r[2];
}

// dividePublic3out3 (RSS.inl:1076-1094) of the opened sum t of the 3-of-3
// shares: u = t - R * 2^k is opened, q = trunc(u / 2^k) exactly; out = r + q
// with q in the x0 slot.  Returns the six output shares in out[p][s].
__device__ __forceinline__ void divide_3out3_m(u64 t, int k, const PrgStream &g, u64 e, u64 out[3][2]) {
    u64 R, r[3];
    trunc_pair_m(g, e, k, R, r);
    const u64 q = trunc_exact_pow2(t - (R << k), k);
#pragma unroll
    for (int p = 0; p < 3; p++) {
        out[p][0] = r[p] + (p == PARTY_A ? q : 0ull);
        out[p][1] = r[next_party(p)] + (p == PARTY_C ? q : 0ull);
    }
}

// dividePublic (RSS.inl:422-436) by 2^k of an RSS value: a -= r' (RSS shares
// of R * 2^k from words 3, 4), each party opens its view t_p, q_p = trunc(t_p /
// 2^k) exactly, out = r + q_p in the x0 slot.
__device__ __forceinline__ Sh3<u64> divide_public_pow2_m(Sh3<u64> v, int k, const PrgStream &g, u64 e) {
    u64 R, r[3];
    trunc_pair_m(g, e, k, R, r);
    const u64 rp1 = prg_word(g, e, 3), rp2 = prg_word(g, e, 4);
    const u64 rp[3] = {(R << k) - rp1 - rp2, rp1, rp2};
#pragma unroll
    for (int p = 0; p < 3; p++) {
        v.s[p][0] -= rp[p];
        v.s[p][1] -= rp[next_party(p)];
    }
    Sh3<u64> o;
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const u64 t = v.s[p][0] + v.s[p][1] + v.s[prev_party(p)][0];
        const u64 q = trunc_exact_pow2(t, k);
        o.s[p][0] = r[p] + (p == PARTY_A ? q : 0ull);
        o.s[p][1] = r[next_party(p)] + (p == PARTY_C ? q : 0ull);
    }
    return o;
}

}  // namespace pir
