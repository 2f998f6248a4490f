// rss_device.h -- per-element device restatement of the 3-party RSS protocol.
//
// Every function here computes, for one element, what each of the three
// parties computes locally in the reference, with the parties' messages
// (reconstruct / reshare / reconstruct3out3) replaced by register moves.  The
// per-party state is explicit, so results are bit-identical to the reference
// semantics for any share values, not only for placeholder-mode ones.
#pragma once

#include "common.h"

namespace pir {

// Share state of one element for all parties: s[p][0] = party p's shareA,
// s[p][1] = party p's shareB.
template <typename T>
struct Sh3 {
    T s[3][2];
};

// RSS *= RSS (RSS.inl:283-320): z_p = x_p (y_p + y_{p+1}) + x_{p+1} y_p, then
// reshare (RSS.inl:1052-1074): out_p = (z_p, z_{next(p)}).
template <typename T>
__device__ __forceinline__ void rss_mult_(Sh3<T> &x, const Sh3<T> &y) {
    T z[3];
#pragma unroll
    for (int p = 0; p < 3; p++) z[p] = (T)(x.s[p][0] * (T)(y.s[p][0] + y.s[p][1]) + x.s[p][1] * y.s[p][0]);
#pragma unroll
    for (int p = 0; p < 3; p++) {
        x.s[p][0] = z[p];
        x.s[p][1] = z[next_party(p)];
    }
}

// RSS &= RSS (RSS.inl:330-345) restricted to the bit positions in `mask`;
// positions outside the mask keep their value (they are outside the strided
// view the reference operates on).
__device__ __forceinline__ void rss_and_masked(u64 xa[3], u64 xb[3], const u64 ya[3], const u64 yb[3], u64 mask) {
    u64 z[3];
#pragma unroll
    for (int p = 0; p < 3; p++) z[p] = (xa[p] & (ya[p] ^ yb[p])) ^ (xb[p] & ya[p]);
#pragma unroll
    for (int p = 0; p < 3; p++) {
        xa[p] = (xa[p] & ~mask) | (z[p] & mask);
        xb[p] = (xb[p] & ~mask) | (z[next_party(p)] & mask);
    }
}

// Bit-position masks of the carry-out tree (RSS.inl:1112-1209).  At level L
// (h = 2^L) the reference's views are gEven = g[h-1 :: 2h], gOdd = g[2h-1 :: 2h],
// pEven = p[0 :: 2h], pOdd = p[h :: 2h] over the 64 bits of one value.
__device__ __forceinline__ u64 tree_mask_geven(int L) {
    switch (L) {
        case 0: return 0x5555555555555555ull;
        case 1: return 0x2222222222222222ull;
        case 2: return 0x0808080808080808ull;
        case 3: return 0x0080008000800080ull;
        case 4: return 0x0000800000008000ull;
        default: return 0x0000000080000000ull;
    }
}
__device__ __forceinline__ u64 tree_mask_peven(int L) {
    switch (L) {
        case 0: return 0x5555555555555555ull;
        case 1: return 0x1111111111111111ull;
        case 2: return 0x0101010101010101ull;
        case 3: return 0x0001000100010001ull;
        case 4: return 0x0000000100000001ull;
        default: return 0x0000000000000001ull;
    }
}

// dReLU (RSS.inl:806-852) on bit-packed shares: the reference expands the
// 64 bits of a value into 64 uint8 shares; AND / XOR are bitwise, so one
// 64-bit word per share holds the same bits.  Returns the result shares in U.
template <typename U>
__device__ __forceinline__ Sh3<U> drelu_(const Sh3<u64> &in) {
    // r = 0 + input; a_p = reconstruct(r)_p + 1 (each party's own view)
    u64 a[3];
#pragma unroll
    for (int p = 0; p < 3; p++) a[p] = in.s[p][0] + in.s[p][1] + in.s[prev_party(p)][0] + 1;
    // rbits = fill(1): A = (1s, 0), B = (0, 0), C = (0, 1s)
    u64 ra[3] = {~0ull, 0ull, 0ull}, rb[3] = {0ull, 0ull, ~0ull};
    const u64 top = 1ull << 63;
    U msba[3], msbb[3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        // setCarryOutMSB (bitwise.cuh:41-64), xor with a63 for A on shareA, C on shareB
        u64 a63 = a[p] >> 63;
        msba[p] = (U)(((ra[p] >> 63) ^ (p == PARTY_A ? a63 : 0ull)) & 1ull);
        msbb[p] = (U)(((rb[p] >> 63) ^ (p == PARTY_C ? a63 : 0ull)) & 1ull);
        a[p] |= top;
        ra[p] &= ~top;
        rb[p] &= ~top;
    }
    // g = rbits & abits on both shares; p = rbits ^ abits on A.sA and C.sB
    u64 ga[3], gb[3], pa[3], pb[3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        ga[p] = ra[p] & a[p];
        gb[p] = rb[p] & a[p];
        pa[p] = ra[p] ^ (p == PARTY_A ? a[p] : 0ull);
        pb[p] = rb[p] ^ (p == PARTY_C ? a[p] : 0ull);
    }
    // carryOut: 6 levels of gEven &= pOdd; pEven &= pOdd; gOdd ^= gEven
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
        rss_and_masked(ga, gb, ya, yb, mge);
#pragma unroll
        for (int p = 0; p < 3; p++) {
            ya[p] = pa[p] >> h;
            yb[p] = pb[p] >> h;
        }
        rss_and_masked(pa, pb, ya, yb, mpe);
#pragma unroll
        for (int p = 0; p < 3; p++) {
            ga[p] ^= (ga[p] & mge) << h;
            gb[p] ^= (gb[p] & mge) << h;
        }
    }
    // preResult = carry bit 63; ^= msb; result = fill(1) - preResult
    Sh3<U> res;
#pragma unroll
    for (int p = 0; p < 3; p++) {
        U prea = (U)((ga[p] >> 63) & 1ull) ^ msba[p];
        U preb = (U)((gb[p] >> 63) & 1ull) ^ msbb[p];
        res.s[p][0] = (U)((U)(p == PARTY_A ? 1 : 0) - prea);
        res.s[p][1] = (U)((U)(p == PARTY_C ? 1 : 0) - preb);
    }
    return res;
}

// selectShare (RSS.inl:542-569): z = ((y - x) * d) + x where d comes from the
// opened bit e_p = reconstruct(b)_p through convex_comb with c = 0
// (RSS.inl:33-61): d_A = (e?1:0, 0), d_B = (0, 0), d_C = (0, e?1:0).
template <typename U>
__device__ __forceinline__ Sh3<u64> select_(const Sh3<u64> &x, const Sh3<u64> &y, const Sh3<U> &b) {
    Sh3<u64> d, r;
#pragma unroll
    for (int p = 0; p < 3; p++) {
        U e = (U)(b.s[p][0] + b.s[p][1] + b.s[prev_party(p)][0]);
        bool one = (e == (U)1);
        d.s[p][0] = (one && p == PARTY_A) ? 1ull : 0ull;
        d.s[p][1] = (one && p == PARTY_C) ? 1ull : 0ull;
        r.s[p][0] = y.s[p][0] - x.s[p][0];
        r.s[p][1] = y.s[p][1] - x.s[p][1];
    }
    rss_mult_(r, d);
#pragma unroll
    for (int p = 0; p < 3; p++) {
        r.s[p][0] += x.s[p][0];
        r.s[p][1] += x.s[p][1];
    }
    return r;
}

// dividePublic by d = 2^k (RSS.inl:422-436, placeholder r = 1, r' = d of
// Precompute.h:107-119): a -= r' (fill(d): A.sA, C.sB); t_p = reconstruct(a)_p;
// a = r + (A.sA, C.sB += t_p / d) with the reference's FP64 truncating division.
__device__ __forceinline__ Sh3<u64> divide_public_pow2_(Sh3<u64> v, int k) {
    const u64 d = (u64)1 << k;
    v.s[PARTY_A][0] -= d;
    v.s[PARTY_C][1] -= d;
    u64 tA = v.s[PARTY_A][0] + v.s[PARTY_A][1] + v.s[PARTY_C][0];
    u64 tC = v.s[PARTY_C][0] + v.s[PARTY_C][1] + v.s[PARTY_B][0];
    Sh3<u64> o;
#pragma unroll
    for (int p = 0; p < 3; p++) o.s[p][0] = o.s[p][1] = 0;
    o.s[PARTY_A][0] = 1 + trunc_div_pow2(tA, k);
    o.s[PARTY_C][1] = 1 + trunc_div_pow2(tC, k);
    return o;
}

// dividePublic by any public d (RSS.inl:422-436 with r = 1, r' = d) with the
// reference's division (S)((double)(S)t / (double)(S)d) of the opened values.
__device__ __forceinline__ Sh3<u64> divide_public_(Sh3<u64> v, u64 d) {
    v.s[PARTY_A][0] -= d;
    v.s[PARTY_C][1] -= d;
    u64 tA = v.s[PARTY_A][0] + v.s[PARTY_A][1] + v.s[PARTY_C][0];
    u64 tC = v.s[PARTY_C][0] + v.s[PARTY_C][1] + v.s[PARTY_B][0];
    Sh3<u64> o;
#pragma unroll
    for (int p = 0; p < 3; p++) o.s[p][0] = o.s[p][1] = 0;
    o.s[PARTY_A][0] = 1 + signed_div_u64(tA, d);
    o.s[PARTY_C][1] = 1 + signed_div_u64(tC, d);
    return o;
}

// dividePublic by per-party public denominators d[p] (RSS.inl:438-455 with
// r = 1, r' = 0 + d): a -= r' (A.sA -= d_A, C.sB -= d_C); t_p = reconstruct(a)_p;
// a = r + (A.sA, C.sB += t_p / d_p).
__device__ __forceinline__ Sh3<u64> divide_public_local_(Sh3<u64> v, const u64 d[3]) {
    v.s[PARTY_A][0] -= d[PARTY_A];
    v.s[PARTY_C][1] -= d[PARTY_C];
    u64 t[3];
#pragma unroll
    for (int p = 0; p < 3; p++) t[p] = v.s[p][0] + v.s[p][1] + v.s[prev_party(p)][0];
    Sh3<u64> o;
#pragma unroll
    for (int p = 0; p < 3; p++) o.s[p][0] = o.s[p][1] = 0;
    o.s[PARTY_A][0] = 1 + signed_div_u64(t[PARTY_A], d[PARTY_A]);
    o.s[PARTY_C][1] = 1 + signed_div_u64(t[PARTY_C], d[PARTY_C]);
    return o;
}

// getPowers (RSS.inl:1211-1235) before its final reconstruct: for bit = 0..63,
// diff = in - 2^bit (A.sA, C.sB), dReLU(diff) in uint8 shares and
// powers = selectShare(powers, fill(bit), dReLU), so that powers ends at the
// largest bit with in - 2^bit >= 0.
__device__ __forceinline__ Sh3<u64> get_powers_(const Sh3<u64> &in) {
    Sh3<u64> powers;
#pragma unroll
    for (int p = 0; p < 3; p++) powers.s[p][0] = powers.s[p][1] = 0;
#pragma unroll 1
    for (int bit = 0; bit < 64; bit++) {
        Sh3<u64> diff = in, cur;
        diff.s[PARTY_A][0] -= (u64)1 << bit;
        diff.s[PARTY_C][1] -= (u64)1 << bit;
        Sh3<u8> cmp = drelu_<u8>(diff);
#pragma unroll
        for (int p = 0; p < 3; p++) cur.s[p][0] = cur.s[p][1] = 0;
        cur.s[PARTY_A][0] = (u64)bit;
        cur.s[PARTY_C][1] = (u64)bit;
        powers = select_(powers, cur, cmp);
    }
    return powers;
}

template <typename T>
__device__ __forceinline__ Sh3<T> load3(const T *__restrict__ base, size_t n, size_t i) {
    Sh3<T> v;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int s = 0; s < 2; s++) v.s[p][s] = base[(size_t)(2 * p + s) * n + i];
    return v;
}

template <typename T>
__device__ __forceinline__ void store3(T *__restrict__ base, size_t n, size_t i, const Sh3<T> &v) {
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int s = 0; s < 2; s++) base[(size_t)(2 * p + s) * n + i] = v.s[p][s];
}

}  // namespace pir
