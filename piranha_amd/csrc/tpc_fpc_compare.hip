// tpc_fpc_compare.hip -- element-wise protocols of the reference's 2-party
// (TPC) and 4-party (FPC) schemes: products and ANDs, dReLU, ReLU,
// selectShare and the maxpool tournament, each one pass per element with
// every party's shares in registers.
#include "tpc_fpc_compare.h"

#include "rss_device.h"

namespace pir {

namespace {

constexpr int BLK = 256;
unsigned grid(size_t n) { return (unsigned)((n + BLK - 1) / BLK); }

// One element's shares: s[0] = party A's, s[1] = party B's.
template <typename T>
struct Tp2 {
    T s[2];
};
template <typename T>
__device__ __forceinline__ Tp2<T> tload(const T *__restrict__ b, size_t n, size_t i) {
    return Tp2<T>{{b[i], b[n + i]}};
}
template <typename T>
__device__ __forceinline__ void tstore(T *__restrict__ b, size_t n, size_t i, const Tp2<T> &v) {
    b[i] = v.s[0];
    b[n + i] = v.s[1];
}

// TPC *= TPC (TPC.inl:242-272) with getBeaverTriples x = y = z = fill(1)
// (Precompute.h:100-105; fill puts 1 on party A only).  Both parties open
// e = x + this and f = y + rhs; then this = z, and party A alone (TPC +=
// DeviceData) adds (f - y) e and (-x) f.
template <typename T>
__device__ __forceinline__ Tp2<T> tpc_mult_(const Tp2<T> &self, const Tp2<T> &rhs) {
    const T e = (T)((T)(1 + self.s[0]) + (T)(0 + self.s[1]));
    const T f = (T)((T)(1 + rhs.s[0]) + (T)(0 + rhs.s[1]));
    Tp2<T> r;
    r.s[0] = (T)((T)(1 + (T)((T)(f - 1) * e)) + (T)((T)(0 - 1) * f));
    r.s[1] = 0;
    return r;
}

// TPC &= TPC (TPC.inl:281-312) with getBooleanBeaverTriples x = y = z =
// fill(1): e = open(x ^ this), f = open(y ^ rhs), opened by addition in the
// element type as reconstruct does; this = z, then party A alone (TPC ^=
// DeviceData) xors in (f ^ y) & e and x & f.
template <typename T>
__device__ __forceinline__ Tp2<T> tpc_and_(const Tp2<T> &self, const Tp2<T> &rhs) {
    const T e = (T)((T)(1 ^ self.s[0]) + (T)(0 ^ self.s[1]));
    const T f = (T)((T)(1 ^ rhs.s[0]) + (T)(0 ^ rhs.s[1]));
    Tp2<T> r;
    r.s[0] = (T)((T)1 ^ (T)((T)(f ^ 1) & e) ^ (T)(1 & f));
    r.s[1] = 0;
    return r;
}

// The dReLU circuit of one party that holds the public value's slot, on
// bit-packed words: `a` is the party's opened input + 1, its rbits are
// fill(1)'s all-ones; the MSB surgery of setCarryOutMSB (bitwise.cuh:41-64)
// or getMSBs / setMSBs (:126-182), then carryOut's six rounds of
// gEven &= pOdd; pEven &= pOdd; gOdd ^= gEven.  Returns preResult = carry ^ msb.
__device__ __forceinline__ u64 public_slot_drelu_pre(u64 a) {
    u64 r = ~0ull;
    const u64 top = 1ull << 63;
    const u64 msb = ((r >> 63) ^ (a >> 63)) & 1ull;
    a |= top;
    r &= ~top;
    u64 g = r & a, p = r ^ a;
#pragma unroll
    for (int L = 0; L < 6; L++) {
        const int h = 1 << L;
        const u64 mge = tree_mask_geven(L), mpe = tree_mask_peven(L);
        g = (g & ~mge) | (g & (p >> 1) & mge);
        p = (p & ~mpe) | (p & (p >> h) & mpe);
        g ^= (g & mge) << h;
    }
    return ((g >> 63) & 1ull) ^ msb;
}

// dReLU (TPC.inl:828-868).  Party B's bit shares are zero throughout: its
// rbits are fill(1)'s 0, `g &= abits` keeps them 0, `p ^= abits` (TPC ^=
// DeviceData) touches party A only, and every AND leaves party B with the
// triple's z share, 0.  So each opened sum of the carry tree is party A's bit,
// the bytes stay 0 / 1, the Beaver AND reduces to A's u & v, and one 64-bit
// word per array carries all 64 bit positions.  result = fill(1) - preResult.
template <typename U>
__device__ __forceinline__ Tp2<U> tpc_drelu_(const Tp2<u64> &in) {
    const U pre = (U)public_slot_drelu_pre(in.s[0] + in.s[1] + 1);
    return Tp2<U>{{(U)(1 - pre), (U)0}};
}

// selectShare (TPC.inl:585-613): e = open(0 ^ b); convex_comb with c = 0
// (tpc_convex_comb_functor, TPC.inl:32-53) gives d = (e == 1, 0); then
// z = (y - x) * d + x.
template <typename U>
__device__ __forceinline__ Tp2<u64> tpc_select_(const Tp2<u64> &x, const Tp2<u64> &y, const Tp2<U> &b) {
    const U e = (U)((U)(0 ^ b.s[0]) + (U)(0 ^ b.s[1]));
    const Tp2<u64> d{{e == (U)1 ? 1ull : 0ull, 0ull}};
    Tp2<u64> r{{y.s[0] - x.s[0], y.s[1] - x.s[1]}};
    r = tpc_mult_(r, d);
    r.s[0] += x.s[0];
    r.s[1] += x.s[1];
    return r;
}

template <typename T>
__global__ __launch_bounds__(BLK) void tpc_mult_kernel(T *__restrict__ x, const T *__restrict__ y, size_t n) {
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) tstore(x, n, i, tpc_mult_(tload(x, n, i), tload(y, n, i)));
}

template <typename T>
__global__ __launch_bounds__(BLK) void tpc_and_kernel(T *__restrict__ x, const T *__restrict__ y, size_t n) {
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) tstore(x, n, i, tpc_and_(tload(x, n, i), tload(y, n, i)));
}

// dReLU, or ReLU (TPC.inl:870-884: dReLU then selectShare(0, input, dresult)) when y is set
template <typename U>
__global__ __launch_bounds__(BLK) void tpc_relu_kernel(const u64 *__restrict__ x, size_t n, U *__restrict__ d,
                                                       u64 *__restrict__ y) {
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const Tp2<u64> v = tload(x, n, i);
    const Tp2<U> b = tpc_drelu_<U>(v);
    tstore(d, n, i, b);
    if (y) tstore(y, n, i, tpc_select_<U>(Tp2<u64>{{0, 0}}, v, b));
}

template <typename U>
__global__ __launch_bounds__(BLK) void tpc_select_kernel(const u64 *__restrict__ x, const u64 *__restrict__ y,
                                                         const U *__restrict__ b, u64 *__restrict__ z, size_t n) {
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) tstore(z, n, i, tpc_select_<U>(tload(x, n, i), tload(y, n, i), tload(b, n, i)));
}

// maxpool (TPC.inl:886-1030) over one window of K gathered values: even / odd
// views at stride 2, 4, ..., K; the even positions are written back.
template <typename U, int K>
__global__ __launch_bounds__(BLK) void tpc_maxpool_kernel(u64 *__restrict__ pools, size_t windows,
                                                          u64 *__restrict__ result, U *__restrict__ dres) {
    size_t w = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (w >= windows) return;
    const size_t n = windows * K;
    Tp2<u64> e[K];
    Tp2<U> d[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        e[j] = tload(pools, n, w * K + j);
        d[j] = Tp2<U>{{(U)1, (U)0}};  // dresult.fill(1)
    }
#pragma unroll
    for (int stride = 2; stride <= K; stride *= 2) {
#pragma unroll
        for (int j = 0; j < K / stride; j++) {
            const Tp2<u64> ev = e[j * stride], od = e[j * stride + stride / 2];
            const Tp2<U> b = tpc_drelu_<U>(Tp2<u64>{{ev.s[0] - od.s[0], ev.s[1] - od.s[1]}});
            e[j * stride] = tpc_select_<U>(od, ev, b);
            // negated = fill(1) - b; dresult &= expandCompare(b, negated)
            const Tp2<U> neg{{(U)(1 - b.s[0]), (U)(0 - b.s[1])}};
#pragma unroll
            for (int t = 0; t < stride; t++) d[j * stride + t] = tpc_and_(d[j * stride + t], t < stride / 2 ? b : neg);
        }
    }
    tstore(result, windows, w, e[0]);
#pragma unroll
    for (int j = 0; j < K; j += 2) tstore(pools, n, w * K + j, e[j]);
#pragma unroll
    for (int j = 0; j < K; j++) tstore(dres, n, w * K + j, d[j]);
}

// ---------------------------------------------------------------------------
// FPC: party q's share s holds component (q + s + 1) % 4 of x0 + x1 + x2 + x3;
// public values sit in x0, party B's shareC, C's shareB and D's shareA
// (FPC.inl:119-124).

template <typename T>
struct Fp4 {
    T s[4][3];
};
template <typename T>
__device__ __forceinline__ Fp4<T> fload(const T *__restrict__ b, size_t n, size_t i) {
    Fp4<T> v;
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int t = 0; t < 3; t++) v.s[q][t] = b[(size_t)(3 * q + t) * n + i];
    return v;
}
template <typename T>
__device__ __forceinline__ void fstore(T *__restrict__ b, size_t n, size_t i, const Fp4<T> &v) {
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int t = 0; t < 3; t++) b[(size_t)(3 * q + t) * n + i] = v.s[q][t];
}
__device__ __forceinline__ constexpr int fpc_public_slot(int q) { return q == 1 ? 2 : q == 2 ? 1 : q == 3 ? 0 : -1; }
__device__ __forceinline__ constexpr int fpc_prev(int q) { return (q + 3) % 4; }

// FPC *= FPC (FPC.inl:321-362) or &= (:374-398): every party's cross terms
// z = A.y1 + B.y0 and z' = A.y2 + C.y0 (^ and & for the AND), the share-wise
// product, then reshareFPC (:1183-1260, zero PRGs) over the pairs (0,1),
// (0,2), (0,3), (1,2), (1,3), (2,3): parties i, j and G add a mask, G the one
// party i sends, z' for the opposite pairs (0,2), (1,3) and z otherwise, into
// their slot of component H, (3 - q + H) % 4.  The reshare adds in the
// element type for the AND as well, as the reference does.
template <typename T, bool BOOLEAN>
__device__ __forceinline__ void fpc_product_(Fp4<T> &x, const Fp4<T> &y) {
    T z[4], zp[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const T a = x.s[q][0], b = x.s[q][1], c = x.s[q][2];
        if (BOOLEAN) {
            z[q] = (T)((T)(a & y.s[q][1]) ^ (T)(b & y.s[q][0]));
            zp[q] = (T)((T)(a & y.s[q][2]) ^ (T)(c & y.s[q][0]));
        } else {
            z[q] = (T)((T)(a * y.s[q][1]) + (T)(b * y.s[q][0]));
            zp[q] = (T)((T)(a * y.s[q][2]) + (T)(c * y.s[q][0]));
        }
#pragma unroll
        for (int t = 0; t < 3; t++) x.s[q][t] = BOOLEAN ? (T)(x.s[q][t] & y.s[q][t]) : (T)(x.s[q][t] * y.s[q][t]);
    }
    constexpr int PI[6] = {0, 0, 0, 1, 1, 2}, PJ[6] = {1, 2, 3, 2, 3, 3}, PH[6] = {3, 3, 2, 0, 0, 1};
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int i = PI[k], j = PJ[k], H = PH[k];
        const bool opp = (j - i) == 2;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (q == H) continue;
            const int src = (q == i || q == j) ? q : i;
            const int slot = (3 - q + H) % 4;
            x.s[q][slot] = (T)(x.s[q][slot] + (opp ? zp[src] : z[src]));
        }
    }
}

// dReLU (FPC.inl:929-989).  Every party's g and p arrays are zero outside its
// public slot: the rbits are fill(1), abits is public, `p ^= abits` touches the
// public slot only, and the AND's cross terms pair two different slots of one
// party, so its reshare masks are 0.  Each of B, C, D runs the one-slot circuit
// on its own opened value; party A (no public slot) ends with zeros.
template <typename U>
__device__ __forceinline__ Fp4<U> fpc_drelu_(const Fp4<u64> &in) {
    Fp4<U> r;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const u64 a = in.s[q][0] + in.s[q][1] + in.s[q][2] + in.s[fpc_prev(q)][0] + 1;
        const U bit = (U)(1 - (U)public_slot_drelu_pre(a));
#pragma unroll
        for (int t = 0; t < 3; t++) r.s[q][t] = t == fpc_public_slot(q) ? bit : (U)0;
    }
    return r;
}

// selectShare (FPC.inl:671-712): e = open(0 ^ b) per party; convex_comb with
// c = 0 (FPC_convex_comb_functor, :30-66) puts e == 1 in the public slot;
// z = (y - x) * d + x.
template <typename U>
__device__ __forceinline__ Fp4<u64> fpc_select_(const Fp4<u64> &x, const Fp4<u64> &y, const Fp4<U> &b) {
    Fp4<u64> d, r;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const U e = (U)(b.s[q][0] + b.s[q][1] + b.s[q][2] + b.s[fpc_prev(q)][0]);
#pragma unroll
        for (int t = 0; t < 3; t++) {
            d.s[q][t] = (t == fpc_public_slot(q) && e == (U)1) ? 1ull : 0ull;
            r.s[q][t] = y.s[q][t] - x.s[q][t];
        }
    }
    fpc_product_<u64, false>(r, d);
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int t = 0; t < 3; t++) r.s[q][t] += x.s[q][t];
    return r;
}

template <typename T, bool BOOLEAN>
__global__ __launch_bounds__(BLK) void fpc_product_kernel(T *__restrict__ x, const T *__restrict__ y, size_t n) {
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    Fp4<T> v = fload(x, n, i);
    fpc_product_<T, BOOLEAN>(v, fload(y, n, i));
    fstore(x, n, i, v);
}

template <typename U>
__global__ __launch_bounds__(BLK) void fpc_relu_kernel(const u64 *__restrict__ x, size_t n, U *__restrict__ d,
                                                       u64 *__restrict__ y) {
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const Fp4<u64> v = fload(x, n, i);
    const Fp4<U> b = fpc_drelu_<U>(v);
    fstore(d, n, i, b);
    if (y) {
        Fp4<u64> zero;
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int t = 0; t < 3; t++) zero.s[q][t] = 0;
        fstore(y, n, i, fpc_select_<U>(zero, v, b));
    }
}

template <typename U>
__global__ __launch_bounds__(BLK) void fpc_select_kernel(const u64 *__restrict__ x, const u64 *__restrict__ y,
                                                         const U *__restrict__ b, u64 *__restrict__ z, size_t n) {
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) fstore(z, n, i, fpc_select_<U>(fload(x, n, i), fload(y, n, i), fload(b, n, i)));
}

// maxpool (FPC.inl:1004-1181) for one window, in place like the reference's
// even views: each round reads the pair from pools, writes the selected value
// to the even position and ANDs the expanded comparison into dres.
template <typename U>
__global__ __launch_bounds__(BLK) void fpc_maxpool_kernel(u64 *__restrict__ pools, size_t windows, int K,
                                                          u64 *__restrict__ result, U *__restrict__ dres) {
    size_t w = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (w >= windows) return;
    const size_t n = windows * (size_t)K, base = w * (size_t)K;
    Fp4<U> one;
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int t = 0; t < 3; t++) one.s[q][t] = t == fpc_public_slot(q) ? (U)1 : (U)0;
    for (int j = 0; j < K; j++) fstore(dres, n, base + j, one);  // dresult.fill(1)
    for (int stride = 2; stride <= K; stride *= 2) {
        for (int j = 0; j < K / stride; j++) {
            const Fp4<u64> ev = fload(pools, n, base + (size_t)j * stride);
            const Fp4<u64> od = fload(pools, n, base + (size_t)j * stride + stride / 2);
            Fp4<u64> diff;
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int t = 0; t < 3; t++) diff.s[q][t] = ev.s[q][t] - od.s[q][t];
            const Fp4<U> b = fpc_drelu_<U>(diff);
            fstore(pools, n, base + (size_t)j * stride, fpc_select_<U>(od, ev, b));
            // negated = fill(1) - b; dresult &= expandCompare(b, negated)
            Fp4<U> neg;
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int t = 0; t < 3; t++) neg.s[q][t] = (U)(one.s[q][t] - b.s[q][t]);
            for (int t = 0; t < stride; t++) {
                const size_t at = base + (size_t)j * stride + t;
                Fp4<U> dv = fload(dres, n, at);
                fpc_product_<U, true>(dv, t < stride / 2 ? b : neg);
                fstore(dres, n, at, dv);
            }
        }
    }
    fstore(result, windows, w, fload(pools, n, base));
}

// ---------------------------------------------------------------------------
// getPowers and taylorSeries (TPC.inl:1202-1290, FPC.inl:1418-1526) through a
// scheme policy: the share algebra of one element for all parties.

struct TpcPolicy {
    static constexpr int P = 2;
    using V = Tp2<u64>;
    using B = Tp2<u8>;
    static __device__ __forceinline__ V load(const u64 *b, size_t n, size_t i) { return tload(b, n, i); }
    static __device__ __forceinline__ void store(u64 *b, size_t n, size_t i, const V &v) { tstore(b, n, i, v); }
    static __device__ __forceinline__ V fill(u64 v) { return V{{v, 0}}; }  // TPC.inl:103-106
    static __device__ __forceinline__ u64 open(const V &x, int) { return x.s[0] + x.s[1]; }
    // TPC -= T and += DeviceData: party A only (TPC.inl:145-189)
    static __device__ __forceinline__ void sub_public(V &x, u64 v) { x.s[0] -= v; }
    static __device__ __forceinline__ void add_public(V &x, const u64 c[P]) { x.s[0] += c[0]; }
    // *= DeviceData, <<= DeviceData: every share (TPC.inl:191-226)
    static __device__ __forceinline__ void scale(V &x, const u64 c[P]) {
        for (int q = 0; q < P; q++) x.s[q] *= c[q];
    }
    static __device__ __forceinline__ void shl(V &x, const u64 c[P]) {
        for (int q = 0; q < P; q++) x.s[q] = shl_clamped(x.s[q], c[q]);
    }
    static __device__ __forceinline__ void add(V &x, const V &y) {
        for (int q = 0; q < P; q++) x.s[q] += y.s[q];
    }
    static __device__ __forceinline__ V mult(const V &x, const V &y) { return tpc_mult_(x, y); }
    static __device__ __forceinline__ B drelu(const V &x) { return tpc_drelu_<u8>(x); }
    static __device__ __forceinline__ V select(const V &x, const V &y, const B &b) { return tpc_select_<u8>(x, y, b); }
    // dividePublic(TPC, DeviceData) (TPC.inl:386-407): r = fill(1), r' = 0 + d
    // (party A), a -= r'; each party opens a and divides by its d; a = r + opened (party A)
    static __device__ __forceinline__ V divide_public(V a, const u64 d[P]) {
        a.s[0] -= d[0];
        return V{{1 + signed_div_u64(open(a, 0), d[0]), 0}};
    }
};

struct FpcPolicy {
    static constexpr int P = 4;
    using V = Fp4<u64>;
    using B = Fp4<u8>;
    static __device__ __forceinline__ V load(const u64 *b, size_t n, size_t i) { return fload(b, n, i); }
    static __device__ __forceinline__ void store(u64 *b, size_t n, size_t i, const V &v) { fstore(b, n, i, v); }
    static __device__ __forceinline__ V fill(u64 v) {  // FPC.inl:119-124
        V x;
        for (int q = 0; q < P; q++)
            for (int t = 0; t < 3; t++) x.s[q][t] = t == fpc_public_slot(q) ? v : 0;
        return x;
    }
    static __device__ __forceinline__ u64 open(const V &x, int q) {  // FPC.inl:582-604
        return x.s[q][0] + x.s[q][1] + x.s[q][2] + x.s[fpc_prev(q)][0];
    }
    // FPC -= T and += DeviceData: the public slot (FPC.inl:184-250)
    static __device__ __forceinline__ void sub_public(V &x, u64 v) {
        for (int q = 1; q < P; q++) x.s[q][fpc_public_slot(q)] -= v;
    }
    static __device__ __forceinline__ void add_public(V &x, const u64 c[P]) {
        for (int q = 1; q < P; q++) x.s[q][fpc_public_slot(q)] += c[q];
    }
    // *= DeviceData, <<= DeviceData: every share (FPC.inl:252-300)
    static __device__ __forceinline__ void scale(V &x, const u64 c[P]) {
        for (int q = 0; q < P; q++)
            for (int t = 0; t < 3; t++) x.s[q][t] *= c[q];
    }
    static __device__ __forceinline__ void shl(V &x, const u64 c[P]) {
        for (int q = 0; q < P; q++)
            for (int t = 0; t < 3; t++) x.s[q][t] = shl_clamped(x.s[q][t], c[q]);
    }
    static __device__ __forceinline__ void add(V &x, const V &y) {
        for (int q = 0; q < P; q++)
            for (int t = 0; t < 3; t++) x.s[q][t] += y.s[q][t];
    }
    static __device__ __forceinline__ V mult(V x, const V &y) {
        fpc_product_<u64, false>(x, y);
        return x;
    }
    static __device__ __forceinline__ B drelu(const V &x) { return fpc_drelu_<u8>(x); }
    static __device__ __forceinline__ V select(const V &x, const V &y, const B &b) { return fpc_select_<u8>(x, y, b); }
    // dividePublic(FPC, DeviceData) (FPC.inl:566-580 with the vector getDividedShares)
    static __device__ __forceinline__ V divide_public(V a, const u64 d[P]) {
        for (int q = 1; q < P; q++) a.s[q][fpc_public_slot(q)] -= d[q];
        V o = fill(0);
        for (int q = 1; q < P; q++) o.s[q][fpc_public_slot(q)] = 1 + signed_div_u64(open(a, q), d[q]);
        return o;
    }
};

// getPowers: 64 rounds of diff = in - 2^bit, dReLU(diff) in uint8 shares,
// powers = selectShare(powers, fill(bit), dReLU).
template <class S>
__device__ __forceinline__ typename S::V scheme_get_powers_(const typename S::V &in) {
    typename S::V powers = S::fill(0);
#pragma unroll 1
    for (int bit = 0; bit < 64; bit++) {
        typename S::V diff = in;
        S::sub_public(diff, (u64)1 << bit);
        powers = S::select(powers, S::fill((u64)bit), S::drelu(diff));
    }
    return powers;
}

template <class S>
__global__ __launch_bounds__(BLK) void scheme_get_powers_kernel(const u64 *__restrict__ in, size_t n,
                                                                u64 *__restrict__ pow) {
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const typename S::V pw = scheme_get_powers_<S>(S::load(in, n, i));
#pragma unroll
    for (int q = 0; q < S::P; q++) pow[(size_t)q * n + i] = S::open(pw, q);
}

// taylorSeries for one element, each party with its own opened power
// pow_q = getPowers_q + 1 and ones_q = 1 << pow_q (see the RSS taylor_kernel).
template <class S>
__global__ __launch_bounds__(BLK) void scheme_taylor_kernel(const u64 *__restrict__ in, size_t n,
                                                            u64 *__restrict__ out, const TaylorTables t) {
    using V = typename S::V;
    constexpr int P = S::P;
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const V x = S::load(in, n, i);
    const V pw = scheme_get_powers_<S>(x);
    u64 pow[P], ones[P], c[P];
#pragma unroll
    for (int q = 0; q < P; q++) {
        pow[q] = S::open(pw, q) + 1;
        ones[q] = shl_clamped(1, pow[q]);
    }
    V o = S::fill(0);
    auto lookup = [&](const TaylorTab &tab) {
#pragma unroll
        for (int q = 0; q < P; q++) c[q] = tab.at(pow[q]);
    };
    if (t.use_a2) {  // scratch = in * in / ones * a2Coeff / ones
        V sc = S::divide_public(S::mult(x, x), ones);
        lookup(t.c2);
        S::scale(sc, c);
        S::add(o, S::divide_public(sc, ones));
    }
    if (t.use_a1) {  // scratch = in * a1Coeff / ones
        V sc = x;
        lookup(t.c1);
        S::scale(sc, c);
        S::add(o, S::divide_public(sc, ones));
    }
    lookup(t.c0);
    S::add_public(o, c);
    lookup(t.f);
    S::scale(o, c);
    o = S::divide_public(o, ones);
    // pow -= FP: the non-negative part divides, the negative part shifts every share left
    u64 neg[P];
#pragma unroll
    for (int q = 0; q < P; q++) {
        const u64 e = pow[q] - (u64)t.fp;
        ones[q] = shl_clamped(1, (s64)e >= 0 ? e : 0);
        neg[q] = (s64)e < 0 ? (u64)0 - e : 0;
    }
    o = S::divide_public(o, ones);
    S::shl(o, neg);
    S::store(out, n, i, o);
}

// dividePublic(Share, DeviceData): d [P][n], each party's public denominators
template <class S>
__global__ __launch_bounds__(BLK) void scheme_divide_public_kernel(u64 *__restrict__ a, size_t n,
                                                                   const u64 *__restrict__ d) {
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    u64 dq[S::P];
#pragma unroll
    for (int q = 0; q < S::P; q++) dq[q] = d[(size_t)q * n + i];
    S::store(a, n, i, S::divide_public(S::load(a, n, i), dq));
}

}  // namespace

template <class S>
static int scheme_divide_public(u64 *a, size_t n, const u64 *d, hipStream_t s) {
    if (n == 0) return OK;
    hipLaunchKernelGGL(scheme_divide_public_kernel<S>, dim3(grid(n)), dim3(BLK), 0, s, a, n, d);
    PIR_LAUNCH_CHECK();
    return OK;
}
int tpc_divide_public_vec(u64 *a, size_t n, const u64 *d, hipStream_t s) {
    return scheme_divide_public<TpcPolicy>(a, n, d, s);
}
int fpc_divide_public_vec(u64 *a, size_t n, const u64 *d, hipStream_t s) {
    return scheme_divide_public<FpcPolicy>(a, n, d, s);
}

template <class S>
static int scheme_get_powers(const u64 *in, size_t n, u64 *pow, hipStream_t s) {
    if (n == 0) return OK;
    hipLaunchKernelGGL(scheme_get_powers_kernel<S>, dim3(grid(n)), dim3(BLK), 0, s, in, n, pow);
    PIR_LAUNCH_CHECK();
    return OK;
}

template <class S>
static int scheme_taylor(const u64 *in, size_t n, u64 *out, const TaylorTables &t, hipStream_t s) {
    if (n == 0) return OK;
    hipLaunchKernelGGL(scheme_taylor_kernel<S>, dim3(grid(n)), dim3(BLK), 0, s, in, n, out, t);
    PIR_LAUNCH_CHECK();
    return OK;
}

int tpc_get_powers(const u64 *in, size_t n, u64 *pow, hipStream_t s) { return scheme_get_powers<TpcPolicy>(in, n, pow, s); }
int fpc_get_powers(const u64 *in, size_t n, u64 *pow, hipStream_t s) { return scheme_get_powers<FpcPolicy>(in, n, pow, s); }
int tpc_taylor(const u64 *in, size_t n, u64 *out, const TaylorTables &t, hipStream_t s) {
    return scheme_taylor<TpcPolicy>(in, n, out, t, s);
}
int fpc_taylor(const u64 *in, size_t n, u64 *out, const TaylorTables &t, hipStream_t s) {
    return scheme_taylor<FpcPolicy>(in, n, out, t, s);
}

int fpc_mult(u64 *x, const u64 *y, size_t n, hipStream_t s) {
    if (n == 0) return OK;
    hipLaunchKernelGGL((fpc_product_kernel<u64, false>), dim3(grid(n)), dim3(BLK), 0, s, x, y, n);
    PIR_LAUNCH_CHECK();
    return OK;
}

int fpc_and(void *x, const void *y, size_t n, int eb, hipStream_t s) {
    PIR_REQUIRE(eb == 1 || eb == 8, "FPC &=: element size must be 1 or 8");
    if (n == 0) return OK;
    if (eb == 1)
        hipLaunchKernelGGL((fpc_product_kernel<u8, true>), dim3(grid(n)), dim3(BLK), 0, s, (u8 *)x, (const u8 *)y, n);
    else
        hipLaunchKernelGGL((fpc_product_kernel<u64, true>), dim3(grid(n)), dim3(BLK), 0, s, (u64 *)x, (const u64 *)y,
                           n);
    PIR_LAUNCH_CHECK();
    return OK;
}

int fpc_relu(const u64 *x, size_t n, void *d, u64 *y, int eb, hipStream_t s) {
    PIR_REQUIRE(eb == 1 || eb == 8, "FPC dReLU: element size must be 1 or 8");
    if (n == 0) return OK;
    if (eb == 1) hipLaunchKernelGGL(fpc_relu_kernel<u8>, dim3(grid(n)), dim3(BLK), 0, s, x, n, (u8 *)d, y);
    else hipLaunchKernelGGL(fpc_relu_kernel<u64>, dim3(grid(n)), dim3(BLK), 0, s, x, n, (u64 *)d, y);
    PIR_LAUNCH_CHECK();
    return OK;
}

int fpc_select(const u64 *x, const u64 *y, const void *b, u64 *z, size_t n, int eb, hipStream_t s) {
    PIR_REQUIRE(eb == 1 || eb == 8, "FPC selectShare: element size must be 1 or 8");
    if (n == 0) return OK;
    if (eb == 1) hipLaunchKernelGGL(fpc_select_kernel<u8>, dim3(grid(n)), dim3(BLK), 0, s, x, y, (const u8 *)b, z, n);
    else hipLaunchKernelGGL(fpc_select_kernel<u64>, dim3(grid(n)), dim3(BLK), 0, s, x, y, (const u64 *)b, z, n);
    PIR_LAUNCH_CHECK();
    return OK;
}

int fpc_maxpool(u64 *pools, size_t n, u64 *result, void *d, int k, int eb, hipStream_t s) {
    PIR_REQUIRE(eb == 1 || eb == 8, "FPC maxpool: element size must be 1 or 8");
    PIR_REQUIRE(k == 2 || k == 4 || k == 8 || k == 16, "FPC maxpool: k must be 2, 4, 8 or 16");
    PIR_REQUIRE(n % (size_t)k == 0, "FPC maxpool: pools size must be a multiple of k");
    const size_t windows = n / k;
    if (windows == 0) return OK;
    if (eb == 1)
        hipLaunchKernelGGL(fpc_maxpool_kernel<u8>, dim3(grid(windows)), dim3(BLK), 0, s, pools, windows, k, result,
                           (u8 *)d);
    else
        hipLaunchKernelGGL(fpc_maxpool_kernel<u64>, dim3(grid(windows)), dim3(BLK), 0, s, pools, windows, k, result,
                           (u64 *)d);
    PIR_LAUNCH_CHECK();
    return OK;
}

int tpc_mult(u64 *x, const u64 *y, size_t n, hipStream_t s) {
    if (n == 0) return OK;
    hipLaunchKernelGGL(tpc_mult_kernel<u64>, dim3(grid(n)), dim3(BLK), 0, s, x, y, n);
    PIR_LAUNCH_CHECK();
    return OK;
}

int tpc_and(void *x, const void *y, size_t n, int eb, hipStream_t s) {
    PIR_REQUIRE(eb == 1 || eb == 8, "TPC &=: element size must be 1 or 8");
    if (n == 0) return OK;
    if (eb == 1) hipLaunchKernelGGL(tpc_and_kernel<u8>, dim3(grid(n)), dim3(BLK), 0, s, (u8 *)x, (const u8 *)y, n);
    else hipLaunchKernelGGL(tpc_and_kernel<u64>, dim3(grid(n)), dim3(BLK), 0, s, (u64 *)x, (const u64 *)y, n);
    PIR_LAUNCH_CHECK();
    return OK;
}

int tpc_relu(const u64 *x, size_t n, void *d, u64 *y, int eb, hipStream_t s) {
    PIR_REQUIRE(eb == 1 || eb == 8, "TPC dReLU: element size must be 1 or 8");
    if (n == 0) return OK;
    if (eb == 1) hipLaunchKernelGGL(tpc_relu_kernel<u8>, dim3(grid(n)), dim3(BLK), 0, s, x, n, (u8 *)d, y);
    else hipLaunchKernelGGL(tpc_relu_kernel<u64>, dim3(grid(n)), dim3(BLK), 0, s, x, n, (u64 *)d, y);
    PIR_LAUNCH_CHECK();
    return OK;
}

int tpc_select(const u64 *x, const u64 *y, const void *b, u64 *z, size_t n, int eb, hipStream_t s) {
    PIR_REQUIRE(eb == 1 || eb == 8, "TPC selectShare: element size must be 1 or 8");
    if (n == 0) return OK;
    if (eb == 1) hipLaunchKernelGGL(tpc_select_kernel<u8>, dim3(grid(n)), dim3(BLK), 0, s, x, y, (const u8 *)b, z, n);
    else hipLaunchKernelGGL(tpc_select_kernel<u64>, dim3(grid(n)), dim3(BLK), 0, s, x, y, (const u64 *)b, z, n);
    PIR_LAUNCH_CHECK();
    return OK;
}

template <typename U>
static int tpc_maxpool_t(u64 *pools, size_t windows, u64 *result, U *d, int k, hipStream_t s) {
    switch (k) {
        case 2: hipLaunchKernelGGL((tpc_maxpool_kernel<U, 2>), dim3(grid(windows)), dim3(BLK), 0, s, pools, windows, result, d); break;
        case 4: hipLaunchKernelGGL((tpc_maxpool_kernel<U, 4>), dim3(grid(windows)), dim3(BLK), 0, s, pools, windows, result, d); break;
        case 8: hipLaunchKernelGGL((tpc_maxpool_kernel<U, 8>), dim3(grid(windows)), dim3(BLK), 0, s, pools, windows, result, d); break;
        case 16: hipLaunchKernelGGL((tpc_maxpool_kernel<U, 16>), dim3(grid(windows)), dim3(BLK), 0, s, pools, windows, result, d); break;
        default: set_error("TPC maxpool: k must be 2, 4, 8 or 16"); return ERR_UNSUPPORTED;
    }
    PIR_LAUNCH_CHECK();
    return OK;
}

int tpc_maxpool(u64 *pools, size_t n, u64 *result, void *d, int k, int eb, hipStream_t s) {
    PIR_REQUIRE(eb == 1 || eb == 8, "TPC maxpool: element size must be 1 or 8");
    PIR_REQUIRE(k > 0 && n % (size_t)k == 0, "TPC maxpool: pools size must be a multiple of k");
    const size_t windows = n / k;
    if (windows == 0) return OK;
    return eb == 1 ? tpc_maxpool_t(pools, windows, result, (u8 *)d, k, s)
                   : tpc_maxpool_t(pools, windows, result, (u64 *)d, k, s);
}

}  // namespace pir
