// tpc_compare.hip -- the reference's 2-party (TPC) element-wise protocols:
// Beaver product and AND, dReLU, ReLU, selectShare and the maxpool tournament,
// each one pass per element with both parties' shares in registers.
#include "tpc_compare.h"

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

// dReLU (TPC.inl:828-868) on bit-packed words.  Party B's bit shares are zero
// throughout: its rbits are fill(1)'s 0, `g &= abits` keeps them 0, `p ^=
// abits` (TPC ^= DeviceData) touches party A only, and every AND leaves party
// B with the triple's z share, 0.  So each opened sum of the carry tree is
// party A's bit, the bytes stay 0 / 1, the Beaver AND reduces to A's
// u & v, and one 64-bit word per array carries all 64 bit positions.
template <typename U>
__device__ __forceinline__ Tp2<U> tpc_drelu_(const Tp2<u64> &in) {
    u64 a = in.s[0] + in.s[1] + 1;  // a = reconstruct(0 + input) + 1, the same for both parties
    u64 r = ~0ull;                  // party A's rbits: fill(1)
    const u64 top = 1ull << 63;
    // setCarryOutMSB (bitwise.cuh:41-64) with the xor of a's MSB for party A
    const U msb = (U)(((r >> 63) ^ (a >> 63)) & 1ull);
    a |= top;
    r &= ~top;
    u64 g = r & a, p = r ^ a;
    // carryOut (TPC.inl:1092-1200): gEven &= pOdd; pEven &= pOdd; gOdd ^= gEven
#pragma unroll
    for (int L = 0; L < 6; L++) {
        const int h = 1 << L;
        const u64 mge = tree_mask_geven(L), mpe = tree_mask_peven(L);
        g = (g & ~mge) | (g & (p >> 1) & mge);
        p = (p & ~mpe) | (p & (p >> h) & mpe);
        g ^= (g & mge) << h;
    }
    // preResult = carry ^ msb; result = fill(1) - preResult (party B: 0 - 0)
    const U pre = (U)((U)((g >> 63) & 1ull) ^ msb);
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

}  // namespace

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
