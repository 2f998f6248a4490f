// fused_rss.hip -- fused, wavefront-level 3-party RSS protocol kernels.
//
// Each kernel reads every party's shares of an element once, runs the whole
// protocol step of the reference for all three parties in registers (the
// inter-party messages are register moves) and writes each output share once.
// Reference functions restated: dReLU / carryOut (RSS.inl:806-852,1112-1209),
// selectShare (RSS.inl:542-569), ReLU (RSS.inl:854-865), maxpool
// (RSS.inl:867-1017) with maxpool_im2row / stride_pad / expandCompare /
// truncate_cols (convolution.cuh:57-108,176-191; bitwise.cuh:66-92;
// matrix.cuh:142-155), reconstruct (RSS.inl:457-479), reshare
// (RSS.inl:1052-1074), RSS *= and &= (RSS.inl:283-345), dividePublic
// (RSS.inl:422-455) and dividePublic3out3 (RSS.inl:1076-1094).
#include "fused_rss.h"
#include "rss_device.h"
#include "rss_masked.h"

namespace pir {

constexpr int EW_BLOCK = 256;

static unsigned ew_grid(size_t n) {
    size_t g = (n + EW_BLOCK - 1) / EW_BLOCK;
    return (unsigned)(g == 0 ? 1 : g);
}

// ---------------------------------------------------------------------------
// dReLU / ReLU

// MASKED: the masked precompute mode (rss_masked.h), element i drawing from
// the call's word stream g.
template <typename U, bool RELU, bool MASKED>
__global__ __launch_bounds__(EW_BLOCK) void relu_kernel(const u64 *__restrict__ x, size_t n, U *__restrict__ dres,
                                                        u64 *__restrict__ y, PrgStream g) {
    size_t i = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    Sh3<u64> in = load3(x, n, i);
    Sh3<U> b = MASKED ? drelu_m<U>(in, g, i, 0) : drelu_<U>(in);
    store3(dres, n, i, b);
    if (RELU) {
        Sh3<u64> zero;
#pragma unroll
        for (int p = 0; p < 3; p++) zero.s[p][0] = zero.s[p][1] = 0;
        store3(y, n, i, MASKED ? select_m<U>(zero, in, b, g, i, PRG_DRELU_WORDS) : select_<U>(zero, in, b));
    }
}

// Wide form: EPT consecutive elements per thread, every share array read and
// written with 16-byte accesses and the bit shares with one EPT-byte (u8) or
// 16-byte (u64) store per array, so a wave keeps more bytes in flight per
// memory instruction.  The host uses it when n % EPT == 0 and every array
// base is aligned for those accesses.
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

template <int EPT>
__device__ __forceinline__ void load3v(const u64 *__restrict__ base, size_t n, size_t i0, Sh3<u64> (&v)[EPT]) {
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const u64x2 *src = (const u64x2 *)(base + (size_t)(2 * p + s) * n + i0);
#pragma unroll
            for (int j = 0; j < EPT / 2; j++) {
                u64x2 t = src[j];
                v[2 * j].s[p][s] = t.x;
                v[2 * j + 1].s[p][s] = t.y;
            }
        }
}

template <int EPT>
__device__ __forceinline__ void store3v(u64 *__restrict__ base, size_t n, size_t i0, const Sh3<u64> (&v)[EPT]) {
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int s = 0; s < 2; s++) {
            u64x2 *dst = (u64x2 *)(base + (size_t)(2 * p + s) * n + i0);
#pragma unroll
            for (int j = 0; j < EPT / 2; j++) dst[j] = (u64x2){v[2 * j].s[p][s], v[2 * j + 1].s[p][s]};
        }
}

template <int EPT>
__device__ __forceinline__ void store3v(u8 *__restrict__ base, size_t n, size_t i0, const Sh3<u8> (&v)[EPT]) {
    static_assert(EPT == 2 || EPT == 4 || EPT == 8, "byte group of 2, 4 or 8");
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int s = 0; s < 2; s++) {
            u64 w = 0;
#pragma unroll
            for (int j = 0; j < EPT; j++) w |= (u64)v[j].s[p][s] << (8 * j);
            u8 *dst = base + (size_t)(2 * p + s) * n + i0;
            if (EPT == 2) *(uint16_t *)dst = (uint16_t)w;
            else if (EPT == 4) *(u32 *)dst = (u32)w;
            else *(u64 *)dst = w;
        }
}

template <typename U, bool RELU, int EPT, bool MASKED>
__global__ __launch_bounds__(EW_BLOCK) void relu_wide_kernel(const u64 *__restrict__ x, size_t n,
                                                             U *__restrict__ dres, u64 *__restrict__ y, PrgStream g) {
    size_t i0 = ((size_t)blockIdx.x * EW_BLOCK + threadIdx.x) * EPT;
    if (i0 >= n) return;
    Sh3<u64> in[EPT];
    load3v<EPT>(x, n, i0, in);
    Sh3<U> b[EPT];
#pragma unroll
    for (int j = 0; j < EPT; j++) b[j] = MASKED ? drelu_m<U>(in[j], g, i0 + j, 0) : drelu_<U>(in[j]);
    store3v<EPT>(dres, n, i0, b);
    if (RELU) {
        Sh3<u64> zero, out[EPT];
#pragma unroll
        for (int p = 0; p < 3; p++) zero.s[p][0] = zero.s[p][1] = 0;
#pragma unroll
        for (int j = 0; j < EPT; j++)
            out[j] = MASKED ? select_m<U>(zero, in[j], b[j], g, i0 + j, PRG_DRELU_WORDS) : select_<U>(zero, in[j], b[j]);
        store3v<EPT>(y, n, i0, out);
    }
}

constexpr int RELU_EPT = 4;

static bool aligned_to(const void *p, size_t a) { return p == nullptr || ((uintptr_t)p % a) == 0; }

template <typename U, bool RELU, bool MASKED>
static void launch_relu_as(const u64 *x, size_t n, U *dres, u64 *y, const PrgStream &g, hipStream_t s) {
    if (n % RELU_EPT == 0 && aligned_to(x, 16) && aligned_to(y, 16) &&
        aligned_to(dres, sizeof(U) == 1 ? RELU_EPT : 16)) {
        hipLaunchKernelGGL((relu_wide_kernel<U, RELU, RELU_EPT, MASKED>), dim3(ew_grid(n / RELU_EPT)),
                           dim3(EW_BLOCK), 0, s, x, n, dres, y, g);
    } else {
        hipLaunchKernelGGL((relu_kernel<U, RELU, MASKED>), dim3(ew_grid(n)), dim3(EW_BLOCK), 0, s, x, n, dres, y, g);
    }
}

template <typename U, bool RELU>
static int launch_relu(const u64 *x, size_t n, U *dres, u64 *y, const PrgStream &g, hipStream_t s) {
    if (n == 0) return OK;
    if (g.on) launch_relu_as<U, RELU, true>(x, n, dres, y, g, s);
    else launch_relu_as<U, RELU, false>(x, n, dres, y, g, s);
    PIR_LAUNCH_CHECK();
    return OK;
}

template <typename U>
int fused_drelu(const u64 *x, size_t n, U *dres, hipStream_t s, const PrgStream &g) {
    return launch_relu<U, false>(x, n, dres, (u64 *)nullptr, g, s);
}

template <typename U>
int fused_relu(const u64 *x, size_t n, u64 *y, U *dres, hipStream_t s, const PrgStream &g) {
    return launch_relu<U, true>(x, n, dres, y, g, s);
}

// ---------------------------------------------------------------------------
// selectShare

template <typename U, bool MASKED>
__global__ __launch_bounds__(EW_BLOCK) void select_kernel(const u64 *__restrict__ x, const u64 *__restrict__ y,
                                                          const U *__restrict__ b, u64 *__restrict__ z, size_t n,
                                                          PrgStream g) {
    size_t i = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    const Sh3<u64> xv = load3(x, n, i), yv = load3(y, n, i);
    const Sh3<U> bv = load3(b, n, i);
    store3(z, n, i, MASKED ? select_m<U>(xv, yv, bv, g, i, 0) : select_<U>(xv, yv, bv));
}

template <typename U>
int fused_select(const u64 *x, const u64 *y, const U *b, u64 *z, size_t n, hipStream_t s, const PrgStream &g) {
    if (n == 0) return OK;
    if (g.on) hipLaunchKernelGGL((select_kernel<U, true>), dim3(ew_grid(n)), dim3(EW_BLOCK), 0, s, x, y, b, z, n, g);
    else hipLaunchKernelGGL((select_kernel<U, false>), dim3(ew_grid(n)), dim3(EW_BLOCK), 0, s, x, y, b, z, n, g);
    PIR_LAUNCH_CHECK();
    return OK;
}

// ---------------------------------------------------------------------------
// maxpool tournament over K (a power of two) gathered values of one window.

template <typename U>
__device__ __forceinline__ Sh3<U> fill1_() {
    Sh3<U> v;
#pragma unroll
    for (int p = 0; p < 3; p++) {
        v.s[p][0] = (U)(p == PARTY_A ? 1 : 0);
        v.s[p][1] = (U)(p == PARTY_C ? 1 : 0);
    }
    return v;
}

template <typename U>
__device__ __forceinline__ void rss_and_elem(Sh3<U> &x, const Sh3<U> &y) {
    U z[3];
#pragma unroll
    for (int p = 0; p < 3; p++) z[p] = (U)((x.s[p][0] & (U)(y.s[p][0] ^ y.s[p][1])) ^ (x.s[p][1] & y.s[p][0]));
#pragma unroll
    for (int p = 0; p < 3; p++) {
        x.s[p][0] = z[p];
        x.s[p][1] = z[next_party(p)];
    }
}

// RSS.inl:867-1017 for one window: even/odd views at stride 2, 4, ..., K.
template <typename U, int K>
__device__ __forceinline__ void tournament(Sh3<u64> (&e)[K], Sh3<U> (&d)[K]) {
#pragma unroll
    for (int j = 0; j < K; j++) d[j] = fill1_<U>();
#pragma unroll
    for (int stride = 2; stride <= K; stride *= 2) {
        const int cnt = K / stride;
        Sh3<U> b[K / 2];
#pragma unroll
        for (int j = 0; j < cnt; j++) {
            const Sh3<u64> &ev = e[j * stride];
            const Sh3<u64> &od = e[j * stride + stride / 2];
            Sh3<u64> diff;
#pragma unroll
            for (int p = 0; p < 3; p++)
#pragma unroll
                for (int s = 0; s < 2; s++) diff.s[p][s] = ev.s[p][s] - od.s[p][s];
            b[j] = drelu_<U>(diff);
            e[j * stride] = select_<U>(od, ev, b[j]);
        }
        // expandCompare(b, 1 - b) then dresult &= expanded
#pragma unroll
        for (int j = 0; j < cnt; j++) {
            Sh3<U> neg;
#pragma unroll
            for (int p = 0; p < 3; p++) {
                neg.s[p][0] = (U)((U)(p == PARTY_A ? 1 : 0) - b[j].s[p][0]);
                neg.s[p][1] = (U)((U)(p == PARTY_C ? 1 : 0) - b[j].s[p][1]);
            }
#pragma unroll
            for (int t = 0; t < stride; t++) rss_and_elem<U>(d[j * stride + t], t < stride / 2 ? b[j] : neg);
        }
    }
}

// The masked form of the tournament (rss_masked.h) for window `win`: compare m
// (in round order) draws words 18m .. 18m + 17 (its dReLU, then its
// selectShare); the AND of round t at pool position pos draws words
// 18(K-1) + 2(tK + pos) and + 1.  1 - b becomes b ^ 1 in the x0 slot.
template <typename U, int K>
__device__ __forceinline__ void tournament_m(Sh3<u64> (&e)[K], Sh3<U> (&d)[K], const PrgStream &g, u64 win) {
#pragma unroll
    for (int j = 0; j < K; j++) d[j] = fill1_<U>();
    u32 m = 0;
    int t = 0;
#pragma unroll
    for (int stride = 2; stride <= K; stride *= 2, t++) {
        const int cnt = K / stride;
        Sh3<U> b[K / 2];
#pragma unroll
        for (int j = 0; j < cnt; j++, m++) {
            const Sh3<u64> &ev = e[j * stride];
            const Sh3<u64> &od = e[j * stride + stride / 2];
            Sh3<u64> diff;
#pragma unroll
            for (int p = 0; p < 3; p++)
#pragma unroll
                for (int s = 0; s < 2; s++) diff.s[p][s] = ev.s[p][s] - od.s[p][s];
            b[j] = drelu_m<U>(diff, g, win, PRG_RELU_WORDS * m);
            e[j * stride] = select_m<U>(od, ev, b[j], g, win, PRG_RELU_WORDS * m + PRG_DRELU_WORDS);
        }
#pragma unroll
        for (int j = 0; j < cnt; j++) {
            Sh3<U> neg = b[j];
            neg.s[PARTY_A][0] ^= (U)1;
            neg.s[PARTY_C][1] ^= (U)1;
#pragma unroll
            for (int tt = 0; tt < stride; tt++) {
                const int pos = j * stride + tt;
                rss_and_m<U>(d[pos], tt < stride / 2 ? b[j] : neg, g, win,
                             PRG_RELU_WORDS * (K - 1) + 2 * (t * K + pos), (U)1);
            }
        }
    }
}

template <typename U, int K, bool MASKED>
__device__ __forceinline__ void tournament_any(Sh3<u64> (&e)[K], Sh3<U> (&d)[K], const PrgStream &g, u64 win) {
    if (MASKED) tournament_m<U, K>(e, d, g, win);
    else tournament<U, K>(e, d);
}

// Free-function form over an already gathered pools buffer [windows][K]; the
// even positions are written back, as the reference's strided views do.
template <typename U, int K, bool MASKED>
__global__ __launch_bounds__(EW_BLOCK) void maxpool_pools_kernel(u64 *__restrict__ pools, size_t windows,
                                                                 u64 *__restrict__ result, U *__restrict__ dres,
                                                                 PrgStream g) {
    size_t w = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (w >= windows) return;
    const size_t n = windows * K;
    Sh3<u64> e[K];
    Sh3<U> d[K];
#pragma unroll
    for (int j = 0; j < K; j++) e[j] = load3(pools, n, w * K + j);
    tournament_any<U, K, MASKED>(e, d, g, w);
    store3(result, windows, w, e[0]);
#pragma unroll
    for (int j = 0; j < K; j += 2) store3(pools, n, w * K + j, e[j]);
#pragma unroll
    for (int j = 0; j < K; j++) store3(dres, n, w * K + j, d[j]);
}

template <typename U, bool MASKED>
static int launch_maxpool_pools(u64 *pools, size_t windows, u64 *result, U *dres, int k, const PrgStream &g,
                                hipStream_t s) {
    const dim3 gr(ew_grid(windows)), bl(EW_BLOCK);
    switch (k) {
        case 2: hipLaunchKernelGGL((maxpool_pools_kernel<U, 2, MASKED>), gr, bl, 0, s, pools, windows, result, dres, g); break;
        case 4: hipLaunchKernelGGL((maxpool_pools_kernel<U, 4, MASKED>), gr, bl, 0, s, pools, windows, result, dres, g); break;
        case 8: hipLaunchKernelGGL((maxpool_pools_kernel<U, 8, MASKED>), gr, bl, 0, s, pools, windows, result, dres, g); break;
        case 16: hipLaunchKernelGGL((maxpool_pools_kernel<U, 16, MASKED>), gr, bl, 0, s, pools, windows, result, dres, g); break;
        default: set_error("maxpool: k must be 2, 4, 8 or 16"); return ERR_UNSUPPORTED;
    }
    PIR_LAUNCH_CHECK();
    return OK;
}

template <typename U>
int fused_maxpool_pools(u64 *pools, size_t n, u64 *result, U *dres, int k, hipStream_t s, const PrgStream &g) {
    PIR_REQUIRE(k > 0 && n % k == 0, "maxpool: pools size must be a multiple of k");
    size_t windows = n / k;
    if (windows == 0) return OK;
    return g.on ? launch_maxpool_pools<U, true>(pools, windows, result, dres, k, g, s)
                : launch_maxpool_pools<U, false>(pools, windows, result, dres, k, g, s);
}

// Layer form (MaxpoolLayer.cu:52-131): gathers each f x f window straight
// from the NHWC activation (maxpool_im2row order: row offset outer, column
// offset inner), pads to K with pad_value in every share (stride_pad; the RSS
// layer uses the same pad value for both shares of every party), runs the
// tournament and keeps the first f*f derivative entries (truncate_cols).
// MASKED: the pool's tournament draws from gp (element = window = output
// index), the fused ReLU from gr (its own call, element = output index).
// F: the window side when fixed at compile time (0: g.f at run time); with it
// the padded positions are known, so the compiler folds the compares of two
// pad values (their difference is zero in every share) and drops the
// derivative entries that truncate_cols discards: a 3x3 window padded to 16
// runs 11 of the tournament's 15 compares.
template <typename U, int K, bool MASKED, int F = 0>
__global__ __launch_bounds__(EW_BLOCK) void maxpool_layer_kernel(const u64 *__restrict__ x, MaxpoolGeom g,
                                                                 u64 *__restrict__ act, U *__restrict__ mprime,
                                                                 u64 *__restrict__ relu_out,
                                                                 u8 *__restrict__ relu_d, PrgStream gp,
                                                                 PrgStream gr) {
    const size_t outN = (size_t)g.batch * g.OH * g.OW * g.C;
    size_t o = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (o >= outN) return;
    const size_t inN = (size_t)g.batch * g.H * g.W * g.C;
    int c = (int)(o % g.C);
    size_t r = o / g.C;
    int ow = (int)(r % g.OW);
    r /= g.OW;
    int oh = (int)(r % g.OH);
    int b = (int)(r / g.OH);
    Sh3<u64> e[K];
    Sh3<U> d[K];
    const int f = F ? F : g.f, f2 = f * f;
#pragma unroll
    for (int j = 0; j < K; j++) {
        if (j < f2) {
            int dh = j / f, dw = j - dh * f;
            int ih = oh * g.stride + dh, iw = ow * g.stride + dw;
            e[j] = load3(x, inN, (((size_t)b * g.H + ih) * g.W + iw) * g.C + c);
        } else {
#pragma unroll
            for (int p = 0; p < 3; p++) e[j].s[p][0] = e[j].s[p][1] = g.pad_value;
        }
    }
    tournament_any<U, K, MASKED>(e, d, gp, o);
    store3(act, outN, o, e[0]);
    if (relu_out) {  // the following ReLU layer on the same element
        Sh3<u8> rb;
        Sh3<u64> ro;
        if (MASKED) {
            ro = relu_m<u8>(e[0], rb, gr, o);
        } else {
            rb = drelu_<u8>(e[0]);
            Sh3<u64> zero;
#pragma unroll
            for (int p = 0; p < 3; p++) zero.s[p][0] = zero.s[p][1] = 0;
            ro = select_<u8>(zero, e[0], rb);
        }
        store3(relu_d, outN, o, rb);
        store3(relu_out, outN, o, ro);
    }
    const size_t mpN = outN * f2;
    if (sizeof(U) == 1 && K == 4 && f2 == 4) {
        // the window's four derivative bytes of each share array in one 4-byte store
#pragma unroll
        for (int p = 0; p < 3; p++)
#pragma unroll
            for (int sh = 0; sh < 2; sh++) {
                u32 v = 0;
#pragma unroll
                for (int j = 0; j < 4; j++) v |= (u32)(u8)d[j].s[p][sh] << (8 * j);
                *(u32 *)(mprime + (size_t)(2 * p + sh) * mpN + o * 4) = v;
            }
        return;
    }
#pragma unroll
    for (int j = 0; j < K; j++)
        if (j < f2) store3(mprime, mpN, o * f2 + j, d[j]);
}

// 2x2 pools with an even channel count: two adjacent channels per thread, so
// every gathered window value of both channels is one 16-byte load per share
// array, the pooled value and the fused ReLU's output one 16-byte store, the
// ReLU's bit shares one 2-byte store and the eight derivative bytes of the two
// windows one 8-byte store per share array.
template <typename U, bool MASKED>
__global__ __launch_bounds__(EW_BLOCK) void maxpool2x2_pair_kernel(const u64 *__restrict__ x, MaxpoolGeom g,
                                                                   u64 *__restrict__ act, U *__restrict__ mprime,
                                                                   u64 *__restrict__ relu_out,
                                                                   u8 *__restrict__ relu_d, PrgStream gp,
                                                                   PrgStream gr) {
    static_assert(sizeof(U) == 1, "byte derivative shares");
    const size_t outN = (size_t)g.batch * g.OH * g.OW * g.C;
    size_t o = ((size_t)blockIdx.x * EW_BLOCK + threadIdx.x) * 2;
    if (o >= outN) return;
    const size_t inN = (size_t)g.batch * g.H * g.W * g.C;
    int c = (int)(o % g.C);
    size_t r = o / g.C;
    int ow = (int)(r % g.OW);
    r /= g.OW;
    int oh = (int)(r % g.OH);
    int b = (int)(r / g.OH);
    Sh3<u64> e0[4], e1[4];
    Sh3<U> d0[4], d1[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int ih = oh * g.stride + (j >> 1), iw = ow * g.stride + (j & 1);
        const size_t idx = (((size_t)b * g.H + ih) * g.W + iw) * g.C + c;
#pragma unroll
        for (int p = 0; p < 3; p++)
#pragma unroll
            for (int sh = 0; sh < 2; sh++) {
                u64x2 t = *(const u64x2 *)(x + (size_t)(2 * p + sh) * inN + idx);
                e0[j].s[p][sh] = t.x;
                e1[j].s[p][sh] = t.y;
            }
    }
    tournament_any<U, 4, MASKED>(e0, d0, gp, o);
    tournament_any<U, 4, MASKED>(e1, d1, gp, o + 1);
    Sh3<u64> pooled[2] = {e0[0], e1[0]};
    store3v<2>(act, outN, o, pooled);
    if (relu_out) {
        Sh3<u8> rb[2];
        Sh3<u64> zero, ro[2];
#pragma unroll
        for (int p = 0; p < 3; p++) zero.s[p][0] = zero.s[p][1] = 0;
        if (MASKED) {
            ro[0] = relu_m<u8>(e0[0], rb[0], gr, o);
            ro[1] = relu_m<u8>(e1[0], rb[1], gr, o + 1);
        } else {
            rb[0] = drelu_<u8>(e0[0]);
            rb[1] = drelu_<u8>(e1[0]);
            ro[0] = select_<u8>(zero, e0[0], rb[0]);
            ro[1] = select_<u8>(zero, e1[0], rb[1]);
        }
        store3v<2>(relu_d, outN, o, rb);
        store3v<2>(relu_out, outN, o, ro);
    }
    const size_t mpN = outN * 4;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int sh = 0; sh < 2; sh++) {
            u64 v = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                v |= (u64)(u8)d0[j].s[p][sh] << (8 * j);
                v |= (u64)(u8)d1[j].s[p][sh] << (8 * (j + 4));
            }
            *(u64 *)(mprime + (size_t)(2 * p + sh) * mpN + o * 4) = v;
        }
}

template <typename U, bool MASKED>
static int launch_maxpool_layer(const u64 *x, const MaxpoolGeom &g, u64 *act, U *mprime, hipStream_t s, u64 *relu_out,
                                u8 *relu_d, const PrgStream &gp, const PrgStream &gr) {
    size_t outN = (size_t)g.batch * g.OH * g.OW * g.C;
    int k = expanded_pool_size(g.f);
    if constexpr (sizeof(U) == 1) {
        if (g.f == 2 && g.C % 2 == 0 && aligned_to(x, 16) && aligned_to(act, 16) && aligned_to(relu_out, 16) &&
            aligned_to(relu_d, 2) && aligned_to(mprime, 8)) {
            hipLaunchKernelGGL((maxpool2x2_pair_kernel<U, MASKED>), dim3(ew_grid(outN / 2)), dim3(EW_BLOCK), 0, s, x,
                               g, act, mprime, relu_out, relu_d, gp, gr);
            PIR_LAUNCH_CHECK();
            return OK;
        }
    }
    const dim3 grid(ew_grid(outN));
    if (k == 4)
        hipLaunchKernelGGL((maxpool_layer_kernel<U, 4, MASKED, 2>), grid, dim3(EW_BLOCK), 0, s, x, g, act, mprime,
                           relu_out, relu_d, gp, gr);
    else if (k == 16 && g.f == 3)
        hipLaunchKernelGGL((maxpool_layer_kernel<U, 16, MASKED, 3>), grid, dim3(EW_BLOCK), 0, s, x, g, act, mprime,
                           relu_out, relu_d, gp, gr);
    else if (k == 16)
        hipLaunchKernelGGL((maxpool_layer_kernel<U, 16, MASKED>), grid, dim3(EW_BLOCK), 0, s, x, g, act, mprime,
                           relu_out, relu_d, gp, gr);
    else {
        set_error("maxpool layer: only 2x2, 3x3 and 4x4 pools are supported");
        return ERR_UNSUPPORTED;
    }
    PIR_LAUNCH_CHECK();
    return OK;
}

template <typename U>
int fused_maxpool_layer(const u64 *x, const MaxpoolGeom &g, u64 *act, U *mprime, hipStream_t s, u64 *relu_out,
                        u8 *relu_d, const PrgStream &gp, const PrgStream &gr) {
    PIR_REQUIRE(!relu_out == !relu_d, "maxpool layer: a fused ReLU needs both outputs");
    PIR_REQUIRE(g.f >= 2 && g.stride >= 1 && g.H >= g.f && g.W >= g.f, "maxpool layer: invalid geometry");
    PIR_REQUIRE(g.OH == (g.H - g.f) / g.stride + 1 && g.OW == (g.W - g.f) / g.stride + 1,
                "maxpool layer: inconsistent output size");
    size_t outN = (size_t)g.batch * g.OH * g.OW * g.C;
    if (outN == 0) return OK;
    PIR_REQUIRE(gp.on == gr.on || !relu_out, "maxpool layer: pool and ReLU streams must agree on the mode");
    return gp.on ? launch_maxpool_layer<U, true>(x, g, act, mprime, s, relu_out, relu_d, gp, gr)
                 : launch_maxpool_layer<U, false>(x, g, act, mprime, s, relu_out, relu_d, gp, gr);
}

// AveragepoolLayer::forward (AveragepoolLayer.cu:49-96): averagepool_im2row,
// reduceSum over each f x f window in every share and dividePublic by f*f, one
// thread per output element; the following ReLU layer is applied to the same
// element when relu_out is set.
__global__ __launch_bounds__(EW_BLOCK) void avgpool_layer_kernel(const u64 *__restrict__ x, MaxpoolGeom g,
                                                                 u64 *__restrict__ act, u64 *__restrict__ relu_out,
                                                                 u8 *__restrict__ relu_d) {
    const size_t outN = (size_t)g.batch * g.OH * g.OW * g.C;
    size_t o = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (o >= outN) return;
    const size_t inN = (size_t)g.batch * g.H * g.W * g.C;
    int c = (int)(o % g.C);
    size_t r = o / g.C;
    int ow = (int)(r % g.OW);
    r /= g.OW;
    int oh = (int)(r % g.OH);
    int b = (int)(r / g.OH);
    Sh3<u64> sum;
#pragma unroll
    for (int p = 0; p < 3; p++) sum.s[p][0] = sum.s[p][1] = 0;
    for (int dh = 0; dh < g.f; dh++)
        for (int dw = 0; dw < g.f; dw++) {
            Sh3<u64> e = load3(x, inN, (((size_t)b * g.H + oh * g.stride + dh) * g.W + ow * g.stride + dw) * g.C + c);
#pragma unroll
            for (int p = 0; p < 3; p++) {
                sum.s[p][0] += e.s[p][0];
                sum.s[p][1] += e.s[p][1];
            }
        }
    Sh3<u64> v = divide_public_(sum, (u64)g.f * (u64)g.f);
    store3(act, outN, o, v);
    if (relu_out) {
        Sh3<u8> rb = drelu_<u8>(v);
        store3(relu_d, outN, o, rb);
        Sh3<u64> zero;
#pragma unroll
        for (int p = 0; p < 3; p++) zero.s[p][0] = zero.s[p][1] = 0;
        store3(relu_out, outN, o, select_<u8>(zero, v, rb));
    }
}

int fused_avgpool_layer(const u64 *x, const MaxpoolGeom &g, u64 *act, hipStream_t s, u64 *relu_out, u8 *relu_d) {
    PIR_REQUIRE(!relu_out == !relu_d, "averagepool layer: a fused ReLU needs both outputs");
    PIR_REQUIRE(g.f >= 1 && g.stride >= 1 && g.H >= g.f && g.W >= g.f, "averagepool layer: invalid geometry");
    PIR_REQUIRE(g.OH == (g.H - g.f) / g.stride + 1 && g.OW == (g.W - g.f) / g.stride + 1,
                "averagepool layer: inconsistent output size");
    size_t outN = (size_t)g.batch * g.OH * g.OW * g.C;
    if (outN == 0) return OK;
    hipLaunchKernelGGL(avgpool_layer_kernel, dim3(ew_grid(outN)), dim3(EW_BLOCK), 0, s, x, g, act, relu_out, relu_d);
    PIR_LAUNCH_CHECK();
    return OK;
}

// ---------------------------------------------------------------------------
// Backward (training) path

template <typename U>
__global__ __launch_bounds__(EW_BLOCK) void relu_backward_kernel(const u64 *__restrict__ delta,
                                                                 const U *__restrict__ rp, size_t n,
                                                                 u64 *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    Sh3<u64> zero;
#pragma unroll
    for (int p = 0; p < 3; p++) zero.s[p][0] = zero.s[p][1] = 0;
    store3(out, n, i, select_<U>(zero, load3(delta, n, i), load3(rp, n, i)));
}

template <typename U>
int fused_relu_backward(const u64 *delta, const U *relu_prime, size_t n, u64 *out, hipStream_t s) {
    if (n == 0) return OK;
    hipLaunchKernelGGL((relu_backward_kernel<U>), dim3(ew_grid(n)), dim3(EW_BLOCK), 0, s, delta, relu_prime, n, out);
    PIR_LAUNCH_CHECK();
    return OK;
}

// One thread per input element (b, h, w, c): every pooling window (oh, ow)
// that covers it contributes selectShare(0, delta[window], maxPrime[window][tap])
// with tap = (h - oh*stride)*f + (w - ow*stride) -- the entry vectorExpand +
// selectShare produce for that position, which maxpool_row2im then adds there.
template <typename U>
__global__ __launch_bounds__(EW_BLOCK) void maxpool_backward_kernel(const u64 *__restrict__ delta,
                                                                    const U *__restrict__ mp, MaxpoolGeom g,
                                                                    u64 *__restrict__ out) {
    const size_t inN = (size_t)g.batch * g.H * g.W * g.C;
    size_t t = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (t >= inN) return;
    int c = (int)(t % g.C);
    size_t r = t / g.C;
    int w = (int)(r % g.W);
    r /= g.W;
    int h = (int)(r % g.H);
    int b = (int)(r / g.H);
    const size_t outN = (size_t)g.batch * g.OH * g.OW * g.C;
    const int f2 = g.f * g.f;
    int oh_lo = h - g.f + 1 > 0 ? (h - g.f + 1 + g.stride - 1) / g.stride : 0;
    int oh_hi = min(g.OH - 1, h / g.stride);
    int ow_lo = w - g.f + 1 > 0 ? (w - g.f + 1 + g.stride - 1) / g.stride : 0;
    int ow_hi = min(g.OW - 1, w / g.stride);
    Sh3<u64> acc, zero;
#pragma unroll
    for (int p = 0; p < 3; p++) acc.s[p][0] = acc.s[p][1] = zero.s[p][0] = zero.s[p][1] = 0;
    for (int oh = oh_lo; oh <= oh_hi; oh++)
        for (int ow = ow_lo; ow <= ow_hi; ow++) {
            size_t o = (((size_t)b * g.OH + oh) * g.OW + ow) * g.C + c;
            int j = (h - oh * g.stride) * g.f + (w - ow * g.stride);
            Sh3<u64> v = select_<U>(zero, load3(delta, outN, o), load3(mp, outN * f2, o * f2 + j));
#pragma unroll
            for (int p = 0; p < 3; p++) {
                acc.s[p][0] += v.s[p][0];
                acc.s[p][1] += v.s[p][1];
            }
        }
    store3(out, inN, t, acc);
}

template <typename U>
int fused_maxpool_backward(const u64 *delta, const U *max_prime, const MaxpoolGeom &g, u64 *out, hipStream_t s) {
    PIR_REQUIRE(g.f >= 1 && g.stride >= 1 && g.H >= g.f && g.W >= g.f, "maxpool backward: invalid geometry");
    PIR_REQUIRE(g.OH == (g.H - g.f) / g.stride + 1 && g.OW == (g.W - g.f) / g.stride + 1,
                "maxpool backward: inconsistent output size");
    size_t inN = (size_t)g.batch * g.H * g.W * g.C;
    if (inN == 0) return OK;
    hipLaunchKernelGGL((maxpool_backward_kernel<U>), dim3(ew_grid(inN)), dim3(EW_BLOCK), 0, s, delta, max_prime, g,
                       out);
    PIR_LAUNCH_CHECK();
    return OK;
}

// FCLayer.cu:205-218 for one bias element o: every share array sums delta
// entries o*batch .. o*batch + batch-1 (reduceSum with reduceRows = false reads
// delta as outDim rows of batch), then dividePublic and the subtraction.
__global__ __launch_bounds__(EW_BLOCK) void fc_bias_update_kernel(const u64 *__restrict__ delta,
                                                                  u64 *__restrict__ bias, int out_dim, int batch,
                                                                  int lr) {
    int o = blockIdx.x * EW_BLOCK + threadIdx.x;
    if (o >= out_dim) return;
    const size_t n = (size_t)out_dim * batch;
    Sh3<u64> db;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int sh = 0; sh < 2; sh++) {
            const u64 *a = delta + (size_t)(2 * p + sh) * n + (size_t)o * batch;
            u64 acc = 0;
            for (int i = 0; i < batch; i++) acc += a[i];
            db.s[p][sh] = acc;
        }
    db = divide_public_pow2_(db, lr);
    Sh3<u64> bv = load3(bias, (size_t)out_dim, (size_t)o);
#pragma unroll
    for (int p = 0; p < 3; p++) {
        bv.s[p][0] -= db.s[p][0];
        bv.s[p][1] -= db.s[p][1];
    }
    store3(bias, (size_t)out_dim, (size_t)o, bv);
}

int fused_fc_bias_update(const u64 *delta, u64 *bias, int out_dim, int batch, int lr, hipStream_t s) {
    PIR_REQUIRE(out_dim >= 0 && batch >= 0 && lr >= 1 && lr <= 62, "FC bias update: invalid arguments");
    if (out_dim == 0) return OK;
    hipLaunchKernelGGL(fc_bias_update_kernel, dim3(ew_grid(out_dim)), dim3(EW_BLOCK), 0, s, delta, bias, out_dim, batch,
                       lr);
    PIR_LAUNCH_CHECK();
    return OK;
}

// Parties A and C open the logits: the only opened values that reach an
// output share (deltas += softmax touches A.sA and C.sB).  out[0] = A's view
// sA + sB + (C's sA), out[1] = C's view sA + sB + (B's sA).
__global__ __launch_bounds__(EW_BLOCK) void open_logits_ac_kernel(const u64 *__restrict__ act, size_t n,
                                                                  u64 *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    const u64 A0 = act[0 * n + i], A1 = act[1 * n + i], B0 = act[2 * n + i];
    const u64 C0 = act[4 * n + i], C1 = act[5 * n + i];
    out[i] = A0 + A1 + C0;
    out[n + i] = C0 + C1 + B0;
}

int fused_open_logits_ac(const u64 *act, size_t n, u64 *out, hipStream_t s) {
    if (n == 0) return OK;
    hipLaunchKernelGGL(open_logits_ac_kernel, dim3(ew_grid(n)), dim3(EW_BLOCK), 0, s, act, n, out);
    PIR_LAUNCH_CHECK();
    return OK;
}

// deltas = (A.sA = sm[0], C.sB = sm[1]) - labels, then dividePublic by 2^lmb
// when lmb > 1 (NeuralNetwork.cu:259-264).
__global__ __launch_bounds__(EW_BLOCK) void softmax_delta_kernel(const u64 *__restrict__ sm,
                                                                 const u64 *__restrict__ labels, size_t n, int lmb,
                                                                 u64 *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    Sh3<u64> d = load3(labels, n, i);
#pragma unroll
    for (int p = 0; p < 3; p++) {
        d.s[p][0] = (u64)0 - d.s[p][0];
        d.s[p][1] = (u64)0 - d.s[p][1];
    }
    d.s[PARTY_A][0] += sm[i];
    d.s[PARTY_C][1] += sm[n + i];
    if (lmb > 1) d = divide_public_pow2_(d, lmb);
    store3(out, n, i, d);
}

int fused_softmax_delta(const u64 *sm, const u64 *labels, size_t n, int lmb, u64 *out, hipStream_t s) {
    PIR_REQUIRE(lmb >= 0 && lmb <= 62, "softmax delta: invalid arguments");
    if (n == 0) return OK;
    hipLaunchKernelGGL(softmax_delta_kernel, dim3(ew_grid(n)), dim3(EW_BLOCK), 0, s, sm, labels, n, lmb, out);
    PIR_LAUNCH_CHECK();
    return OK;
}

// AveragepoolLayer::backward (AveragepoolLayer.cu:98-125): the expanded delta
// (averagepool_expand_delta, convolution.cuh:194-207: deltas entry
// b*C*P + c + j*C takes delta[b*C + c] for j < P, b over image-window pairs;
// entries past n_in * P stay zero), then dividePublic by P over all of deltas.
// One thread per deltas entry.
__global__ __launch_bounds__(EW_BLOCK) void avgpool_backward_kernel(const u64 *__restrict__ delta, size_t n_in, int C,
                                                                    int P, size_t n_out, u64 *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n_out) return;
    Sh3<u64> v;
    if (i < n_in * (size_t)P) {
        size_t b = i / ((size_t)C * P);
        v = load3(delta, n_in, b * C + (size_t)(i % (size_t)C));
    } else {
#pragma unroll
        for (int p = 0; p < 3; p++) v.s[p][0] = v.s[p][1] = 0;
    }
    store3(out, n_out, i, divide_public_(v, (u64)P));
}

int fused_avgpool_backward(const u64 *delta, size_t n_in, int C, int P, u64 *out, size_t n_out, hipStream_t s) {
    PIR_REQUIRE(C > 0 && P > 0 && n_in % (size_t)C == 0, "averagepool backward: invalid arguments");
    PIR_REQUIRE(n_in * (size_t)P <= n_out, "averagepool backward: the expanded delta does not fit the input size");
    if (n_out == 0) return OK;
    hipLaunchKernelGGL(avgpool_backward_kernel, dim3(ew_grid(n_out)), dim3(EW_BLOCK), 0, s, delta, n_in, C, P, n_out,
                       out);
    PIR_LAUNCH_CHECK();
    return OK;
}

template int fused_relu_backward<u8>(const u64 *, const u8 *, size_t, u64 *, hipStream_t);
template int fused_maxpool_backward<u8>(const u64 *, const u8 *, const MaxpoolGeom &, u64 *, hipStream_t);

int expanded_pool_size(int f) {
    // pow(ceil(log2(f*f)), 2) (convolution.cuh:288, MaxpoolLayer.cu:69)
    int f2 = f * f, c = 0;
    while ((1 << c) < f2) c++;
    return c * c;
}

// ---------------------------------------------------------------------------
// Share algebra with communication.

template <typename T>
__global__ __launch_bounds__(EW_BLOCK) void reconstruct_kernel(const T *__restrict__ x, size_t n, T *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    Sh3<T> v = load3(x, n, i);
#pragma unroll
    for (int p = 0; p < 3; p++) out[(size_t)p * n + i] = (T)(v.s[p][0] + v.s[p][1] + v.s[prev_party(p)][0]);
}

template <typename T>
int fused_reconstruct(const T *x, size_t n, T *out, hipStream_t s) {
    if (n == 0) return OK;
    hipLaunchKernelGGL((reconstruct_kernel<T>), dim3(ew_grid(n)), dim3(EW_BLOCK), 0, s, x, n, out);
    PIR_LAUNCH_CHECK();
    return OK;
}

template <typename T>
__global__ __launch_bounds__(EW_BLOCK) void reshare_kernel(const T *__restrict__ c, size_t n, T *__restrict__ out,
                                                           PrgStream g) {
    size_t i = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    T z[3];
#pragma unroll
    for (int p = 0; p < 3; p++) z[p] = c[(size_t)p * n + i];
    if (g.on) {  // masked: the arithmetic zero sharing of words 0, 1
        store3(out, n, i, reshare_m<T, false>(z, g, i, PRG_RESHARE, (T)~(T)0));
        return;
    }
#pragma unroll
    for (int p = 0; p < 3; p++) {
        out[(size_t)(2 * p) * n + i] = z[p];
        out[(size_t)(2 * p + 1) * n + i] = z[next_party(p)];
    }
}

int fused_reshare(const u64 *c, size_t n, u64 *out, hipStream_t s, const PrgStream &g) {
    if (n == 0) return OK;
    hipLaunchKernelGGL((reshare_kernel<u64>), dim3(ew_grid(n)), dim3(EW_BLOCK), 0, s, c, n, out, g);
    PIR_LAUNCH_CHECK();
    return OK;
}

template <typename T, bool AND>
__global__ __launch_bounds__(EW_BLOCK) void rss_mul_kernel(T *__restrict__ x, const T *__restrict__ y, size_t n,
                                                           PrgStream g) {
    size_t i = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    Sh3<T> a = load3(x, n, i), b = load3(y, n, i);
    if (g.on) {  // masked: full-width zero sharing of words 0, 1
        if (AND) rss_and_m<T>(a, b, g, i, PRG_RESHARE, (T)~(T)0);
        else rss_mult_m<T>(a, b, g, i, PRG_RESHARE);
    } else {
        if (AND) rss_and_elem<T>(a, b);
        else rss_mult_<T>(a, b);
    }
    store3(x, n, i, a);
}

int fused_rss_mult(u64 *x, const u64 *y, size_t n, hipStream_t s, const PrgStream &g) {
    if (n == 0) return OK;
    hipLaunchKernelGGL((rss_mul_kernel<u64, false>), dim3(ew_grid(n)), dim3(EW_BLOCK), 0, s, x, y, n, g);
    PIR_LAUNCH_CHECK();
    return OK;
}

template <typename T>
int fused_rss_and(T *x, const T *y, size_t n, hipStream_t s, const PrgStream &g) {
    if (n == 0) return OK;
    hipLaunchKernelGGL((rss_mul_kernel<T, true>), dim3(ew_grid(n)), dim3(EW_BLOCK), 0, s, x, y, n, g);
    PIR_LAUNCH_CHECK();
    return OK;
}

// dividePublic (RSS.inl:422-455 with the placeholder r = 1, r' = d of
// Precompute.h:107-134).  Per party p: a -= r'; t_p = reconstruct(a)_p;
// a = r + (A.sA, C.sB += t_p / d) with the reference's division
// (S)((double)(S)t / (double)(S)d).  `dv` (per-party local denominators [3][n])
// overrides the scalar `d` when set.
__global__ __launch_bounds__(EW_BLOCK) void divide_public_kernel(u64 *__restrict__ a, size_t n, u64 d,
                                                                 const u64 *__restrict__ dv, PrgStream g, int k) {
    size_t i = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    Sh3<u64> v = load3(a, n, i);
    if (g.on) {  // masked: d = 2^k, the random truncation pair of element i
        store3(a, n, i, divide_public_pow2_m(v, k, g, i));
        return;
    }
    u64 dd[3];
#pragma unroll
    for (int p = 0; p < 3; p++) dd[p] = dv ? dv[(size_t)p * n + i] : d;
    // rPrime: scalar -> fill(d) (A.sA = C.sB = d); vector -> 0 + d on A.sA and C.sB,
    // the same shares either way
    store3(a, n, i, divide_public_local_(v, dd));
}

int fused_divide_public(u64 *a, size_t n, u64 d, const u64 *dv, hipStream_t s, const PrgStream &g) {
    int k = 0;
    if (g.on) {
        PIR_REQUIRE(!dv && d != 0 && (d & (d - 1)) == 0,
                    "dividePublic: the masked precompute supports one public power-of-two denominator");
        while (((u64)1 << k) != d) k++;
        PIR_REQUIRE(k >= 1 && k <= 62, "dividePublic: the masked precompute needs d = 2^k, 1 <= k <= 62");
    }
    if (n == 0) return OK;
    hipLaunchKernelGGL(divide_public_kernel, dim3(ew_grid(n)), dim3(EW_BLOCK), 0, s, a, n, d, dv, g, k);
    PIR_LAUNCH_CHECK();
    return OK;
}

// dividePublic3out3 over 3-of-3 shares a [3][n] -> out [3][2][n]
__global__ __launch_bounds__(EW_BLOCK) void divide_3out3_kernel(const u64 *__restrict__ a, size_t n, int k,
                                                                u64 *__restrict__ out, PrgStream g) {
    size_t i = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    if (g.on) {  // masked: the random truncation pair of element i
        Sh3<u64> o;
        divide_3out3_m(a[i] + a[n + i] + a[2 * n + i], k, g, i, o.s);
        store3(out, n, i, o);
        return;
    }
    u64 t = a[i] + a[n + i] + a[2 * n + i] - ((u64)1 << k);
    u64 base = 1 + trunc_div_pow2(t, k);
    Sh3<u64> o;
#pragma unroll
    for (int p = 0; p < 3; p++) o.s[p][0] = o.s[p][1] = 0;
    o.s[PARTY_A][0] = base;
    o.s[PARTY_C][1] = base;
    store3(out, n, i, o);
}

int fused_divide_3out3(const u64 *a, size_t n, int k, u64 *out, hipStream_t s, const PrgStream &g) {
    PIR_REQUIRE(k >= 0 && k <= 62, "dividePublic3out3: denominator must be 2^k, 0 <= k <= 62");
    PIR_REQUIRE(!g.on || k >= 1, "dividePublic3out3: the masked precompute needs k >= 1");
    if (n == 0) return OK;
    hipLaunchKernelGGL(divide_3out3_kernel, dim3(ew_grid(n)), dim3(EW_BLOCK), 0, s, a, n, k, out, g);
    PIR_LAUNCH_CHECK();
    return OK;
}

// getPowers (RSS.inl:1211-1235): the 64 compare-and-select rounds on one
// element's shares in registers, then each party's opened power -> pow [3][n].
__global__ __launch_bounds__(EW_BLOCK) void get_powers_kernel(const u64 *__restrict__ in, size_t n,
                                                              u64 *__restrict__ pow) {
    size_t i = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    Sh3<u64> pw = get_powers_(load3(in, n, i));
#pragma unroll
    for (int p = 0; p < 3; p++) pow[(size_t)p * n + i] = pw.s[p][0] + pw.s[p][1] + pw.s[prev_party(p)][0];
}

int fused_get_powers(const u64 *in, size_t n, u64 *pow, hipStream_t s) {
    if (n == 0) return OK;
    hipLaunchKernelGGL(get_powers_kernel, dim3(ew_grid(n)), dim3(EW_BLOCK), 0, s, in, n, pow);
    PIR_LAUNCH_CHECK();
    return OK;
}

// taylorSeries (RSS.inl:1237-1320) for one element, every party with its own
// opened power pow_p = getPowers_p + 1 and ones_p = 1 << pow_p:
//   out = [(in * in / ones) * c2 / ones] + [in * c1 / ones] + c0
//   out = out * f / ones / (1 << max(pow - FP, 0)) << max(FP - pow, 0)
// where each division is dividePublic by the party's public denominator.
__global__ __launch_bounds__(EW_BLOCK) void taylor_kernel(const u64 *__restrict__ in, size_t n,
                                                          u64 *__restrict__ out, const TaylorTables t) {
    size_t i = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    const Sh3<u64> x = load3(in, n, i);
    const Sh3<u64> pw = get_powers_(x);
    u64 pow[3], ones[3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        pow[p] = pw.s[p][0] + pw.s[p][1] + pw.s[prev_party(p)][0] + 1;
        ones[p] = shl_clamped(1, pow[p]);
    }
    Sh3<u64> o;
#pragma unroll
    for (int p = 0; p < 3; p++) o.s[p][0] = o.s[p][1] = 0;
    // scratch *= c (RSS *= DeviceData: both shares); out += scratch (RSS += RSS)
    auto scale_add = [&](Sh3<u64> sc, const TaylorTab &c, bool square) {
        if (square) {
            rss_mult_(sc, x);
            sc = divide_public_local_(sc, ones);
        }
#pragma unroll
        for (int p = 0; p < 3; p++) {
            const u64 cp = c.at(pow[p]);
            sc.s[p][0] *= cp;
            sc.s[p][1] *= cp;
        }
        sc = divide_public_local_(sc, ones);
#pragma unroll
        for (int p = 0; p < 3; p++) {
            o.s[p][0] += sc.s[p][0];
            o.s[p][1] += sc.s[p][1];
        }
    };
    if (t.use_a2) scale_add(x, t.c2, true);
    if (t.use_a1) scale_add(x, t.c1, false);
    // out += a0Coeff (RSS += DeviceData: A.sA and C.sB)
    o.s[PARTY_A][0] += t.c0.at(pow[PARTY_A]);
    o.s[PARTY_C][1] += t.c0.at(pow[PARTY_C]);
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const u64 fp_ = t.f.at(pow[p]);
        o.s[p][0] *= fp_;
        o.s[p][1] *= fp_;
    }
    o = divide_public_local_(o, ones);
    // pow -= FP; the non-negative part divides, the negative part shifts left
    u64 pos_ones[3], neg[3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        const u64 e = pow[p] - (u64)t.fp;
        pos_ones[p] = shl_clamped(1, (s64)e >= 0 ? e : 0);
        neg[p] = (s64)e < 0 ? (u64)0 - e : 0;
    }
    o = divide_public_local_(o, pos_ones);
#pragma unroll
    for (int p = 0; p < 3; p++) {
        o.s[p][0] = shl_clamped(o.s[p][0], neg[p]);
        o.s[p][1] = shl_clamped(o.s[p][1], neg[p]);
    }
    store3(out, n, i, o);
}

int fused_taylor(const u64 *in, size_t n, u64 *out, const TaylorTables &t, hipStream_t s) {
    if (n == 0) return OK;
    hipLaunchKernelGGL(taylor_kernel, dim3(ew_grid(n)), dim3(EW_BLOCK), 0, s, in, n, out, t);
    PIR_LAUNCH_CHECK();
    return OK;
}

// convex_comb (RSS.inl:33-61, 1321-1333) per party: d = e ? (A: 1-c0, -c1;
// B: -c0, -c1; C: -c0, 1-c1) : c, with e the party's opened bit.
template <typename U>
__global__ __launch_bounds__(EW_BLOCK) void convex_comb_kernel(u64 *__restrict__ d, const u64 *__restrict__ c,
                                                               const U *__restrict__ e, size_t n) {
    size_t i = (size_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    Sh3<u64> cv = load3(c, n, i), dv;
#pragma unroll
    for (int p = 0; p < 3; p++) {
        if (e[(size_t)p * n + i] == (U)1) {
            dv.s[p][0] = (p == PARTY_A ? 1ull : 0ull) - cv.s[p][0];
            dv.s[p][1] = (p == PARTY_C ? 1ull : 0ull) - cv.s[p][1];
        } else {
            dv.s[p][0] = cv.s[p][0];
            dv.s[p][1] = cv.s[p][1];
        }
    }
    store3(d, n, i, dv);
}

template <typename U>
int fused_convex_comb(u64 *d, const u64 *c, const U *e, size_t n, hipStream_t s) {
    if (n == 0) return OK;
    hipLaunchKernelGGL((convex_comb_kernel<U>), dim3(ew_grid(n)), dim3(EW_BLOCK), 0, s, d, c, e, n);
    PIR_LAUNCH_CHECK();
    return OK;
}

// Explicit instantiations
template int fused_convex_comb<u8>(u64 *, const u64 *, const u8 *, size_t, hipStream_t);
template int fused_convex_comb<u64>(u64 *, const u64 *, const u64 *, size_t, hipStream_t);
template int fused_drelu<u8>(const u64 *, size_t, u8 *, hipStream_t, const PrgStream &);
template int fused_drelu<u64>(const u64 *, size_t, u64 *, hipStream_t, const PrgStream &);
template int fused_relu<u8>(const u64 *, size_t, u64 *, u8 *, hipStream_t, const PrgStream &);
template int fused_relu<u64>(const u64 *, size_t, u64 *, u64 *, hipStream_t, const PrgStream &);
template int fused_select<u8>(const u64 *, const u64 *, const u8 *, u64 *, size_t, hipStream_t, const PrgStream &);
template int fused_select<u64>(const u64 *, const u64 *, const u64 *, u64 *, size_t, hipStream_t, const PrgStream &);
template int fused_maxpool_pools<u8>(u64 *, size_t, u64 *, u8 *, int, hipStream_t, const PrgStream &);
template int fused_maxpool_pools<u64>(u64 *, size_t, u64 *, u64 *, int, hipStream_t, const PrgStream &);
template int fused_maxpool_layer<u8>(const u64 *, const MaxpoolGeom &, u64 *, u8 *, hipStream_t, u64 *, u8 *,
                                     const PrgStream &, const PrgStream &);
template int fused_reconstruct<u64>(const u64 *, size_t, u64 *, hipStream_t);
template int fused_reconstruct<u8>(const u8 *, size_t, u8 *, hipStream_t);
template int fused_rss_and<u8>(u8 *, const u8 *, size_t, hipStream_t, const PrgStream &);
template int fused_rss_and<u64>(u64 *, const u64 *, size_t, hipStream_t, const PrgStream &);

}  // namespace pir
