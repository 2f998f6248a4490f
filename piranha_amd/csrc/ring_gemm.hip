// ring_gemm.hip -- Z_2^64 ring GEMM and implicit-GEMM convolution for gfx950.
//
// Reference interface replaced: gpu::gemm (src/gpu/gemm.cuh:64-112, CUTLASS
// SIMT uint64) and gpu::conv_fprop (src/gpu/conv.cuh:359-401, CUTLASS implicit
// GEMM, NHWC x KRSC -> NHWC, cross-correlation).  Both are restated here as two
// exact CDNA4 forms (see ring_gemm.h): a 64-bit VALU kernel and the signed-digit
// int8-limb MFMA kernel.
#include "ring_gemm.h"

#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "rss_device.h"
#include "rss_masked.h"
#include "runtime.h"

namespace pir {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// Shared helpers

// Signed-digit recoding: the 8 bytes of the result, read as int8, are digits
// d_i in [-128, 127] with sum_i d_i 2^(8i) == v (mod 2^64).
__device__ __forceinline__ u64 recode_digits(u64 v) {
    const u64 k80 = 0x8080808080808080ull;
    return (v + k80) ^ k80;
}

__device__ __forceinline__ u32 perm(u32 s0, u32 s1, u32 sel) { return __builtin_amdgcn_perm(s0, s1, sel); }

// 4x4 byte transpose: rows a,b,c,d (bytes = columns) -> o[i] = (a_i, b_i, c_i, d_i).
__device__ __forceinline__ void transpose4(u32 a, u32 b, u32 c, u32 d, u32 &o0, u32 &o1, u32 &o2, u32 &o3) {
    u32 t0 = perm(b, a, 0x06020400u);  // a0 b0 a2 b2
    u32 t1 = perm(b, a, 0x07030501u);  // a1 b1 a3 b3
    u32 t2 = perm(d, c, 0x06020400u);  // c0 d0 c2 d2
    u32 t3 = perm(d, c, 0x07030501u);  // c1 d1 c3 d3
    o0 = perm(t2, t0, 0x05040100u);
    o2 = perm(t2, t0, 0x07060302u);
    o1 = perm(t3, t1, 0x05040100u);
    o3 = perm(t3, t1, 0x07060302u);
}

// 8x8 byte transpose: w[k] holds digit bytes (limb 0..7) of element k;
// out[i] holds limb i of elements k = 0..7 (byte k).
__device__ __forceinline__ void transpose8x8(const u64 w[8], u64 out[8]) {
    u32 lo[8], hi[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        lo[k] = (u32)w[k];
        hi[k] = (u32)(w[k] >> 32);
    }
    u32 a0, a1, a2, a3, b0, b1, b2, b3, c0, c1, c2, c3, d0, d1, d2, d3;
    transpose4(lo[0], lo[1], lo[2], lo[3], a0, a1, a2, a3);  // limbs 0-3, k 0-3
    transpose4(lo[4], lo[5], lo[6], lo[7], b0, b1, b2, b3);  // limbs 0-3, k 4-7
    transpose4(hi[0], hi[1], hi[2], hi[3], c0, c1, c2, c3);  // limbs 4-7, k 0-3
    transpose4(hi[4], hi[5], hi[6], hi[7], d0, d1, d2, d3);  // limbs 4-7, k 4-7
    out[0] = (u64)a0 | ((u64)b0 << 32);
    out[1] = (u64)a1 | ((u64)b1 << 32);
    out[2] = (u64)a2 | ((u64)b2 << 32);
    out[3] = (u64)a3 | ((u64)b3 << 32);
    out[4] = (u64)c0 | ((u64)d0 << 32);
    out[5] = (u64)c1 | ((u64)d1 << 32);
    out[6] = (u64)c2 | ((u64)d2 << 32);
    out[7] = (u64)c3 | ((u64)d3 << 32);
}

typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

// The placeholder truncation of one raw product (see Epilogue): 1 + q, then
// the optional second division.
__device__ __forceinline__ u64 trunc_base(const Epilogue &e, u64 raw) {
    const int k = e.trunc_bits;
    u64 base = (u64)1 + trunc_div_pow2(raw - ((u64)1 << k), k);
    if (e.div_bits) base = (u64)1 + trunc_div_pow2(base - ((u64)1 << e.div_bits), e.div_bits);
    return base;
}

// Truncation + bias epilogue for one output element (see Epilogue).
__device__ __forceinline__ void trunc_store(const Epilogue &e, size_t MN, int N, int m, int n, u64 raw) {
    const u64 base = trunc_base(e, raw);
    size_t idx = e.tc ? (size_t)m * N + n : (size_t)n * (MN / N) + m;
    if (e.sub_out) {
        for (int j = 0; j < e.nout; j++)
            if ((e.base_mask >> j) & 1u) e.out[(size_t)j * MN + idx] -= base;
        return;
    }
    if (e.relu_out) {
        // the following ReLU layer on the same element: dReLU + selectShare
        Sh3<u64> x;
#pragma unroll
        for (int j = 0; j < 6; j++) {
            u64 v = ((e.base_mask >> j) & 1u) ? base : (u64)0;
            if (e.bias) v += e.bias[(size_t)j * N + n];
            e.out[(size_t)j * MN + idx] = v;
            x.s[j >> 1][j & 1] = v;
        }
        Sh3<u8> b = drelu_<u8>(x);
        store3(e.relu_d, MN, idx, b);
        Sh3<u64> zero;
#pragma unroll
        for (int p = 0; p < 3; p++) zero.s[p][0] = zero.s[p][1] = 0;
        store3(e.relu_out, MN, idx, select_<u8>(zero, x, b));
        return;
    }
    if (e.bw_relu_out) {
        // the preceding ReLU layer's backward on the same element
        Sh3<u64> x, zero;
#pragma unroll
        for (int j = 0; j < 6; j++) {
            u64 v = ((e.base_mask >> j) & 1u) ? base : (u64)0;
            if (e.bias) v += e.bias[(size_t)j * N + n];
            e.out[(size_t)j * MN + idx] = v;
            x.s[j >> 1][j & 1] = v;
            zero.s[j >> 1][j & 1] = 0;
        }
        store3(e.bw_relu_out, MN, idx, select_<u8>(zero, x, load3(e.bw_relu_d, MN, idx)));
        return;
    }
    for (int j = 0; j < e.nout; j++) {
        u64 v = ((e.base_mask >> j) & 1u) ? base : (u64)0;
        if (e.bias) v += e.bias[(size_t)j * N + n];
        e.out[(size_t)j * MN + idx] = v;
    }
}

// The masked form of trunc_store for the forward epilogue (no division,
// subtraction or backward ReLU): the output shares are the random truncation
// pair's r plus q in the x0 slot (dividePublic3out3 with r' = r * 2^k), the
// bias shares, and the fused ReLU drawing from relu_prg.
__device__ __forceinline__ void trunc_store_m(const Epilogue &e, size_t MN, int N, int m, int n, u64 raw) {
    size_t idx = e.tc ? (size_t)m * N + n : (size_t)n * (MN / N) + m;
    Sh3<u64> x;
    divide_3out3_m(raw, e.trunc_bits, e.prg, idx, x.s);
    if (e.relu_out) {
#pragma unroll
        for (int j = 0; j < 6; j++) {
            u64 v = x.s[j >> 1][j & 1];
            if (e.bias) v += e.bias[(size_t)j * N + n];
            e.out[(size_t)j * MN + idx] = v;
            x.s[j >> 1][j & 1] = v;
        }
        Sh3<u8> b;
        Sh3<u64> r = relu_m<u8>(x, b, e.relu_prg, idx);
        store3(e.relu_d, MN, idx, b);
        store3(e.relu_out, MN, idx, r);
        return;
    }
    for (int j = 0; j < e.nout; j++) {
        u64 v = j < 6 ? x.s[j >> 1][j & 1] : (u64)0;
        if (e.bias) v += e.bias[(size_t)j * N + n];
        e.out[(size_t)j * MN + idx] = v;
    }
}

__device__ __forceinline__ void epilogue_store(const Epilogue &e, int M, int N, int m, int n, u64 raw) {
    if (e.mode == EPI_RAW) {
        if (e.tc) e.C[(size_t)m * N + n] = raw;
        else e.C[(size_t)n * M + m] = raw;
    } else {
        trunc_store(e, (size_t)M * N, N, m, n, raw);
    }
}

// trunc_store for the two adjacent row-major outputs (m, n) and (m, n + 1)
// with 16-byte stores of the 64-bit arrays and 2-byte stores of the bit
// arrays.  Valid when Epilogue::pairs is set (tc, no sub_out, N and n even,
// arrays aligned).
__device__ __forceinline__ void trunc_store2(const Epilogue &e, size_t MN, int N, int m, int n, u64 raw0,
                                             u64 raw1) {
    const u64 base0 = trunc_base(e, raw0), base1 = trunc_base(e, raw1);
    const size_t idx = (size_t)m * N + n;
    if (e.relu_out || e.bw_relu_out) {
        Sh3<u64> x0, x1, zero;
#pragma unroll
        for (int j = 0; j < 6; j++) {
            const bool on = (e.base_mask >> j) & 1u;
            u64 v0 = on ? base0 : (u64)0, v1 = on ? base1 : (u64)0;
            if (e.bias) {
                const u64x2 b = *(const u64x2 *)(e.bias + (size_t)j * N + n);
                v0 += b.x;
                v1 += b.y;
            }
            *(u64x2 *)(e.out + (size_t)j * MN + idx) = (u64x2){v0, v1};
            x0.s[j >> 1][j & 1] = v0;
            x1.s[j >> 1][j & 1] = v1;
            zero.s[j >> 1][j & 1] = 0;
        }
        Sh3<u8> b0, b1;
        u64 *dst;
        if (e.relu_out) {
            // the following ReLU layer on the same elements: dReLU + selectShare
            b0 = drelu_<u8>(x0);
            b1 = drelu_<u8>(x1);
#pragma unroll
            for (int j = 0; j < 6; j++)
                *(unsigned short *)(e.relu_d + (size_t)j * MN + idx) =
                    (unsigned short)(b0.s[j >> 1][j & 1] | ((unsigned)b1.s[j >> 1][j & 1] << 8));
            dst = e.relu_out;
        } else {
            // the preceding ReLU layer's backward on the same elements
#pragma unroll
            for (int j = 0; j < 6; j++) {
                const unsigned d = *(const unsigned short *)(e.bw_relu_d + (size_t)j * MN + idx);
                b0.s[j >> 1][j & 1] = (u8)d;
                b1.s[j >> 1][j & 1] = (u8)(d >> 8);
            }
            dst = e.bw_relu_out;
        }
        const Sh3<u64> r0 = select_<u8>(zero, x0, b0), r1 = select_<u8>(zero, x1, b1);
#pragma unroll
        for (int j = 0; j < 6; j++)
            *(u64x2 *)(dst + (size_t)j * MN + idx) = (u64x2){r0.s[j >> 1][j & 1], r1.s[j >> 1][j & 1]};
        return;
    }
    for (int j = 0; j < e.nout; j++) {
        const bool on = (e.base_mask >> j) & 1u;
        u64 v0 = on ? base0 : (u64)0, v1 = on ? base1 : (u64)0;
        if (e.bias) {
            const u64x2 b = *(const u64x2 *)(e.bias + (size_t)j * N + n);
            v0 += b.x;
            v1 += b.y;
        }
        *(u64x2 *)(e.out + (size_t)j * MN + idx) = (u64x2){v0, v1};
    }
}

__device__ __forceinline__ void epilogue_store2(const Epilogue &e, int M, int N, int m, int n, u64 raw0, u64 raw1) {
    if (e.mode == EPI_RAW) *(u64x2 *)(e.C + (size_t)m * N + n) = (u64x2){raw0, raw1};
    else trunc_store2(e, (size_t)M * N, N, m, n, raw0, raw1);
}

// ---------------------------------------------------------------------------
// B packing: Bp[nt][ks][limb][64 n][32 k] int8, zero padded in n and k.
// One workgroup per (n-tile, K-step) block of 64 x 32 values: the block is read
// along the operand's contiguous dimension (n for row-major and the dgrad
// filter view, k for column-major), recoded and staged in LDS, then each
// thread byte-transposes 8 k of one n and writes its 8 limb words; every
// plane of the block is 2 KB of contiguous output.

__device__ __forceinline__ u64 pack_src_value(const LimbBSrc &src, int seg, int k, int n) {
    if (n >= src.N || k >= src.K) return 0;
    size_t idx;
    if (src.dgrad_filter) {
        int fwc = src.kfw * src.kc;
        int r = k / fwc, rem = k - r * fwc;
        int sc = rem / src.kc, c = rem - sc * src.kc;
        idx = (((size_t)c * src.kfh + r) * src.kfw + sc) * src.N + n;
    } else {
        idx = src.row_major ? (size_t)k * src.N + n : (size_t)n * src.K + k;
    }
    u64 v = 0;
#pragma unroll
    for (int t = 0; t < LIMB_MAX_TERMS; t++)
        if (src.b[seg][t]) v += (u64)(s64)src.coef[seg][t] * src.b[seg][t][idx];
    return v;
}

__global__ __launch_bounds__(256) void limb_pack_b_kernel(LimbBSrc src, int8_t *__restrict__ out, int nks) {
    __shared__ u64 tile[LIMB_BN][LIMB_BK + 1];
    const int ks = blockIdx.x, nt = blockIdx.y, tid = threadIdx.x;
    const int kps = src.Kp / LIMB_BK;  // K-steps per segment
    const int seg = ks / kps;
    const int k0 = (ks % kps) * LIMB_BK, n0 = nt * LIMB_BN;
    if (src.row_major || src.dgrad_filter) {
        // n contiguous: a wave reads 64 consecutive n of one k
        const int nl = tid % LIMB_BN, kl0 = tid / LIMB_BN;
#pragma unroll
        for (int j = 0; j < LIMB_BK / 4; j++) {
            int kl = kl0 + 4 * j;
            tile[nl][kl] = recode_digits(pack_src_value(src, seg, k0 + kl, n0 + nl));
        }
    } else {
        // k contiguous: half a wave reads 32 consecutive k of one n
        const int kl = tid % LIMB_BK, nl0 = tid / LIMB_BK;
#pragma unroll
        for (int j = 0; j < LIMB_BN / 8; j++) {
            int nl = nl0 + 8 * j;
            tile[nl][kl] = recode_digits(pack_src_value(src, seg, k0 + kl, n0 + nl));
        }
    }
    __syncthreads();
    const int nl = tid / (LIMB_BK / 8), kq = tid % (LIMB_BK / 8);
    u64 w[8], limbs[8];
#pragma unroll
    for (int j = 0; j < 8; j++) w[j] = tile[nl][kq * 8 + j];
    transpose8x8(w, limbs);
    int8_t *blk = out + ((size_t)nt * nks + ks) * (8 * LIMB_BN * LIMB_BK);
#pragma unroll
    for (int i = 0; i < 8; i++) *(u64 *)(blk + (size_t)i * LIMB_BN * LIMB_BK + nl * LIMB_BK + kq * 8) = limbs[i];
}

int limb_pack_b(const LimbBSrc &src, int8_t *packed, hipStream_t s) {
    PIR_REQUIRE(src.nseg >= 1 && src.nseg <= LIMB_MAX_SEG, "limb_pack_b: nseg must be 1..12");
    PIR_REQUIRE(!src.dgrad_filter || (src.kfh > 0 && src.kfw > 0 && src.kc > 0 && src.K == src.kfh * src.kfw * src.kc),
                "limb_pack_b: dgrad filter view needs K = kfh * kfw * kc");
    PIR_REQUIRE(src.Kp % LIMB_BK == 0 && src.Kp >= src.K, "limb_pack_b: Kp must be a multiple of 32 >= K");
    int nks = src.nseg * src.Kp / LIMB_BK;
    const int ntiles = (src.N + LIMB_BN - 1) / LIMB_BN;
    if (ntiles == 0 || nks == 0) return OK;
    hipLaunchKernelGGL(limb_pack_b_kernel, dim3(nks, ntiles), dim3(256), 0, s, src, packed, nks);
    PIR_LAUNCH_CHECK();
    return OK;
}

// ---------------------------------------------------------------------------
// Limb MFMA GEMM.  256 threads = 4 waves, each owning a 32x32 output block
// computed with v_mfma_i32_32x32x32_i8; K-step 32.  Two workgroup shapes:
//   WM = 2: 64x64 tile, waves 2 (M) x 2 (N);
//   WM = 1: 32x128 tile, waves 1 x 4, used when the packed-N-tile count is
//           even: every A row is loaded and recoded once per 128 columns.
// A is recoded and byte-transposed into 8 limb planes in LDS (shared by the
// waves of a row block), double-buffered so a K-step needs one barrier.  B is
// pre-packed in the MFMA operand layout, so each wave loads its own 16 bytes
// per plane straight from global memory into registers; only A passes through
// LDS.  The next K-step's A and B loads are issued before the current step's
// MFMAs, and a wave raises its issue priority while it issues MFMAs, so the
// other workgroup's wave on the same SIMD stages in the MFMA shadow.

constexpr int LDS_ROW = 32;

// A-operand loader: each thread owns KPT consecutive k of one row of the tile.
// Loads are branch-free (indices are clamped into the operand and the value is
// masked to zero), so all of them are in flight together.  The K-step position
// (segment, step within segment) is uniform across the workgroup.
//   AMODE 0: dense row-major A, 1: dense column-major A,
//   AMODE 2: implicit im2col, any Cin (one decode per K-step, then a step along the filter row),
//   AMODE 3: implicit im2col with Cin % 32 == 0 (one (r, s) per K-step, decoded
//            once per workgroup and advanced incrementally),
//   AMODE 4: dgrad gather of dy, any channel count (per-element decode),
//   AMODE 5: dgrad gather with Cin % 32 == 0 (one (r, s) per K-step),
//   AMODE 6: wgrad transposed im2col: the row fixes (r, s, c), the thread's KPT
//            consecutive k are consecutive output pixels (b, p, q).
template <int AMODE, int KPT>
struct ALoader {
    const LimbA &a;
    int am, akq, kps;
    bool row_ok;
    int seg, kk;          // uniform: segment and K-step within the segment
    const u64 *src;       // a.seg[seg], reloaded only when the segment changes
    int cb, ih0, iw0;     // fprop: image and top-left input coordinate of this row's pixel;
                          // dgrad: image and (h + ph, w + pw) of this row's pixel
    int r, s, c0;         // AMODE 3/5: k = kk*32 = (r*fw + s)*Cin + c0; AMODE 6: the row's (r, s, c)
    // AMODE 2: k = rr * fw*Cin + j reads element pix + rr * W*Cin + j of the
    // segment (j = s*Cin + c runs along one input row), inside the image iff
    // rr in [rlo, rhi) and j in [jlo, jhi)
    long long pix;
    int rlo, rhi, jlo, jhi;

    __device__ ALoader(const LimbA &a_, int am_, int akq_) : a(a_), am(am_), akq(akq_) {
        kps = a.Kp / LIMB_BK;
        row_ok = am < a.M;
        cb = ih0 = iw0 = 0;
        r = s = c0 = 0;
        pix = 0;
        rlo = rhi = jlo = jhi = 0;
        const int m = row_ok ? am : 0;
        if (AMODE == 2 || AMODE == 3) {
            int ohw = a.OH * a.OW;
            cb = m / ohw;
            int rem = m - cb * ohw;
            int oh = rem / a.OW, ow = rem - (rem / a.OW) * a.OW;
            ih0 = oh * a.stride - a.ph;
            iw0 = ow * a.stride - a.pw;
            if (AMODE == 2) {
                pix = (((long long)cb * a.H + ih0) * a.W + iw0) * a.Cin;
                rlo = max(0, -ih0);
                rhi = row_ok ? min(a.fh, a.H - ih0) : 0;  // empty range: the row reads nothing
                jlo = max(0, -iw0) * a.Cin;
                jhi = min(a.fw, a.W - iw0) * a.Cin;
            }
        } else if (AMODE == 4 || AMODE == 5) {
            int hw = a.H * a.W;
            cb = m / hw;
            int rem = m - cb * hw;
            int h = rem / a.W, w = rem - (rem / a.W) * a.W;
            ih0 = h + a.ph;
            iw0 = w + a.pw;
        } else if (AMODE == 6) {
            int fwc = a.fw * a.Cin;
            r = m / fwc;
            int rem = m - r * fwc;
            s = rem / a.Cin;
            c0 = rem - s * a.Cin;
        }
    }
    __device__ void decode() {
        if (AMODE == 3 || AMODE == 5) {
            int k = kk * LIMB_BK, fwc = a.fw * a.Cin;
            r = k / fwc;
            int rem = k - r * fwc;
            s = rem / a.Cin;
            c0 = rem - s * a.Cin;
        }
    }
    __device__ void init(int ks) {
        seg = ks / kps;
        kk = ks - seg * kps;
        src = a.seg[seg];
        decode();
    }
    __device__ void advance() {
        if (++kk == kps) {
            kk = 0;
            seg++;
            src = a.seg[seg < a.nseg ? seg : a.nseg - 1];
            decode();
        } else if (AMODE == 3 || AMODE == 5) {
            c0 += LIMB_BK;
            if (c0 == a.Cin) {
                c0 = 0;
                if (++s == a.fw) {
                    s = 0;
                    ++r;
                }
            }
        }
    }
    // dgrad: output pixel (p, q) of dy that tap (rr, ss) of this row's pixel reads
    __device__ __forceinline__ bool dgrad_src(int rr, int ss, int &p, int &q) const {
        int hp = ih0 - rr, wq = iw0 - ss;
        bool ok = hp >= 0 && wq >= 0;
        if (a.stride == 1) {
            p = hp;
            q = wq;
        } else {
            ok = ok && (hp % a.stride) == 0 && (wq % a.stride) == 0;
            p = hp / a.stride;
            q = wq / a.stride;
        }
        return ok && p < a.OH && q < a.OW;
    }
    // Loads the thread's KPT words of the current K-step into w (raw: an
    // out-of-range element reads a valid clamped address) and sets bit e of
    // ok for each element inside the operand.  The mask is applied when the
    // words are staged, after the MFMAs, so no wait on these loads precedes them.
    __device__ void load(u64 w[KPT], u32 &ok_bits) const {
        const int kb = kk * LIMB_BK + akq * KPT;
        if (AMODE == 3) {
            int ih = ih0 + r, iw = iw0 + s;
            bool ok = row_ok && kb < a.K && ih >= 0 && ih < a.H && iw >= 0 && iw < a.W;
            const u64 *p = src + (((size_t)cb * a.H + (ok ? ih : 0)) * a.W + (ok ? iw : 0)) * a.Cin +
                           (ok ? c0 + akq * KPT : 0);
            ok_bits = ok ? (1u << KPT) - 1 : 0u;
#pragma unroll
            for (int j = 0; j < KPT; j++) w[j] = p[j];
        } else if (AMODE == 5) {
            int pp, qq;
            bool ok = dgrad_src(r, s, pp, qq) && row_ok && kb < a.K;
            const u64 *p = src + (((size_t)cb * a.OH + (ok ? pp : 0)) * a.OW + (ok ? qq : 0)) * a.Cin +
                           (ok ? c0 + akq * KPT : 0);
            ok_bits = ok ? (1u << KPT) - 1 : 0u;
#pragma unroll
            for (int j = 0; j < KPT; j++) w[j] = p[j];
        } else if (AMODE == 2) {
            // decode the thread's first k once, then step along the filter row
            const int fwc = a.fw * a.Cin, wc = a.W * a.Cin;
            int rr = (int)a.fd_fwc.div((u32)kb);
            int j = kb - rr * fwc;
            int d = rr * wc + j;
            ok_bits = 0;
#pragma unroll
            for (int e = 0; e < KPT; e++) {
                // rr < rhi <= fh also keeps k < K (the segment's last k is fh*fwc - 1)
                const bool ok = rr >= rlo && rr < rhi && j >= jlo && j < jhi;
                w[e] = src[ok ? pix + d : 0];
                ok_bits |= ok ? 1u << e : 0u;
                const bool wrap = ++j == fwc;
                j = wrap ? 0 : j;
                rr += wrap ? 1 : 0;
                d += wrap ? 1 + wc - fwc : 1;
            }
        } else if (AMODE == 4) {
            const int fwc = a.fw * a.Cin;
            ok_bits = 0;
#pragma unroll
            for (int j = 0; j < KPT; j++) {
                int k = kb + j;
                int rr = (int)a.fd_fwc.div((u32)k);
                int rem = k - rr * fwc;
                int ss = (int)a.fd_cin.div((u32)rem);
                int cc = rem - ss * a.Cin;
                int pp, qq;
                const bool ok = row_ok && k < a.K && dgrad_src(rr, ss, pp, qq);
                const size_t idx = ok ? (((size_t)cb * a.OH + pp) * a.OW + qq) * a.Cin + cc : 0;
                w[j] = src[idx];
                ok_bits |= ok ? 1u << j : 0u;
            }
        } else if (AMODE == 6) {
            // k = (n, p, q): decode the first of the thread's KPT pixels, then
            // step along the output row with additions only: the element
            // offset is rowoff (row ih of image n, channel c) + iw * Cin, and a
            // new output row or image moves rowoff by a fixed step
            const int ohw = a.OH * a.OW;
            int n = (int)a.fd_ohw.div((u32)kb);
            int rem = kb - n * ohw;
            int p = (int)a.fd_ow.div((u32)rem);
            int q = rem - p * a.OW;
            int ih = p * a.stride - a.ph + r, iw = q * a.stride - a.pw + s;
            long long rowoff = (((long long)n * a.H + ih) * a.W) * a.Cin + c0;  // used only when in range
            if (a.OW % KPT == 0) {
                // the thread's KPT pixels lie in one output row (kb is a
                // multiple of KPT, and so are OW and the segment's K): one
                // row check, then only the column moves
                const bool rin = row_ok && kb < a.K && ih >= 0 && ih < a.H;
                ok_bits = 0;
#pragma unroll
                for (int j = 0; j < KPT; j++) {
                    const int iwj = iw + j * a.stride;
                    const bool ok = rin && iwj >= 0 && iwj < a.W;
                    w[j] = src[ok ? rowoff + (long long)iwj * a.Cin : 0];
                    ok_bits |= ok ? 1u << j : 0u;
                }
                return;
            }
            int off = iw * a.Cin;
            const int col_step = a.stride * a.Cin;
            const long long row_step = (long long)a.stride * a.W * a.Cin;
            const long long img_step = (long long)a.H * a.W * a.Cin - (long long)a.OH * row_step;
            ok_bits = 0;
#pragma unroll
            for (int j = 0; j < KPT; j++) {
                // kb + j < K keeps n < batch
                const bool ok = row_ok && kb + j < a.K && ih >= 0 && ih < a.H && iw >= 0 && iw < a.W;
                w[j] = src[ok ? rowoff + off : 0];
                ok_bits |= ok ? 1u << j : 0u;
                iw += a.stride;
                off += col_step;
                if (++q == a.OW) {
                    q = 0;
                    iw = s - a.pw;
                    off = iw * a.Cin;
                    ih += a.stride;
                    rowoff += row_step;
                    if (++p == a.OH) {
                        p = 0;
                        ih = r - a.ph;
                        rowoff += img_step;
                    }
                }
            }
        } else if (AMODE == 0 && (a.K % KPT) == 0) {
            // a thread's KPT consecutive k are all inside K or all outside: one contiguous run
            bool ok = row_ok && kb < a.K;
            const u64 *p = src + (size_t)(row_ok ? am : 0) * a.lda + (ok ? kb : 0);
            ok_bits = ok ? (1u << KPT) - 1 : 0u;
#pragma unroll
            for (int j = 0; j < KPT; j++) w[j] = p[j];
        } else {
            const int m = row_ok ? am : 0;
            ok_bits = 0;
#pragma unroll
            for (int j = 0; j < KPT; j++) {
                int k = kb + j;
                bool ok = row_ok && k < a.K;
                int kc = ok ? k : 0;
                size_t idx = AMODE == 0 ? (size_t)m * a.lda + kc : (size_t)kc * a.lda + m;
                w[j] = src[idx];
                ok_bits |= ok ? 1u << j : 0u;
            }
        }
    }
};

// Recode KPT consecutive k of one row (element j zero unless bit j of ok_bits
// is set) and store limb plane i's bytes for them at
// As + i * plane + arow * LDS_ROW + akq * KPT.
template <int KPT>
__device__ __forceinline__ void stage_a(int8_t *As, int plane, int arow, int akq, const u64 w[KPT], u32 ok_bits) {
    u64 d[KPT];
#pragma unroll
    for (int j = 0; j < KPT; j++) d[j] = recode_digits(((ok_bits >> j) & 1u) ? w[j] : 0ull);
    int8_t *dst = As + arow * LDS_ROW + akq * KPT;
    if (KPT == 8) {
        u64 limbs[8];
        transpose8x8(d, limbs);
#pragma unroll
        for (int i = 0; i < 8; i++) *(u64 *)(dst + i * plane) = limbs[i];
    } else {
        u32 l[8];
        transpose4((u32)d[0], (u32)d[1], (u32)d[2], (u32)d[3], l[0], l[1], l[2], l[3]);
        transpose4((u32)(d[0] >> 32), (u32)(d[1] >> 32), (u32)(d[2] >> 32), (u32)(d[3] >> 32), l[4], l[5], l[6],
                   l[7]);
#pragma unroll
        for (int i = 0; i < 8; i++) *(u32 *)(dst + i * plane) = l[i];
    }
}

// One workgroup covers at most LIMB_FLUSH_STEPS K-steps (8192 k), so the eight
// int32 group accumulators never overflow and no 64-bit running sum is needed
// inside the loop; longer K is split across workgroups (split-K) and the 64-bit
// partials are summed by split_reduce_kernel.  With <= 256 registers per lane
// (128 accumulators + operands) two workgroups share a CU, so one's staging
// overlaps the other's MFMAs.
template <int AMODE, int WM, bool HELP>
__device__ __forceinline__ void limb_gemm_body(const LimbA &a, const int8_t *__restrict__ Bp, int N, int nks,
                                               int steps_per_split, const Epilogue &epi, u64 *__restrict__ part) {
    constexpr int TM = 32 * WM, TN = 128 / WM, KPT = 4 * WM;
    constexpr int KG = LIMB_BK / KPT;  // threads per row for staging
    constexpr int PLANE = TM * LDS_ROW;
    constexpr int STEP = 8 * LIMB_BN * LIMB_BK;  // packed B bytes per n-tile per K-step
    // software-pipelined K-loops: the 32x128 tile loads A two K-steps ahead;
    // the 64x64 tile (twice the A words per thread) one step ahead in the
    // fprop / dgrad forms, while its dense and wgrad forms keep the plain loop
    // (they stream the A planes from LDS, which measures faster for them)
    constexpr bool PIPE = WM == 1;
    constexpr bool PIPE1 = WM == 2 && AMODE >= 2 && AMODE <= 5;
    constexpr int TS = TN + 2;                         // row stride (u64) of the staged output tile
    constexpr int A_BYTES = 2 * 8 * PLANE, T_BYTES = TM * TS * 8;
    __shared__ __attribute__((aligned(16))) int8_t smem[A_BYTES > T_BYTES ? A_BYTES : T_BYTES];
    int8_t(*As2)[8 * PLANE] = (int8_t(*)[8 * PLANE])smem;
    int cur = 0;

    // XCD-aware tile order: the dispatcher deals consecutive workgroup ids
    // round-robin over the 8 XCDs, so give each XCD a contiguous run of logical
    // tiles (M fastest): workgroups that share a B slice then share one L2.
    int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    {
        const int total = gridDim.x * gridDim.y * gridDim.z, per_xcd = total / 8;
        if (bid < per_xcd * 8) bid = (bid % 8) * per_xcd + bid / 8;
    }
    // Dense operands (no filter-tap reuse of A in L2) take the tiles in groups
    // of 8 row tiles, M fastest within a group: an XCD's 64 resident
    // workgroups then cover 8 row x 8 column tiles, not 64 x 1, which cuts the
    // distinct A rows per K-step eightfold for eight times the B slices (which
    // the column-sharing waves of several XCDs fetch anyway).
    // Measured on the 3-term ring product: n = 8192 +10 %, 4096 +3 %, 2048
    // unchanged (groups of 2, 4, 6, 16 and 32 were slower).
    constexpr int DENSE_GM = 8;
    const int tiles2d = gridDim.x * gridDim.y, bz = bid / tiles2d, t2 = bid - bz * tiles2d;
    int bx, by;
    if (AMODE <= 1) {
        const int per_group = DENSE_GM * gridDim.y, g = t2 / per_group, first = g * DENSE_GM;
        const int gm = min(DENSE_GM, (int)gridDim.x - first), r = t2 - g * per_group;
        bx = first + r % gm;
        by = r / gm;
    } else {
        bx = t2 % gridDim.x;
        by = t2 / gridDim.x;
    }

    // the wave index is uniform across the wave: read it into a scalar, so the
    // per-wave quantities below (and the branches on them) are scalar
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave / (4 / WM), wn = wave % (4 / WM);
    const int m0 = bx * TM, n0 = by * TN + wn * 32;
    // a wave whose 32 columns lie past N (N not a multiple of the tile width)
    // still stages A and meets the barriers, but loads no B and issues no MFMA
    const bool live = n0 < N;
    const int M = a.M;
    const int ks_begin = bz * steps_per_split;
    const int ks_end = min(nks, ks_begin + steps_per_split);

    // column-major forms: consecutive lanes take consecutive rows (contiguous in memory)
    const int arow = (AMODE == 1 || AMODE == 6) ? tid % TM : tid / KG;
    const int akq = (AMODE == 1 || AMODE == 6) ? tid / TM : tid % KG;
    ALoader<AMODE, KPT> L(a, m0 + arow, akq);

    // this lane's 16 bytes of each B limb plane: column n0 + lane % 32, k half lane / 32
    const int8_t *bw = Bp + (size_t)(n0 / LIMB_BN) * nks * STEP + ((n0 % LIMB_BN) + (lane & 31)) * LIMB_BK +
                       (lane >> 5) * 16;
    const int a_off = (wm * 32 + (lane & 31)) * LDS_ROW + (lane >> 5) * 16;
    v16i acc[8];
#pragma unroll
    for (int q = 0; q < 8; q++) acc[q] = (v16i){0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

    static_assert(!HELP || WM == 1, "the helper form is a 32x128 tile form");
    if constexpr (HELP) {
        // 32x128 tile with three column blocks (64 < N <= 96): waves 0-2 sum
        // the limb products of s = 0..6 for their own block (28 MFMAs a
        // K-step), and wave 3, which has no columns of its own, sums s = 7 for
        // all three blocks (3 x 8 MFMAs), so every SIMD's matrix pipe takes a
        // share instead of one SIMD idling.  acc[s] holds s in waves 0-2,
        // acc[c] block c's s = 7 in wave 3.  Wave 3 keeps all three blocks' B
        // planes, each reloaded for the next K-step as soon as its last MFMA
        // has issued, so no load is waited on within a step.  A is staged by
        // every wave, loaded two K-steps ahead and staged interleaved with the
        // MFMAs, as in the loop below (four or five staging instructions per
        // MFMA here: two or three measured 2-3 % slower).  Measured on the
        // AlexNet-CIFAR first conv over 986 images (N = 96): 673 -> 610 us.
        // Splitting s = 0 and 7 off to the fourth wave instead, which then
        // needed block 2's B planes within the step they were loaded in, made
        // it wait on the step's A loads too and ran 790 us.
        u64 wa[KPT], wb[KPT];
        u32 oka = 0, okb = 0;
        if (ks_begin < ks_end) {
            L.init(ks_begin);
            L.load(wa, oka);
            stage_a<KPT>(As2[0], PLANE, arow, akq, wa, oka);
            if (ks_begin + 1 < ks_end) {
                L.advance();
                L.load(wa, oka);  // A(ks_begin + 1)
            }
        }
        constexpr int PL = LIMB_BN * LIMB_BK;  // bytes per packed B plane
        if (wave < 3) {
            v4i bv[7];  // plane 7 only meets s = 7
            if (ks_begin < ks_end) {
                const int8_t *bp = bw + (size_t)ks_begin * STEP;
#pragma unroll
                for (int j = 0; j < 7; j++) bv[j] = *(const v4i *)(bp + j * PL);
            }
            __syncthreads();
            auto step = [&](int ks, u64(&ws)[KPT], u32 &oks, u64(&wl)[KPT], u32 &okl) __attribute__((always_inline)) {
                v4i ap[7];
                const int8_t *As = As2[cur];
#pragma unroll
                for (int i = 0; i < 7; i++) ap[i] = *(const v4i *)(As + i * PLANE + a_off);
                const bool more = ks + 1 < ks_end;
                if (ks + 2 < ks_end) L.advance();
                L.load(wl, okl);
                const int8_t *bp = bw + (size_t)(more ? ks + 1 : ks) * STEP;
                __builtin_amdgcn_sched_barrier(0);
                __builtin_amdgcn_s_setprio(1);
#pragma unroll
                for (int j = 0; j < 7; j++) {
#pragma unroll
                    for (int i = 0; i + j <= 6; i++)
                        acc[i + j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ap[i], bv[j], acc[i + j], 0, 0, 0);
                    bv[j] = *(const v4i *)(bp + j * PL);
                }
                stage_a<KPT>(As2[cur ^ 1], PLANE, arow, akq, ws, oks);
#pragma unroll
                for (int g = 0; g < 20; g++) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
                    __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);  // VALU
                }
#pragma unroll
                for (int g = 0; g < 8; g++) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
                    __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // LDS write
                }
                __builtin_amdgcn_s_setprio(0);
                cur ^= 1;
                __syncthreads();
            };
            for (int ks = ks_begin; ks < ks_end; ks += 2) {
                step(ks, wa, oka, wb, okb);
                if (ks + 1 < ks_end) step(ks + 1, wb, okb, wa, oka);
            }
        } else {
            // this lane's bytes of block c's B planes (block c = columns 32c..32c+31)
            const int8_t *bwc[3];
#pragma unroll
            for (int c = 0; c < 3; c++)
                bwc[c] = Bp + (size_t)(c / 2) * nks * STEP + ((c % 2) * 32 + (lane & 31)) * LIMB_BK + (lane >> 5) * 16;
            v4i bb[3][8];
            if (ks_begin < ks_end) {
#pragma unroll
                for (int c = 0; c < 3; c++)
#pragma unroll
                    for (int j = 0; j < 8; j++) bb[c][j] = *(const v4i *)(bwc[c] + (size_t)ks_begin * STEP + j * PL);
            }
            __syncthreads();
            auto step = [&](int ks, u64(&ws)[KPT], u32 &oks, u64(&wl)[KPT], u32 &okl) __attribute__((always_inline)) {
                v4i ap[8];
                const int8_t *As = As2[cur];
#pragma unroll
                for (int i = 0; i < 8; i++) ap[i] = *(const v4i *)(As + i * PLANE + a_off);
                const bool more = ks + 1 < ks_end;
                if (ks + 2 < ks_end) L.advance();
                L.load(wl, okl);
                const size_t nxt = (size_t)(more ? ks + 1 : ks) * STEP;
                __builtin_amdgcn_sched_barrier(0);
                __builtin_amdgcn_s_setprio(1);
                // the three blocks' s = 7 chains interleaved
#pragma unroll
                for (int i = 0; i < 8; i++)
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        acc[c] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ap[i], bb[c][7 - i], acc[c], 0, 0, 0);
                        bb[c][7 - i] = *(const v4i *)(bwc[c] + nxt + (7 - i) * PL);
                    }
                stage_a<KPT>(As2[cur ^ 1], PLANE, arow, akq, ws, oks);
#pragma unroll
                for (int g = 0; g < 16; g++) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
                    __builtin_amdgcn_sched_group_barrier(0x002, 5, 0);  // VALU
                }
#pragma unroll
                for (int g = 0; g < 8; g++) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
                    __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // LDS write
                }
                __builtin_amdgcn_s_setprio(0);
                cur ^= 1;
                __syncthreads();
            };
            for (int ks = ks_begin; ks < ks_end; ks += 2) {
                step(ks, wa, oka, wb, okb);
                if (ks + 1 < ks_end) step(ks + 1, wb, okb, wa, oka);
            }
        }
    } else if constexpr (PIPE) {
        // Software-pipelined K-loop of the 32x128 tile (every operand form):
        // A is loaded two K-steps ahead, so the words staged during a step
        // were loaded one step earlier and need no wait, and the staging of
        // A(ks + 1) is interleaved with the step's own 36 MFMAs: each wave
        // keeps its matrix pipe busy while it stages (measured: two staging
        // instructions per MFMA, the LDS writes over the last eight MFMAs,
        // beats one per MFMA, three per MFMA and the compiler's own order by
        // 2-3 %).  The eight A planes of the step are read right after the
        // barrier, before any MFMA.  The MFMAs run B-plane by
        // B-plane, and plane j of B(ks + 1) is loaded into plane j's registers
        // as soon as its last MFMA of the step has issued, so B needs one
        // register set.  Unrolled by two so the two A register sets trade roles
        // without copies (a copy would wait on loads still in flight).
        u64 wa[KPT], wb[KPT];
        u32 oka = 0, okb = 0;
        v4i bv[8];
        if (ks_begin < ks_end) {
            L.init(ks_begin);
            L.load(wa, oka);
            const int8_t *bp = bw + (size_t)ks_begin * STEP;
            if (live)
#pragma unroll
                for (int j = 0; j < 8; j++) bv[j] = *(const v4i *)(bp + j * LIMB_BN * LIMB_BK);
            stage_a<KPT>(As2[0], PLANE, arow, akq, wa, oka);
            if (ks_begin + 1 < ks_end) {
                L.advance();
                L.load(wa, oka);  // A(ks_begin + 1)
            }
        }
        __syncthreads();
        // one K-step: MFMAs on A(ks) (LDS buffer cur) and B(ks); loads A(ks + 2)
        // into wl and B(ks + 1) into the B registers; stages A(ks + 1), held in
        // ws, into the other buffer
        auto step = [&](int ks, u64(&ws)[KPT], u32 &oks, u64(&wl)[KPT], u32 &okl) __attribute__((always_inline)) {
            v4i ap[8];
            const int8_t *As = As2[cur];
#pragma unroll
            for (int i = 0; i < 8; i++) ap[i] = *(const v4i *)(As + i * PLANE + a_off);
            const bool more = ks + 1 < ks_end;
            if (ks + 2 < ks_end) L.advance();
            L.load(wl, okl);  // past the end: re-reads the last step
            const int8_t *bp = bw + (size_t)(more ? ks + 1 : ks) * STEP;
            __builtin_amdgcn_sched_barrier(0);
            if (live) {
                __builtin_amdgcn_s_setprio(1);
#pragma unroll
                for (int j = 0; j < 8; j++) {
#pragma unroll
                    for (int i = 0; i < 8 - j; i++)
                        acc[i + j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ap[i], bv[j], acc[i + j], 0, 0, 0);
                    bv[j] = *(const v4i *)(bp + j * LIMB_BN * LIMB_BK);
                }
                stage_a<KPT>(As2[cur ^ 1], PLANE, arow, akq, ws, oks);
                // one or two staging instructions between consecutive MFMAs,
                // the LDS writes towards the end of the run
#pragma unroll
                for (int g = 0; g < 28; g++) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
                    __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);  // VALU
                }
#pragma unroll
                for (int g = 0; g < 8; g++) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
                    __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // LDS write
                }
                __builtin_amdgcn_s_setprio(0);
            } else {
                stage_a<KPT>(As2[cur ^ 1], PLANE, arow, akq, ws, oks);
            }
            cur ^= 1;
            __syncthreads();
        };
        for (int ks = ks_begin; ks < ks_end; ks += 2) {
            step(ks, wa, oka, wb, okb);
            if (ks + 1 < ks_end) step(ks + 1, wb, okb, wa, oka);
        }
    } else if constexpr (PIPE1) {
        // 64x64 tile, fprop / dgrad forms: A(ks + 1) is loaded at the start
        // of step ks into the one register set there is room for (8 words per
        // thread; two sets spill) and staged interleaved with the step's last
        // 30 MFMAs, four staging instructions per MFMA, once the loads have
        // had the first six MFMAs' time to land; B is reloaded plane by plane
        // as in the 32x128 loop.  Measured against staging after the MFMA run:
        // VGG16's second conv 385 -> 370 us, its first conv 104 -> 92 us (a
        // lead of 2, 4, 8 or 10 MFMAs and 3-5 instructions per MFMA were
        // within 1 % of this, but for a lead of 2).
        u64 w[KPT];
        u32 wok = 0;
        v4i bv[8];
        if (ks_begin < ks_end) {
            L.init(ks_begin);
            L.load(w, wok);
            const int8_t *bp = bw + (size_t)ks_begin * STEP;
            if (live)
#pragma unroll
                for (int j = 0; j < 8; j++) bv[j] = *(const v4i *)(bp + j * LIMB_BN * LIMB_BK);
            stage_a<KPT>(As2[0], PLANE, arow, akq, w, wok);
        }
        __syncthreads();
        for (int ks = ks_begin; ks < ks_end; ++ks) {
            v4i ap[8];
            const int8_t *As = As2[cur];
#pragma unroll
            for (int i = 0; i < 8; i++) ap[i] = *(const v4i *)(As + i * PLANE + a_off);
            const bool more = ks + 1 < ks_end;
            if (more) L.advance();
            L.load(w, wok);  // past the end: re-reads the last step
            const int8_t *bp = bw + (size_t)(more ? ks + 1 : ks) * STEP;
            __builtin_amdgcn_sched_barrier(0);
            if (live) {
                __builtin_amdgcn_s_setprio(1);
#pragma unroll
                for (int j = 0; j < 8; j++) {
#pragma unroll
                    for (int i = 0; i < 8 - j; i++)
                        acc[i + j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ap[i], bv[j], acc[i + j], 0, 0, 0);
                    bv[j] = *(const v4i *)(bp + j * LIMB_BN * LIMB_BK);
                }
                stage_a<KPT>(As2[cur ^ 1], PLANE, arow, akq, w, wok);
#pragma unroll
                for (int g = 0; g < 6; g++) __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
#pragma unroll
                for (int g = 0; g < 22; g++) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
                    __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);  // VALU
                }
#pragma unroll
                for (int g = 0; g < 8; g++) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
                    __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // LDS write
                }
                __builtin_amdgcn_s_setprio(0);
            } else {
                stage_a<KPT>(As2[cur ^ 1], PLANE, arow, akq, w, wok);
            }
            cur ^= 1;
            __syncthreads();
        }
    } else {
    u64 w[KPT];
    u32 wok = 0;
    v4i bv[8], bn[8];
    if (ks_begin < ks_end) {
        L.init(ks_begin);
        L.load(w, wok);
        const int8_t *bp = bw + (size_t)ks_begin * STEP;
        if (live)
#pragma unroll
            for (int j = 0; j < 8; j++) bv[j] = *(const v4i *)(bp + j * LIMB_BN * LIMB_BK);
        stage_a<KPT>(As2[0], PLANE, arow, akq, w, wok);
    }
    __syncthreads();

    // 64x64 tile, dense and wgrad forms: A(ks + 1) and B(ks + 1) are loaded
    // during step ks, the A planes are streamed from LDS one plane ahead of
    // their MFMAs (measured faster for these forms than reading all eight
    // first), and A(ks + 1) is staged after the MFMA run.
    for (int ks = ks_begin; ks < ks_end; ++ks) {
        // prefetch the next K-step into registers (the last step re-reads itself)
        const bool more = ks + 1 < ks_end;
        if (more) L.advance();
        L.load(w, wok);
        const int8_t *bp = bw + (size_t)(more ? ks + 1 : ks) * STEP;
        if (live)
#pragma unroll
            for (int j = 0; j < 8; j++) bn[j] = *(const v4i *)(bp + j * LIMB_BN * LIMB_BK);

        // acc[i+j] += A_i . B_j for i + j < 8
        if (live) {
            const int8_t *As = As2[cur];
            v4i av = *(const v4i *)(As + a_off);
            __builtin_amdgcn_s_setprio(1);
#pragma unroll
            for (int i = 0; i < 8; i++) {
                v4i an = av;
                if (i < 7) an = *(const v4i *)(As + (i + 1) * PLANE + a_off);
#pragma unroll
                for (int j = 0; j < 8 - i; j++)
                    acc[i + j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, bv[j], acc[i + j], 0, 0, 0);
                av = an;
            }
            __builtin_amdgcn_s_setprio(0);
        }
        // the other buffer was last read in the previous step, before the barrier
        stage_a<KPT>(As2[cur ^ 1], PLANE, arow, akq, w, wok);
        cur ^= 1;
#pragma unroll
        for (int j = 0; j < 8; j++) bv[j] = bn[j];
        __syncthreads();
    }
    }

    // ---- epilogue: C = sum_s acc_s << 8s (32x32 MFMA layout: lane -> column,
    // register -> row) is staged in LDS, whose A buffers every wave has left
    // behind the loop's last barrier; a rolled loop then stores the tile with
    // consecutive lanes on consecutive elements of the output's contiguous
    // dimension, two adjacent columns (16 bytes) per lane where the output
    // allows.  One copy of the epilogue's code serves the whole tile.
    u64 *T = (u64 *)smem;
    if (!HELP || wave < 3) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            u64 raw = 0;
#pragma unroll
            for (int q = 0; q < (HELP ? 7 : 8); q++) raw += ((u64)(s64)acc[q][r]) << (8 * q);
            T[(wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)) * TS + wn * 32 + (lane & 31)] = raw;
        }
    }
    __syncthreads();
    if constexpr (HELP) {
        // wave 3 adds its s = 7 sums into the three blocks
        if (wave == 3) {
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < 16; r++)
                    T[(8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)) * TS + c * 32 + (lane & 31)] +=
                        (u64)(s64)acc[c][r] << 56;
        }
        __syncthreads();
    }
    const int nt0 = by * TN;  // first column of the workgroup tile
    if (epi.pairs) {
        constexpr int HALF = TN / 2;
        for (int p = tid; p < TM * HALF; p += 256) {
            const int ml = p / HALF, nl = 2 * (p - ml * HALF);
            const int m = m0 + ml, n = nt0 + nl;
            if (m >= M || n >= N) continue;
            const u64x2 v = *(const u64x2 *)(T + ml * TS + nl);
            if (part) *(u64x2 *)(part + ((size_t)bz * M + m) * N + n) = v;
            else epilogue_store2(epi, M, N, m, n, v.x, v.y);
        }
    } else {
        const bool rows = part || epi.tc;  // the output is row-major
        for (int q = tid; q < TM * TN; q += 256) {
            const int ml = rows ? q / TN : q % TM, nl = rows ? q % TN : q / TM;
            const int m = m0 + ml, n = nt0 + nl;
            if (m >= M || n >= N) continue;
            const u64 v = T[ml * TS + nl];
            if (part) part[((size_t)bz * M + m) * N + n] = v;
            else epilogue_store(epi, M, N, m, n, v);
        }
    }
}

template <int AMODE, int WM>
__global__ __launch_bounds__(256, 2) void limb_gemm_kernel(LimbA a, const int8_t *__restrict__ Bp, int N, int nks,
                                                           int steps_per_split, Epilogue epi, u64 *__restrict__ part) {
    limb_gemm_body<AMODE, WM, false>(a, Bp, N, nks, steps_per_split, epi, part);
}

// The 32x128 tile with three column blocks (64 < N <= 96), fprop forms: the
// fourth wave takes a share of the other three's MFMAs.
template <int AMODE>
__global__ __launch_bounds__(256, 2) void limb_gemm_help_kernel(LimbA a, const int8_t *__restrict__ Bp, int N,
                                                                int nks, int steps_per_split, Epilogue epi,
                                                                u64 *__restrict__ part) {
    limb_gemm_body<AMODE, 1, true>(a, Bp, N, nks, steps_per_split, epi, part);
}

// Sum of the split-K partials of output element(s) t (PAIRS: t and t + 1 as one
// 16-byte load per partial); the partial loads are issued eight (then four) at
// a time, so a thread waits on memory once per batch, not once per partial.
template <bool PAIRS>
__device__ __forceinline__ u64x2 sum_partials(const u64 *__restrict__ part, int split, size_t MN, size_t t) {
    u64x2 acc = (u64x2){0, 0};
    int z = 0;
    // eight partials in flight per wait where the split allows (the training
    // step's wgrad reductions run splits of up to ~60: 11.94 -> 11.89 ms)
    for (; z + 8 <= split; z += 8) {
        u64x2 v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const u64 *p = part + (size_t)(z + q) * MN + t;
            v[q] = PAIRS ? *(const u64x2 *)p : (u64x2){*p, 0};
        }
        acc += ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
    }
    for (; z + 4 <= split; z += 4) {
        u64x2 v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const u64 *p = part + (size_t)(z + q) * MN + t;
            v[q] = PAIRS ? *(const u64x2 *)p : (u64x2){*p, 0};
        }
        acc += (v[0] + v[1]) + (v[2] + v[3]);
    }
    for (; z < split; z++) {
        const u64 *p = part + (size_t)z * MN + t;
        acc += PAIRS ? *(const u64x2 *)p : (u64x2){*p, 0};
    }
    return acc;
}

// The split-K reduction with the product's epilogue; PAIRS: each thread
// takes two adjacent row-major outputs (Epilogue::pairs conditions, partials
// 16-byte aligned).
template <bool MASKED, bool PAIRS>
__global__ __launch_bounds__(256) void split_reduce_kernel(const u64 *__restrict__ part, int split, int M, int N,
                                                           Epilogue epi) {
    const size_t MN = (size_t)M * N;
    const size_t t = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * (PAIRS ? 2 : 1);
    if (t >= MN) return;
    const u64x2 raw = sum_partials<PAIRS>(part, split, MN, t);
    const int m = (int)(t / N), n = (int)(t % N);
    if (MASKED) {
        trunc_store_m(epi, MN, N, m, n, raw.x);
        if (PAIRS) trunc_store_m(epi, MN, N, m, n + 1, raw.y);
    } else if (PAIRS) {
        epilogue_store2(epi, M, N, m, n, raw.x, raw.y);
    } else {
        epilogue_store(epi, M, N, m, n, raw.x);
    }
}

// Whether an epilogue may store two adjacent columns per lane: row-major
// output, N even, no subtraction, the 64-bit arrays 16-byte and the bit
// arrays 2-byte aligned.
static bool epilogue_pairs_ok(const Epilogue &e, int N) {
    auto al = [](const void *q, uintptr_t b) { return ((uintptr_t)q % b) == 0; };
    if (N % 2 != 0 || !e.tc) return false;
    if (e.mode == EPI_RAW) return al(e.C, 16);
    return !e.sub_out && al(e.out, 16) && (!e.bias || al(e.bias, 16)) &&
           (!e.relu_out || (al(e.relu_out, 16) && al(e.relu_d, 2))) &&
           (!e.bw_relu_out || (al(e.bw_relu_out, 16) && al(e.bw_relu_d, 2)));
}

int rss_trunc_epilogue(const u64 *raw, int M, int N, int nsum, const Epilogue &epi, hipStream_t s) {
    size_t MN = (size_t)M * N;
    if (MN == 0) return OK;
    const bool masked = epi.mode == EPI_TRUNC && epi.prg.on;
    PIR_REQUIRE(!masked || (epi.trunc_bits >= 1 && epi.nout <= 6 && !epi.div_bits && !epi.sub_out &&
                            !epi.bw_relu_out && (!epi.relu_out || epi.relu_prg.on)),
                "masked precompute: the product epilogue supports the forward truncation only");
    // pairs: the partials are row-major M x N, so a pair never crosses a row when N is even
    const bool pairs = N % 2 == 0 && ((uintptr_t)raw % 16) == 0 && (masked || epilogue_pairs_ok(epi, N));
    const size_t threads = pairs ? MN / 2 : MN;
    const dim3 grid(grid_for(threads, 256));
    if (masked && pairs) hipLaunchKernelGGL((split_reduce_kernel<true, true>), grid, dim3(256), 0, s, raw, nsum, M, N, epi);
    else if (masked) hipLaunchKernelGGL((split_reduce_kernel<true, false>), grid, dim3(256), 0, s, raw, nsum, M, N, epi);
    else if (pairs) hipLaunchKernelGGL((split_reduce_kernel<false, true>), grid, dim3(256), 0, s, raw, nsum, M, N, epi);
    else hipLaunchKernelGGL((split_reduce_kernel<false, false>), grid, dim3(256), 0, s, raw, nsum, M, N, epi);
    PIR_LAUNCH_CHECK();
    return OK;
}

// ---------------------------------------------------------------------------
// Launch plans and the autotune table

namespace {

// Heuristic split-K factor from a round model: 2 workgroups per CU on 256 CUs
// run 512 at a time, so a launch of W workgroups takes ceil(W / 512) rounds of
// (K-steps per workgroup + a fixed prologue / epilogue cost of ~8 K-steps).
// The split with the fewest rounds x steps wins (the smaller on ties); split-K
// partials are cheap next to that for the shapes of the network.
int heuristic_split(int M, int N, int nks, bool wide) {
    const long tiles = wide ? (long)((M + 31) / 32) * ((N + 127) / 128) : (long)((M + 63) / 64) * ((N + 63) / 64);
    const int lo = (nks + LIMB_FLUSH_STEPS - 1) / LIMB_FLUSH_STEPS;
    int hi = nks / 4;
    if (hi > LIMB_MAX_SPLIT) hi = LIMB_MAX_SPLIT;
    if (hi < lo) hi = lo;
    int best = lo;
    long best_cost = -1;
    for (int sp = lo; sp <= hi; sp++) {
        long rounds = (tiles * sp + 511) / 512;
        long cost = rounds * ((nks + sp - 1) / sp + 8);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = sp;
        }
    }
    return best;
}

bool can_be_wide(int N) { return (((N + LIMB_BN - 1) / LIMB_BN) % 2) == 0; }

std::string shape_key(const LimbA &a, int N) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "%d %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d", a.M, N, a.K, a.Kp,
                  a.nseg, a.lda, a.row_major, a.conv, a.H, a.W, a.Cin, a.fh, a.fw, a.ph, a.pw, a.stride, a.OH, a.OW);
    return buf;
}

std::mutex g_tune_mu;
std::map<std::string, LimbPlan> g_tuned;

bool tuned_plan(const std::string &key, LimbPlan *p) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it == g_tuned.end()) return false;
    *p = it->second;
    return true;
}

}  // namespace

LimbPlan limb_plan(const LimbA &a, int N, int split_hint) {
    const int nks = a.nseg * a.Kp / LIMB_BK;
    LimbPlan p;
    if (split_hint <= 0 && tuned_plan(shape_key(a, N), &p)) return p;
    p.wide = can_be_wide(N) ? 1 : 0;
    p.split = limb_effective_split(nks, split_hint > 0 ? split_hint : heuristic_split(a.M, N, nks, p.wide != 0));
    return p;
}

size_t limb_workspace_bytes(const LimbPlan &p, int M, int N, bool masked) {
    return (p.split > 1 || masked) ? (size_t)(p.split > 1 ? p.split : 1) * M * N * sizeof(u64) : 0;
}

int limb_tune_entries() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return (int)g_tuned.size();
}

void limb_tune_clear() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.clear();
}

// Text format, one shape per line: the 18 integers of the shape key (M N K Kp nseg lda
// row_major conv H W Cin fh fw ph pw stride OH OW), ':', wide, split.
int limb_tune_save(const char *path) {
    PIR_REQUIRE(path, "autotune save: null path");
    std::FILE *f = std::fopen(path, "w");
    PIR_REQUIRE(f, "autotune save: cannot open file");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto &kv : g_tuned) std::fprintf(f, "%s : %d %d\n", kv.first.c_str(), kv.second.wide, kv.second.split);
    std::fclose(f);
    return OK;
}

int limb_tune_load(const char *path) {
    PIR_REQUIRE(path, "autotune load: null path");
    std::FILE *f = std::fopen(path, "r");
    PIR_REQUIRE(f, "autotune load: cannot open file");
    char line[512];
    std::map<std::string, LimbPlan> got;
    bool bad = false;
    while (std::fgets(line, sizeof line, f)) {
        std::string l(line);
        size_t c = l.find(':');
        if (c == std::string::npos) continue;
        int v[18], wide = 0, split = 0;
        if (std::sscanf(l.c_str(), "%d %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d", &v[0], &v[1], &v[2],
                        &v[3], &v[4], &v[5], &v[6], &v[7], &v[8], &v[9], &v[10], &v[11], &v[12], &v[13], &v[14],
                        &v[15], &v[16], &v[17]) != 18 ||
            std::sscanf(l.c_str() + c + 1, "%d %d", &wide, &split) != 2 || split < 1 || (wide != 0 && wide != 1) ||
            (wide && !can_be_wide(v[1]))) {
            bad = true;
            break;
        }
        got[l.substr(0, c - 1)] = LimbPlan{wide, split};
    }
    std::fclose(f);
    PIR_REQUIRE(!bad, "autotune load: malformed entry");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto &kv : got) g_tuned[kv.first] = kv.second;
    return OK;
}

int limb_gemm(const LimbA &a_in, const int8_t *packedB, int N, const Epilogue &epi, const LimbPlan &plan,
              u64 *workspace, hipStream_t s) {
    LimbA a = a_in;
    if (a.conv != LIMB_DENSE) {
        PIR_REQUIRE(a.fw > 0 && a.Cin > 0 && a.OH > 0 && a.OW > 0, "limb_gemm: invalid convolution geometry");
        a.fd_fwc = FastDiv((u32)(a.fw * a.Cin));
        a.fd_cin = FastDiv((u32)a.Cin);
        a.fd_ohw = FastDiv((u32)(a.OH * a.OW));
        a.fd_ow = FastDiv((u32)a.OW);
    }
    PIR_REQUIRE(a.nseg >= 1 && a.nseg <= LIMB_MAX_SEG, "limb_gemm: nseg must be 1..12");
    PIR_REQUIRE(a.Kp % LIMB_BK == 0 && a.Kp >= a.K, "limb_gemm: Kp must be a multiple of 32 >= K");
    PIR_REQUIRE(epi.mode != EPI_TRUNC || (epi.trunc_bits >= 1 && epi.trunc_bits <= 62),
                "limb_gemm: truncation bits must be in [1, 62]");
    PIR_REQUIRE(epi.mode != EPI_TRUNC || (epi.nout >= 1 && epi.nout <= 32), "limb_gemm: bad output share count");
    PIR_REQUIRE(!epi.relu_out || (epi.mode == EPI_TRUNC && epi.nout == 6 && epi.relu_d),
                "limb_gemm: a fused ReLU needs the six RSS output arrays");
    PIR_REQUIRE(!epi.sub_out || (epi.mode == EPI_TRUNC && !epi.bias && !epi.relu_out && !epi.bw_relu_out),
                "limb_gemm: the subtracting epilogue takes no bias and no fused ReLU");
    PIR_REQUIRE(!epi.bw_relu_out || (epi.mode == EPI_TRUNC && epi.nout == 6 && epi.bw_relu_d && !epi.relu_out),
                "limb_gemm: a fused ReLU backward needs the six RSS output arrays and the derivative");
    PIR_REQUIRE(epi.div_bits >= 0 && epi.div_bits <= 62, "limb_gemm: division bits must be in [0, 62]");
    if (a.M == 0 || N == 0) return OK;
    const int nks = a.nseg * a.Kp / LIMB_BK;
    const int split = limb_effective_split(nks, plan.split);
    const int per = (nks + split - 1) / split;
    const bool masked = epi.mode == EPI_TRUNC && epi.prg.on;
    PIR_REQUIRE((split == 1 && !masked) || workspace, "limb_gemm: split-K or a masked epilogue needs a workspace");
    PIR_REQUIRE(!plan.wide || can_be_wide(N), "limb_gemm: the 32x128 tile needs an even number of 64-column tiles");
    const int wide = plan.wide;
    dim3 grid(wide ? (a.M + 31) / 32 : (a.M + 63) / 64, wide ? (N + 127) / 128 : (N + 63) / 64, split);
    u64 *part = (split > 1 || masked) ? workspace : nullptr;
    int amode;
    switch (a.conv) {
        case LIMB_FPROP: amode = (a.Cin % LIMB_BK) == 0 ? 3 : 2; break;
        case LIMB_DGRAD: amode = (a.Cin % LIMB_BK) == 0 ? 5 : 4; break;
        case LIMB_WGRAD: amode = 6; break;
        default: amode = a.row_major ? 0 : 1; break;
    }
    const int variant = amode + (wide ? LIMB_VARIANT_WIDE : 0);
    Epilogue ek = epi;
    ek.pairs = (part ? N % 2 == 0 && ((uintptr_t)part % 16) == 0 : epilogue_pairs_ok(epi, N)) ? 1 : 0;
    GemmTimer &timer = ctx().timer;
    GemmTimer::Rec rec{nullptr, nullptr, 2.0 * 36 * a.M * (double)N * a.nseg * a.K, variant};
    if (timer.on) {
        PIR_HIP(hipEventCreate(&rec.start));
        PIR_HIP(hipEventCreate(&rec.stop));
        PIR_HIP(hipEventRecord(rec.start, s));
    }
#define PIR_LIMB_LAUNCH(AM, WM)                                                                                \
    hipLaunchKernelGGL((limb_gemm_kernel<AM, WM>), grid, dim3(256), 0, s, a, packedB, N, nks, per, ek, part)
    // three column blocks in the 32x128 tile: the fourth wave helps (fprop forms)
    const bool help = wide && (N + 31) / 32 == 3 && (amode == 2 || amode == 3);
    if (help && amode == 2)
        hipLaunchKernelGGL((limb_gemm_help_kernel<2>), grid, dim3(256), 0, s, a, packedB, N, nks, per, ek, part);
    else if (help)
        hipLaunchKernelGGL((limb_gemm_help_kernel<3>), grid, dim3(256), 0, s, a, packedB, N, nks, per, ek, part);
    else switch (variant) {
        case 0: PIR_LIMB_LAUNCH(0, 2); break;
        case 1: PIR_LIMB_LAUNCH(1, 2); break;
        case 2: PIR_LIMB_LAUNCH(2, 2); break;
        case 3: PIR_LIMB_LAUNCH(3, 2); break;
        case 4: PIR_LIMB_LAUNCH(4, 2); break;
        case 5: PIR_LIMB_LAUNCH(5, 2); break;
        case 6: PIR_LIMB_LAUNCH(6, 2); break;
        case LIMB_VARIANT_WIDE + 0: PIR_LIMB_LAUNCH(0, 1); break;
        case LIMB_VARIANT_WIDE + 1: PIR_LIMB_LAUNCH(1, 1); break;
        case LIMB_VARIANT_WIDE + 2: PIR_LIMB_LAUNCH(2, 1); break;
        case LIMB_VARIANT_WIDE + 3: PIR_LIMB_LAUNCH(3, 1); break;
        case LIMB_VARIANT_WIDE + 4: PIR_LIMB_LAUNCH(4, 1); break;
        case LIMB_VARIANT_WIDE + 5: PIR_LIMB_LAUNCH(5, 1); break;
        default: PIR_LIMB_LAUNCH(6, 1); break;
    }
#undef PIR_LIMB_LAUNCH
    PIR_LAUNCH_CHECK();
    if (timer.on) {
        PIR_HIP(hipEventRecord(rec.stop, s));
        timer.recs.push_back(rec);
    }
    if (split > 1 || masked) PIR_TRY(rss_trunc_epilogue(workspace, a.M, N, split, epi, s));
    return OK;
}

int limb_run(const LimbA &a, const int8_t *packedB, int N, const Epilogue &epi, int split_hint, hipStream_t s) {
    if (a.M == 0 || N == 0) return OK;
    Context &c = ctx();
    const std::string key = shape_key(a, N);
    LimbPlan dummy;
    if (c.tune && split_hint <= 0 && !c.alloc.capturing && !tuned_plan(key, &dummy)) {
        // candidates: both tile shapes, split factors around the heuristic
        const int nks = a.nseg * a.Kp / LIMB_BK;
        std::vector<LimbPlan> cand;
        for (int wide = can_be_wide(N) ? 1 : 0; wide >= 0; wide--) {
            const int h = heuristic_split(a.M, N, nks, wide != 0);
            for (int f : {1, h / 4, h / 2, h - 1, h, h + 1, 2 * h, 4 * h}) {
                if (f < 1 || f > LIMB_MAX_SPLIT) continue;
                LimbPlan p{wide, limb_effective_split(nks, f)};
                bool dup = false;
                for (auto &q : cand) dup |= (q.wide == p.wide && q.split == p.split);
                if (!dup) cand.push_back(p);
            }
        }
        hipEvent_t e0, e1;
        PIR_HIP(hipEventCreate(&e0));
        PIR_HIP(hipEventCreate(&e1));
        LimbPlan best = cand[0];
        float best_ms = 1e30f;
        int rc = OK;
        for (auto &p : cand) {
            DevBuf ws(limb_workspace_bytes(p, a.M, N, epi.mode == EPI_TRUNC && epi.prg.on));
            rc = limb_gemm(a, packedB, N, epi, p, (u64 *)ws.get(), s);  // warm-up
            float ms = 1e30f;
            for (int r = 0; r < 2 && rc == OK; r++) {
                if (hipEventRecord(e0, s) != hipSuccess) rc = ERR_HIP;
                if (rc == OK) rc = limb_gemm(a, packedB, N, epi, p, (u64 *)ws.get(), s);
                float t = 0;
                if (rc == OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                 hipEventElapsedTime(&t, e0, e1) != hipSuccess))
                    rc = ERR_HIP;
                if (rc == OK && t < ms) ms = t;
            }
            if (rc != OK) break;
            if (ms < best_ms) {
                best_ms = ms;
                best = p;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc != OK) {
            set_error("limb GEMM autotune: a candidate launch failed");
            return rc;
        }
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[key] = best;
    }
    LimbPlan p = limb_plan(a, N, split_hint);
    DevBuf ws(limb_workspace_bytes(p, a.M, N, epi.mode == EPI_TRUNC && epi.prg.on));
    return limb_gemm(a, packedB, N, epi, p, (u64 *)ws.get(), s);
}

int limb_ring_gemm(int M, int N, int K, const u64 *A, bool ta, const u64 *B, bool tb, u64 *C, bool tc,
                   hipStream_t s) {
    PIR_REQUIRE(M >= 0 && N >= 0 && K > 0, "limb gemm: invalid shape");
    LimbBSrc src{};
    src.b[0][0] = B;
    src.coef[0][0] = 1;
    src.nseg = 1;
    src.K = K;
    src.N = N;
    src.Kp = round_up(K, LIMB_BK);
    src.row_major = tb ? 1 : 0;
    DevBuf packed(limb_b_bytes(N, 1, src.Kp));
    PIR_TRY(limb_pack_b(src, (int8_t *)packed.get(), s));
    LimbA a{};
    a.seg[0] = A;
    a.nseg = 1;
    a.M = M;
    a.K = K;
    a.Kp = src.Kp;
    a.lda = ta ? K : M;
    a.row_major = ta ? 1 : 0;
    Epilogue e{};
    e.mode = EPI_RAW;
    e.C = C;
    e.tc = tc ? 1 : 0;
    return limb_run(a, (const int8_t *)packed.get(), N, e, 0, s);
}

int limb_ring_conv(const u64 *x, const u64 *w, u64 *y, int b, int H, int W, int Din, int Dout, int fh, int fw,
                   int ph, int pw, int stride, hipStream_t s) {
    PIR_REQUIRE(b > 0 && H > 0 && W > 0 && Din > 0 && Dout > 0 && fh > 0 && fw > 0 && stride > 0 && ph >= 0 &&
                    pw >= 0,
                "limb conv: invalid geometry");
    const int OH = (H + 2 * ph - fh) / stride + 1, OW = (W + 2 * pw - fw) / stride + 1;
    PIR_REQUIRE(H + 2 * ph >= fh && W + 2 * pw >= fw && OH > 0 && OW > 0, "limb conv: empty output");
    const int K = fh * fw * Din, M = b * OH * OW;
    LimbBSrc src{};
    src.b[0][0] = w;
    src.coef[0][0] = 1;
    src.nseg = 1;
    src.K = K;
    src.N = Dout;
    src.Kp = round_up(K, LIMB_BK);
    src.row_major = 0;  // KRSC weights: B(k, n) = w[n * K + k]
    DevBuf packed(limb_b_bytes(Dout, 1, src.Kp));
    PIR_TRY(limb_pack_b(src, (int8_t *)packed.get(), s));
    LimbA a{};
    a.seg[0] = x;
    a.nseg = 1;
    a.M = M;
    a.K = K;
    a.Kp = src.Kp;
    a.conv = LIMB_FPROP;
    a.row_major = 1;
    a.H = H;
    a.W = W;
    a.Cin = Din;
    a.fh = fh;
    a.fw = fw;
    a.ph = ph;
    a.pw = pw;
    a.stride = stride;
    a.OH = OH;
    a.OW = OW;
    Epilogue e{};
    e.mode = EPI_RAW;
    e.C = y;
    e.tc = 1;
    return limb_run(a, (const int8_t *)packed.get(), Dout, e, 0, s);
}

// gpu::conv_dgrad: the limb GEMM M = b*H*W, N = Din, K = fh*fw*Dout over the
// dgrad gather of dy and the filter read in (r, s, c)-major order.
int limb_ring_dgrad(const u64 *dy, const u64 *w, u64 *dx, int b, int OH, int OW, int Dout, int fh, int fw, int Din,
                    int ph, int pw, int stride, int H, int W, hipStream_t s) {
    PIR_REQUIRE(b > 0 && H > 0 && W > 0 && OH > 0 && OW > 0 && Din > 0 && Dout > 0 && fh > 0 && fw > 0 &&
                    stride > 0 && ph >= 0 && pw >= 0,
                "limb dgrad: invalid geometry");
    const int K = fh * fw * Dout;
    LimbBSrc src{};
    src.b[0][0] = w;
    src.coef[0][0] = 1;
    src.nseg = 1;
    src.K = K;
    src.N = Din;
    src.Kp = round_up(K, LIMB_BK);
    src.dgrad_filter = 1;
    src.kfh = fh;
    src.kfw = fw;
    src.kc = Dout;
    DevBuf packed(limb_b_bytes(Din, 1, src.Kp));
    PIR_TRY(limb_pack_b(src, (int8_t *)packed.get(), s));
    LimbA a{};
    a.seg[0] = dy;
    a.nseg = 1;
    a.M = b * H * W;
    a.K = K;
    a.Kp = src.Kp;
    a.conv = LIMB_DGRAD;
    a.H = H;
    a.W = W;
    a.Cin = Dout;
    a.fh = fh;
    a.fw = fw;
    a.ph = ph;
    a.pw = pw;
    a.stride = stride;
    a.OH = OH;
    a.OW = OW;
    Epilogue e{};
    e.mode = EPI_RAW;
    e.C = dx;
    e.tc = 1;
    return limb_run(a, (const int8_t *)packed.get(), Din, e, 0, s);
}

// gpu::conv_wgrad: the limb GEMM M = fh*fw*Din, N = Dout, K = b*OH*OW over the
// transposed im2col of x and dy as a row-major K x N operand; the M x N result
// is written column-major, which is the [Dout][fh][fw][Din] filter layout.
int limb_ring_wgrad(const u64 *dy, const u64 *x, u64 *dw, int b, int OH, int OW, int Dout, int H, int W, int Din,
                    int fh, int fw, int ph, int pw, int stride, hipStream_t s) {
    PIR_REQUIRE(b > 0 && H > 0 && W > 0 && OH > 0 && OW > 0 && Din > 0 && Dout > 0 && fh > 0 && fw > 0 &&
                    stride > 0 && ph >= 0 && pw >= 0,
                "limb wgrad: invalid geometry");
    const int K = b * OH * OW;
    LimbBSrc src{};
    src.b[0][0] = dy;
    src.coef[0][0] = 1;
    src.nseg = 1;
    src.K = K;
    src.N = Dout;
    src.Kp = round_up(K, LIMB_BK);
    src.row_major = 1;
    DevBuf packed(limb_b_bytes(Dout, 1, src.Kp));
    PIR_TRY(limb_pack_b(src, (int8_t *)packed.get(), s));
    LimbA a{};
    a.seg[0] = x;
    a.nseg = 1;
    a.M = fh * fw * Din;
    a.K = K;
    a.Kp = src.Kp;
    a.conv = LIMB_WGRAD;
    a.H = H;
    a.W = W;
    a.Cin = Din;
    a.fh = fh;
    a.fw = fw;
    a.ph = ph;
    a.pw = pw;
    a.stride = stride;
    a.OH = OH;
    a.OW = OW;
    Epilogue e{};
    e.mode = EPI_RAW;
    e.C = dw;
    e.tc = 0;
    return limb_run(a, (const int8_t *)packed.get(), Dout, e, 0, s);
}

// ---------------------------------------------------------------------------
// Plain 64-bit VALU GEMM: 64x64 tile, K-step 16, 4x4 outputs per thread.
// blockIdx.z selects a K-chunk of `kchunk` k; with more than one chunk every
// chunk writes its own M x N partial slab (C + z * M * N), summed afterwards.

struct PlainConv {
    int H, W, Cin, fh, fw, ph, pw, stride, dil, OH, OW;
};

// AMODE 0: A column-major, 1: row-major, 2: fprop im2col, 3: dgrad gather
// (m = (b, h, w), k = (r, s, c) over dy [b][OH][OW][Cin]), 4: wgrad transposed
// im2col (m = (r, s, c), k = (b, p, q)).  BMODE 0: B(k, n) = B[n*K + k],
// 1: B[k*N + n], 2: the dgrad filter [c][fh][fw][N] at k = (r*fw + s)*Cin + c.
template <int AMODE, int BMODE, bool TC>
__global__ __launch_bounds__(256) void plain_gemm_kernel(int M, int N, int K, int kchunk, const u64 *__restrict__ A,
                                                         const u64 *__restrict__ B, u64 *__restrict__ C,
                                                         PlainConv cv) {
    __shared__ u64 As[16][64 + 1];
    __shared__ u64 Bs[16][64 + 1];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    const int kbeg = blockIdx.z * kchunk, kend = min(K, kbeg + kchunk);
    C += (size_t)blockIdx.z * M * N;
    u64 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0;

    for (int k0 = kbeg; k0 < kend; k0 += 16) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            int idx = tid + e * 256;
            int mm, kk;
            if (AMODE == 0 || AMODE == 4) {
                kk = idx >> 6;
                mm = idx & 63;
            } else {
                mm = idx >> 4;
                kk = idx & 15;
            }
            int m = m0 + mm, k = k0 + kk;
            u64 v = 0;
            if (m < M && k < kend) {
                if (AMODE == 0) v = A[(size_t)k * M + m];
                else if (AMODE == 1) v = A[(size_t)m * K + k];
                else if (AMODE == 2) {
                    int ohw = cv.OH * cv.OW;
                    int b = m / ohw, rem = m - b * ohw;
                    int oh = rem / cv.OW, ow = rem - oh * cv.OW;
                    int fwc = cv.fw * cv.Cin;
                    int r = k / fwc, rm = k - r * fwc;
                    int s = rm / cv.Cin, c = rm - s * cv.Cin;
                    int ih = oh * cv.stride - cv.ph + r * cv.dil, iw = ow * cv.stride - cv.pw + s * cv.dil;
                    if (ih >= 0 && ih < cv.H && iw >= 0 && iw < cv.W)
                        v = A[(((size_t)b * cv.H + ih) * cv.W + iw) * cv.Cin + c];
                } else if (AMODE == 3) {
                    int hw = cv.H * cv.W;
                    int b = m / hw, rem = m - b * hw;
                    int h = rem / cv.W, w = rem - h * cv.W;
                    int fwc = cv.fw * cv.Cin;
                    int r = k / fwc, rm = k - r * fwc;
                    int s = rm / cv.Cin, c = rm - s * cv.Cin;
                    int hp = h + cv.ph - r * cv.dil, wq = w + cv.pw - s * cv.dil;
                    if (hp >= 0 && wq >= 0 && hp % cv.stride == 0 && wq % cv.stride == 0) {
                        int p = hp / cv.stride, q = wq / cv.stride;
                        if (p < cv.OH && q < cv.OW) v = A[(((size_t)b * cv.OH + p) * cv.OW + q) * cv.Cin + c];
                    }
                } else {
                    int fwc = cv.fw * cv.Cin;
                    int r = m / fwc, rm = m - r * fwc;
                    int s = rm / cv.Cin, c = rm - s * cv.Cin;
                    int ohw = cv.OH * cv.OW;
                    int b = k / ohw, rem = k - b * ohw;
                    int p = rem / cv.OW, q = rem - p * cv.OW;
                    int ih = p * cv.stride - cv.ph + r * cv.dil, iw = q * cv.stride - cv.pw + s * cv.dil;
                    if (ih >= 0 && ih < cv.H && iw >= 0 && iw < cv.W)
                        v = A[(((size_t)b * cv.H + ih) * cv.W + iw) * cv.Cin + c];
                }
            }
            As[kk][mm] = v;
            int nn;
            if (BMODE != 0) {
                kk = idx >> 6;
                nn = idx & 63;
            } else {
                nn = idx >> 4;
                kk = idx & 15;
            }
            int n = n0 + nn;
            k = k0 + kk;
            u64 bvv = 0;
            if (n < N && k < kend) {
                if (BMODE == 0) bvv = B[(size_t)n * K + k];
                else if (BMODE == 1) bvv = B[(size_t)k * N + n];
                else {
                    int fwc = cv.fw * cv.Cin;
                    int r = k / fwc, rm = k - r * fwc;
                    int s = rm / cv.Cin, c = rm - s * cv.Cin;
                    bvv = B[(((size_t)c * cv.fh + r) * cv.fw + s) * N + n];
                }
            }
            Bs[kk][nn] = bvv;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
            u64 av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) av[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; j++) bv[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] += av[i] * bv[j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int m = m0 + ty + 16 * i, n = n0 + tx + 16 * j;
            if (m < M && n < N) {
                if (TC) C[(size_t)m * N + n] = acc[i][j];
                else C[(size_t)n * M + m] = acc[i][j];
            }
        }
}

__global__ __launch_bounds__(256) void plain_split_sum_kernel(const u64 *__restrict__ part, int split, size_t MN,
                                                              u64 *__restrict__ C) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= MN) return;
    u64 v = 0;
    for (int z = 0; z < split; z++) v += part[(size_t)z * MN + t];
    C[t] = v;
}

template <int AMODE, int BMODE>
static void launch_plain(dim3 g, hipStream_t s, int M, int N, int K, int kchunk, const u64 *A, const u64 *B, u64 *C,
                         bool tc, PlainConv cv) {
    if (tc) hipLaunchKernelGGL((plain_gemm_kernel<AMODE, BMODE, true>), g, dim3(256), 0, s, M, N, K, kchunk, A, B, C, cv);
    else hipLaunchKernelGGL((plain_gemm_kernel<AMODE, BMODE, false>), g, dim3(256), 0, s, M, N, K, kchunk, A, B, C, cv);
}

// K-chunking for small outputs with long K (wgrad): enough workgroups to cover
// the chip, each chunk a multiple of the K-step.
static int plain_split(int M, int N, int K) {
    long tiles = (long)((M + 63) / 64) * ((N + 63) / 64);
    if (tiles >= 256 || K < 1024) return 1;
    long want = (512 + tiles - 1) / tiles;
    long cap = K / 256;
    if (want > cap) want = cap;
    return want < 1 ? 1 : (int)want;
}

template <int AMODE, int BMODE>
static int run_plain(int M, int N, int K, const u64 *A, const u64 *B, u64 *C, bool tc, PlainConv cv, hipStream_t s) {
    if (M == 0 || N == 0) return OK;
    int split = plain_split(M, N, K);
    int kchunk = round_up((K + split - 1) / split, 16);
    if (kchunk < 16) kchunk = 16;
    split = (K + kchunk - 1) / kchunk;
    if (split < 1) split = 1;
    dim3 g((M + 63) / 64, (N + 63) / 64, split);
    if (split == 1) {
        launch_plain<AMODE, BMODE>(g, s, M, N, K, K > 0 ? K : 16, A, B, C, tc, cv);
        PIR_LAUNCH_CHECK();
        return OK;
    }
    size_t MN = (size_t)M * N;
    DevBuf part(MN * split * sizeof(u64));
    launch_plain<AMODE, BMODE>(g, s, M, N, K, kchunk, A, B, (u64 *)part.get(), tc, cv);
    PIR_LAUNCH_CHECK();
    hipLaunchKernelGGL(plain_split_sum_kernel, dim3(grid_for(MN, 256)), dim3(256), 0, s, (const u64 *)part.get(),
                       split, MN, C);
    PIR_LAUNCH_CHECK();
    return OK;
}

int plain_gemm(int M, int N, int K, const u64 *A, bool ta, const u64 *B, bool tb, u64 *C, bool tc, hipStream_t s) {
    PIR_REQUIRE(M >= 0 && N >= 0 && K >= 0, "gemm: negative dimension");
    PlainConv cv{};
    if (ta) return tb ? run_plain<1, 1>(M, N, K, A, B, C, tc, cv, s) : run_plain<1, 0>(M, N, K, A, B, C, tc, cv, s);
    return tb ? run_plain<0, 1>(M, N, K, A, B, C, tc, cv, s) : run_plain<0, 0>(M, N, K, A, B, C, tc, cv, s);
}

int plain_conv_fprop(const u64 *x, const u64 *w, u64 *y, int b, int H, int W, int Din, int Dout, int fh, int fw,
                     int ph, int pw, int stride, int dil, hipStream_t s) {
    PIR_REQUIRE(b > 0 && H > 0 && W > 0 && Din > 0 && Dout > 0 && fh > 0 && fw > 0 && stride > 0 && dil > 0,
                "conv_fprop: invalid geometry");
    PlainConv cv{H, W, Din, fh, fw, ph, pw, stride, dil, 0, 0};
    cv.OH = (H + 2 * ph - ((fh - 1) * dil + 1)) / stride + 1;
    cv.OW = (W + 2 * pw - ((fw - 1) * dil + 1)) / stride + 1;
    PIR_REQUIRE(cv.OH > 0 && cv.OW > 0, "conv_fprop: empty output");
    // filter [Dout][fh][fw][Din] is B(k, n) = w[n*K + k]: column-major K x N
    return run_plain<2, 0>(b * cv.OH * cv.OW, Dout, fh * fw * Din, x, w, y, true, cv, s);
}

// gpu::conv_dgrad (conv.cuh:405-446): dx = sum over taps of dy . w, as the GEMM
// M = b*H*W, N = Din, K = fh*fw*Dout with the dgrad gather of dy.
int plain_conv_dgrad(const u64 *dy, const u64 *w, u64 *dx, int b, int OH, int OW, int Dout, int fh, int fw, int Din,
                     int ph, int pw, int stride, int dil, int H, int W, hipStream_t s) {
    PIR_REQUIRE(b > 0 && H > 0 && W > 0 && OH > 0 && OW > 0 && Din > 0 && Dout > 0 && fh > 0 && fw > 0 &&
                    stride > 0 && dil > 0 && ph >= 0 && pw >= 0,
                "conv_dgrad: invalid geometry");
    PlainConv cv{H, W, Dout, fh, fw, ph, pw, stride, dil, OH, OW};
    return run_plain<3, 2>(b * H * W, Din, fh * fw * Dout, dy, w, dx, true, cv, s);
}

// gpu::conv_wgrad (conv.cuh:448-490): dw^T = im2col(x)^T . dy, the GEMM
// M = fh*fw*Din, N = Dout, K = b*OH*OW, written column-major into [Dout][fh][fw][Din].
int plain_conv_wgrad(const u64 *dy, const u64 *x, u64 *dw, int b, int OH, int OW, int Dout, int H, int W, int Din,
                     int fh, int fw, int ph, int pw, int stride, int dil, hipStream_t s) {
    PIR_REQUIRE(b > 0 && H > 0 && W > 0 && OH > 0 && OW > 0 && Din > 0 && Dout > 0 && fh > 0 && fw > 0 &&
                    stride > 0 && dil > 0 && ph >= 0 && pw >= 0,
                "conv_wgrad: invalid geometry");
    PlainConv cv{H, W, Din, fh, fw, ph, pw, stride, dil, OH, OW};
    return run_plain<4, 1>(fh * fw * Din, Dout, b * OH * OW, x, dy, dw, false, cv, s);
}

}  // namespace pir
