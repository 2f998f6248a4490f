// ring_gemm.h -- Z_2^64 ring GEMM / implicit-GEMM convolution on gfx950.
//
// Two exact forms of C = A.B mod 2^64:
//   * plain:  64-bit VALU multiply-add, LDS-tiled (the reference's CUTLASS
//     SIMT uint64 GEMM restated for CDNA4 wave64);
//   * limb:   every 64-bit operand is recoded into 8 signed 8-bit digits
//     a = sum_i d_i 2^(8i), d_i in [-128, 127] (digits are the bytes of
//     (a + 0x8080..80) ^ 0x8080..80), and C = sum_{i+j<8} 2^(8(i+j)) A_i.B_j
//     is evaluated as 36 int8 MFMA products with int32 accumulation, grouped
//     by s = i+j into 8 accumulators, flushed into 64-bit sums every 8192 k
//     (|partial| <= 8192 * 8 * 2^14 = 2^30, so no int32 overflow can occur).
#pragma once

#include "common.h"
#include "prg.h"

namespace pir {

constexpr int LIMB_MAX_SEG = 12;  // K-segments per product (9 for the 4-party matmul)
constexpr int LIMB_MAX_TERMS = 4;  // share arrays summed into one packed B segment

// Operand A description for the limb GEMM: up to LIMB_MAX_SEG K-segments (one
// per share array), each either a dense matrix or an implicit-GEMM view of an
// NHWC tensor.
enum LimbConv {
    LIMB_DENSE = 0,
    // fprop im2col of x [b][H][W][Cin]: m = (b, oh, ow), k = (r*fw + s)*Cin + c
    LIMB_FPROP = 1,
    // dgrad gather of dy [b][OH][OW][Cin] (Cin = the layer's output channels):
    // m = (b, h, w) of the H x W input image, k = (r*fw + s)*Cin + c,
    // A(m, k) = dy[b][p][q][c] where h + ph - r = p*stride and w + pw - s = q*stride
    LIMB_DGRAD = 2,
    // wgrad transposed im2col of x [b][H][W][Cin]: m = (r*fw + s)*Cin + c,
    // k = (b, p, q) over the OH x OW output, A(m, k) = x[b][p*stride-ph+r][q*stride-pw+s][c]
    LIMB_WGRAD = 3,
};

struct LimbA {
    const u64 *seg[LIMB_MAX_SEG];
    int nseg;
    int M, K;        // per-segment logical shape
    int Kp;          // K rounded up to the K-step (32)
    int lda;         // dense: leading dimension
    int row_major;   // dense: A(m,k) = row_major ? p[m*lda+k] : p[k*lda+m]
    int conv;        // LimbConv
    int H, W, Cin, fh, fw, ph, pw, stride, OH, OW;
    // divisors of the per-element index decode, set by limb_gemm
    FastDiv fd_fwc, fd_cin, fd_ohw, fd_ow;
};

// Operand B (weights) description for the packing kernel: segment s is the
// mod-2^64 combination sum_t coef[s][t] * b[s][t] of up to LIMB_MAX_TERMS share
// arrays (null pointers are skipped); B(k,n) = row_major ? p[k*N+n] : p[n*K+k],
// or, with dgrad_filter set, the filter [kc][kfh][kfw][N] read as
// B((r*kfw + s)*kc + c, n) = p[((c*kfh + r)*kfw + s)*N + n].
struct LimbBSrc {
    const u64 *b[LIMB_MAX_SEG][LIMB_MAX_TERMS];
    int coef[LIMB_MAX_SEG][LIMB_MAX_TERMS];
    int nseg;
    int K, N, Kp;
    int row_major;
    int dgrad_filter, kfh, kfw, kc;
};

// Epilogue modes.
enum EpiMode { EPI_RAW = 0, EPI_TRUNC = 1 };

struct Epilogue {
    int mode;
    // C(m,n) / out(m,n) stored row-major (ldc = N) if tc, else column-major (ldc = M)
    u64 *C;
    int tc;
    // EPI_TRUNC: out holds `nout` share arrays of M*N; array j receives
    // (bit j of base_mask ? 1 + q : 0) + (bias ? bias[j*N + n] : 0) with
    // q = trunc((raw - 2^k) / 2^k), the dividePublic of the placeholder
    // precompute (r = 1, r' = 2^k): RSS.inl:1076-1094 (arrays A.sA, C.sB),
    // TPC.inl:370-384 (party A), FPC.inl:566-580 (one party's share 0/1/2).
    u64 *out;
    int nout;
    unsigned base_mask;
    int trunc_bits;
    const u64 *bias;  // [nout][N] or null
    // div_bits > 0: the truncated value v = 1 + q is then divided by
    // 2^div_bits as dividePublic does on those shares (RSS.inl:422-436):
    // v' = 1 + trunc((v - 2^div_bits) / 2^div_bits).
    int div_bits;
    // sub_out: subtract the result shares from `out` instead of storing them
    // (a weight update W -= dW in the GEMM that produces dW); bias must be null.
    int sub_out;
    // Optional ReLU of the output (ReLU layer fused into its producer; the six
    // RSS arrays only, nout == 6): relu_out [3][2][M*N] = x * [x >= 0] and
    // relu_d [3][2][M*N] = dReLU bits (RSS.inl:806-865), both in out's layout.
    u64 *relu_out;
    u8 *relu_d;
    // Optional backward of the ReLU layer that precedes this product's layer
    // (nout == 6): bw_relu_out [3][2][M*N] = selectShare(0, out, bw_relu_d)
    // with bw_relu_d that ReLU layer's reluPrime (ReLULayer.cu:91-123).
    const u8 *bw_relu_d;
    u64 *bw_relu_out;
    // Masked precompute mode (EPI_TRUNC): prg.on draws the truncation pair of
    // output element idx (its index in out's layout) from prg, and the fused
    // ReLU from relu_prg; the product is then written raw and the truncation
    // runs in the reduction kernel (limb_gemm).  No div_bits, sub_out or
    // fused ReLU backward in this mode.
    PrgStream prg;
    PrgStream relu_prg;
    // Set by limb_gemm, not by callers: the limb kernel may store two adjacent
    // columns per lane (row-major output or split-K partials, N even, the
    // 64-bit arrays 16-byte and the bit arrays 2-byte aligned, no sub_out).
    int pairs;
};

constexpr int LIMB_BM = 64, LIMB_BN = 64, LIMB_BK = 32;
constexpr int LIMB_MAX_SPLIT = 512;     // split-K factors the planner considers
constexpr int LIMB_VARIANT_WIDE = 8;    // kernel variant = A-operand form (0..6) + 8 for the 32x128 tile
constexpr int LIMB_VARIANTS = 16;
constexpr int LIMB_FLUSH_STEPS = 256;  // 256 * 32 = 8192 k per int32 accumulation chunk

inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Split-K factor actually used by limb_gemm for `nks` K-steps: at least
// ceil(nks / LIMB_FLUSH_STEPS) (each workgroup covers <= 8192 k), at most nks,
// and rounded so that every split is non-empty.  Callers size the workspace
// (split * M * N * 8 bytes when > 1) with this value.
inline int limb_effective_split(int nks, int split) {
    int lo = (nks + LIMB_FLUSH_STEPS - 1) / LIMB_FLUSH_STEPS;
    if (split < lo) split = lo;
    if (split > nks) split = nks;
    if (split < 1) split = 1;
    int per = (nks + split - 1) / split;
    return (nks + per - 1) / per;
}

// Size in bytes of the packed B operand.
inline size_t limb_b_bytes(int N, int nseg, int Kp) {
    size_t nt = (size_t)(N + LIMB_BN - 1) / LIMB_BN;
    size_t ks = (size_t)nseg * Kp / LIMB_BK;
    return nt * ks * 8 * LIMB_BN * LIMB_BK;
}

int limb_pack_b(const LimbBSrc &src, int8_t *packed, hipStream_t s);

// Launch plan of one limb GEMM: workgroup tile (wide: 32x128, only when the
// packed 64-column B tiles pair up; else 64x64) and split-K factor.
struct LimbPlan {
    int wide;
    int split;
};

// The plan for this shape: the autotuned entry when one exists, else the
// heuristic (split_hint > 0 forces that split factor).
LimbPlan limb_plan(const LimbA &a, int N, int split_hint = 0);
// Workspace bytes a plan needs (split * M * N * 8 when split > 1, or when the
// epilogue is masked, which always reduces from the workspace).
size_t limb_workspace_bytes(const LimbPlan &p, int M, int N, bool masked = false);

// Runs the limb GEMM with `plan`.  plan.split > 1 partitions the K-steps
// across workgroups and reduces the 64-bit partials (exact: addition mod 2^64
// commutes) from `workspace`.
int limb_gemm(const LimbA &a, const int8_t *packedB, int N, const Epilogue &epi, const LimbPlan &plan,
              u64 *workspace, hipStream_t s);

// Plan + workspace (context allocator) + launch.  With autotuning switched on
// in the context (and no graph capture open), an untuned shape first times
// every candidate plan on the real operands (each run overwrites the same
// output) and records the fastest in the process-wide table.
int limb_run(const LimbA &a, const int8_t *packedB, int N, const Epilogue &epi, int split_hint, hipStream_t s);

// gpu::gemm (gemm.cuh:64-112) and gpu::conv_fprop (conv.cuh:380-401) in the
// limb form: B packed into a temporary, plan, launch with a raw epilogue.
// Transpose flags as in gpu::gemm (set = row-major storage).
int limb_ring_gemm(int M, int N, int K, const u64 *A, bool ta, const u64 *B, bool tb, u64 *C, bool tc,
                   hipStream_t s);
int limb_ring_conv(const u64 *x, const u64 *w, u64 *y, int b, int H, int W, int Din, int Dout, int fh, int fw,
                   int ph, int pw, int stride, hipStream_t s);
// gpu::conv_dgrad / conv_wgrad (conv.cuh:405-490) in the limb form (dil = 1).
int limb_ring_dgrad(const u64 *dy, const u64 *w, u64 *dx, int b, int OH, int OW, int Dout, int fh, int fw, int Din,
                    int ph, int pw, int stride, int H, int W, hipStream_t s);
int limb_ring_wgrad(const u64 *dy, const u64 *x, u64 *dw, int b, int OH, int OW, int Dout, int H, int W, int Din,
                    int fh, int fw, int ph, int pw, int stride, hipStream_t s);

// Process-wide autotune table.
int limb_tune_entries();
void limb_tune_clear();
int limb_tune_save(const char *path);
int limb_tune_load(const char *path);

// Plain VALU forms with the reference's layout conventions (gemm.cuh:64-112,
// conv.cuh:359-401).
int plain_gemm(int M, int N, int K, const u64 *A, bool ta, const u64 *B, bool tb, u64 *C, bool tc, hipStream_t s);
int plain_conv_fprop(const u64 *x, const u64 *w, u64 *y, int b, int H, int W, int Din, int Dout, int fh, int fw,
                     int ph, int pw, int stride, int dil, hipStream_t s);
int plain_conv_dgrad(const u64 *dy, const u64 *w, u64 *dx, int b, int OH, int OW, int Dout, int fh, int fw, int Din,
                     int ph, int pw, int stride, int dil, int H, int W, hipStream_t s);
int plain_conv_wgrad(const u64 *dy, const u64 *x, u64 *dw, int b, int OH, int OW, int Dout, int H, int W, int Din,
                     int fh, int fw, int ph, int pw, int stride, int dil, hipStream_t s);

// Standalone truncation epilogue over a raw (3-of-3 summed) product, used by
// the split-K path and by callers that already hold Sum_p raw_p.
int rss_trunc_epilogue(const u64 *raw, int M, int N, int nsum, const Epilogue &epi, hipStream_t s);

}  // namespace pir
