// tpc_fpc.h -- matmul, convolution (kFprop), reconstruct and dividePublic of
// the reference's 2-party (TPC) and 4-party (FPC) schemes, all parties at once.
//
// Layouts (device memory, uint64):
//   TPC value  [2 parties][n]           one share per party (TPC.h:15-60);
//   FPC value  [4 parties][3 shares][n] party q's share s holds component
//              (q + s + 1) % 4 of x = x0 + x1 + x2 + x3 (FPC.inl:119-124,497-506).
//
// Each function runs in the context's execution mode: MODE_PLAIN restates the
// reference's op-by-op sequence (one launch per DeviceData op, one GEMM per
// product and party); MODE_FUSED with the limb GEMM evaluates each party's
// opened 3-of-3 (or 2-of-2) sum as one K-concatenated limb product with the
// dividePublic truncation in its epilogue.  Both give bit-identical shares.
#pragma once

#include "protocol.h"

namespace pir {

// reconstruct, TPC.inl:551-572: out [2][n], out_q = x_q + x_{other(q)}
void tpc_reconstruct(const u64 *x, size_t n, u64 *out);
// dividePublic(TPC, T), TPC.inl:370-384 (placeholder r = 1, r' = d)
void tpc_divide_public(u64 *a, size_t n, u64 d);
// matmul, TPC.inl:574-583 with localMatMul :1059-1090; a [2][M*K], b [2][K*N], c [2][M*N]
void tpc_matmul(const u64 *a, const u64 *b, u64 *c, int M, int N, int K, bool ta, bool tb, bool tc, int trunc);
// convolution(kFprop), TPC.inl:644-690, 788-826 (NHWC x KRSC -> NHWC)
void tpc_convolution(const u64 *A, const u64 *B, u64 *C, int batch, int H, int W, int f, int Din, int Dout,
                     int stride, int pad, int trunc);

// reconstruct, FPC.inl:582-604: out [4][n] = own three shares + previous party's share 0
void fpc_reconstruct(const u64 *x, size_t n, u64 *out);
// dividePublic(FPC, T), FPC.inl:566-580
void fpc_divide_public(u64 *a, size_t n, u64 d);
// reshareFPC, FPC.inl:1183-1260 (zero PRGs): z, zp [4][n] per party, added into out [4][3][n]
void fpc_reshare(const u64 *z, const u64 *zp, size_t n, int i, int j, u64 *out);
// matmul, FPC.inl:635-666; a [4][3][M*K], b [4][3][K*N], c [4][3][M*N]
void fpc_matmul(const u64 *a, const u64 *b, u64 *c, int M, int N, int K, bool ta, bool tb, bool tc, int trunc);
// convolution(kFprop), FPC.inl:744-782, 861-927
void fpc_convolution(const u64 *A, const u64 *B, u64 *C, int batch, int H, int W, int f, int Din, int Dout,
                     int stride, int pad, int trunc);

}  // namespace pir
