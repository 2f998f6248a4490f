/*
 * pir.h -- C ABI of the MI355X-native secret-share compute path
 * (libpir_hip.so).  Plain pointers and sizes only; every call returns 0 on
 * success or a nonzero status, with a message available from pir_last_error().
 * All device pointers are HIP device memory; `stream` is a hipStream_t (NULL =
 * the default stream).  Calls are asynchronous on `stream` unless stated.
 *
 * Share layouts
 *   RSS tensor (all three parties): [3 parties][2 shares][n] uint64, party p
 *   holds (shareA, shareB) = (x_p, x_{p+1}) as in src/mpc/RSS.h:97-100.
 *   Party-local tensor (one DeviceData per party): [3][n].
 *   Bit / derivative shares use uint8 (the layers' Share<uint8_t>) or uint64.
 *
 * The reference interface each entry point replaces is cited as file:line of
 * the reference repository (ucbrise/piranha).
 */
#ifndef PIR_H
#define PIR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status / plumbing ------------------------------------------------- */
int pir_version(void);
const char *pir_last_error(void);
int pir_malloc(void **ptr, size_t bytes);                   /* thrust::device_vector storage, DeviceData.h:231-277 */
int pir_free(void *ptr);
int pir_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream); /* synchronous w.r.t. the host */
int pir_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream); /* synchronous w.r.t. the host */
int pir_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
int pir_fill_u64(uint64_t *dst, uint64_t value, size_t n, void *stream); /* DeviceData::fill, DeviceData.h:83-85 */
int pir_fill_u8(uint8_t *dst, uint8_t value, size_t n, void *stream);
/* Streams and threads: a call queues its device work on the stream it is
 * given (NULL: the null stream).  Temporary device buffers are cached per stream and
 * reused only on the same stream, so successive calls may alternate streams
 * without synchronising.  The free-standing calls keep their state (and their
 * cache of temporaries) per calling thread; a pir_net belongs to one thread at
 * a time and keeps its own cache until pir_net_destroy.  pir_set_mode is
 * process-wide.  pir_stream_destroy synchronises the stream and returns the
 * calling thread's cached temporaries of that stream to HIP. */
int pir_stream_create(void **stream);
int pir_stream_destroy(void *stream);
/* No counterpart in the reference (thrust frees every temporary at once):
 * synchronise the device and return all of the calling thread's cached
 * temporaries to HIP; the bytes the calling thread's cache holds. */
int pir_release_cached(void);
size_t pir_cached_bytes(void);
int pir_stream_synchronize(void *stream);
int pir_device_synchronize(void);

/* ---- ring GEMM / convolution (gpu:: layer) ------------------------------- */
/* algo: 0 = plain 64-bit VALU, 1 = int8-limb MFMA (exact, same result). */
enum { PIR_ALGO_PLAIN = 0, PIR_ALGO_LIMB = 1 };

/* gpu::gemm, src/gpu/gemm.cuh:64-112: C = A.B mod 2^64, transpose_x != 0 means
 * row-major storage of x (lda = K, ldb = N, ldc = N), else column-major. */
int pir_gemm(int M, int N, int K, const uint64_t *A, int transpose_a, const uint64_t *B, int transpose_b,
             uint64_t *C, int transpose_c, int algo, void *stream);

/* gpu::conv_fprop, src/gpu/conv.cuh:380-401: NHWC x[b][H][W][Din] (*) KRSC
 * w[Dout][fh][fw][Din] -> NHWC y[b][OH][OW][Dout], cross-correlation. */
int pir_conv_fprop(const uint64_t *x, const uint64_t *w, uint64_t *y, int b, int H, int W, int Din, int Dout,
                   int fh, int fw, int ph, int pw, int stride, int dilation, int algo, void *stream);

/* gpu::conv_dgrad, src/gpu/conv.cuh:405-446: dy[b][OH][OW][Dout] and
 * w[Dout][fh][fw][Din] -> dx[b][H][W][Din] = the gradient of conv_fprop's input.
 * The limb form needs dilation 1. */
int pir_conv_dgrad(const uint64_t *dy, const uint64_t *w, uint64_t *dx, int b, int OH, int OW, int Dout, int fh,
                   int fw, int Din, int ph, int pw, int stride, int dilation, int H, int W, int algo, void *stream);

/* gpu::conv_wgrad, src/gpu/conv.cuh:448-490: dy[b][OH][OW][Dout] and
 * x[b][H][W][Din] -> dw[Dout][fh][fw][Din] = the gradient of conv_fprop's filter. */
int pir_conv_wgrad(const uint64_t *dy, const uint64_t *x, uint64_t *dw, int b, int OH, int OW, int Dout, int H,
                   int W, int Din, int fh, int fw, int ph, int pw, int stride, int dilation, int algo, void *stream);

/* The RSS local product of one party, localMatMul (src/mpc/RSS.inl:1096-1110):
 * C = A0.B0 + A0.B1 + A1.B0 mod 2^64, all row-major (A: MxK, B: KxN, C: MxN).
 * algo 0 runs the reference's form (three plain GEMMs, one launch each, plus
 * the two accumulations); algo 1 the K-concatenated limb-MFMA form
 * A0.(B0+B1) + A1.B0 (equal mod 2^64), including the operand packing. */
int pir_ring_gemm3(int algo, int M, int N, int K, const uint64_t *A0, const uint64_t *A1, const uint64_t *B0,
                   const uint64_t *B1, uint64_t *C, void *stream);

/* No counterpart in the reference: pir_ring_gemm3's limb form (algo 1) with
 * nothing allocated, for callers that cannot allocate device memory (code
 * under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The packed operand and the split-K partials live in a fixed
 * device array of PIR_NOALLOC_SCRATCH_BYTES on each GPU, held from the
 * library's load on that GPU until the process ends; a shape that needs more
 * (about 16 * K * N bytes, plus 8 * M * N per extra K split) fails with status
 * 1 and a message.  Every call on one GPU shares that array, so calls that
 * overlap on one GPU (other streams or threads) race: the caller orders them. */
#define PIR_NOALLOC_SCRATCH_BYTES (128u << 20)
int pir_ring_gemm3_noalloc(int M, int N, int K, const uint64_t *A0, const uint64_t *A1, const uint64_t *B0,
                           const uint64_t *B1, uint64_t *C, void *stream);

/* The limb form with explicit operands, no counterpart in the reference (it is
 * how the CUTLASS gemm above is re-expressed): C = sum_s A_s . B_s over nseg
 * (1..12) K-segments, B_s(k,n) = b0[s] + (b1 ? b1[s] : 0) packed once (weights) by
 * pir_limb_pack into pir_limb_packed_bytes() bytes.  A_s is M x K, row-major if
 * a_row_major; C is row-major if tc.  split <= 0 picks a split-K factor. */
size_t pir_limb_packed_bytes(int K, int N, int nseg);
int pir_limb_pack(const uint64_t *const *b0, const uint64_t *const *b1, int nseg, int K, int N, int b_row_major,
                  int8_t *packed, void *stream);
int pir_limb_gemm(const uint64_t *const *a, int nseg, int M, int K, int a_row_major, const int8_t *packed, int N,
                  uint64_t *C, int tc, int split, void *stream);

/* ---- bit / layout kernels (src/gpu/bitwise.cuh, matrix.cuh, convolution.cuh) */
int pir_bitexpand(const uint64_t *a, size_t n, void *bits, int bits_elem_bytes, void *stream); /* bitwise.cuh:98-124 */
int pir_set_carry_out_msb(void *rbits, void *abits, void *msb, size_t n, int bit_width, int xor_msb,
                          int elem_bytes, void *stream);                                   /* bitwise.cuh:184-213 */
int pir_expand_compare(const void *b, const void *inv_b, void *expanded, size_t n, size_t expanded_n,
                       int elem_bytes, void *stream);                                      /* bitwise.cuh:215-246 */
int pir_element_vector_add(uint64_t *a, const uint64_t *b, int rowwise, int rows, int cols, void *stream); /* matrix.cuh:253-277 */
int pir_truncate_cols(const void *a, size_t a_n, void *b, int start_cols, int end_cols, int elem_bytes,
                      void *stream);                                                       /* matrix.cuh:382-400 */
int pir_stride_pad(const uint64_t *in, uint64_t *out, size_t n_in, int stride, int pad, uint64_t pad_value,
                   void *stream);                                                          /* convolution.cuh:306-337 */
/* maxpool_im2row, convolution.cuh:252-304; out must hold pir_maxpool_im2row_size() values */
size_t pir_maxpool_im2row_size(int H, int W, int f, int Din, int batch, int stride, int padding);
int pir_maxpool_im2row(const uint64_t *im, uint64_t *out, int H, int W, int f, int Din, int batch, int stride,
                       int padding, uint64_t pad_value, void *stream);
/* gpu::averagepool_im2row, convolution.cuh:339-371 (the reference host's
 * argument order, width before height); out must hold
 * pir_averagepool_im2row_size() values: f*f per (image, window, channel) */
size_t pir_averagepool_im2row_size(int image_width, int image_height, int f, int Din, int batch, int stride,
                                   int padding);
int pir_averagepool_im2row(const uint64_t *im, uint64_t *out, int image_width, int image_height, int f, int Din,
                           int batch, int stride, int padding, void *stream);
/* gpu::averagepool_expand_delta, convolution.cuh:194-207, 412-437: in[k]
 * (k = b*Din + c) fills out[b*Din*pool_size + c + j*Din] for j < pool_size;
 * out_capacity (entries of out) must be at least n_in * pool_size */
int pir_averagepool_expand_delta(const uint64_t *in, size_t n_in, uint64_t *out, size_t out_capacity, int Din,
                                 int pool_size, void *stream);
/* gpu::transpose, matrix.cuh:43-59, 202-221: b[row*cols + col] = a[col*rows + row];
 * elem_bytes 8 or 1 */
int pir_transpose(const void *a, void *b, int rows, int cols, int elem_bytes, void *stream);
/* gpu::getMSBs, bitwise.cuh:23-30, 126-152: msbs[i] = bits[i*bit_width + bit_width - 1]
 * for i < n_msbs; gpu::setMSBs, bitwise.cuh:32-39, 154-182: that element of each of
 * the n_bits / bit_width groups = val.  elem_bytes 8 or 1 */
int pir_get_msbs(const void *bits, void *msbs, size_t n_msbs, size_t bit_width, int elem_bytes, void *stream);
int pir_set_msbs(void *bits, size_t n_bits, uint64_t val, size_t bit_width, int elem_bytes, void *stream);
/* gpu::reduceSum, matrix.cuh:109-126, 331-351: b[j] += sum_i (reduce_rows ?
 * a[cols*i + j] (j < cols) : a[cols*j + i] (j < rows)) */
int pir_reduce_sum(const uint64_t *a, uint64_t *b, int reduce_rows, int rows, int cols, void *stream);
/* gpu::vectorExpand, matrix.cuh:128-139, 354-380: b[i*factor + j] = a[i] */
int pir_vector_expand(const uint64_t *a, size_t n, uint64_t *b, int factor, void *stream);
/* gpu::maxpool_row2im, convolution.cuh:110-151, 376-410: adds each pool entry
 * rows[window*f*f + v] into its image position of out (parameter order and the
 * height / width naming of the reference's host function) */
int pir_maxpool_row2im(const uint64_t *rows, uint64_t *out, int image_height, int image_width, int f, int Din,
                       int batch, int stride, void *stream);

/* ---- DeviceData element ops (src/gpu/DeviceData.h:131-209) ----------------- */
enum { PIR_OP_ADD = 0, PIR_OP_SUB, PIR_OP_MUL, PIR_OP_XOR, PIR_OP_AND, PIR_OP_SHR, PIR_OP_SHL, PIR_OP_DIV };
/* x[i*sx] = x[i*sx] op y[i*sy] (y == NULL: op scalar).  elem_bytes 8 or 1.  DIV
 * is the signed division of functors.cuh:33-45 / 94-105, (S)((double)(S)x /
 * (double)(S)y), restated in integer arithmetic (binary64 rounding of both
 * operands, round-to-nearest-even quotient, truncation toward zero; a quotient
 * outside the signed range clamps, 0/0 gives 0). */
int pir_dd_op(int op, int elem_bytes, void *x, long long x_stride, const void *y, long long y_stride,
              uint64_t scalar, size_t n, void *stream);

/* ---- RSS protocol, all three parties (src/mpc/RSS.h:103-145) -------------- */
/* Execution form of the pir_rss_* calls: PIR_MODE_FUSED (one fused kernel per
 * protocol step, default) or PIR_MODE_PLAIN (the reference's op-by-op launches).
 * Both give bit-identical shares. */
int pir_set_mode(int mode);
int pir_rss_reconstruct(const void *x, size_t n, void *out, int elem_bytes, void *stream); /* RSS.inl:457-479 */
int pir_rss_reshare(const uint64_t *c, size_t n, uint64_t *out, void *stream);             /* RSS.inl:1052-1074 */
int pir_rss_mult(uint64_t *x, const uint64_t *y, size_t n, void *stream);                 /* RSS.inl:283-320 */
int pir_rss_and(void *x, const void *y, size_t n, int elem_bytes, void *stream);          /* RSS.inl:330-345 */
/* dividePublic, RSS.inl:422-455: d (scalar) or d_local ([3][n] per-party) denominators */
int pir_rss_divide_public(uint64_t *a, size_t n, uint64_t d, const uint64_t *d_local, void *stream);
/* dividePublic3out3, RSS.inl:1076-1094: a = 3-of-3 shares [3][n], denominator 2^k */
int pir_rss_divide_public_3out3(const uint64_t *a, size_t n, int k, uint64_t *out, void *stream);
/* matmul, RSS.inl:481-540 (transpose flags as in pir_gemm); truncation by 2^trunc */
int pir_rss_matmul(const uint64_t *a, const uint64_t *b, uint64_t *c, int M, int N, int K, int ta, int tb, int tc,
                   int trunc, int algo, void *stream);
/* convolution(kFprop), RSS.inl:705-803 */
int pir_rss_convolution(const uint64_t *a, const uint64_t *w, uint64_t *c, int batch, int H, int W, int f, int Din,
                        int Dout, int stride, int pad, int trunc, int algo, void *stream);
/* convolution(A, B, C, op, ...), RSS.inl:641-803, for cutlass::conv::Operator
 * kFprop (A = input, B = filter, C = output), kDgrad (A = delta [b][OH][OW][Dout],
 * B = filter, C = input gradient [b][H][W][Din]) or kWgrad (A = delta, B = forward
 * input, C = filter gradient [Dout][f][f][Din]); truncation by 2^trunc */
enum { PIR_CONV_FPROP = 0, PIR_CONV_DGRAD = 1, PIR_CONV_WGRAD = 2 };
int pir_rss_convolution_op(int op, const uint64_t *a, const uint64_t *b, uint64_t *c, int batch, int H, int W, int f,
                           int Din, int Dout, int stride, int pad, int trunc, int algo, void *stream);
int pir_rss_drelu(const uint64_t *x, size_t n, void *result, int elem_bytes, void *stream);  /* RSS.inl:806-852 */
int pir_rss_relu(const uint64_t *x, size_t n, uint64_t *y, void *dresult, int elem_bytes, void *stream); /* :854-865 */
int pir_rss_select_share(const uint64_t *x, const uint64_t *y, const void *b, uint64_t *z, size_t n, int elem_bytes,
                         void *stream);                                                    /* RSS.inl:542-569 */
/* maxpool over gathered pools [n/k windows][k] (modified in place like the reference), RSS.inl:867-1017 */
int pir_rss_maxpool(uint64_t *pools, size_t n, uint64_t *result, void *dresult, int k, int elem_bytes, void *stream);
/* getPowers, RSS.inl:1211-1235: each party's opened power of x [3][2][n] -> pow [3][n] */
int pir_rss_get_powers(const uint64_t *x, size_t n, uint64_t *pow, void *stream);
/* taylorSeries, RSS.inl:1237-1320: a0 + a1 x + a2 x^2 on the range-reduced x,
 * rescaled by fn (PIR_TAYLOR_SQRT / _INVERSE / _SIGMOID, the functors of
 * functors.h:33-49) at fp fractional bits (the reference's FLOAT_PRECISION);
 * out [3][2][n] must not overlap x */
enum { PIR_TAYLOR_SQRT = 0, PIR_TAYLOR_INVERSE = 1, PIR_TAYLOR_SIGMOID = 2 };
int pir_rss_taylor_series(const uint64_t *x, size_t n, uint64_t *out, double a0, double a1, double a2, int fn,
                          int fp, void *stream);
/* sqrt / inverse / sigmoid, RSS.inl:570-612: taylorSeries with the reference's coefficients */
int pir_rss_sqrt(const uint64_t *x, size_t n, uint64_t *out, int fp, void *stream);
int pir_rss_inverse(const uint64_t *x, size_t n, uint64_t *out, int fp, void *stream);
int pir_rss_sigmoid(const uint64_t *x, size_t n, uint64_t *out, int fp, void *stream);

/* ---- masked precompute mode and the seeded PRG ---------------------------
 * The reference's correlated randomness is placeholder constants
 * (src/mpc/Precompute.h:73-134: r = 1, r' = d, Beaver triples of ones; no mask
 * in reshare, RSS.inl:1052-1074; r = 0 and rbits = 1 in dReLU, RSS.inl:806-852;
 * c = cbits = 0 in selectShare, RSS.inl:542-569), and its AES PRG
 * (src/mpc/AESObject.cpp:11-90) is never instantiated (src/main.cu:96-104).
 * PIR_PRECOMPUTE_MASKED deals that randomness from Philox4x32-10 (Salmon et
 * al., SC'11) keyed by `seed`: every protocol call that uses randomness takes
 * the next 32-bit call number, and random word j of element e of that call is
 * half (j & 1) of Philox4x32-10(key = seed, counter = (e_lo, e_hi, call, j >> 1)).
 * In that mode dReLU ends with XOR 1 instead of 1 - x, selectShare opens its
 * bit by XOR, and truncation returns floor or ceil of t / 2^k (the reference's
 * r' = r * d plan, Precompute.h:107).  It runs the RSS inference protocols in
 * the fused form: reshare, mult, and, dividePublic (2^k), dividePublic3out3,
 * matmul, convolution (fprop), dReLU, ReLU, selectShare, maxpool and the FC /
 * CNN / ReLU / Maxpool layers; everything else refuses it (status 3). */
enum { PIR_PRECOMPUTE_PLACEHOLDER = 0, PIR_PRECOMPUTE_MASKED = 1 };
/* the calling thread's free-standing calls: mode, key and next call number */
int pir_set_precompute(int mode, uint64_t seed, uint32_t call);
int pir_get_prg_call(uint32_t *call);
/* Philox4x32-10 blocks: out[4i..4i+3] of counter ctr[4i..4i+3], key key[2i..2i+1]
 * (device arrays of n blocks) */
int pir_prg_philox(const uint32_t *ctr, const uint32_t *key, size_t n, uint32_t *out, void *stream);
/* out[2b + h] = word h of element b of call `call` under `seed` (n words) */
int pir_prg_fill(uint64_t seed, uint32_t call, size_t n, uint64_t *out, void *stream);
/* The dealer: a random consistent RSS sharing shares [3][2][n] of the ring
 * values[n] (x1, x2 = words 0, 1 of element i of the call, x0 = v - x1 - x2);
 * pir_rss_rerandomize re-deals the value party A opens from shares_in (may
 * equal shares_out). */
int pir_rss_deal(const uint64_t *values, size_t n, uint64_t seed, uint32_t call, uint64_t *shares, void *stream);
int pir_rss_rerandomize(const uint64_t *shares_in, size_t n, uint64_t seed, uint32_t call, uint64_t *shares_out,
                        void *stream);

/* ---- TPC (2 parties) and FPC (4 parties), src/mpc/TPC.inl, FPC.inl ------- */
/* TPC value: [2 parties][n] (one share each).  FPC value: [4 parties][3 shares][n],
 * party q's share s holding component (q + s + 1) % 4.  Placeholder precompute
 * as in the reference (Precompute.h).  pir_set_mode selects the fused form
 * (one limb product per party with the truncation in its epilogue) or the
 * reference's op-by-op form; algo picks the GEMM of the op-by-op form. */
int pir_tpc_reconstruct(const uint64_t *x, size_t n, uint64_t *out, void *stream); /* TPC.inl:551-572, out [2][n] */
int pir_tpc_divide_public(uint64_t *a, size_t n, uint64_t d, void *stream);       /* TPC.inl:370-384 */
/* TPC element-wise protocols, both parties' shares [2][n]; bit values uint8
 * (elem_bytes 1) or uint64 (8); Beaver triples x = y = z = fill(1) as the
 * reference's placeholder precompute (Precompute.h:93-105) */
int pir_tpc_mult(uint64_t *x, const uint64_t *y, size_t n, void *stream);        /* TPC *= TPC, TPC.inl:242-272 */
int pir_tpc_and(void *x, const void *y, size_t n, int elem_bytes, void *stream); /* TPC &= TPC, TPC.inl:281-312 */
int pir_tpc_drelu(const uint64_t *x, size_t n, void *result, int elem_bytes, void *stream); /* TPC.inl:828-868 */
int pir_tpc_relu(const uint64_t *x, size_t n, uint64_t *y, void *dresult, int elem_bytes, void *stream); /* :870-884 */
int pir_tpc_select_share(const uint64_t *x, const uint64_t *y, const void *b, uint64_t *z, size_t n, int elem_bytes,
                         void *stream);                                          /* TPC.inl:585-613 */
/* maxpool over gathered pools [n/k windows][k], k in {2, 4, 8, 16} (pools modified like the reference), TPC.inl:886-1030 */
int pir_tpc_maxpool(uint64_t *pools, size_t n, uint64_t *result, void *dresult, int k, int elem_bytes, void *stream);
/* matmul, TPC.inl:574-583 + localMatMul :1059-1090 (transpose flags as in pir_gemm) */
int pir_tpc_matmul(const uint64_t *a, const uint64_t *b, uint64_t *c, int M, int N, int K, int ta, int tb, int tc,
                   int trunc, int algo, void *stream);
/* convolution(kFprop), TPC.inl:644-690, 788-826 */
int pir_tpc_convolution(const uint64_t *a, const uint64_t *w, uint64_t *c, int batch, int H, int W, int f, int Din,
                        int Dout, int stride, int pad, int trunc, int algo, void *stream);
int pir_fpc_reconstruct(const uint64_t *x, size_t n, uint64_t *out, void *stream); /* FPC.inl:582-604, out [4][n] */
int pir_fpc_divide_public(uint64_t *a, size_t n, uint64_t d, void *stream);       /* FPC.inl:566-580 */
/* FPC element-wise protocols on [4][3][n] values, bit values uint8 (elem_bytes 1)
 * or uint64 (8); reshareFPC with zero PRGs as the reference's placeholder */
int pir_fpc_mult(uint64_t *x, const uint64_t *y, size_t n, void *stream);        /* FPC *= FPC, FPC.inl:321-362 */
int pir_fpc_and(void *x, const void *y, size_t n, int elem_bytes, void *stream); /* FPC &= FPC, FPC.inl:374-398 */
int pir_fpc_drelu(const uint64_t *x, size_t n, void *result, int elem_bytes, void *stream); /* FPC.inl:929-989 */
int pir_fpc_relu(const uint64_t *x, size_t n, uint64_t *y, void *dresult, int elem_bytes, void *stream); /* :991-1002 */
int pir_fpc_select_share(const uint64_t *x, const uint64_t *y, const void *b, uint64_t *z, size_t n, int elem_bytes,
                         void *stream);                                          /* FPC.inl:671-712 */
/* maxpool over gathered pools [n/k windows][k], k in {2, 4, 8, 16} (pools modified like the reference), FPC.inl:1004-1181 */
int pir_fpc_maxpool(uint64_t *pools, size_t n, uint64_t *result, void *dresult, int k, int elem_bytes, void *stream);
/* dividePublic(Share, DeviceData) (TPC.inl:386-407, FPC.inl:566-580): per-party
 * public denominators d_local [2][n] (TPC) or [4][n] (FPC) */
int pir_tpc_divide_public_local(uint64_t *a, size_t n, const uint64_t *d_local, void *stream);
int pir_fpc_divide_public_local(uint64_t *a, size_t n, const uint64_t *d_local, void *stream);
/* getPowers (TPC.inl:1202-1226, FPC.inl:1418-1441): each party's opened power,
 * pow [2][n] / [4][n]; taylorSeries (TPC.inl:1228-1290, FPC.inl:1443-1526) and
 * sqrt / inverse / sigmoid (TPC.inl:615-642, FPC.inl:714-741) as pir_rss_taylor_series */
int pir_tpc_get_powers(const uint64_t *x, size_t n, uint64_t *pow, void *stream);
int pir_fpc_get_powers(const uint64_t *x, size_t n, uint64_t *pow, void *stream);
int pir_tpc_taylor_series(const uint64_t *x, size_t n, uint64_t *out, double a0, double a1, double a2, int fn,
                          int fp, void *stream);
int pir_fpc_taylor_series(const uint64_t *x, size_t n, uint64_t *out, double a0, double a1, double a2, int fn,
                          int fp, void *stream);
int pir_tpc_sqrt(const uint64_t *x, size_t n, uint64_t *out, int fp, void *stream);
int pir_tpc_inverse(const uint64_t *x, size_t n, uint64_t *out, int fp, void *stream);
int pir_tpc_sigmoid(const uint64_t *x, size_t n, uint64_t *out, int fp, void *stream);
int pir_fpc_sqrt(const uint64_t *x, size_t n, uint64_t *out, int fp, void *stream);
int pir_fpc_inverse(const uint64_t *x, size_t n, uint64_t *out, int fp, void *stream);
int pir_fpc_sigmoid(const uint64_t *x, size_t n, uint64_t *out, int fp, void *stream);
/* reshareFPC, FPC.inl:1183-1260: z, zp [4][n] per party, added into out [4][3][n]; 0 <= i < j < 4 */
int pir_fpc_reshare(const uint64_t *z, const uint64_t *zp, size_t n, int i, int j, uint64_t *out, void *stream);
/* matmul, FPC.inl:635-666 */
int pir_fpc_matmul(const uint64_t *a, const uint64_t *b, uint64_t *c, int M, int N, int K, int ta, int tb, int tc,
                   int trunc, int algo, void *stream);
/* convolution(kFprop), FPC.inl:744-782, 861-927 */
int pir_fpc_convolution(const uint64_t *a, const uint64_t *w, uint64_t *c, int batch, int H, int W, int f, int Din,
                        int Dout, int stride, int pad, int trunc, int algo, void *stream);

/* ---- layers and network (src/nn/FCLayer.cu etc., NeuralNetwork.cu) -------------- */
typedef struct pir_net pir_net;
enum { PIR_LAYER_FC = 0, PIR_LAYER_CNN = 1, PIR_LAYER_RELU = 2, PIR_LAYER_MAXPOOL = 3, PIR_LAYER_AVGPOOL = 4,
       PIR_LAYER_LN = 5, PIR_LAYER_RES = 6 };
enum { PIR_MODE_FUSED = 0, PIR_MODE_PLAIN = 1 };
/* spec: per layer 8 ints {type, a0..a6}:
 *   FC      {0, inputDim, outputDim}                                (FCConfig.h:16)
 *   CNN     {1, H, W, Cin, Cout, filter, stride, padding}           (CNNConfig.h:27)
 *   ReLU    {2, inputDim}                                           (ReLUConfig.h:12)
 *   Maxpool {3, H, W, C, pool, stride}                              (MaxpoolConfig.h:16)
 *   Avgpool {4, H, W, C, pool, stride}                              (AveragepoolConfig.h:15)
 *   LN      {5, inputDim}                                           (LNConfig.h:13)
 *   Res     {6, H, W, in_planes, planes, num_blocks, stride, expansion} (ResLayerConfig.h:19; expansion 1)
 * Each layer's input size must equal the previous layer's output size.
 * Weights are initialised as in FCLayer.cu:31-48 / CNNLayer.cu:36-50 with
 * std::default_random_engine(seed + layer) unless set with pir_net_set_weights. */
int pir_net_create(pir_net **net, const int *spec, int num_layers, int batch, int frac_bits, int seed, int mode,
                   void *stream);
int pir_net_destroy(pir_net *net);
int pir_net_num_layers(const pir_net *net);
/* ---- model files and snapshots ---------------------------------------------- */
/* loadModel (model.cu:24-115) without building the network: the file's
 * input_size, num_classes, batch_size, the layer count (*num_layers) and up to
 * max_layers layer specs (8 ints each, the pir_net_create form).  Layer types
 * fc, relu, cnn, maxpool, averagepool, ln, res. */
int pir_model_read(const char *network_filename, int *input_size, int *num_classes, int *batch_size, int *num_layers,
                   int *spec, int max_layers);
/* loadModel + the network constructor: batch <= 0 takes the file's batch_size */
int pir_net_load_model(pir_net **net, const char *network_filename, int batch, int frac_bits, int seed, int mode,
                       void *stream);
int pir_net_batch_size(const pir_net *net);
int pir_net_layer_spec(const pir_net *net, int layer, int *spec); /* spec: 8 ints */
/* NeuralNetwork::saveSnapshot (NeuralNetwork.cu:85-104, util.cuh:291-303):
 * path/weight{i} and path/bias{i} (FC), path/gamma{i} and path/beta{i} (LN) as
 * space-separated reconstructed doubles
 * in the default stream format; creates the directory */
int pir_net_save_snapshot(pir_net *net, const char *path);
/* NeuralNetwork::loadSnapshot (NeuralNetwork.cu:74-83, util.cuh:271-289):
 * reads those files into public shares; missing values stay zero */
int pir_net_load_snapshot(pir_net *net, const char *path);
/* setPublic of weights (and FC biases) from doubles, RSS.inl:122-142 */
int pir_net_set_weights(pir_net *net, int layer, const double *weights, size_t n_weights, const double *biases,
                        size_t n_biases);
/* NeuralNetwork::forward (NeuralNetwork.cu:108-133): input setPublic from doubles */
int pir_net_forward(pir_net *net, const double *input, size_t n);
/* forward from device-resident input shares [3][2][batch*input_size]
 * (NULL: the shares already in the network's input buffer) */
int pir_net_forward_shares(pir_net *net, const uint64_t *input_shares);
/* the network's own input share buffer and its per-share element count */
int pir_net_input(pir_net *net, uint64_t **shares, size_t *n);
/* wait for all work queued by the network */
int pir_net_synchronize(pir_net *net);
/* device pointer to a layer's activation shares [3][2][n] and its size */
int pir_net_activation(pir_net *net, int layer, const uint64_t **shares, size_t *n);
/* device pointer to a ReLU / Maxpool layer's derivative shares (uint8) */
int pir_net_derivative(pir_net *net, int layer, const uint8_t **shares, size_t *n);
int pir_net_weights(pir_net *net, int layer, const uint64_t **shares, size_t *n, const uint64_t **bias_shares,
                    size_t *n_bias);
/* weights (and FC biases) from host share arrays [3][2][n] */
int pir_net_set_weight_shares(pir_net *net, int layer, const uint64_t *weight_shares, const uint64_t *bias_shares);
/* Every parameter tensor of a layer, in the order and with the names of its
 * snapshot files (weight{n} / bias{n}, gamma{n} / beta{n} for LN, a Res layer's
 * inner layers block by block, ResLayer.cu:134-167): the count, one tensor's
 * device shares [3][2][n] and name, and setting one from host shares. */
int pir_net_num_params(pir_net *net, int layer);
int pir_net_param(pir_net *net, int layer, int index, const uint64_t **shares, size_t *n);
int pir_net_param_name(pir_net *net, int layer, int index, char *name, size_t capacity);
int pir_net_set_param_shares(pir_net *net, int layer, int index, const uint64_t *shares);
/* hipGraph capture of the whole forward (input from pir_net_forward_shares' buffer).
 * A weight change (set_weights, a backward pass) drops the graph; capture again. */
int pir_net_capture(pir_net *net, const uint64_t *input_shares);
int pir_net_replay(pir_net *net);

/* ---- training (NeuralNetwork.cu:136-149, 207-267, 435-467) ----------------
 * log_learning_rate (main.cu:48, default 5): updates divide by 2^log_lr. */
int pir_net_set_log_learning_rate(pir_net *net, int log_lr);
/* NeuralNetwork::backward after a forward: labels (batch x classes doubles,
 * one-hot) become public shares; the reveal-softmax delta, then every layer's
 * Layer::backward in reverse order with its weight update. */
int pir_net_backward(pir_net *net, const double *labels, size_t n);
/* the same with label shares already on the device [3][2][batch*classes] */
int pir_net_backward_shares(pir_net *net, const uint64_t *label_shares);
/* Layer::backward(delta, forwardInput) of one layer (FCLayer.cu:133-224,
 * CNNLayer.cu:118-195, ReLULayer.cu:91-123, MaxpoolLayer.cu:133-180): delta and
 * forward_input are device shares [3][2][n] of the layer's output / input size */
int pir_net_layer_backward(pir_net *net, int layer, const uint64_t *delta, const uint64_t *forward_input);
/* Layer::getDelta: device shares of dL/d(input of layer) after a backward pass;
 * layer == number of layers gives the reveal-softmax delta (dL/d network output) */
int pir_net_delta(pir_net *net, int layer, const uint64_t **shares, size_t *n);
/* Training-step API (the loop body of main.cu:306-340: forward + backward):
 * labels are kept on the device (setPublic of batch x classes doubles, or
 * written directly into pir_net_train_labels' buffer), the input is the
 * network's input buffer (pir_net_input).  capture_train runs one eager step,
 * then captures forward + backward into a hipGraph replayed by replay_train;
 * any eager weight change or backward drops it. */
int pir_net_set_train_labels(pir_net *net, const double *labels, size_t n);
int pir_net_train_labels(pir_net *net, uint64_t **shares, size_t *n);
int pir_net_train_step(pir_net *net);
int pir_net_capture_train(pir_net *net);
int pir_net_replay_train(pir_net *net);
/* autotune every limb GEMM shape of a training step / time them like
 * pir_net_profile_gemm (each runs one eager training step) */
int pir_net_autotune_train(pir_net *net);
int pir_net_profile_train_gemm(pir_net *net, double *ms, double *ops, int *launches);
/* one eager training step with HIP events: ms[i] (i < L) = layer i's forward,
 * ms[L] = the softmax delta, ms[L + 1 + i] = layer i's backward (count = 2L + 1;
 * in the fused form a ReLU layer runs inside its neighbour and reports ~0 ms) */
int pir_net_profile_train_layers(pir_net *net, float *ms, int count);
/* one eager forward with a HIP event after each layer on the network's stream;
 * ms[i] = device time of layer i (count = number of layers).  In the fused form
 * a ReLU layer is computed inside the layer before it and reports ~0 ms. */
int pir_net_profile_layers(pir_net *net, float *ms, int count);
/* one eager forward with HIP events around every limb GEMM kernel launch on
 * the network's stream; per kernel variant v: ms[v] = summed kernel time,
 * ops[v] = summed algorithmic int8 ops (2*36*M*N*nseg*K), launches[v] = launch
 * count.  v = A-operand form (PIR_LIMB_DENSE_ROW .. PIR_LIMB_WGRAD), plus
 * PIR_LIMB_WIDE for the 32x128 workgroup tile (else 64x64); the kernel is
 * limb_gemm_kernel<form, wide ? 1 : 2>, or limb_gemm_help_kernel<form> for an
 * fprop product with 64 < N <= 96 on the 32x128 tile (counted with that
 * variant).  Arrays hold PIR_LIMB_VARIANTS entries. */
enum { PIR_LIMB_DENSE_ROW = 0, PIR_LIMB_DENSE_COL, PIR_LIMB_CONV, PIR_LIMB_CONV_C32, PIR_LIMB_DGRAD,
       PIR_LIMB_DGRAD_C32, PIR_LIMB_WGRAD, PIR_LIMB_WIDE = 8, PIR_LIMB_VARIANTS = 16 };
int pir_net_profile_gemm(pir_net *net, double *ms, double *ops, int *launches);
int pir_net_set_mode(pir_net *net, int mode);
/* The network's precompute mode (see pir_set_precompute): in the masked mode
 * layer i of a forward uses call number base + i, and each forward advances
 * the base by the layer count.  A captured forward replays the call numbers of
 * its capture. */
int pir_net_set_precompute(pir_net *net, int mode, uint64_t seed, uint32_t call);
/* Re-deals every parameter tensor (pir_net_param order, call k for the k-th)
 * and the input buffer (the next call) as random sharings of the values they
 * hold, from `seed` (the dealer of pir_rss_rerandomize). */
int pir_net_rerandomize(pir_net *net, uint64_t seed);

/* ---- limb GEMM autotuner (no counterpart in the reference: CUTLASS tile
 * shapes are fixed at compile time there) ------------------------------------
 * Each limb GEMM shape runs with a plan: workgroup tile (32x128 or 64x64) and
 * split-K factor.  With tuning on, the first eager launch of an untuned shape
 * times the candidate plans on the real operands and records the fastest in a
 * process-wide table; later launches, and hipGraph captures, use the table.
 * pir_net_autotune runs one eager forward of the network with tuning on. */
int pir_net_autotune(pir_net *net);
int pir_autotune_set(int on); /* for the free-standing pir_* calls */
int pir_autotune_entries(void);
int pir_autotune_clear(void);
int pir_autotune_save(const char *path); /* text, one shape per line */
int pir_autotune_load(const char *path);

/* ---- measured peaks ----------------------------------------------------------
 * No counterpart in the reference (its GEMM profiling harness,
 * src/test/MatMulProfiling.cu:192-529, times the CUTLASS kernels only): the
 * measured denominators of the roofline.  which:
 *   PIR_PEAK_HBM_COPY   GB/s (bytes read + written) of a 16-byte-per-lane copy of `bytes`
 *   PIR_PEAK_HBM_READ   GB/s of a read-only stream over `bytes` (writes nothing)
 *   PIR_PEAK_HBM_WRITE  GB/s of a write-only fill of `bytes`
 *   PIR_PEAK_MFMA_I8    int8 TOP/s of back-to-back v_mfma_i32_32x32x32_i8 (bytes ignored)
 *   PIR_PEAK_MFMA_I8_RANDOM  the same on full-entropy operand bytes (the MFMA array draws more
 *                       power on them and sustains a lower rate; the digit planes of uniformly
 *                       random shares are such operands)
 *   PIR_PEAK_VALU_U64   G 64-bit multiply-adds per second on the vector ALUs (bytes ignored)
 * rate = the measured rate, ms = average launch time over `reps` launches after
 * one warm-up.  Allocates and frees its buffers; synchronises the host. */
enum { PIR_PEAK_HBM_COPY = 0, PIR_PEAK_HBM_READ = 1, PIR_PEAK_HBM_WRITE = 2, PIR_PEAK_MFMA_I8 = 3,
       PIR_PEAK_VALU_U64 = 4, PIR_PEAK_MFMA_I8_RANDOM = 5 };
int pir_measure_peak(int which, size_t bytes, int reps, double *rate, double *ms, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PIR_H */
