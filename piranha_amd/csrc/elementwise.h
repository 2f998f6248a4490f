// elementwise.h -- DeviceData element ops and reference bit / layout kernels.
#pragma once

#include "common.h"

namespace pir {

enum EwOp { OP_ADD = 0, OP_SUB, OP_MUL, OP_XOR, OP_AND, OP_SHR, OP_SHL, OP_DIV, OP_SET };

// Up to MULTI_OP_MAX strided element-op pairs in one launch:
// x[a][i*xs] op= y[a][i*ys] (or op= scalar[a] when y[a] is null).
constexpr int MULTI_OP_MAX = 12;  // every share array of a 4-party value
struct MultiOp {
    void *x[MULTI_OP_MAX];
    long long xs[MULTI_OP_MAX];
    const void *y[MULTI_OP_MAX];
    long long ys[MULTI_OP_MAX];
    u64 scalar[MULTI_OP_MAX];
    int count;
    int op;
    size_t n;
};

int multi_op(const MultiOp &m, int elem_bytes, hipStream_t s);
int dd_op(int op, int elem_bytes, void *x, long long xs, const void *y, long long ys, u64 scalar, size_t n,
          hipStream_t s);
int fill_u64(u64 *p, u64 v, size_t n, hipStream_t s);
int fill_u8(u8 *p, u8 v, size_t n, hipStream_t s);
int bitexpand(const u64 *a, size_t n, void *bits, int elem_bytes, hipStream_t s);
int set_carry_out_msb(void *rbits, void *abits, void *msb, size_t n, int bw, int xr, int eb, hipStream_t s);
int expand_compare(const void *b, const void *inv, void *out, size_t n, size_t out_n, int eb, hipStream_t s);
int element_vector_add(u64 *a, const u64 *b, int rowwise, int rows, int cols, hipStream_t s);
// gpu::transpose (matrix.cuh:43-59, 202-221): b[row*cols + col] = a[col*rows + row]
int transpose(const void *a, void *b, int rows, int cols, int elem_bytes, hipStream_t s);
// gpu::getMSBs / setMSBs (bitwise.cuh:23-39, 126-182): element bit_width - 1 of
// each group of bit_width elements
int get_msbs(const void *bits, void *msbs, size_t n_msbs, size_t bit_width, int elem_bytes, hipStream_t s);
int set_msbs(void *bits, size_t n_bits, u64 val, size_t bit_width, int elem_bytes, hipStream_t s);
// gpu::elementVectorSubtract (matrix.cuh:97-107): a[row*cols + col] -= b[rowwise ? col : row]
int element_vector_subtract(u64 *a, const u64 *b, int rowwise, int rows, int cols, hipStream_t s);
// out[i] = tab.at(x[i]): a tabulated coefficient functor of taylorSeries
// (functors.cuh:107-116, functors.h:22-31) applied through thrust::transform.
int taylor_lookup(const u64 *x, size_t n, const TaylorTab &tab, u64 *out, hipStream_t s);
// filter_positive_powers / filter_negative_powers (functors.h:7-20) on one array:
// pos[i] = (S)x >= 0 ? x : 0, neg[i] = (S)x < 0 ? -(S)x : 0.
int filter_powers(const u64 *x, size_t n, u64 *pos, u64 *neg, hipStream_t s);
int truncate_cols(const void *a, size_t a_n, void *b, int sc, int ec, int eb, hipStream_t s);
int stride_pad(const u64 *in, u64 *out, size_t n_in, int stride, int pad, u64 pv, hipStream_t s);
size_t maxpool_im2row_size(int H, int W, int f, int Din, int batch, int stride, int padding);
int maxpool_im2row(const u64 *im, u64 *out, int H, int W, int f, int Din, int batch, int stride, int padding, u64 pv,
                   hipStream_t s);
// gpu::averagepool_im2row (convolution.cuh:339-371): the maxpool_im2row gather
// with the same (height, width) swap and no padding to a power of two; output
// f*f x windows x batch x Din entries.
size_t averagepool_im2row_size(int imageWidth, int imageHeight, int f, int Din, int batch, int stride, int padding);
int averagepool_im2row(const u64 *im, u64 *out, int imageWidth, int imageHeight, int f, int Din, int batch, int stride,
                       int padding, hipStream_t s);
// gpu::averagepool_expand_delta (convolution.cuh:194-207, 412-437): writes the
// first n_in * pool_size entries of out (which must hold that many).
int averagepool_expand_delta(const u64 *in, size_t n_in, u64 *out, size_t out_capacity, int Din, int pool_size,
                             hipStream_t s);
// gpu::reduceSum (matrix.cuh:109-126, 331-351): b[col] += sum over the other axis.
int reduce_sum(const u64 *a, u64 *b, int reduce_rows, int rows, int cols, hipStream_t s);
// gpu::vectorExpand (matrix.cuh:128-139, 354-380): b[i*f + j] = a[i].
int vector_expand(const u64 *a, size_t n, u64 *b, int factor, hipStream_t s);
// gpu::maxpool_row2im (convolution.cuh:110-151, 376-410): adds every pool
// entry into its image position (the kernel-side height/width naming).
int maxpool_row2im(const u64 *rows, u64 *out, int imageHeight, int imageWidth, int f, int Din, int batch, int stride,
                   hipStream_t s);

}  // namespace pir
