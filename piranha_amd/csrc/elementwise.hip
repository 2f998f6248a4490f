// elementwise.hip -- DeviceData element ops and the reference's bit / layout
// kernels, one launch per op (the reference's thrust::transform and custom
// kernels of src/gpu/{DeviceData.h,bitwise.cuh,matrix.cuh,convolution.cuh}).
// These are the drop-in gpu:: boundary and the building blocks of the
// op-by-op ("plain") execution mode; the fused mode uses fused_rss.hip.
#include "elementwise.h"

namespace pir {

constexpr int BLK = 256;

static unsigned g1(size_t n) {
    size_t g = (n + BLK - 1) / BLK;
    return (unsigned)(g == 0 ? 1 : g);
}

template <typename T>
__device__ __forceinline__ T ew_apply(int op, T x, T y) {
    typedef typename std::make_signed<T>::type S;
    switch (op) {
        case OP_ADD: return (T)(x + y);
        case OP_SUB: return (T)(x - y);
        case OP_MUL: return (T)(x * y);
        case OP_XOR: return (T)(x ^ y);
        case OP_AND: return (T)(x & y);
        // functors.cuh:46-92 evaluated on the reference's GPU, whose shifts
        // clamp the amount to the register width: >> by >= width fills with the
        // sign, << by >= width gives 0 (this unit masks the amount instead)
        case OP_SHR: return y >= (T)(8 * sizeof(T)) ? (T)((S)x < 0 ? ~(T)0 : (T)0) : (T)((S)x >> y);
        case OP_SHL: return y >= (T)(8 * sizeof(T)) ? (T)0 : (T)(x << y);
        case OP_DIV: {
            // functors.cuh:33-45 / 94-105: (S)((double)(S)x / (double)(S)y)
            if constexpr (sizeof(T) == 8) {
                return (T)signed_div_u64((u64)x, (u64)y);
            } else {
                // narrow types: the double quotient of two small integers truncates
                // to the integer quotient (C division truncates toward zero too),
                // clamped to the type's range as the 64-bit form is
                const int lo = -(1 << (sizeof(T) * 8 - 1)), hi = (1 << (sizeof(T) * 8 - 1)) - 1;
                int xs = (int)(S)x, ys = (int)(S)y;
                if (ys == 0) return (T)(S)(xs == 0 ? 0 : (xs < 0 ? lo : hi));
                int q = xs / ys;
                return (T)(S)(q > hi ? hi : q);
            }
        }
        case OP_SET: return y;
    }
    return x;
}

// Up to MULTI_OP_MAX independent (x, y) array pairs per launch (one per party
// share), so an op over every share of every party is one launch.
template <typename T>
__global__ __launch_bounds__(BLK) void multi_op_kernel(MultiOp m) {
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    const int a = blockIdx.y;
    if (i >= m.n) return;
    T *x = (T *)m.x[a] + (long long)i * m.xs[a];
    T y = m.y[a] ? *((const T *)m.y[a] + (long long)i * m.ys[a]) : (T)m.scalar[a];
    *x = ew_apply<T>(m.op, *x, y);
}

int multi_op(const MultiOp &m, int elem_bytes, hipStream_t s) {
    PIR_REQUIRE(m.count >= 1 && m.count <= MULTI_OP_MAX, "multi_op: 1..12 arrays");
    if (m.n == 0) return OK;
    dim3 g(g1(m.n), m.count);
    if (elem_bytes == 8) hipLaunchKernelGGL(multi_op_kernel<u64>, g, dim3(BLK), 0, s, m);
    else if (elem_bytes == 1) hipLaunchKernelGGL(multi_op_kernel<u8>, g, dim3(BLK), 0, s, m);
    else {
        set_error("multi_op: element size must be 1 or 8");
        return ERR_ARG;
    }
    PIR_LAUNCH_CHECK();
    return OK;
}

int dd_op(int op, int elem_bytes, void *x, long long xs, const void *y, long long ys, u64 scalar, size_t n,
          hipStream_t s) {
    MultiOp m{};
    m.count = 1;
    m.op = op;
    m.n = n;
    m.x[0] = x;
    m.xs[0] = xs;
    m.y[0] = y;
    m.ys[0] = ys;
    m.scalar[0] = scalar;
    return multi_op(m, elem_bytes, s);
}

template <typename T>
__global__ __launch_bounds__(BLK) void fill_kernel(T *p, T v, size_t n) {
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) p[i] = v;
}

int fill_u64(u64 *p, u64 v, size_t n, hipStream_t s) {
    if (n == 0) return OK;
    hipLaunchKernelGGL(fill_kernel<u64>, dim3(g1(n)), dim3(BLK), 0, s, p, v, n);
    PIR_LAUNCH_CHECK();
    return OK;
}

int fill_u8(u8 *p, u8 v, size_t n, hipStream_t s) {
    if (n == 0) return OK;
    hipLaunchKernelGGL(fill_kernel<u8>, dim3(g1(n)), dim3(BLK), 0, s, p, v, n);
    PIR_LAUNCH_CHECK();
    return OK;
}

// bitwise.cuh:11-21 -- b[i*64 + j] = (a[i] >> j) & 1, one thread per bit.
template <typename U>
__global__ __launch_bounds__(BLK) void bitexpand_kernel(const u64 *__restrict__ a, size_t n, U *__restrict__ b) {
    size_t t = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (t >= n * 64) return;
    b[t] = (U)((a[t >> 6] >> (t & 63)) & 1);
}

int bitexpand(const u64 *a, size_t n, void *bits, int elem_bytes, hipStream_t s) {
    if (n == 0) return OK;
    if (elem_bytes == 1) hipLaunchKernelGGL(bitexpand_kernel<u8>, dim3(g1(n * 64)), dim3(BLK), 0, s, a, n, (u8 *)bits);
    else if (elem_bytes == 8) hipLaunchKernelGGL(bitexpand_kernel<u64>, dim3(g1(n * 64)), dim3(BLK), 0, s, a, n, (u64 *)bits);
    else {
        set_error("bitexpand: element size must be 1 or 8");
        return ERR_ARG;
    }
    PIR_LAUNCH_CHECK();
    return OK;
}

// bitwise.cuh:41-64
template <typename U>
__global__ __launch_bounds__(BLK) void carry_msb_kernel(U *rbits, U *abits, U *msb, size_t n, int bw, int xr) {
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    size_t bi = i * bw + (bw - 1);
    U m = rbits[bi];
    if (xr) m ^= abits[bi];
    msb[i] = m;
    abits[bi] = 1;
    rbits[bi] = 0;
}

int set_carry_out_msb(void *rbits, void *abits, void *msb, size_t n, int bw, int xr, int eb, hipStream_t s) {
    if (n == 0) return OK;
    if (eb == 1) hipLaunchKernelGGL(carry_msb_kernel<u8>, dim3(g1(n)), dim3(BLK), 0, s, (u8 *)rbits, (u8 *)abits, (u8 *)msb, n, bw, xr);
    else if (eb == 8) hipLaunchKernelGGL(carry_msb_kernel<u64>, dim3(g1(n)), dim3(BLK), 0, s, (u64 *)rbits, (u64 *)abits, (u64 *)msb, n, bw, xr);
    else {
        set_error("setCarryOutMSB: element size must be 1 or 8");
        return ERR_ARG;
    }
    PIR_LAUNCH_CHECK();
    return OK;
}

// bitwise.cuh:66-92 (within bounds): expanded[i*2f + e] = e < f ? b[i] : inv[i]
template <typename U>
__global__ __launch_bounds__(BLK) void expand_compare_kernel(const U *b, const U *inv, U *out, size_t n, size_t f) {
    size_t t = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (t >= n * 2 * f) return;
    size_t i = t / (2 * f), e = t % (2 * f);
    out[t] = e < f ? b[i] : inv[i];
}

int expand_compare(const void *b, const void *inv, void *out, size_t n, size_t out_n, int eb, hipStream_t s) {
    if (n == 0 && out_n == 0) return OK;  // empty ranges: nothing to expand
    PIR_REQUIRE(n > 0 && out_n % (2 * n) == 0, "expandCompare: expanded size must be a multiple of 2*b.size");
    size_t f = (out_n / n) / 2;
    if (eb == 1) hipLaunchKernelGGL(expand_compare_kernel<u8>, dim3(g1(out_n)), dim3(BLK), 0, s, (const u8 *)b, (const u8 *)inv, (u8 *)out, n, f);
    else if (eb == 8) hipLaunchKernelGGL(expand_compare_kernel<u64>, dim3(g1(out_n)), dim3(BLK), 0, s, (const u64 *)b, (const u64 *)inv, (u64 *)out, n, f);
    else {
        set_error("expandCompare: element size must be 1 or 8");
        return ERR_ARG;
    }
    PIR_LAUNCH_CHECK();
    return OK;
}

// matrix.cuh:76-84 -- a[col*rows + row] += b[rowwise ? col : row]
__global__ __launch_bounds__(BLK) void evadd_kernel(u64 *a, const u64 *b, int rowwise, int rows, int cols) {
    size_t t = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (t >= (size_t)rows * cols) return;
    int col = (int)(t / rows), row = (int)(t % rows);
    a[t] += b[rowwise ? col : row];
}

int element_vector_add(u64 *a, const u64 *b, int rowwise, int rows, int cols, hipStream_t s) {
    size_t n = (size_t)rows * cols;
    if (n == 0) return OK;
    hipLaunchKernelGGL(evadd_kernel, dim3(g1(n)), dim3(BLK), 0, s, a, b, rowwise, rows, cols);
    PIR_LAUNCH_CHECK();
    return OK;
}

// matrix.cuh:43-59 (transpose): b[row*cols + col] = a[col*rows + row]; one
// thread per output element, so the writes are contiguous
template <typename T>
__global__ __launch_bounds__(BLK) void transpose_kernel(const T *a, T *b, int rows, int cols) {
    size_t t = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (t >= (size_t)rows * cols) return;
    size_t row = t / (size_t)cols, col = t % (size_t)cols;
    b[t] = a[col * (size_t)rows + row];
}

int transpose(const void *a, void *b, int rows, int cols, int eb, hipStream_t s) {
    PIR_REQUIRE(rows >= 0 && cols >= 0 && (eb == 1 || eb == 8), "transpose: invalid arguments");
    size_t n = (size_t)rows * cols;
    if (n == 0) return OK;
    if (eb == 8) hipLaunchKernelGGL(transpose_kernel<u64>, dim3(g1(n)), dim3(BLK), 0, s, (const u64 *)a, (u64 *)b, rows, cols);
    else hipLaunchKernelGGL(transpose_kernel<u8>, dim3(g1(n)), dim3(BLK), 0, s, (const u8 *)a, (u8 *)b, rows, cols);
    PIR_LAUNCH_CHECK();
    return OK;
}

// bitwise.cuh:23-39 (getMSBs / setMSBs): element bitWidth - 1 of every group
// of bitWidth is the most significant bit of one value
template <typename T>
__global__ __launch_bounds__(BLK) void get_msbs_kernel(const T *bits, T *msbs, size_t n, size_t bw) {
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) msbs[i] = bits[i * bw + (bw - 1)];
}
template <typename T>
__global__ __launch_bounds__(BLK) void set_msbs_kernel(T *bits, T v, size_t n, size_t bw) {
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) bits[i * bw + (bw - 1)] = v;
}

int get_msbs(const void *bits, void *msbs, size_t n_msbs, size_t bit_width, int eb, hipStream_t s) {
    PIR_REQUIRE(bit_width >= 1 && (eb == 1 || eb == 8), "getMSBs: invalid arguments");
    if (n_msbs == 0) return OK;
    if (eb == 8) hipLaunchKernelGGL(get_msbs_kernel<u64>, dim3(g1(n_msbs)), dim3(BLK), 0, s, (const u64 *)bits, (u64 *)msbs, n_msbs, bit_width);
    else hipLaunchKernelGGL(get_msbs_kernel<u8>, dim3(g1(n_msbs)), dim3(BLK), 0, s, (const u8 *)bits, (u8 *)msbs, n_msbs, bit_width);
    PIR_LAUNCH_CHECK();
    return OK;
}

int set_msbs(void *bits, size_t n_bits, u64 val, size_t bit_width, int eb, hipStream_t s) {
    PIR_REQUIRE(bit_width >= 1 && (eb == 1 || eb == 8), "setMSBs: invalid arguments");
    size_t n = n_bits / bit_width;  // bits.size() / bitWidth, bitwise.cuh:156
    if (n == 0) return OK;
    if (eb == 8) hipLaunchKernelGGL(set_msbs_kernel<u64>, dim3(g1(n)), dim3(BLK), 0, s, (u64 *)bits, (u64)val, n, bit_width);
    else hipLaunchKernelGGL(set_msbs_kernel<u8>, dim3(g1(n)), dim3(BLK), 0, s, (u8 *)bits, (u8)val, n, bit_width);
    PIR_LAUNCH_CHECK();
    return OK;
}

__global__ __launch_bounds__(BLK) void taylor_lookup_kernel(const u64 *__restrict__ x, size_t n, const TaylorTab tab,
                                                             u64 *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) out[i] = tab.at(x[i]);
}

int taylor_lookup(const u64 *x, size_t n, const TaylorTab &tab, u64 *out, hipStream_t s) {
    if (n == 0) return OK;
    hipLaunchKernelGGL(taylor_lookup_kernel, dim3(g1(n)), dim3(BLK), 0, s, x, n, tab, out);
    PIR_LAUNCH_CHECK();
    return OK;
}

__global__ __launch_bounds__(BLK) void filter_powers_kernel(const u64 *__restrict__ x, size_t n, u64 *__restrict__ pos,
                                                            u64 *__restrict__ neg) {
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const u64 v = x[i];
    pos[i] = (s64)v >= 0 ? v : 0;
    neg[i] = (s64)v < 0 ? (u64)0 - v : 0;
}

int filter_powers(const u64 *x, size_t n, u64 *pos, u64 *neg, hipStream_t s) {
    if (n == 0) return OK;
    hipLaunchKernelGGL(filter_powers_kernel, dim3(g1(n)), dim3(BLK), 0, s, x, n, pos, neg);
    PIR_LAUNCH_CHECK();
    return OK;
}

// matrix.cuh:97-107 (elementVectorSubtract): a[row*cols + col] -= b[rowwise ? col : row]
__global__ __launch_bounds__(BLK) void evsub_kernel(u64 *a, const u64 *b, int rowwise, int rows, int cols) {
    size_t t = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (t >= (size_t)rows * cols) return;
    size_t row = t / (size_t)cols, col = t % (size_t)cols;
    a[t] -= b[rowwise ? col : row];
}

int element_vector_subtract(u64 *a, const u64 *b, int rowwise, int rows, int cols, hipStream_t s) {
    PIR_REQUIRE(rows >= 0 && cols >= 0, "elementVectorSubtract: negative dimension");
    size_t n = (size_t)rows * cols;
    if (n == 0) return OK;
    hipLaunchKernelGGL(evsub_kernel, dim3(g1(n)), dim3(BLK), 0, s, a, b, rowwise, rows, cols);
    PIR_LAUNCH_CHECK();
    return OK;
}

// matrix.cuh:142-155
template <typename T>
__global__ __launch_bounds__(BLK) void truncate_cols_kernel(const T *a, T *b, size_t rows, int sc, int ec) {
    size_t t = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (t >= rows * ec) return;
    size_t r = t / ec, c = t % ec;
    b[t] = a[r * sc + c];
}

int truncate_cols(const void *a, size_t a_n, void *b, int sc, int ec, int eb, hipStream_t s) {
    PIR_REQUIRE(sc > 0 && ec >= 0 && ec <= sc, "truncate_cols: invalid column counts");
    size_t rows = a_n / sc;
    if (rows * ec == 0) return OK;
    if (eb == 1) hipLaunchKernelGGL(truncate_cols_kernel<u8>, dim3(g1(rows * ec)), dim3(BLK), 0, s, (const u8 *)a, (u8 *)b, rows, sc, ec);
    else if (eb == 8) hipLaunchKernelGGL(truncate_cols_kernel<u64>, dim3(g1(rows * ec)), dim3(BLK), 0, s, (const u64 *)a, (u64 *)b, rows, sc, ec);
    else {
        set_error("truncate_cols: element size must be 1 or 8");
        return ERR_ARG;
    }
    PIR_LAUNCH_CHECK();
    return OK;
}

// convolution.cuh:176-191 -- each block of `stride` values gets `pad` pad values
__global__ __launch_bounds__(BLK) void stride_pad_kernel(const u64 *in, u64 *out, size_t blocks, int stride, int pad,
                                                         u64 pv) {
    size_t t = (size_t)blockIdx.x * BLK + threadIdx.x;
    size_t w = (size_t)stride + pad;
    if (t >= blocks * w) return;
    size_t j = t / w, i = t % w;
    out[t] = i < (size_t)stride ? in[j * stride + i] : pv;
}

int stride_pad(const u64 *in, u64 *out, size_t n_in, int stride, int pad, u64 pv, hipStream_t s) {
    PIR_REQUIRE(stride > 0 && pad >= 0, "stride_pad: invalid stride/pad");
    size_t blocks = n_in / stride;
    if (blocks == 0) return OK;
    hipLaunchKernelGGL(stride_pad_kernel, dim3(g1(blocks * (stride + pad))), dim3(BLK), 0, s, in, out, blocks, stride,
                       pad, pv);
    PIR_LAUNCH_CHECK();
    return OK;
}

// maxpool_im2row, convolution.cuh:57-108 / :252-304.  One thread per
// (image-channel, window); taps outside the image are skipped without
// advancing the output index, the rest of the pool stays as initialised.
__global__ __launch_bounds__(BLK) void im2row_kernel(const u64 *im, u64 *out, int kW, int kH, int f, int Din,
                                                     int batch, int stride, int padding, int wK, int hK, int padded,
                                                     u64 pv) {
    size_t windows = (size_t)wK * hK;
    size_t t = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (t >= windows * batch * Din) return;
    size_t wi = t % windows;
    int ic = (int)(t / windows);
    int bb = ic / Din, ch = ic % Din;
    int baseCol = (int)(wi / hK) * stride - padding;
    int baseRow = (int)(wi % hK) * stride - padding;
    size_t o = (size_t)bb * (windows * Din * padded) + wi * ((size_t)Din * padded) + (size_t)ch * padded;
    int f2 = f * f, written = 0;
    for (int c = 0; c < f; c++)
        for (int r = 0; r < f; r++) {
            int y = baseRow + r, x = baseCol + c;
            if (y < 0 || y >= kH || x < 0 || x >= kW) continue;
            size_t src = (size_t)bb * kH * kW * Din + (size_t)x * kH * Din + (size_t)y * Din + ch;
            out[o + written++] = im[src];
        }
    // the unpadded buffer is zero-initialised; stride_pad appends pad values
    for (int j = written; j < f2; j++) out[o + j] = 0;
    for (int j = f2; j < padded; j++) out[o + j] = pv;
}

size_t maxpool_im2row_size(int H, int W, int f, int Din, int batch, int stride, int padding) {
    int wK = ((H - f + 2 * padding) / stride) + 1, hK = ((W - f + 2 * padding) / stride) + 1;
    if (wK <= 0 || hK <= 0) return 0;
    int padded = 0, c = 0;
    while ((1 << c) < f * f) c++;
    padded = c * c;
    return (size_t)padded * wK * hK * batch * Din;
}

int maxpool_im2row(const u64 *im, u64 *out, int H, int W, int f, int Din, int batch, int stride, int padding, u64 pv,
                   hipStream_t s) {
    PIR_REQUIRE(f >= 2 && stride >= 1 && Din > 0 && batch > 0, "maxpool_im2row: invalid geometry");
    int kW = H, kH = W;  // the host passes (imageHeight, imageWidth) into the kernel's (width, height)
    int wK = ((kW - f + 2 * padding) / stride) + 1, hK = ((kH - f + 2 * padding) / stride) + 1;
    PIR_REQUIRE(wK > 0 && hK > 0, "maxpool_im2row: empty output");
    int c = 0;
    while ((1 << c) < f * f) c++;
    int padded = c * c;
    size_t threads = (size_t)wK * hK * batch * Din;
    hipLaunchKernelGGL(im2row_kernel, dim3(g1(threads)), dim3(BLK), 0, s, im, out, kW, kH, f, Din, batch, stride,
                       padding, wK, hK, padded, pv);
    PIR_LAUNCH_CHECK();
    return OK;
}

size_t averagepool_im2row_size(int imageWidth, int imageHeight, int f, int Din, int batch, int stride, int padding) {
    if (f < 1 || stride < 1) return 0;
    int wK = ((imageWidth - f + 2 * padding) / stride) + 1, hK = ((imageHeight - f + 2 * padding) / stride) + 1;
    if (wK <= 0 || hK <= 0 || Din <= 0 || batch <= 0) return 0;
    return (size_t)f * f * wK * hK * batch * Din;
}

int averagepool_im2row(const u64 *im, u64 *out, int imageWidth, int imageHeight, int f, int Din, int batch, int stride,
                       int padding, hipStream_t s) {
    PIR_REQUIRE(f >= 1 && stride >= 1 && Din > 0 && batch > 0 && padding >= 0, "averagepool_im2row: invalid geometry");
    // the host passes (imageHeight, imageWidth) into maxpool_im2row's (width, height)
    int kW = imageHeight, kH = imageWidth;
    int wK = ((kW - f + 2 * padding) / stride) + 1, hK = ((kH - f + 2 * padding) / stride) + 1;
    PIR_REQUIRE(wK > 0 && hK > 0, "averagepool_im2row: empty output");
    size_t threads = (size_t)wK * hK * batch * Din;
    hipLaunchKernelGGL(im2row_kernel, dim3(g1(threads)), dim3(BLK), 0, s, im, out, kW, kH, f, Din, batch, stride,
                       padding, wK, hK, f * f, (u64)0);
    PIR_LAUNCH_CHECK();
    return OK;
}

// averagepool_expand_delta, convolution.cuh:194-207 / :412-437: input entry
// k = b*Din + c (b over image-window pairs) fills out[b*Din*P + c + j*Din] for
// j < P.  Gathered: one thread per written output entry.
__global__ __launch_bounds__(BLK) void avgpool_expand_kernel(const u64 *in, u64 *out, size_t n_out, int Din, int P) {
    size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n_out) return;
    size_t b = i / ((size_t)Din * P);
    int c = (int)(i % (size_t)Din);
    out[i] = in[b * Din + c];
}

int averagepool_expand_delta(const u64 *in, size_t n_in, u64 *out, size_t out_capacity, int Din, int pool_size,
                             hipStream_t s) {
    PIR_REQUIRE(Din > 0 && pool_size > 0 && n_in % (size_t)Din == 0,
                "averagepool_expand_delta: the input must hold whole Din-channel groups");
    PIR_REQUIRE(n_in * (size_t)pool_size <= out_capacity,
                "averagepool_expand_delta: the output must hold input size x poolSize entries");
    size_t n_out = n_in * (size_t)pool_size;
    if (n_out == 0) return OK;
    hipLaunchKernelGGL(avgpool_expand_kernel, dim3(g1(n_out)), dim3(BLK), 0, s, in, out, n_out, Din, pool_size);
    PIR_LAUNCH_CHECK();
    return OK;
}

// matrix.cuh:109-126 -- b[col] += sum_i (reduceRows ? a[cols*i + col] : a[cols*col + i])
__global__ __launch_bounds__(BLK) void reduce_sum_kernel(const u64 *a, u64 *b, int reduceRows, int rows, int cols) {
    int col = blockIdx.x * BLK + threadIdx.x;
    int out = reduceRows ? cols : rows;
    if (col >= out) return;
    u64 acc = b[col];
    if (reduceRows)
        for (int i = 0; i < rows; i++) acc += a[(size_t)cols * i + col];
    else
        for (int i = 0; i < cols; i++) acc += a[(size_t)cols * col + i];
    b[col] = acc;
}

int reduce_sum(const u64 *a, u64 *b, int reduce_rows, int rows, int cols, hipStream_t s) {
    PIR_REQUIRE(rows >= 0 && cols >= 0, "reduceSum: negative dimension");
    int out = reduce_rows ? cols : rows;
    if (out == 0) return OK;
    hipLaunchKernelGGL(reduce_sum_kernel, dim3(g1(out)), dim3(BLK), 0, s, a, b, reduce_rows, rows, cols);
    PIR_LAUNCH_CHECK();
    return OK;
}

// matrix.cuh:128-139, one thread per output element
__global__ __launch_bounds__(BLK) void vector_expand_kernel(const u64 *a, u64 *b, size_t total, int f) {
    size_t t = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (t < total) b[t] = a[t / f];
}

int vector_expand(const u64 *a, size_t n, u64 *b, int factor, hipStream_t s) {
    PIR_REQUIRE(factor >= 1, "vectorExpand: factor must be >= 1");
    size_t total = n * (size_t)factor;
    if (total == 0) return OK;
    hipLaunchKernelGGL(vector_expand_kernel, dim3(g1(total)), dim3(BLK), 0, s, a, b, total, factor);
    PIR_LAUNCH_CHECK();
    return OK;
}

// convolution.cuh:110-151 gathered per output position: the reference adds
// rows[window * f*f + v] into output[b][valCol][valRow][c] with atomics; here
// each position sums the pool entries that land on it (the same sum mod 2^64)
// and adds it once, so no atomics are needed.  Kernel-side names: windowCol
// runs over widthKernels = (imageWidth - f)/stride + 1, windowRow over
// heightKernels = (imageHeight - f)/stride + 1, and
// outputIdx = b*H*W*Din + valCol*imageHeight*Din + valRow*Din + c.
__global__ __launch_bounds__(BLK) void row2im_kernel(const u64 *rows, u64 *out, int iH, int iW, int f, int Din,
                                                     int batch, int stride, int wK, int hK) {
    size_t total = (size_t)batch * iH * iW * Din;
    size_t t = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (t >= total) return;
    int c = (int)(t % Din);
    size_t r = t / Din;
    int valRow = (int)(r % iH);
    r /= iH;
    int valCol = (int)(r % iW);
    int b = (int)(r / iW);
    int wc_lo = valCol - f + 1 > 0 ? (valCol - f + 1 + stride - 1) / stride : 0;
    int wc_hi = min(wK - 1, valCol / stride);
    int wr_lo = valRow - f + 1 > 0 ? (valRow - f + 1 + stride - 1) / stride : 0;
    int wr_hi = min(hK - 1, valRow / stride);
    u64 acc = 0;
    for (int wc = wc_lo; wc <= wc_hi; wc++)
        for (int wr = wr_lo; wr <= wr_hi; wr++) {
            int v = (valCol - wc * stride) * f + (valRow - wr * stride);
            size_t win = (size_t)b * wK * hK * Din + ((size_t)wc * hK + wr) * Din + c;
            acc += rows[win * f * f + v];
        }
    out[t] += acc;
}

int maxpool_row2im(const u64 *rows, u64 *out, int imageHeight, int imageWidth, int f, int Din, int batch, int stride,
                   hipStream_t s) {
    PIR_REQUIRE(f >= 1 && stride >= 1 && Din > 0 && batch > 0 && imageHeight >= f && imageWidth >= f,
                "maxpool_row2im: invalid geometry");
    int wK = (imageWidth - f) / stride + 1, hK = (imageHeight - f) / stride + 1;
    size_t total = (size_t)batch * imageHeight * imageWidth * Din;
    hipLaunchKernelGGL(row2im_kernel, dim3(g1(total)), dim3(BLK), 0, s, rows, out, imageHeight, imageWidth, f, Din,
                       batch, stride, wK, hK);
    PIR_LAUNCH_CHECK();
    return OK;
}

}  // namespace pir
