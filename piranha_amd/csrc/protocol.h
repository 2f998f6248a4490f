// protocol.h -- 3-party RSS protocol functions with the reference's names and
// argument order (src/mpc/RSS.h:103-145), operating on all parties at once.
//
// Each function runs in one of two execution modes (Context::mode):
//   MODE_FUSED  one fused kernel per protocol step (fused_rss.hip) and the
//               int8-limb MFMA ring GEMM with the truncation fused into its
//               epilogue; needs contiguous share tensors.
//   MODE_PLAIN  the reference's op-by-op structure, one launch per DeviceData /
//               RSS operation, with uint8-per-bit dReLU shares and the plain
//               64-bit VALU GEMM (the baseline form).
// Both produce bit-identical shares.
#pragma once

#include "device_data.h"
#include "ring_gemm.h"

namespace pir {

typedef RSS<u64> RSS64;

// cutlass::conv::Operator (RSS.inl:705-790)
enum ConvOperator { kFprop = 0, kDgrad = 1, kWgrad = 2 };

template <typename T>
void reconstruct(const RSS<T> &in, PartyData<T> &out);
template <typename T>
void reshare(const PartyData<T> &c, RSS<T> &out);
void reconstruct3out3(const PartyData<u64> &in, PartyData<u64> &out);
void dividePublic(RSS64 &a, u64 denominator);
void dividePublic(RSS64 &a, PartyData<u64> &denominators);
void dividePublic3out3(PartyData<u64> &a, u64 denominator, RSS64 &out);
void localMatMul(const RSS64 &a, const RSS64 &b, PartyData<u64> &c, int M, int N, int K, bool transpose_a,
                 bool transpose_b, bool transpose_c);
void matmul(const RSS64 &a, const RSS64 &b, RSS64 &c, int M, int N, int K, bool transpose_a, bool transpose_b,
            bool transpose_c, u64 truncation);
void localFprop(const RSS64 &A, const RSS64 &B, PartyData<u64> &C, int batchSize, int imageHeight, int imageWidth,
                int Din, int Dout, int filterHeight, int filterWidth, int paddingHeight, int paddingWidth, int stride,
                int dilation);
// RSS.inl:641-703
void localDgrad(const RSS64 &A, const RSS64 &B, PartyData<u64> &C, int batchSize, int outputHeight, int outputWidth,
                int Dout, int filterHeight, int filterWidth, int Din, int paddingHeight, int paddingWidth, int stride,
                int dilation, int imageHeight, int imageWidth);
void localWgrad(const RSS64 &A, const RSS64 &B, PartyData<u64> &C, int batchSize, int outputHeight, int outputWidth,
                int Dout, int imageHeight, int imageWidth, int Din, int filterHeight, int filterWidth,
                int paddingHeight, int paddingWidth, int stride, int dilation);
// kFprop: A = input, B = filter, C = output; kDgrad: A = delta, B = filter,
// C = input gradient; kWgrad: A = delta, B = forward input, C = filter gradient.
void convolution(const RSS64 &A, const RSS64 &B, RSS64 &C, ConvOperator op, int batchSize, int imageHeight,
                 int imageWidth, int filterSize, int Din, int Dout, int stride, int padding, int truncation);
template <typename U>
void dReLU(const RSS64 &input, RSS<U> &result);
template <typename U>
void ReLU(const RSS64 &input, RSS64 &result, RSS<U> &dresult);
template <typename U>
void selectShare(const RSS64 &x, const RSS64 &y, const RSS<U> &b, RSS64 &z);
template <typename U>
void maxpool(RSS64 &input, RSS64 &result, RSS<U> &dresult, int k);
template <typename T>
void carryOut(RSS<T> &p, RSS<T> &g, int k, RSS<T> &out);

// getPowers (RSS.inl:1211-1235): each party's opened power, the largest bit b
// with in - 2^b >= 0 as a signed value (0 when there is none).
void getPowers(const RSS64 &in, PartyData<u64> &pow);
// taylorSeries (RSS.inl:1237-1320): a0 + a1 x + a2 x^2 on x range-reduced by
// each element's power, rescaled by the functor fn at FP = fp fractional bits.
// `out` must not overlap `in`.
enum TaylorFn { TAYLOR_SQRT = 0, TAYLOR_INVERSE = 1, TAYLOR_SIGMOID = 2 };
TaylorTables taylor_tables(double a0, double a1, double a2, int fn, int fp);
void taylorSeries(const RSS64 &in, RSS64 &out, double a0, double a1, double a2, int fn, int fp);
// sqrt / inverse / sigmoid (RSS.inl:570-612).  The reference's sqrt is
// rss_sqrt here so that ::sqrt stays visible inside namespace pir.
void rss_sqrt(const RSS64 &in, RSS64 &out, int fp);
void inverse(const RSS64 &in, RSS64 &out, int fp);
void sigmoid(const RSS64 &in, RSS64 &out, int fp);

template <typename T>
void rss_mult_impl(RSS<T> &x, const RSS<T> &y);
template <typename T>
void rss_and_impl(RSS<T> &x, const RSS<T> &y);

template <typename T>
RSS<T> &RSS<T>::operator*=(const RSS<T> &r) {
    rss_mult_impl(*this, r);
    return *this;
}
template <typename T>
RSS<T> &RSS<T>::operator&=(const RSS<T> &r) {
    rss_and_impl(*this, r);
    return *this;
}

// Limb-GEMM helpers shared by the protocol and the layers.
// Packs the RSS weight operand of all parties into the 6-segment K-concatenated
// limb form [B_p0 + B_p1 ; B_p0] for p = A, B, C.
int pack_rss_weights(const RSS64 &B, int K, int N, bool b_row_major, int8_t *packed);
size_t packed_rss_weights_bytes(int K, int N);
// Same 6-segment packing with the operand geometry given by `geom` (K, N, Kp,
// row_major or the dgrad filter view; the share pointers are filled in here).
int pack_rss_operand(const RSS64 &B, const LimbBSrc &geom, int8_t *packed);
// Fused 3-party product over an A view `a` (geometry and conv form set by the
// caller; the six share segments are filled in here from A) into the
// epilogue `e` (mode EPI_TRUNC with the RSS output layout).
int rss_limb_run(const RSS64 &A, LimbA a, const int8_t *packed, int N, const Epilogue &e, int split = 0);
// The 3out3 output epilogue of the fused product: out [3][2][M*N] receives
// A(1+q, 0), B(0, 0), C(0, 1+q).
Epilogue rss_trunc_epi(u64 *out, int trunc_bits, bool tc);
// convolution(kDgrad / kWgrad) in the fused limb form (one 6-segment product
// with the dividePublic3out3 epilogue); for kDgrad, bw_relu_d / bw_relu_out
// (optional) add the preceding ReLU layer's backward to the epilogue.
void rss_conv_grad_limb(ConvOperator op, const RSS64 &A, const RSS64 &B, RSS64 &C, int batch, int H, int W, int f,
                        int Din, int Dout, int stride, int pad, int trunc, const u8 *bw_relu_d = nullptr,
                        u64 *bw_relu_out = nullptr);
// Geometry of the dgrad / wgrad limb views of a convolution layer.
LimbA limb_dgrad_view(int batch, int H, int W, int f, int Din, int Dout, int stride, int pad);
LimbA limb_wgrad_view(int batch, int H, int W, int f, int Din, int Dout, int stride, int pad);
// Runs the fused 3-party product + truncation (+ bias) on packed weights.
// conv == nullptr: dense A (row-major if ta).  out layout row-major if tc.
struct ConvShape {
    int batch, H, W, Cin, f, stride, pad, OH, OW;
};
// relu_out / relu_d (optional, contiguous [3][2][M*N]): the following ReLU
// layer evaluated on the product's output in the same epilogue.
// g / gr: the masked-mode streams of the product and of the fused ReLU (off in
// the placeholder mode).
int rss_limb_product(const RSS64 &A, bool ta, const ConvShape *conv, int M, int N, int K, const int8_t *packed,
                     int trunc_bits, const RSS64 *bias, RSS64 &out, bool tc, int split = 0, u64 *relu_out = nullptr,
                     u8 *relu_d = nullptr, const PrgStream &g = PrgStream{}, const PrgStream &gr = PrgStream{});

}  // namespace pir
