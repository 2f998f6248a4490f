// fused_rss.h -- fused 3-party RSS protocol kernels (host launchers).
// All share tensors are [3 parties][2 shares][n] arrays; party-local tensors
// are [3][n].  Every launcher is asynchronous on the given stream.
#pragma once

#include "common.h"
#include "prg.h"

namespace pir {

struct MaxpoolGeom {
    int batch, H, W, C, f, stride, OH, OW;
    u64 pad_value;
};

int expanded_pool_size(int f);

// Launchers taking a PrgStream run the masked precompute mode when it is on
// (rss_masked.h; element i of the call draws from its own word range) and the
// reference's placeholder constants otherwise.
template <typename U>
int fused_drelu(const u64 *x, size_t n, U *dres, hipStream_t s, const PrgStream &g = PrgStream{});
template <typename U>
int fused_relu(const u64 *x, size_t n, u64 *y, U *dres, hipStream_t s, const PrgStream &g = PrgStream{});
template <typename U>
int fused_select(const u64 *x, const u64 *y, const U *b, u64 *z, size_t n, hipStream_t s,
                 const PrgStream &g = PrgStream{});
template <typename U>
int fused_maxpool_pools(u64 *pools, size_t n, u64 *result, U *dres, int k, hipStream_t s,
                        const PrgStream &g = PrgStream{});
// relu_out / relu_d (optional): the following ReLU layer computed on the pooled
// values in the same pass (ReLU output shares and dReLU bits, [3][2][outN]);
// gp is the pool's stream, gr the ReLU layer's.
template <typename U>
int fused_maxpool_layer(const u64 *x, const MaxpoolGeom &g, u64 *act, U *mprime, hipStream_t s,
                        u64 *relu_out = nullptr, u8 *relu_d = nullptr, const PrgStream &gp = PrgStream{},
                        const PrgStream &gr = PrgStream{});
template <typename T>
int fused_reconstruct(const T *x, size_t n, T *out, hipStream_t s);
int fused_reshare(const u64 *c, size_t n, u64 *out, hipStream_t s, const PrgStream &g = PrgStream{});
int fused_rss_mult(u64 *x, const u64 *y, size_t n, hipStream_t s, const PrgStream &g = PrgStream{});
template <typename T>
int fused_rss_and(T *x, const T *y, size_t n, hipStream_t s, const PrgStream &g = PrgStream{});
int fused_divide_public(u64 *a, size_t n, u64 d, const u64 *dv, hipStream_t s, const PrgStream &g = PrgStream{});
template <typename U>
int fused_convex_comb(u64 *d, const u64 *c, const U *e, size_t n, hipStream_t s);
int fused_divide_3out3(const u64 *a, size_t n, int k, u64 *out, hipStream_t s, const PrgStream &g = PrgStream{});
// getPowers (RSS.inl:1211-1235): in [3][2][n] -> each party's opened power [3][n].
int fused_get_powers(const u64 *in, size_t n, u64 *pow, hipStream_t s);
// taylorSeries (RSS.inl:1237-1320) in one pass: in [3][2][n] -> out [3][2][n]
// (out must not overlap in); t holds the tabulated coefficient functors.
int fused_taylor(const u64 *in, size_t n, u64 *out, const TaylorTables &t, hipStream_t s);

// ---- backward (training) path ----
// ReLULayer::backward (ReLULayer.cu:91-123): out = selectShare(0, delta, reluPrime).
template <typename U>
int fused_relu_backward(const u64 *delta, const U *relu_prime, size_t n, u64 *out, hipStream_t s);
// AveragepoolLayer::forward (AveragepoolLayer.cu:49-96) in one pass (pad_value
// of g unused), optionally with the following ReLU layer.
int fused_avgpool_layer(const u64 *x, const MaxpoolGeom &g, u64 *act, hipStream_t s, u64 *relu_out = nullptr,
                        u8 *relu_d = nullptr);
// AveragepoolLayer::backward (AveragepoolLayer.cu:98-125): delta [3][2][n_in]
// -> out [3][2][n_out] (n_out >= n_in * P).
int fused_avgpool_backward(const u64 *delta, size_t n_in, int C, int P, u64 *out, size_t n_out, hipStream_t s);
// MaxpoolLayer::backward (MaxpoolLayer.cu:133-180): vectorExpand, selectShare
// with maxPrime and maxpool_row2im in one pass, gathered per input element.
template <typename U>
int fused_maxpool_backward(const u64 *delta, const U *max_prime, const MaxpoolGeom &g, u64 *out, hipStream_t s);
// FCLayer.cu:205-218: db = reduceSum(delta, false, outDim, batch) per share,
// dividePublic(db, 2^lr), biases -= db.
int fused_fc_bias_update(const u64 *delta, u64 *bias, int out_dim, int batch, int lr, hipStream_t s);
// NeuralNetwork::_reveal_softmax_grad (NeuralNetwork.cu:207-267), the device
// halves around the host softmax: parties A and C open the logits
// (out [2][n]), and deltas = (A.sA, C.sB = softmax values [2][n]) - labels,
// then dividePublic by 2^lmb when lmb > 1.
int fused_open_logits_ac(const u64 *act, size_t n, u64 *out, hipStream_t s);
int fused_softmax_delta(const u64 *sm, const u64 *labels, size_t n, int lmb, u64 *out, hipStream_t s);

}  // namespace pir
