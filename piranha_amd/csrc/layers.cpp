// layers.cpp -- FC / CNN / ReLU / Maxpool layers (forward and backward) and the network.
#include "layers.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "elementwise.h"
#include "fused_rss.h"

namespace pir {

// (uint64)f as the reference's GPU converts: clamped to the uint64 range,
// NaN to 0 (to_fixed_functor, inv_sqrt_fixed_point_functor).
static u64 to_u64_clamped(double f) {
    if (!(f > 0.0)) return 0;  // NaN, zero or negative
    if (f >= 18446744073709551616.0) return ~0ull;
    return (u64)f;
}


static bool fused_limb() { return ctx().mode == MODE_FUSED && ctx().algo == ALGO_LIMB; }

// ---------------------------------------------------------------------------
// FC (FCLayer.cu:20-130)

FCLayer::FCLayer(const FCConfig &c, int n, int seed)
    : Layer(n), conf(c), activations(c.batchSize * c.outputDim), weights(c.inputDim * c.outputDim),
      biases(c.outputDim) {
    initialize(seed);
}

// Kaiming initialisation with std::default_random_engine(seed) and bias 0.01
// (FCLayer.cu:31-48).
void FCLayer::initialize(int seed) {
    std::default_random_engine generator(seed);
    double std_dev = std::sqrt((double)2 / (double)conf.inputDim);
    std::normal_distribution<double> distribution(0.0, std_dev);
    std::vector<double> w(weights.size());
    for (size_t i = 0; i < w.size(); i++) w[i] = distribution(generator);
    weights.setPublic(w, ctx().frac_bits);
    std::vector<double> b(biases.size(), 0.01);
    biases.setPublic(b, ctx().frac_bits);
    packed_valid = false;
}

void FCLayer::printLayer() {
    printf("----------------------------------------------\n(%d) FC Layer\t\t  %zu x %zu\n\t\t\t  %zu\t\t (Batch Size)\n",
           layerNum + 1, conf.inputDim, conf.outputDim, conf.batchSize);
}

// matmul + dividePublic3out3 + elementVectorAdd of both bias shares (and the
// next ReLU layer when relu_out is given), one limb-GEMM launch
void FCLayer::limbForward(const RSS64 &input, u64 *relu_out, u8 *relu_d, const PrgStream &g, const PrgStream &gr) {
    const int M = (int)conf.batchSize, N = (int)conf.outputDim, K = (int)conf.inputDim;
    if (!packed_valid) {
        // repacked in place: a captured graph keeps pointing at the same buffer
        if (packed.bytes() != packed_rss_weights_bytes(K, N)) packed = DevBuf(packed_rss_weights_bytes(K, N));
        check(pack_rss_weights(weights, K, N, true, (int8_t *)packed.get()));
        packed_valid = true;
    }
    check(rss_limb_product(input, true, nullptr, M, N, K, (const int8_t *)packed.get(), ctx().frac_bits, &biases,
                           activations, true, 0, relu_out, relu_d, g, gr));
}

// In the masked precompute mode the product draws from this layer's call and
// the fused ReLU from the ReLU layer's (the next call number).
bool FCLayer::forwardWithReLU(const RSS64 &input, ReLULayer &relu) {
    if (!(fused_limb() && input.base() && relu.activations.base() && relu.reluPrime.base())) return false;
    const PrgStream g = call_prg("FC forward"), gr = call_prg("ReLU forward");
    limbForward(input, relu.activations.base(), relu.reluPrime.base(), g, gr);
    return true;
}

void FCLayer::forward(const RSS64 &input) {
    const int M = (int)conf.batchSize, N = (int)conf.outputDim, K = (int)conf.inputDim;
    const int fp = ctx().frac_bits;
    if (fused_limb() && input.base()) {
        limbForward(input, nullptr, nullptr, call_prg("FC forward"), PrgStream{});
        return;
    }
    activations.zero();
    matmul(input, weights, activations, M, N, K, true, true, true, (u64)fp);
    for (int q = 0; q < 3; q++)
        for (int i = 0; i < 2; i++)
            check(element_vector_add(activations.getShare(q, i)->data(), biases.getShare(q, i)->data(), 0, N, M,
                                     ctx().stream));
}

// FCLayer::backward, FCLayer.cu:133-224: deltas = delta . W^T (truncated by
// 2^FP), W -= delta^T . X (truncated by 2^(FP + log_lr)), and the bias update
// of reduceSum / dividePublic.  The fused form runs the two products as limb
// GEMMs, the second with the subtraction in its epilogue, and the bias update
// as one kernel.
void FCLayer::backward(const RSS64 &delta, const RSS64 &forwardInput) {
    require_placeholder("layer backward (training)");
    const int B = (int)conf.batchSize, in = (int)conf.inputDim, out = (int)conf.outputDim;
    const int fp = ctx().frac_bits, lr = ctx().log_lr;
    ensureDeltas();
    if (fused_limb() && delta.base() && forwardInput.base()) {
        limbBackward(delta, forwardInput, nullptr);
        return;
    }
    deltas.zero();
    matmul(delta, weights, deltas, B, in, out, true, false, true, (u64)fp);
    RSS64 dW((size_t)out * in);
    matmul(delta, forwardInput, dW, out, in, B, false, true, false, (u64)(fp + lr));
    weights -= dW;
    RSS64 db((size_t)out);
    for (int q = 0; q < 3; q++)
        for (int i = 0; i < 2; i++)
            check(reduce_sum(delta.getShare(q, i)->data(), db.getShare(q, i)->data(), 0, out, B, ctx().stream));
    dividePublic(db, (u64)1 << lr);
    biases -= db;
    packed_valid = false;
}

// The fused form of FCLayer::backward; relu (optional) is the ReLU layer
// before this one, whose backward runs in the deltas product's epilogue.
void FCLayer::limbBackward(const RSS64 &delta, const RSS64 &forwardInput, ReLULayer *relu) {
    const int B = (int)conf.batchSize, in = (int)conf.inputDim, out = (int)conf.outputDim;
    const int fp = ctx().frac_bits, lr = ctx().log_lr;
    // (1) deltas: M = batch, N = in, K = out; B(k = o, n = i) = W[i*out + o]
    DevBuf pw(packed_rss_weights_bytes(out, in));
    check(pack_rss_weights(weights, out, in, false, (int8_t *)pw.get()));
    LimbA a1{};
    a1.M = B;
    a1.K = out;
    a1.Kp = round_up(out, LIMB_BK);
    a1.lda = out;
    a1.row_major = 1;
    Epilogue e1 = rss_trunc_epi(deltas.base(), fp, true);
    if (relu) {
        e1.bw_relu_d = relu->reluPrime.base();
        e1.bw_relu_out = relu->getDelta()->base();
    }
    check(rss_limb_run(delta, a1, (const int8_t *)pw.get(), in, e1));
    // (2) M = out, N = in, K = batch; A(m = o, k = b) = delta[b*out + o], B(k, n) = X[b*in + i],
    // the M x N result column-major is W's [in][out] layout
    DevBuf px(packed_rss_weights_bytes(B, in));
    check(pack_rss_weights(forwardInput, B, in, true, (int8_t *)px.get()));
    LimbA a{};
    a.M = out;
    a.K = B;
    a.Kp = round_up(B, LIMB_BK);
    a.lda = out;
    a.row_major = 0;
    Epilogue e = rss_trunc_epi(weights.base(), fp + lr, false);
    e.sub_out = 1;
    check(rss_limb_run(delta, a, (const int8_t *)px.get(), in, e));
    // (3)
    check(fused_fc_bias_update(delta.base(), biases.base(), out, B, lr, ctx().stream));
    packed_valid = false;
}

bool FCLayer::backwardWithReLU(const RSS64 &delta, const RSS64 &forwardInput, ReLULayer &relu) {
    if (!(fused_limb() && delta.base() && forwardInput.base() && relu.reluPrime.base() &&
          relu.inputSize() == inputSize()))
        return false;
    ensureDeltas();
    relu.allocDelta();
    limbBackward(delta, forwardInput, &relu);
    return true;
}

// ---------------------------------------------------------------------------
// CNN (CNNLayer.cu:23-115); weights are [Dout][f][f][Din], no bias.

CNNLayer::CNNLayer(const CNNConfig &c, int n, int seed)
    : Layer(n), conf(c), weights(c.filterSize * c.filterSize * c.inputFeatures * c.outputFeatures),
      activations(c.batchSize * c.outputFeatures *
                  ((c.imageWidth - c.filterSize + 2 * c.padding) / c.stride + 1) *
                  ((c.imageHeight - c.filterSize + 2 * c.padding) / c.stride + 1)) {
    initialize(seed);
}

void CNNLayer::initialize(int seed) {
    std::default_random_engine generator(seed);
    double std_dev = std::sqrt((double)2 / (double)(conf.filterSize * conf.filterSize * conf.inputFeatures));
    std::normal_distribution<double> distribution(0.0, std_dev);
    std::vector<double> w(weights.size());
    for (size_t i = 0; i < w.size(); i++) w[i] = distribution(generator);
    weights.setPublic(w, ctx().frac_bits);
    packed_valid = false;
}

void CNNLayer::printLayer() {
    printf("----------------------------------------------\n(%d) CNN Layer\t\t  %zu x %zu x %zu\n\t\t\t  %zu x %zu  \t(Filter Size)\n"
           "\t\t\t  %zu , %zu \t(Stride, padding)\n\t\t\t  %zu\t\t(Batch Size)\n\t\t\t  %d x %d x %zu \t(Output)\n",
           layerNum + 1, conf.imageHeight, conf.imageWidth, conf.inputFeatures, conf.filterSize, conf.filterSize,
           conf.stride, conf.padding, conf.batchSize, outW(), outH(), conf.outputFeatures);
}

// implicit-GEMM convolution + dividePublic3out3 (and the next ReLU layer when
// relu_out is given), one limb-GEMM launch
void CNNLayer::limbForward(const RSS64 &input, u64 *relu_out, u8 *relu_d, const PrgStream &g, const PrgStream &gr) {
    const int f = (int)conf.filterSize, Din = (int)conf.inputFeatures, Dout = (int)conf.outputFeatures;
    const int K = f * f * Din;
    if (!packed_valid) {
        if (packed.bytes() != packed_rss_weights_bytes(K, Dout)) packed = DevBuf(packed_rss_weights_bytes(K, Dout));
        check(pack_rss_weights(weights, K, Dout, false, (int8_t *)packed.get()));
        packed_valid = true;
    }
    ConvShape cs{(int)conf.batchSize, (int)conf.imageHeight, (int)conf.imageWidth, Din, f, (int)conf.stride,
                 (int)conf.padding, outH(), outW()};
    check(rss_limb_product(input, true, &cs, cs.batch * cs.OH * cs.OW, Dout, K, (const int8_t *)packed.get(),
                           ctx().frac_bits, nullptr, activations, true, 0, relu_out, relu_d, g, gr));
}

bool CNNLayer::forwardWithReLU(const RSS64 &input, ReLULayer &relu) {
    if (!(fused_limb() && input.base() && relu.activations.base() && relu.reluPrime.base())) return false;
    const PrgStream g = call_prg("CNN forward"), gr = call_prg("ReLU forward");
    limbForward(input, relu.activations.base(), relu.reluPrime.base(), g, gr);
    return true;
}

void CNNLayer::forward(const RSS64 &input) {
    const int fp = ctx().frac_bits;
    const int f = (int)conf.filterSize, Din = (int)conf.inputFeatures, Dout = (int)conf.outputFeatures;
    if (fused_limb() && input.base()) {
        limbForward(input, nullptr, nullptr, call_prg("CNN forward"), PrgStream{});
        return;
    }
    activations.zero();
    convolution(input, weights, activations, kFprop, (int)conf.batchSize, (int)conf.imageHeight,
                (int)conf.imageWidth, f, Din, Dout, (int)conf.stride, (int)conf.padding, fp);
}

// CNNLayer::backward, CNNLayer.cu:118-195: deltas = convolution(kDgrad),
// dF = convolution(kWgrad), dividePublic(dF, 2^log_lr), W -= dF (the
// reference's bias update is commented out).  The fused form runs dgrad and
// wgrad as limb GEMMs, the dividePublic and the subtraction in wgrad's epilogue.
void CNNLayer::backward(const RSS64 &delta, const RSS64 &forwardInput) {
    require_placeholder("layer backward (training)");
    const int f = (int)conf.filterSize, Din = (int)conf.inputFeatures, Dout = (int)conf.outputFeatures;
    const int B = (int)conf.batchSize, H = (int)conf.imageHeight, W = (int)conf.imageWidth;
    const int st = (int)conf.stride, pad = (int)conf.padding, fp = ctx().frac_bits, lr = ctx().log_lr;
    ensureDeltas();
    if (fused_limb() && delta.base() && forwardInput.base()) {
        limbBackward(delta, forwardInput, nullptr);
        return;
    }
    deltas.zero();
    convolution(delta, weights, deltas, kDgrad, B, H, W, f, Din, Dout, st, pad, fp);
    RSS64 dF(weights.size());
    convolution(delta, forwardInput, dF, kWgrad, B, H, W, f, Din, Dout, st, pad, fp);
    dividePublic(dF, (u64)1 << lr);
    weights -= dF;
    packed_valid = false;
}

// The fused form of CNNLayer::backward; relu (optional) is the ReLU layer
// before this one, whose backward runs in the dgrad product's epilogue.
void CNNLayer::limbBackward(const RSS64 &delta, const RSS64 &forwardInput, ReLULayer *relu) {
    const int f = (int)conf.filterSize, Din = (int)conf.inputFeatures, Dout = (int)conf.outputFeatures;
    const int B = (int)conf.batchSize, H = (int)conf.imageHeight, W = (int)conf.imageWidth;
    const int st = (int)conf.stride, pad = (int)conf.padding, fp = ctx().frac_bits, lr = ctx().log_lr;
    rss_conv_grad_limb(kDgrad, delta, weights, deltas, B, H, W, f, Din, Dout, st, pad, fp,
                       relu ? relu->reluPrime.base() : nullptr, relu ? relu->getDelta()->base() : nullptr);
    LimbA a = limb_wgrad_view(B, H, W, f, Din, Dout, st, pad);
    LimbBSrc g{};
    g.K = a.K;
    g.N = Dout;
    g.Kp = a.Kp;
    g.row_major = 1;
    DevBuf packed(limb_b_bytes(Dout, 6, g.Kp));
    check(pack_rss_operand(delta, g, (int8_t *)packed.get()));
    Epilogue e = rss_trunc_epi(weights.base(), fp, false);
    e.div_bits = lr;
    e.sub_out = 1;
    check(rss_limb_run(forwardInput, a, (const int8_t *)packed.get(), Dout, e));
    packed_valid = false;
}

bool CNNLayer::backwardWithReLU(const RSS64 &delta, const RSS64 &forwardInput, ReLULayer &relu) {
    if (!(fused_limb() && delta.base() && forwardInput.base() && relu.reluPrime.base() &&
          relu.inputSize() == inputSize()))
        return false;
    ensureDeltas();
    relu.allocDelta();
    limbBackward(delta, forwardInput, &relu);
    return true;
}

// ---------------------------------------------------------------------------
// ReLU (ReLULayer.cu:21-88)

ReLULayer::ReLULayer(const ReLUConfig &c, int n, int)
    : Layer(n), conf(c), activations(c.batchSize * c.inputDim), reluPrime(c.batchSize * c.inputDim) {}

void ReLULayer::printLayer() {
    printf("----------------------------------------------\n(%d) ReLU Layer\t\t  %zu x %zu\n", layerNum + 1,
           conf.batchSize, conf.inputDim);
}

void ReLULayer::forward(const RSS64 &input) {
    if (ctx().mode != MODE_FUSED) activations.zero();
    ReLU(input, activations, reluPrime);
}

// ReLULayer::backward, ReLULayer.cu:91-123: deltas = selectShare(0, delta, reluPrime)
void ReLULayer::backward(const RSS64 &delta, const RSS64 &) {
    require_placeholder("layer backward (training)");
    ensureDeltas();
    if (ctx().mode == MODE_FUSED && delta.base()) {
        check(fused_relu_backward<u8>(delta.base(), reluPrime.base(), delta.size(), deltas.base(), ctx().stream));
        return;
    }
    deltas.zero();
    RSS64 zeros(delta.size());
    selectShare(zeros, delta, reluPrime, deltas);
}

// ---------------------------------------------------------------------------
// Maxpool (MaxpoolLayer.cu:17-131)

MaxpoolLayer::MaxpoolLayer(const MaxpoolConfig &c, int n, int)
    : Layer(n), conf(c),
      activations(c.batchSize * c.features * ((c.imageWidth - c.poolSize) / c.stride + 1) *
                  ((c.imageHeight - c.poolSize) / c.stride + 1)),
      maxPrime(((c.imageWidth - c.poolSize) / c.stride + 1) * ((c.imageHeight - c.poolSize) / c.stride + 1) *
               c.features * c.batchSize * c.poolSize * c.poolSize) {}

void MaxpoolLayer::printLayer() {
    printf("----------------------------------------------\n(%d) Maxpool Layer\t  %zu x %zu x %zu\n\t\t\t  %zu  \t\t(Pooling Size)\n"
           "\t\t\t  %zu \t\t(Stride)\n\t\t\t  %zu\t\t(Batch Size)\n",
           layerNum + 1, conf.imageHeight, conf.imageWidth, conf.features, conf.poolSize, conf.stride, conf.batchSize);
}

bool MaxpoolLayer::forwardWithReLU(const RSS64 &input, ReLULayer &relu) {
    if (!(ctx().mode == MODE_FUSED && input.base() && relu.activations.base() && relu.reluPrime.base()))
        return false;
    const u64 pad_value = (u64)(s64)(-10LL * (1LL << ctx().frac_bits));
    MaxpoolGeom g{(int)conf.batchSize, (int)conf.imageHeight, (int)conf.imageWidth, (int)conf.features,
                  (int)conf.poolSize, (int)conf.stride, outH(), outW(), pad_value};
    const PrgStream gp = call_prg("maxpool forward"), gr = call_prg("ReLU forward");
    check(fused_maxpool_layer<u8>(input.base(), g, activations.base(), maxPrime.base(), ctx().stream,
                                  relu.activations.base(), relu.reluPrime.base(), gp, gr));
    return true;
}

void MaxpoolLayer::forward(const RSS64 &input) {
    const int fp = ctx().frac_bits;
    // (T)(-10 * (1 << FLOAT_PRECISION)) in every share (RSS has two shares, so the
    // 4-share special case of MaxpoolLayer.cu:77-92 never applies)
    const u64 pad_value = (u64)(s64)(-10LL * (1LL << fp));
    const int f = (int)conf.poolSize;
    if (ctx().mode == MODE_FUSED && input.base()) {
        MaxpoolGeom g{(int)conf.batchSize, (int)conf.imageHeight, (int)conf.imageWidth, (int)conf.features, f,
                      (int)conf.stride, outH(), outW(), pad_value};
        check(fused_maxpool_layer<u8>(input.base(), g, activations.base(), maxPrime.base(), ctx().stream, nullptr,
                                      nullptr, call_prg("maxpool forward")));
        return;
    }
    activations.zero();
    maxPrime.zero();
    int eps = expanded_pool_size(f);
    size_t poolsN = maxpool_im2row_size((int)conf.imageHeight, (int)conf.imageWidth, f, (int)conf.features,
                                        (int)conf.batchSize, (int)conf.stride, 0);
    RSS64 pools(poolsN);
    for (int q = 0; q < 3; q++)
        for (int i = 0; i < 2; i++)
            check(maxpool_im2row(input.getShare(q, i)->data(), pools.getShare(q, i)->data(), (int)conf.imageHeight,
                                 (int)conf.imageWidth, f, (int)conf.features, (int)conf.batchSize, (int)conf.stride, 0,
                                 pad_value, ctx().stream));
    RSS<u8> emp(pools.size());
    maxpool(pools, activations, emp, eps);
    for (int q = 0; q < 3; q++)
        for (int i = 0; i < 2; i++)
            check(truncate_cols(emp.getShare(q, i)->data(), emp.size(), maxPrime.getShare(q, i)->data(), eps, f * f, 1,
                                ctx().stream));
}

// MaxpoolLayer::backward, MaxpoolLayer.cu:133-180: vectorExpand of delta over
// each pool, selectShare with maxPrime, maxpool_row2im into deltas.
void MaxpoolLayer::backward(const RSS64 &delta, const RSS64 &) {
    require_placeholder("layer backward (training)");
    ensureDeltas();
    const int f = (int)conf.poolSize;
    if (ctx().mode == MODE_FUSED && delta.base()) {
        MaxpoolGeom g{(int)conf.batchSize, (int)conf.imageHeight, (int)conf.imageWidth, (int)conf.features, f,
                      (int)conf.stride, outH(), outW(), 0};
        check(fused_maxpool_backward<u8>(delta.base(), maxPrime.base(), g, deltas.base(), ctx().stream));
        return;
    }
    deltas.zero();
    RSS64 deltaPools(maxPrime.size());
    for (int q = 0; q < 3; q++)
        for (int i = 0; i < 2; i++)
            check(vector_expand(delta.getShare(q, i)->data(), delta.size(), deltaPools.getShare(q, i)->data(), f * f,
                                ctx().stream));
    RSS64 zeros(deltaPools.size());
    selectShare(zeros, deltaPools, maxPrime, deltaPools);
    // the layer passes (imageWidth, imageHeight) into row2im's (imageHeight, imageWidth)
    for (int q = 0; q < 3; q++)
        for (int i = 0; i < 2; i++)
            check(maxpool_row2im(deltaPools.getShare(q, i)->data(), deltas.getShare(q, i)->data(),
                                 (int)conf.imageWidth, (int)conf.imageHeight, f, (int)conf.features,
                                 (int)conf.batchSize, (int)conf.stride, ctx().stream));
}

// ---------------------------------------------------------------------------
// AveragepoolLayer.cu

AveragepoolLayer::AveragepoolLayer(const AveragepoolConfig &c, int n, int)
    : Layer(n), conf(c),
      activations(c.batchSize * c.features * ((c.imageWidth - c.poolSize) / c.stride + 1) *
                  ((c.imageHeight - c.poolSize) / c.stride + 1)) {}

void AveragepoolLayer::printLayer() {
    printf("----------------------------------------------\n(%d) Averagepool Layer\t  %zu x %zu x %zu\n\t\t\t  %zu  \t\t(Pooling Size)\n"
           "\t\t\t  %zu \t\t(Stride)\n\t\t\t  %zu\t\t(Batch Size)\n",
           layerNum + 1, conf.imageHeight, conf.imageWidth, conf.features, conf.poolSize, conf.stride, conf.batchSize);
}

MaxpoolGeom AveragepoolLayer::geom() const {
    return MaxpoolGeom{(int)conf.batchSize, (int)conf.imageHeight, (int)conf.imageWidth, (int)conf.features,
                       (int)conf.poolSize, (int)conf.stride, outH(), outW(), 0};
}

bool AveragepoolLayer::forwardWithReLU(const RSS64 &input, ReLULayer &relu) {
    require_placeholder("Averagepool layer");
    if (!(ctx().mode == MODE_FUSED && input.base() && relu.activations.base() && relu.reluPrime.base()))
        return false;
    check(fused_avgpool_layer(input.base(), geom(), activations.base(), ctx().stream, relu.activations.base(),
                              relu.reluPrime.base()));
    return true;
}

// AveragepoolLayer.cu:49-96
void AveragepoolLayer::forward(const RSS64 &input) {
    require_placeholder("Averagepool layer");
    const int f = (int)conf.poolSize;
    if (ctx().mode == MODE_FUSED && input.base()) {
        check(fused_avgpool_layer(input.base(), geom(), activations.base(), ctx().stream));
        return;
    }
    activations.zero();
    RSS64 pools(averagepool_im2row_size((int)conf.imageWidth, (int)conf.imageHeight, f, (int)conf.features,
                                        (int)conf.batchSize, (int)conf.stride, 0));
    for (int q = 0; q < 3; q++)
        for (int i = 0; i < 2; i++)
            check(averagepool_im2row(input.getShare(q, i)->data(), pools.getShare(q, i)->data(), (int)conf.imageWidth,
                                     (int)conf.imageHeight, f, (int)conf.features, (int)conf.batchSize,
                                     (int)conf.stride, 0, ctx().stream));
    for (int q = 0; q < 3; q++)
        for (int i = 0; i < 2; i++)
            check(reduce_sum(pools.getShare(q, i)->data(), activations.getShare(q, i)->data(), 0,
                             (int)activations.size(), f * f, ctx().stream));
    dividePublic(activations, (u64)f * (u64)f);
}

// AveragepoolLayer.cu:98-125
void AveragepoolLayer::backward(const RSS64 &delta, const RSS64 &) {
    ensureDeltas();
    const int P = (int)(conf.poolSize * conf.poolSize);
    if (ctx().mode == MODE_FUSED && delta.base()) {
        check(fused_avgpool_backward(delta.base(), delta.size(), (int)conf.features, P, deltas.base(), deltas.size(),
                                     ctx().stream));
        return;
    }
    deltas.zero();
    for (int q = 0; q < 3; q++)
        for (int i = 0; i < 2; i++)
            check(averagepool_expand_delta(delta.getShare(q, i)->data(), delta.size(), deltas.getShare(q, i)->data(),
                                           deltas.size(), (int)conf.features, P, ctx().stream));
    dividePublic(deltas, (u64)P);
}

// ---------------------------------------------------------------------------
// LNLayer.cu

LNLayer::LNLayer(const LNConfig &c, int n, int)
    : Layer(n), conf(c), gamma(c.inputSize), beta(c.inputSize), xhat(c.numBatches * c.inputSize),
      invSigma(c.numBatches), activations(c.inputSize * c.numBatches), revealed_(c.numBatches),
      invVals_(c.numBatches) {
    // LNLayer.cu:28-35
    gamma.setPublic(std::vector<double>(c.inputSize, 1.0), ctx().frac_bits);
    beta.setPublic(std::vector<double>(c.inputSize, 0.0), ctx().frac_bits);
    invSqrt_.prepare(3 * c.numBatches);
}

void LNLayer::printLayer() {
    printf("----------------------------------------------\n(%d) LN Layer\t\t  %zu x %zu\n", layerNum + 1,
           conf.inputSize, conf.numBatches);
}

// v = (double)(int64)x / 2^fp; v = 1 / sqrt(v); (uint64)(v * 2^fp) with the
// clamping conversion of the reference's GPU.  sqrt and the division are
// correctly rounded IEEE-754 operations on the host as on any conforming unit.
void LNLayer::InvSqrtStage::compute(const u64 *in, u64 *out, size_t n) const {
    const double scale = (double)(1 << fp);
    for (size_t i = 0; i < n; i++) {
        double v = (double)(s64)in[i] / scale;
        v = 1 / std::sqrt(v);
        out[i] = to_u64_clamped(v * scale);
    }
}

// LNLayer.cu:66-162
void LNLayer::forward(const RSS64 &input) {
    require_placeholder("LN layer");
    const int fp = ctx().frac_bits;
    if (fp < 8) {  // epsilon = 1 << (FLOAT_PRECISION - 8), LNLayer.cu:119
        set_error("LN layer: needs at least 8 fractional bits");
        throw PirError(ERR_UNSUPPORTED);
    }
    const int nb = (int)conf.numBatches;
    const size_t n = input.size();
    const int bs = (int)(n / (size_t)nb);
    hipStream_t s = ctx().stream;
    activations.zero();
    xhat.zero();
    invSigma.zero();

    RSS64 mu(nb);
    for (int q = 0; q < 3; q++)
        for (int i = 0; i < 2; i++)
            check(reduce_sum(input.getShare(q, i)->data(), mu.getShare(q, i)->data(), 0, nb, bs, s));
    dividePublic(mu, (u64)bs);
    xhat += input;
    for (int q = 0; q < 3; q++)
        for (int i = 0; i < 2; i++)
            check(element_vector_subtract(xhat.getShare(q, i)->data(), mu.getShare(q, i)->data(), 0, nb, bs, s));

    RSS64 std2(n);
    std2 += xhat;
    std2 *= xhat;
    dividePublic(std2, (u64)1 << fp);
    RSS64 variance(nb);
    for (int q = 0; q < 3; q++)
        for (int i = 0; i < 2; i++)
            check(reduce_sum(std2.getShare(q, i)->data(), variance.getShare(q, i)->data(), 0, nb, bs, s));
    dividePublic(variance, (u64)bs);
    variance += (u64)(1 << (fp - 8));  // epsilon

    reconstruct(variance, revealed_);
    invSqrt_.fp = fp;
    invSqrt_.enqueue(revealed_.base(), invVals_.base(), s);
    invSigma += invVals_;

    RSS64 expInv(n);
    for (int q = 0; q < 3; q++)
        for (int i = 0; i < 2; i++)
            check(vector_expand(invSigma.getShare(q, i)->data(), invSigma.size(), expInv.getShare(q, i)->data(),
                                (int)(n / invSigma.size()), s));
    xhat *= expInv;
    dividePublic(xhat, (u64)1 << fp);

    RSS64 expGamma(n), expBeta(n);
    for (int q = 0; q < 3; q++)
        for (int i = 0; i < 2; i++) {
            check(vector_expand(gamma.getShare(q, i)->data(), gamma.size(), expGamma.getShare(q, i)->data(),
                                (int)(n / gamma.size()), s));
            check(vector_expand(beta.getShare(q, i)->data(), beta.size(), expBeta.getShare(q, i)->data(),
                                (int)(n / gamma.size()), s));
        }
    activations += xhat;
    activations *= expGamma;
    dividePublic(activations, (u64)1 << fp);
    activations += expBeta;
}

// LNLayer.cu:166-284 (xhat is overwritten, as in the reference)
void LNLayer::backward(const RSS64 &delta, const RSS64 &) {
    ensureDeltas();
    const int fp = ctx().frac_bits, lr = ctx().log_lr;
    const int nb = (int)conf.numBatches;
    const size_t n = delta.size();
    const int bs = (int)(n / (size_t)nb);
    hipStream_t s = ctx().stream;
    auto per_share = [](auto f) {
        for (int q = 0; q < 3; q++)
            for (int i = 0; i < 2; i++) f(q, i);
    };
    auto expand = [&](const RSS64 &a, RSS64 &b, int q, int i) {
        check(vector_expand(a.getShare(q, i)->data(), a.size(), b.getShare(q, i)->data(), (int)(b.size() / a.size()),
                            s));
    };
    deltas.zero();

    RSS64 dBeta(beta.size());
    per_share([&](int q, int i) {
        check(reduce_sum(delta.getShare(q, i)->data(), dBeta.getShare(q, i)->data(), 1, nb, bs, s));
    });
    dividePublic(dBeta, (u64)1 << lr);

    RSS64 dGamma(gamma.size()), temp(n);
    temp += delta;
    temp *= xhat;
    per_share([&](int q, int i) {
        check(reduce_sum(temp.getShare(q, i)->data(), dGamma.getShare(q, i)->data(), 1, nb, bs, s));
    });
    dividePublic(dGamma, (u64)1 << (lr + fp));

    RSS64 dxhat(n), expGamma(n);
    dxhat += delta;
    per_share([&](int q, int i) { expand(gamma, expGamma, q, i); });
    dxhat *= expGamma;
    dividePublic(dxhat, (u64)1 << fp);

    RSS64 green(n), greenTemp(nb);
    green += dxhat;
    green *= xhat;
    per_share([&](int q, int i) {
        check(reduce_sum(green.getShare(q, i)->data(), greenTemp.getShare(q, i)->data(), 0, nb, bs, s));
    });
    dividePublic(greenTemp, (u64)1 << fp);
    per_share([&](int q, int i) { expand(greenTemp, green, q, i); });
    xhat *= green;
    dividePublic(xhat, (u64)1 << fp);

    RSS64 blue(nb);
    per_share([&](int q, int i) {
        check(reduce_sum(dxhat.getShare(q, i)->data(), blue.getShare(q, i)->data(), 0, nb, bs, s));
    });
    deltas += dxhat;
    deltas *= (u64)bs;
    RSS64 expBlue(n), expInv(n);
    per_share([&](int q, int i) {
        expand(blue, expBlue, q, i);
        expand(invSigma, expInv, q, i);
    });
    deltas -= expBlue;
    deltas -= xhat;
    deltas *= expInv;
    dividePublic(deltas, ((u64)1 << fp) * (u64)bs);

    gamma -= dGamma;
    beta -= dBeta;
}

// ---------------------------------------------------------------------------
// ResLayer.cu

ResLayer::ResLayer(const ResLayerConfig &c, int n, int) : Layer(n), conf(c) {
    if (c.expansion != 1) {
        // the block output has `planes` channels and a shortcut expansion * planes:
        // the residual sum needs them equal
        set_error("ResLayer: only expansion 1 is supported");
        throw PirError(ERR_UNSUPPORTED);
    }
    size_t ih = c.imageHeight, iw = c.imageWidth;
    int layer_ctr = n * 1000;
    appendBlock(ih, iw, c.in_planes, c.planes, c.stride, layer_ctr);
    for (size_t i = 1; i < c.num_blocks; i++) appendBlock(ih, iw, c.planes * c.expansion, c.planes, 1, layer_ctr);
}

// ResLayer.cu:42-103
void ResLayer::appendBlock(size_t &ih, size_t &iw, size_t in_planes, size_t planes, size_t stride, int &layer_ctr) {
    const int base_seed = layerNum * 1000 + (int)blocks.size() * 10;
    const size_t b = conf.batchSize;
    Block block, shortcut;
    block.emplace_back(new CNNLayer(CNNConfig(ih, iw, in_planes, planes, 3, stride, 1, b), layer_ctr++, base_seed + 0));
    if (stride != 1 || in_planes != conf.expansion * planes) {
        shortcut.emplace_back(
            new CNNLayer(CNNConfig(ih, iw, in_planes, conf.expansion * planes, 1, stride, 0, b), layer_ctr++, base_seed + 2));
        size_t oh = (ih + 2 - 3) / stride + 1, ow = (iw + 2 - 3) / stride + 1;
        shortcut.emplace_back(new LNLayer(LNConfig(oh * ow * conf.expansion * planes, b), layer_ctr++, 0));
    }
    ih = (ih + 2 - 3) / stride + 1;
    iw = (iw + 2 - 3) / stride + 1;
    block.emplace_back(new LNLayer(LNConfig(iw * ih * planes, b), layer_ctr++, 0));
    block.emplace_back(new ReLULayer(ReLUConfig(iw * ih * planes, b), layer_ctr++, 0));
    block.emplace_back(new CNNLayer(CNNConfig(ih, iw, planes, planes, 3, 1, 1, b), layer_ctr++, base_seed + 1));
    block.emplace_back(new LNLayer(LNConfig(iw * ih * planes, b), layer_ctr++, 0));
    block.emplace_back(new ReLULayer(ReLUConfig(iw * ih * planes, b), layer_ctr++, 0));
    blocks.push_back(std::move(block));
    shortcuts.push_back(std::move(shortcut));
}

void ResLayer::printLayer() {
    printf("----------------------------------------------\n(%d) Res Layer\n\n\t\t\t  %zu\t\t(BatchSize)\t\t%zu\t\t(Num "
           "Blocks)\n\t\t\t  %zu\t\t(In Planes)\t\t%zu\t\t(Planes)\n\t\t\t  %zu\t\t(Stride)\t\t%zu\t\t(Expansion)\n",
           layerNum + 1, conf.batchSize, conf.num_blocks, conf.in_planes, conf.planes, conf.stride, conf.expansion);
    for (size_t i = 0; i < blocks.size(); i++) {
        for (size_t k = 0; k + 1 < blocks[i].size(); k++) blocks[i][k]->printLayer();
        for (auto &l : shortcuts[i]) l->printLayer();
        blocks[i].back()->printLayer();
    }
}

// ResLayer.cu:196-219: the block's layers in order; the last one takes the
// penultimate activation plus the shortcut's output (or the block input).
void ResLayer::forwardBlock(size_t idx, Block &b, const RSS64 *input, bool isShortcut) {
    b[0]->forward(*input);
    for (size_t i = 1; i + 1 < b.size(); i++) b[i]->forward(*b[i - 1]->getActivation());
    const RSS64 *pen = b[b.size() - 2]->getActivation();
    RSS64 lastLayerInput(pen->size());
    lastLayerInput += *pen;
    if (!isShortcut) {
        if (!shortcuts[idx].empty()) {
            forwardBlock(0, shortcuts[idx], input, true);
            lastLayerInput += *shortcuts[idx].back()->getActivation();
        } else {
            lastLayerInput += *input;
        }
    }
    b.back()->forward(lastLayerInput);
}

// ResLayer.cu:227-236
void ResLayer::forward(const RSS64 &input) {
    require_placeholder("Res layer");
    forwardBlock(0, blocks[0], &input, false);
    for (size_t i = 1; i < blocks.size(); i++) forwardBlock(i, blocks[i], blocks[i - 1].back()->getActivation(), false);
}

// ResLayer.cu:250-286.  Without a shortcut the reference adds the scalar 1 to
// the first layer's delta (RSS += 1: A.sA and C.sB), which is kept.
void ResLayer::backwardBlock(size_t idx, Block &b, const RSS64 *deltas, const RSS64 *forwardInput, bool isShortcut) {
    const size_t last = b.size() - 1;
    const RSS64 *pen = b[last - 1]->getActivation();
    RSS64 lastLayerInput(pen->size());
    lastLayerInput += *pen;
    if (!isShortcut) {
        if (!shortcuts[idx].empty()) lastLayerInput += *shortcuts[idx].back()->getActivation();
        else lastLayerInput += *forwardInput;
    }
    b[last]->backward(*deltas, lastLayerInput);
    for (size_t i = last - 1; i > 0; i--) b[i]->backward(*b[i + 1]->getDelta(), *b[i - 1]->getActivation());
    if (b.size() > 1) b[0]->backward(*b[1]->getDelta(), *forwardInput);
    if (!isShortcut) {
        if (!shortcuts[idx].empty()) {
            backwardBlock(0, shortcuts[idx], b[last]->getDelta(), forwardInput, true);
            *b[0]->getDelta() += *shortcuts[idx][0]->getDelta();
        } else {
            *b[0]->getDelta() += (u64)1;
        }
    }
}

// ResLayer.cu:288-307
void ResLayer::backward(const RSS64 &deltas, const RSS64 &forwardInput) {
    const size_t L = blocks.size() - 1;
    if (blocks.size() > 1) backwardBlock(L, blocks[L], &deltas, blocks[L - 1].back()->getActivation(), false);
    else backwardBlock(L, blocks[L], &deltas, &forwardInput, false);
    for (size_t i = blocks.size() - 2; blocks.size() > 2 && i > 0; i--)
        backwardBlock(i, blocks[i], blocks[i + 1].front()->getDelta(), blocks[i - 1].back()->getActivation(), false);
    if (blocks.size() > 1) backwardBlock(0, blocks[0], blocks[1].front()->getDelta(), &forwardInput, false);
}

// ResLayer.cu:134-167 (saveSnapshotBlock order: block layers, then the shortcut)
std::vector<std::pair<std::string, RSS64 *>> ResLayer::snapshotTensors() {
    std::vector<std::pair<std::string, RSS64 *>> t;
    for (size_t i = 0; i < blocks.size(); i++) {
        for (auto &l : blocks[i])
            for (auto &e : l->snapshotTensors()) t.push_back(e);
        for (auto &l : shortcuts[i])
            for (auto &e : l->snapshotTensors()) t.push_back(e);
    }
    return t;
}

void ResLayer::allocDelta() {
    for (size_t i = 0; i < blocks.size(); i++) {
        for (auto &l : blocks[i]) l->allocDelta();
        for (auto &l : shortcuts[i]) l->allocDelta();
    }
}

void ResLayer::weightsChanged() {
    for (size_t i = 0; i < blocks.size(); i++) {
        for (auto &l : blocks[i]) l->weightsChanged();
        for (auto &l : shortcuts[i]) l->weightsChanged();
    }
}

// ---------------------------------------------------------------------------
// Network (NeuralNetwork.cu:32-133)

// Geometry of one layer spec row: every size positive, windows inside the
// (padded) image, so that the layer's size arithmetic and kernels stay in range.
static bool spec_valid(const int *v) {
    switch (v[0]) {
        case 0: return v[1] > 0 && v[2] > 0;
        case 1:
            return v[1] > 0 && v[2] > 0 && v[3] > 0 && v[4] > 0 && v[5] > 0 && v[6] > 0 && v[7] >= 0 &&
                   v[1] + 2 * v[7] >= v[5] && v[2] + 2 * v[7] >= v[5];
        case 2: return v[1] > 0;
        case 5: return v[1] > 0;
        case 6:
            return v[1] > 0 && v[2] > 0 && v[3] > 0 && v[4] > 0 && v[5] > 0 && v[6] > 0 && v[7] > 0 && v[1] + 2 >= 3 &&
                   v[2] + 2 >= 3;
        case 3:
        case 4: return v[1] > 0 && v[2] > 0 && v[3] > 0 && v[4] > 0 && v[5] > 0 && v[1] >= v[4] && v[2] >= v[4];
        default: return true;  // rejected below
    }
}

NeuralNetwork::NeuralNetwork(const std::vector<LayerSpec> &spec, int b, int seed) : batch(b) {
    for (size_t i = 0; i < spec.size(); i++) {
        const int *v = spec[i].v;
        int s = seed + (int)i;
        if (!spec_valid(v)) {
            set_error("network: layer " + std::to_string(i) + " has an invalid geometry");
            throw PirError(ERR_ARG);
        }
        switch (v[0]) {
            case 0: layers.emplace_back(new FCLayer(FCConfig(v[1], b, v[2]), (int)i, s)); break;
            case 1: layers.emplace_back(new CNNLayer(CNNConfig(v[1], v[2], v[3], v[4], v[5], v[6], v[7], b), (int)i, s)); break;
            case 2: layers.emplace_back(new ReLULayer(ReLUConfig(v[1], b), (int)i, s)); break;
            case 3: layers.emplace_back(new MaxpoolLayer(MaxpoolConfig(v[1], v[2], v[3], v[4], v[5], b), (int)i, s)); break;
            case 4:
                layers.emplace_back(new AveragepoolLayer(AveragepoolConfig(v[1], v[2], v[3], v[4], v[5], b), (int)i, s));
                break;
            case 5: layers.emplace_back(new LNLayer(LNConfig(v[1], b), (int)i, s)); break;
            case 6:
                layers.emplace_back(
                    new ResLayer(ResLayerConfig(b, v[1], v[2], v[3], v[4], v[5], v[6], v[7]), (int)i, s));
                break;
            default:
                set_error("network: only FC, CNN, ReLU, Maxpool, Averagepool, LN and Res layers are supported");
                throw PirError(ERR_UNSUPPORTED);
        }
    }
    for (size_t i = 1; i < layers.size(); i++)
        if (layers[i]->inputSize() != layers[i - 1]->getActivation()->size()) {
            set_error("network: layer " + std::to_string(i) + " input size does not match layer " +
                      std::to_string(i - 1) + "'s output");
            throw PirError(ERR_ARG);
        }
    if (!layers.empty()) input.resize(layers[0]->inputSize());
}

NeuralNetwork::~NeuralNetwork() = default;

void NeuralNetwork::forward(const std::vector<double> &data) {
    input.zero();
    input.setPublic(data, ctx().frac_bits);
    run();
}

// In the masked precompute mode layer i of a forward owns call number
// base + i of the context's stream (a ReLU layer computed inside the layer
// before it keeps its own number), and the forward advances the counter by
// the layer count.
void NeuralNetwork::run(const std::function<void(size_t)> &after_layer) {
    Context &c = ctx();
    const u32 base = c.prg_call;
    for (size_t i = 0; i < layers.size(); i++) {
        c.prg_call = base + (u32)i;
        const RSS64 &in = i == 0 ? input : *layers[i - 1]->getActivation();
        ReLULayer *next = i + 1 < layers.size() ? dynamic_cast<ReLULayer *>(layers[i + 1].get()) : nullptr;
        if (next && next->inputSize() == layers[i]->getActivation()->size() && layers[i]->forwardWithReLU(in, *next)) {
            if (after_layer) after_layer(i);
            i++;
        } else {
            layers[i]->forward(in);
        }
        if (after_layer) after_layer(i);
    }
    c.prg_call = base + (u32)layers.size();
}

void NeuralNetwork::backward(const std::vector<double> &labels) {
    RSS64 labelShare(labels.size());
    labelShare.setPublic(labels, ctx().frac_bits);
    backwardShares(labelShare);
}

void NeuralNetwork::backwardShares(const RSS64 &labelShare, const std::function<void(size_t)> &after_step) {
    require_placeholder("network backward (training)");
    if (layers.size() < 2) {
        set_error("network backward: needs at least two layers");
        throw PirError(ERR_UNSUPPORTED);
    }
    if (labelShare.size() != layers.back()->getActivation()->size() || batch <= 0 ||
        labelShare.size() % (size_t)batch != 0) {
        set_error("network backward: labels must have batch x classes entries");
        throw PirError(ERR_ARG);
    }
    if (outputDelta.size() != labelShare.size()) outputDelta.resize(labelShare.size());
    _reveal_softmax_grad(labelShare, outputDelta);
    if (after_step) after_step(0);
    _backward_pass(outputDelta, after_step);
}

void NeuralNetwork::prepareBackward() {
    if (layers.empty()) return;
    size_t n = layers.back()->getActivation()->size();
    if (outputDelta.size() != n) outputDelta.resize(n);
    for (auto &l : layers) l->allocDelta();
    _prepare_softmax();
}

HostStage::~HostStage() {
    if (in_) (void)hipHostFree(in_);
    if (out_) (void)hipHostFree(out_);
}

void HostStage::prepare(size_t n) {
    if (n == n_ && in_) return;
    if (in_) hip_check(hipHostFree(in_));
    if (out_) hip_check(hipHostFree(out_));
    in_ = out_ = nullptr;
    n_ = 0;
    hip_check(hipHostMalloc((void **)&in_, std::max<size_t>(n, 1) * sizeof(u64), hipHostMallocDefault));
    hip_check(hipHostMalloc((void **)&out_, std::max<size_t>(n, 1) * sizeof(u64), hipHostMallocDefault));
    n_ = n;
}

void HostStage::enqueue(const u64 *dev_in, u64 *dev_out, hipStream_t s) {
    if (!in_) {
        set_error("host stage used before prepare()");
        throw PirError(ERR_STATE);
    }
    if (n_ == 0) return;
    const size_t bytes = n_ * sizeof(u64);
    hip_check(hipMemcpyAsync(in_, dev_in, bytes, hipMemcpyDeviceToHost, s));
    hip_check(hipLaunchHostFunc(s, &HostStage::run, this));
    hip_check(hipMemcpyAsync(dev_out, out_, bytes, hipMemcpyHostToDevice, s));
}

void HostStage::run(void *self) {
    const HostStage &h = *(const HostStage *)self;
    h.compute(h.in_, h.out_, h.n_);
}

// NeuralNetwork.cu:220-257 for each view: x = (double)(int64)v / 2^fp
// (to_double_functor, functors.h:105-114), exp summed in element order per row,
// each value divided by its row's sum, back with (uint64)(x * 2^fp)
// (to_fixed_functor, functors.h:116-123).
// Runs on the runtime's callback thread, so it allocates nothing: exp is
// evaluated twice per element (row-sum pass, output pass), the same function
// of the same input both times.
void NeuralNetwork::SoftmaxStage::compute(const u64 *in, u64 *out, size_t n) const {
    const double scale = (double)(1 << fp);
    const size_t per_view = n / (size_t)views;
    for (int v = 0; v < views; v++)
        for (int r = 0; r < rows; r++) {
            const size_t base = (size_t)v * per_view + (size_t)r * nc;
            double sum = 0.0;
            for (int i = 0; i < nc; i++) sum += std::exp((double)(s64)in[base + i] / scale);
            for (int i = 0; i < nc; i++) {
                double q = std::exp((double)(s64)in[base + i] / scale) / sum;
                out[base + i] = to_u64_clamped(q * scale);
            }
        }
}

void NeuralNetwork::_prepare_softmax() {
    if (layers.empty() || batch <= 0) return;
    const size_t n = layers.back()->getActivation()->size();
    const int nc = (int)(n / (size_t)batch), fp = ctx().frac_bits;
    if (softmaxOpen_.size() != n) {
        softmaxOpen_.resize(n);
        softmaxVals_.resize(n);
    }
    for (SoftmaxStage *st : {&softmaxAC_, &softmaxAll_}) {
        st->views = st == &softmaxAC_ ? 2 : 3;
        st->rows = batch;
        st->nc = nc;
        st->fp = fp;
        st->prepare((size_t)st->views * n);
    }
}

void NeuralNetwork::_reveal_softmax_grad(const RSS64 &labels, RSS64 &deltas) {
    const RSS64 &act = *layers.back()->getActivation();
    const size_t n = act.size();
    const int lmb = (int)std::log2((double)batch);  // LOG_MINI_BATCH, model.cu:39
    _prepare_softmax();
    if (ctx().mode == MODE_FUSED && act.base() && labels.base() && deltas.base()) {
        check(fused_open_logits_ac(act.base(), n, softmaxOpen_.base(), ctx().stream));
        softmaxAC_.enqueue(softmaxOpen_.base(), softmaxVals_.base(), ctx().stream);
        check(fused_softmax_delta(softmaxVals_.base(), labels.base(), n, lmb, deltas.base(), ctx().stream));
        return;
    }
    RSS64 x(n);
    x += act;
    reconstruct(x, softmaxOpen_);
    softmaxAll_.enqueue(softmaxOpen_.base(), softmaxVals_.base(), ctx().stream);
    deltas.zero();
    deltas += softmaxVals_;
    deltas -= labels;
    if (lmb > 1) dividePublic(deltas, (u64)1 << lmb);
}

// Layer L-1 takes the output delta, layer i the delta of layer i+1, each with
// the forward input of its own forward pass (the network input for layer 0).
// In the fused form a ReLU layer's backward runs inside the product of the
// layer after it (which reads that ReLU's activation as its forward input).
void NeuralNetwork::_backward_pass(const RSS64 &deltas, const std::function<void(size_t)> &after_step) {
    const RSS64 *d = &deltas;
    size_t i = layers.size();
    while (i > 0) {
        i--;
        const RSS64 &in = i == 0 ? input : *layers[i - 1]->getActivation();
        ReLULayer *prev = i > 0 ? dynamic_cast<ReLULayer *>(layers[i - 1].get()) : nullptr;
        if (prev && layers[i]->backwardWithReLU(*d, in, *prev)) {
            if (after_step) after_step(1 + i);
            i--;  // the ReLU layer i-1 is done
            d = prev->getDelta();
        } else {
            layers[i]->backward(*d, in);
            d = layers[i]->getDelta();
        }
        if (after_step) after_step(1 + i);
    }
}

void NeuralNetwork::printNetwork() {
    for (auto &l : layers) l->printLayer();
}

}  // namespace pir
