// layers.h -- layers and network for 3-party RSS inference and training, with
// the reference's layer API (src/nn/Layer.h:10-30; FCLayer.cu, CNNLayer.cu,
// ReLULayer.cu, MaxpoolLayer.cu; NeuralNetwork.cu:32-149, 207-267, 435-467).
#pragma once

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "fused_rss.h"
#include "protocol.h"

namespace pir {

// Layer configurations, same fields and argument order as the reference
// (FCConfig.h:16, CNNConfig.h:27, ReLUConfig.h:12, MaxpoolConfig.h:16,
// AveragepoolConfig.h:15).
struct FCConfig {
    size_t inputDim, batchSize, outputDim;
    FCConfig(size_t in, size_t batch, size_t out) : inputDim(in), batchSize(batch), outputDim(out) {}
};
struct CNNConfig {
    size_t imageHeight, imageWidth, inputFeatures, outputFeatures, filterSize, stride, padding, batchSize;
    CNNConfig(size_t h, size_t w, size_t cin, size_t cout, size_t f, size_t s, size_t p, size_t b)
        : imageHeight(h), imageWidth(w), inputFeatures(cin), outputFeatures(cout), filterSize(f), stride(s),
          padding(p), batchSize(b) {}
};
struct ReLUConfig {
    size_t inputDim, batchSize;
    ReLUConfig(size_t in, size_t b) : inputDim(in), batchSize(b) {}
};
struct MaxpoolConfig {
    size_t imageHeight, imageWidth, features, poolSize, stride, batchSize;
    MaxpoolConfig(size_t h, size_t w, size_t c, size_t pool, size_t s, size_t b)
        : imageHeight(h), imageWidth(w), features(c), poolSize(pool), stride(s), batchSize(b) {}
};

struct AveragepoolConfig {
    size_t imageHeight, imageWidth, features, poolSize, stride, batchSize;
    AveragepoolConfig(size_t h, size_t w, size_t c, size_t pool, size_t s, size_t b)
        : imageHeight(h), imageWidth(w), features(c), poolSize(pool), stride(s), batchSize(b) {}
};

// A host computation in stream order between two device buffers: the device
// values are copied to pinned host memory, a host function maps them and the
// results are copied back; a captured graph records the function as a host
// node.  Used where the reference computes on revealed (public) values on the
// host.  prepare() allocates the pinned buffers and must run outside any
// capture; they keep their addresses until the stage is destroyed.
class HostStage {
  public:
    HostStage() = default;
    HostStage(const HostStage &) = delete;
    HostStage &operator=(const HostStage &) = delete;
    virtual ~HostStage();
    void prepare(size_t n);  // n u64 values in and out
    void enqueue(const u64 *dev_in, u64 *dev_out, hipStream_t s);

  protected:
    virtual void compute(const u64 *in, u64 *out, size_t n) const = 0;

  private:
    static void run(void *self);
    u64 *in_ = nullptr, *out_ = nullptr;
    size_t n_ = 0;
};

struct LNConfig {
    size_t inputSize, numBatches;
    LNConfig(size_t in, size_t nb) : inputSize(in), numBatches(nb) {}
};

class ReLULayer;

class Layer {
  public:
    int layerNum = 0;
    explicit Layer(int n) : layerNum(n) {}
    virtual ~Layer() = default;
    virtual void printLayer() = 0;
    virtual void forward(const RSS64 &input) = 0;
    // Layer::backward: `delta` is dL/d(this layer's output), `forwardInput` the
    // input of the forward pass; writes dL/d(input) into getDelta() and applies
    // the weight update (learning rate 2^-log_lr of the context).
    virtual void backward(const RSS64 &delta, const RSS64 &forwardInput) = 0;
    virtual RSS64 *getActivation() = 0;
    virtual RSS64 *getDelta() { return &deltas; }
    virtual RSS64 *getWeights() { return nullptr; }
    virtual RSS64 *getBiases() { return nullptr; }
    // the parameter tensors of saveSnapshot / loadSnapshot with their file
    // names (weight{n} / bias{n}; gamma{n} / beta{n} for LN; a ResLayer lists
    // those of its inner layers, numbered layerNum * 1000 + k)
    virtual std::vector<std::pair<std::string, RSS64 *>> snapshotTensors() {
        std::vector<std::pair<std::string, RSS64 *>> t;
        if (getWeights()) t.emplace_back("weight" + std::to_string(layerNum), getWeights());
        if (getBiases()) t.emplace_back("bias" + std::to_string(layerNum), getBiases());
        return t;
    }
    // reluPrime / maxPrime of ReLU and Maxpool layers (Share<uint8_t>)
    virtual RSS<u8> *getDerivative() { return nullptr; }
    virtual size_t inputSize() const = 0;
    // Called after the weights were written (invalidates packed forms).
    virtual void weightsChanged() {}
    // Fused form only: forward() plus the ReLU layer that follows, in one pass
    // (the ReLU layer's activations and reluPrime are written here, bit for bit
    // what its own forward would write).  Returns false when this layer cannot.
    virtual bool forwardWithReLU(const RSS64 &input, ReLULayer &relu) { return false; }
    // Fused form only: backward() plus the backward of the ReLU layer that
    // precedes this one (its input is that ReLU's activation), in one pass; the
    // ReLU layer's deltas receive exactly what its own backward would write.
    virtual bool backwardWithReLU(const RSS64 &delta, const RSS64 &forwardInput, ReLULayer &relu) { return false; }
    virtual void allocDelta() { ensureDeltas(); }

  protected:
    // allocated on the first backward pass (inference never needs it)
    void ensureDeltas() {
        if (deltas.size() != inputSize()) deltas.resize(inputSize());
    }
    RSS64 deltas;
};

class FCLayer : public Layer {
  public:
    FCLayer(const FCConfig &conf, int layerNum, int seed);
    void printLayer() override;
    void forward(const RSS64 &input) override;
    void backward(const RSS64 &delta, const RSS64 &forwardInput) override;
    bool forwardWithReLU(const RSS64 &input, ReLULayer &relu) override;
    bool backwardWithReLU(const RSS64 &delta, const RSS64 &forwardInput, ReLULayer &relu) override;
    RSS64 *getActivation() override { return &activations; }
    RSS64 *getWeights() override { return &weights; }
    RSS64 *getBiases() override { return &biases; }
    size_t inputSize() const override { return conf.batchSize * conf.inputDim; }
    void weightsChanged() override { packed_valid = false; }

    FCConfig conf;
    RSS64 activations, weights, biases;

  private:
    void limbForward(const RSS64 &input, u64 *relu_out, u8 *relu_d, const PrgStream &g, const PrgStream &gr);
    void limbBackward(const RSS64 &delta, const RSS64 &forwardInput, ReLULayer *relu);
    void initialize(int seed);
    DevBuf packed;
    bool packed_valid = false;
};

class CNNLayer : public Layer {
  public:
    CNNLayer(const CNNConfig &conf, int layerNum, int seed);
    void printLayer() override;
    void forward(const RSS64 &input) override;
    void backward(const RSS64 &delta, const RSS64 &forwardInput) override;
    bool forwardWithReLU(const RSS64 &input, ReLULayer &relu) override;
    bool backwardWithReLU(const RSS64 &delta, const RSS64 &forwardInput, ReLULayer &relu) override;
    RSS64 *getActivation() override { return &activations; }
    RSS64 *getWeights() override { return &weights; }
    size_t inputSize() const override {
        return conf.batchSize * conf.imageHeight * conf.imageWidth * conf.inputFeatures;
    }
    void weightsChanged() override { packed_valid = false; }
    int outH() const { return (int)((conf.imageHeight - conf.filterSize + 2 * conf.padding) / conf.stride + 1); }
    int outW() const { return (int)((conf.imageWidth - conf.filterSize + 2 * conf.padding) / conf.stride + 1); }

    CNNConfig conf;
    RSS64 weights, activations;

  private:
    void limbForward(const RSS64 &input, u64 *relu_out, u8 *relu_d, const PrgStream &g, const PrgStream &gr);
    void limbBackward(const RSS64 &delta, const RSS64 &forwardInput, ReLULayer *relu);
    void initialize(int seed);
    DevBuf packed;
    bool packed_valid = false;
};

class ReLULayer : public Layer {
  public:
    ReLULayer(const ReLUConfig &conf, int layerNum, int seed);
    void printLayer() override;
    void forward(const RSS64 &input) override;
    void backward(const RSS64 &delta, const RSS64 &forwardInput) override;
    RSS64 *getActivation() override { return &activations; }
    RSS<u8> *getDerivative() override { return &reluPrime; }
    size_t inputSize() const override { return conf.batchSize * conf.inputDim; }

    ReLUConfig conf;
    RSS64 activations;
    RSS<u8> reluPrime;
};

class MaxpoolLayer : public Layer {
  public:
    MaxpoolLayer(const MaxpoolConfig &conf, int layerNum, int seed);
    void printLayer() override;
    void forward(const RSS64 &input) override;
    void backward(const RSS64 &delta, const RSS64 &forwardInput) override;
    bool forwardWithReLU(const RSS64 &input, ReLULayer &relu) override;
    RSS64 *getActivation() override { return &activations; }
    RSS<u8> *getDerivative() override { return &maxPrime; }
    size_t inputSize() const override {
        return conf.batchSize * conf.imageHeight * conf.imageWidth * conf.features;
    }
    int outH() const { return (int)((conf.imageHeight - conf.poolSize) / conf.stride + 1); }
    int outW() const { return (int)((conf.imageWidth - conf.poolSize) / conf.stride + 1); }

    MaxpoolConfig conf;
    RSS64 activations;
    RSS<u8> maxPrime;
};

// AveragepoolLayer.cu: window sums in every share, dividePublic by f*f; the
// backward expands each delta over its window (in the reference's own
// averagepool_expand_delta order) and divides by f*f again.
class AveragepoolLayer : public Layer {
  public:
    AveragepoolLayer(const AveragepoolConfig &conf, int layerNum, int seed);
    void printLayer() override;
    void forward(const RSS64 &input) override;
    void backward(const RSS64 &delta, const RSS64 &forwardInput) override;
    bool forwardWithReLU(const RSS64 &input, ReLULayer &relu) override;
    RSS64 *getActivation() override { return &activations; }
    size_t inputSize() const override {
        return conf.batchSize * conf.imageHeight * conf.imageWidth * conf.features;
    }
    int outH() const { return (int)((conf.imageHeight - conf.poolSize) / conf.stride + 1); }
    int outW() const { return (int)((conf.imageWidth - conf.poolSize) / conf.stride + 1); }

    AveragepoolConfig conf;
    RSS64 activations;

  private:
    MaxpoolGeom geom() const;
};

// LNLayer.cu: layer normalisation over each image's inputSize features, with
// the variance revealed and its inverse square root taken on the host
// (LNLayer.cu:124-132), as a composition of the protocol operations.
class LNLayer : public Layer {
  public:
    LNLayer(const LNConfig &conf, int layerNum, int seed);
    void printLayer() override;
    void forward(const RSS64 &input) override;
    void backward(const RSS64 &delta, const RSS64 &forwardInput) override;
    RSS64 *getActivation() override { return &activations; }
    RSS64 *getWeights() override { return &gamma; }  // LNLayer.h:40-41
    RSS64 *getBiases() override { return &beta; }
    std::vector<std::pair<std::string, RSS64 *>> snapshotTensors() override {
        return {{"gamma" + std::to_string(layerNum), &gamma}, {"beta" + std::to_string(layerNum), &beta}};
    }
    size_t inputSize() const override { return conf.inputSize * conf.numBatches; }

    LNConfig conf;
    RSS64 gamma, beta, xhat, invSigma, activations;

  private:
    // inv_sqrt_fixed_point_functor (functors.h:92-103) on revealed values
    class InvSqrtStage : public HostStage {
      public:
        int fp = 0;

      protected:
        void compute(const u64 *in, u64 *out, size_t n) const override;
    };
    InvSqrtStage invSqrt_;
    PartyData<u64> revealed_, invVals_;
};

struct ResLayerConfig {
    size_t batchSize, imageHeight, imageWidth, in_planes, planes, num_blocks, stride, expansion;
    ResLayerConfig(size_t b, size_t h, size_t w, size_t in, size_t p, size_t nb, size_t s, size_t e)
        : batchSize(b), imageHeight(h), imageWidth(w), in_planes(in), planes(p), num_blocks(nb), stride(s),
          expansion(e) {}
};

// ResLayer.cu: num_blocks residual blocks, each CNN 3x3 (stride) - LN - ReLU -
// CNN 3x3 - LN - ReLU, the last ReLU applied to the sum of its input and the
// block input (or, when the shape changes, of a 1x1 CNN - LN shortcut).  Inner
// layers are numbered layerNum * 1000 + k and seeded layerNum * 1000 +
// 10 * block + {0, 1, 2} (ResLayer.cu:23-103).
class ResLayer : public Layer {
  public:
    ResLayer(const ResLayerConfig &conf, int layerNum, int seed);
    void printLayer() override;
    void forward(const RSS64 &input) override;
    void backward(const RSS64 &delta, const RSS64 &forwardInput) override;
    RSS64 *getActivation() override { return blocks.back().back()->getActivation(); }
    RSS64 *getDelta() override { return blocks.front().front()->getDelta(); }
    size_t inputSize() const override {
        return conf.batchSize * conf.imageHeight * conf.imageWidth * conf.in_planes;
    }
    std::vector<std::pair<std::string, RSS64 *>> snapshotTensors() override;
    void allocDelta() override;
    void weightsChanged() override;

    using Block = std::vector<std::unique_ptr<Layer>>;
    ResLayerConfig conf;
    std::vector<Block> blocks, shortcuts;  // an empty shortcut: the block input is added

  private:
    void appendBlock(size_t &ih, size_t &iw, size_t in_planes, size_t planes, size_t stride, int &layer_ctr);
    void forwardBlock(size_t idx, Block &b, const RSS64 *input, bool isShortcut);
    void backwardBlock(size_t idx, Block &b, const RSS64 *deltas, const RSS64 *forwardInput, bool isShortcut);
};

// Layer description used by the C ABI: {type, a0..a6} (see include/pir.h).
struct LayerSpec {
    int v[8];
};

class NeuralNetwork {
  public:
    NeuralNetwork(const std::vector<LayerSpec> &spec, int batch, int seed);
    ~NeuralNetwork();
    // NeuralNetwork::forward, NeuralNetwork.cu:108-133
    void forward(const std::vector<double> &data);
    // forward from the input shares already held in `input`; in the fused form a
    // ReLU layer is evaluated inside the layer before it.  after_layer(i), when
    // given, is called once every layer i has been issued (in order).
    void run(const std::function<void(size_t)> &after_layer = nullptr);
    // NeuralNetwork::backward, NeuralNetwork.cu:136-149: labels (one-hot,
    // batch x classes doubles) become public shares, then the reveal-softmax
    // delta and the backward pass with weight updates.
    void backward(const std::vector<double> &labels);
    // the same from label shares already on the device ([3][2][batch*classes]);
    // after_step(k), when given, is called once backward step k has been issued
    // (k = 0: the softmax delta, k = 1 + i: layer i's backward; in the fused form
    // a ReLU layer's backward runs inside the layer after it).
    void backwardShares(const RSS64 &labelShare, const std::function<void(size_t)> &after_step = nullptr);
    // Allocates every buffer a backward pass keeps (each layer's deltas and the
    // output delta) up front, so that the temporaries of a training step come
    // from the allocator cache in the same order on every step (a captured
    // step then replays the eager step's allocations exactly).
    void prepareBackward();
    void printNetwork();
    // dL/d(network output) of the last backward (input of the last layer's backward)
    RSS64 outputDelta;
    std::vector<std::unique_ptr<Layer>> layers;
    RSS64 input;
    int batch;

  private:
    // The reveal-softmax is computed on the host, as in the reference, whose
    // softmax loop reads the opened logits one by one into host doubles
    // (NeuralNetwork.cu:220-257).
    class SoftmaxStage : public HostStage {
      public:
        int views = 0, rows = 0, nc = 0, fp = 0;

      protected:
        void compute(const u64 *in, u64 *out, size_t n) const override;
    };
    SoftmaxStage softmaxAC_, softmaxAll_;  // fused form (views A, C), op-by-op form (3 views)
    PartyData<u64> softmaxOpen_, softmaxVals_;
    void _prepare_softmax();
    void _reveal_softmax_grad(const RSS64 &labels, RSS64 &deltas);  // NeuralNetwork.cu:207-267
    void _backward_pass(const RSS64 &deltas, const std::function<void(size_t)> &after_step);  // :446-467
};

}  // namespace pir
