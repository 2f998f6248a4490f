// layers.h -- forward layers and network for 3-party RSS inference, with the
// reference's layer API (src/nn/Layer.h:10-30; FCLayer.cu, CNNLayer.cu,
// ReLULayer.cu, MaxpoolLayer.cu; NeuralNetwork.cu:32-133).
#pragma once

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "protocol.h"

namespace pir {

// Layer configurations, same fields and argument order as the reference
// (FCConfig.h:16, CNNConfig.h:27, ReLUConfig.h:12, MaxpoolConfig.h:16).
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

class ReLULayer;

class Layer {
  public:
    int layerNum = 0;
    explicit Layer(int n) : layerNum(n) {}
    virtual ~Layer() = default;
    virtual void printLayer() = 0;
    virtual void forward(const RSS64 &input) = 0;
    virtual RSS64 *getActivation() = 0;
    virtual RSS64 *getWeights() { return nullptr; }
    virtual RSS64 *getBiases() { return nullptr; }
    // reluPrime / maxPrime of ReLU and Maxpool layers (Share<uint8_t>)
    virtual RSS<u8> *getDerivative() { return nullptr; }
    virtual size_t inputSize() const = 0;
    // Called after the weights were written (invalidates packed forms).
    virtual void weightsChanged() {}
    // Fused form only: forward() plus the ReLU layer that follows, in one pass
    // (the ReLU layer's activations and reluPrime are written here, bit for bit
    // what its own forward would write).  Returns false when this layer cannot.
    virtual bool forwardWithReLU(const RSS64 &input, ReLULayer &relu) { return false; }
};

class FCLayer : public Layer {
  public:
    FCLayer(const FCConfig &conf, int layerNum, int seed);
    void printLayer() override;
    void forward(const RSS64 &input) override;
    bool forwardWithReLU(const RSS64 &input, ReLULayer &relu) override;
    RSS64 *getActivation() override { return &activations; }
    RSS64 *getWeights() override { return &weights; }
    RSS64 *getBiases() override { return &biases; }
    size_t inputSize() const override { return conf.batchSize * conf.inputDim; }
    void weightsChanged() override { packed_valid = false; }

    FCConfig conf;
    RSS64 activations, weights, biases;

  private:
    void limbForward(const RSS64 &input, u64 *relu_out, u8 *relu_d);
    void initialize(int seed);
    DevBuf packed;
    bool packed_valid = false;
};

class CNNLayer : public Layer {
  public:
    CNNLayer(const CNNConfig &conf, int layerNum, int seed);
    void printLayer() override;
    void forward(const RSS64 &input) override;
    bool forwardWithReLU(const RSS64 &input, ReLULayer &relu) override;
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
    void limbForward(const RSS64 &input, u64 *relu_out, u8 *relu_d);
    void initialize(int seed);
    DevBuf packed;
    bool packed_valid = false;
};

class ReLULayer : public Layer {
  public:
    ReLULayer(const ReLUConfig &conf, int layerNum, int seed);
    void printLayer() override;
    void forward(const RSS64 &input) override;
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
    void printNetwork();
    std::vector<std::unique_ptr<Layer>> layers;
    RSS64 input;
    int batch;
};

}  // namespace pir
