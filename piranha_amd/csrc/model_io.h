// model_io.h -- the reference's model files and snapshots: loadModel
// (src/util/model.cu:24-115) reads a network description in JSON, and
// NeuralNetwork::saveSnapshot / loadSnapshot (NeuralNetwork.cu:74-104 with
// FCLayer.cu:51-69, CNNLayer.cu:53-64 and util.cuh:271-303) keep the weights as
// text files of reconstructed doubles.
#pragma once

#include <string>
#include <vector>

#include "layers.h"

namespace pir {

struct ModelFile {
    int input_size = 0, num_classes = 0, batch_size = 0;
    std::string dataset;
    std::vector<LayerSpec> layers;
};

// Parses a model file; throws PirError (ERR_ARG for malformed JSON, missing
// fields or an unknown layer type).
ModelFile load_model_file(const std::string &filename);

// Writes path/weight{i} (and path/bias{i} for FC layers; path/gamma{i} and
// path/beta{i} for LN layers; FCLayer.cu:51-69, CNNLayer.cu:53-64,
// LNLayer.cu:47-63) for every layer with parameters: the reconstructed
// fixed-point values as doubles, each followed by a space, in the C++ stream's
// default format (6 significant digits).  Creates the directory when missing.
void save_snapshot(NeuralNetwork &net, const std::string &path);
// Reads the same files back into public shares (setPublic); a file with fewer
// values than the layer holds leaves the rest at zero, a missing file gives
// zeros (an empty stream), as the reference's readers do.
void load_snapshot(NeuralNetwork &net, const std::string &path);

}  // namespace pir
