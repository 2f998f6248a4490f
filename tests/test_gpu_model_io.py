"""Model files and weight snapshots on the GPU: networks built by loadModel
(model.cu:24-115) from the reference's model descriptions, and
NeuralNetwork::saveSnapshot / loadSnapshot (NeuralNetwork.cu:74-104,
util.cuh:271-303) -- the reference's SecureML weights, written in the
reference's own snapshot text format, loaded and run forward against the
oracle and against the reference's recorded layer outputs."""
import os

import numpy as np
import pytest

import oracle as O
import piranha_amd as P
from piranha_amd import configs as C

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
MODELS = os.path.join(HERE, "golden", "models")
FP = 13


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(HERE, "golden", "fixtures.npz"))


@pytest.mark.parametrize("name,batch", [("lenet.json", 2), ("vgg16-cifar10.json", 1), ("secureml.json", 0)])
def test_network_from_model_file(name, batch):
    path = os.path.join(MODELS, name)
    net = P.Network.from_model_file(path, FP, batch=batch, seed=4)
    _, _, file_batch, spec = P.read_model(path)
    assert net.batch == (batch or file_batch)
    assert net.spec == spec
    # the same network from explicit layer specs: identical weights and outputs
    layers = C.from_model_json(path)
    ref = P.Network(C.to_spec(layers), net.batch, FP, seed=4)
    x = np.random.default_rng(1).uniform(0, 1, net.batch * C.input_size(layers))
    net.forward(x)
    ref.forward(x)
    for i in range(len(layers)):
        assert np.array_equal(net.activation(i), ref.activation(i)), f"layer {i}"


def write_reference_format(path, vals):
    # files/test/secureml/*: one value per line, printed with 19 significant digits
    with open(path, "w") as f:
        f.write("".join("%.18e\n" % v for v in vals))


def test_load_reference_snapshot_and_forward(tmp_path, gold):
    for k in (0, 2, 4):
        write_reference_format(tmp_path / f"weight{k}", gold[f"secureml_weight{k}"])
        write_reference_format(tmp_path / f"bias{k}", gold[f"secureml_bias{k}"])
    fp = 26  # the reference tests' FLOAT_PRECISION (globals.h:19-21)
    net = P.Network(C.to_spec(C.SECUREML), 128, fp, seed=1)
    net.load_snapshot(str(tmp_path))
    params = {}
    for k in (0, 2, 4):
        W, b = net.weights(k)
        assert np.array_equal(W, O.set_public(gold[f"secureml_weight{k}"], fp))
        assert np.array_equal(b, O.set_public(gold[f"secureml_bias{k}"], fp))
        params[k] = (W, b)
    x = gold["secureml_input"]
    net.forward(x)
    want = O.net_forward(C.SECUREML, 128, fp, params, O.set_public(x, fp))
    for i, (act, _) in enumerate(want):
        assert np.array_equal(net.activation(i), act), f"layer {i}"
        got = O.reconstruct(act)[0].view(np.int64) / float(1 << fp)
        # ASSERT_EPSILON (util.cuh:19) against files/test/secureml/outputlayer{i}
        assert np.allclose(got, gold[f"secureml_out{i}"], atol=1e-3), f"layer {i} vs outputlayer{i}"


def cpp_default_format(vals):
    # std::ostream << double with the default flags and precision 6, then " "
    return "".join("%g " % v for v in vals)


def test_save_snapshot_format_and_round_trip(tmp_path):
    layers = C.LENET
    net = P.Network(C.to_spec(layers), 2, FP, seed=8)
    out = tmp_path / "snap"
    net.save_snapshot(str(out))
    names = sorted(os.listdir(out))
    want_names = sorted([f"weight{i}" for i, l in enumerate(layers) if l[0] in ("fc", "cnn")] +
                        [f"bias{i}" for i, l in enumerate(layers) if l[0] == "fc"])
    assert names == want_names
    fresh = P.Network(C.to_spec(layers), 2, FP, seed=99)
    fresh.load_snapshot(str(out))
    for i, l in enumerate(layers):
        if l[0] not in ("fc", "cnn"):
            continue
        W, b = net.weights(i)
        vals = O.reconstruct(W)[0].view(np.int64) / float(1 << FP)
        text = (out / f"weight{i}").read_text()
        assert text == cpp_default_format(vals)
        parsed = np.array([float(t) for t in text.split()])
        assert np.array_equal(fresh.weights(i)[0], O.set_public(parsed, FP))
        if b is not None:
            bt = (out / f"bias{i}").read_text()
            assert bt == cpp_default_format(O.reconstruct(b)[0].view(np.int64) / float(1 << FP))
            assert np.array_equal(fresh.weights(i)[1], O.set_public(np.array([float(t) for t in bt.split()]), FP))


def test_load_short_and_missing_files(tmp_path):
    net = P.Network([(P.LAYER_FC, 4, 3)], 1, FP, seed=2)
    (tmp_path / "weight0").write_text("0.5 -1.25\n3")  # 3 of 12 values; no bias0 file
    net.load_snapshot(str(tmp_path))
    W, b = net.weights(0)
    assert np.array_equal(W, O.set_public(np.array([0.5, -1.25, 3.0] + [0.0] * 9), FP))
    assert np.array_equal(b, O.set_public(np.zeros(3), FP))


def test_load_model_errors(tmp_path):
    bottleneck = tmp_path / "res.json"  # ResLayer blocks with expansion 4 (the residual sum needs 1)
    bottleneck.write_text('{"input_size": 48, "num_classes": 8, "batch_size": 1, "dataset": "x", "model": ['
                          '{"layer": "res", "input_hw": [4, 4], "in_planes": 3, "out_planes": 2, "num_blocks": 1, '
                          '"stride": 1, "expansion": 4}]}')
    with pytest.raises(P.PirError, match="expansion"):
        P.Network.from_model_file(str(bottleneck), FP)
    bad = tmp_path / "bad.json"
    bad.write_text('{"input_size": 4, "num_classes": 3, "batch_size": 1, "dataset": "x", "model": ['
                   '{"layer": "fc", "input_dim": 4, "output_dim": 3}, {"layer": "relu", "input_dim": 4}]}')
    with pytest.raises(P.PirError, match="input size does not match"):
        P.Network.from_model_file(str(bad), FP)
