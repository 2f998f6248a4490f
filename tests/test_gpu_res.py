"""ResNet blocks (ResLayer.cu) on the GPU against the oracle's composition of
its own CNN, LN and ReLU layers, bit for bit: a Res layer's forward and
backward with every inner weight update (with and without shortcut blocks,
including the reference's scalar +1 on the delta of a block without one),
fused and op-by-op; the inner layers' numbering and snapshot order; and the
reference's ResNet18-CIFAR model file, forward and training."""
import os

import numpy as np
import pytest

import oracle as O
import piranha_amd as P
from piranha_amd import configs as C

pytestmark = pytest.mark.gpu

MODES = [P.MODE_FUSED, P.MODE_PLAIN]
MODELS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models")


def flat_params(net, layer):
    return [a for _, a in net.params(layer)]


@pytest.mark.parametrize("mode", MODES, ids=["fused", "plain"])
@pytest.mark.parametrize("spec,batch", [((8, 8, 4, 8, 2, 2, 1), 2),   # shortcut block, then a plain one
                                        ((6, 6, 8, 8, 1, 1, 1), 3),   # no shortcut at all: the +1 quirk
                                        ((5, 7, 3, 6, 1, 1, 1), 2)])  # channel change at stride 1
def test_res_layer(rng, mode, spec, batch):
    fp, lr = 20, 5
    net = P.Network([(P.LAYER_RES,) + spec], batch, fp, seed=3, mode=mode)
    net.set_log_learning_rate(lr)
    H, W, cin = spec[:3]
    x = O.set_public(rng.uniform(-1, 1, batch * H * W * cin), fp)
    params = flat_params(net, 0)
    dx = P.to_dev(x)
    net.forward_shares(dx.ptr)
    want, _ = O.res_forward(x, batch, fp, spec, params)
    assert np.array_equal(net.activation(0), want)
    delta = O.set_public(rng.normal(0, 0.05, want.shape[-1]), fp)
    dd = P.to_dev(delta)
    net.layer_backward(0, dd.ptr, dx.ptr)
    want_d, want_p = O.res_backward(delta, x, batch, fp, lr, spec, params)
    assert np.array_equal(net.delta(0), want_d), "layer delta differs"
    got_p = flat_params(net, 0)
    assert len(got_p) == len(want_p)
    for k, (a, b) in enumerate(zip(got_p, want_p)):
        assert np.array_equal(a, b), f"parameter {k} ({net.params(0)[k][0]}) differs"


def test_inner_layer_names():
    # layer 2 of a network: inner layers numbered 2000 + k in creation order
    # (CNN1, shortcut CNN3 and LN3, LN1, ReLU1, CNN2, LN2, ReLU2), listed block
    # layers first, then the shortcut (ResLayer.cu:42-103, 134-167)
    layers = [("relu", 8 * 8 * 4), ("relu", 8 * 8 * 4), ("res", 8, 8, 4, 8, 2, 2, 1)]
    net = P.Network(C.to_spec(layers), 1, 13, seed=1)
    names = [n for n, _ in net.params(2)]
    assert names == ["weight2000", "gamma2003", "beta2003", "weight2005", "gamma2006", "beta2006",
                     "weight2001", "gamma2002", "beta2002",
                     "weight2008", "gamma2009", "beta2009", "weight2011", "gamma2012", "beta2012"]


def one_hot(rng, batch, n_classes):
    lab = np.zeros((batch, n_classes))
    lab[np.arange(batch), rng.integers(0, n_classes, batch)] = 1.0
    return lab.ravel()


def test_resnet18_cifar_training_step():
    layers = C.from_model_json(os.path.join(MODELS, "resnet18-cifar10.json"))
    batch, fp, lr = 1, 20, 5
    rng = np.random.default_rng(11)
    net = P.Network(C.to_spec(layers), batch, fp, seed=2)
    net.set_log_learning_rate(lr)
    params = []
    for i, l in enumerate(layers):
        if l[0] == "res":
            params.append(flat_params(net, i))
        elif l[0] in ("fc", "cnn", "ln"):
            params.append(net.weights(i))
        else:
            params.append(None)
    x = rng.uniform(0.0, 1.0, batch * C.input_size(layers))
    net.forward(x)
    fwd = O.net_forward(layers, batch, fp, {i: p for i, p in enumerate(params) if p is not None},
                        O.set_public(x, fp))
    for i, (act, _) in enumerate(fwd):
        assert np.array_equal(net.activation(i), act), f"layer {i} ({layers[i][0]}) activation differs"
    labels = one_hot(rng, batch, 10)
    net.backward(labels)
    deltas, new_params = O.net_backward(layers, batch, fp, params, O.set_public(x, fp), fwd, labels, lr)
    for i in range(len(layers) - 1, -1, -1):
        assert np.array_equal(net.delta(i), deltas[i]), f"layer {i} ({layers[i][0]}) delta differs"
        if layers[i][0] == "res":
            for k, (a, b) in enumerate(zip(flat_params(net, i), new_params[i])):
                assert np.array_equal(a, b), f"layer {i} parameter {k} differs"
        elif params[i] is not None:
            assert np.array_equal(net.weights(i)[0], new_params[i][0]), f"layer {i} weights differ"


def test_res_snapshot_round_trip(tmp_path, rng):
    layers = [("res", 6, 6, 2, 4, 1, 2, 1)]
    fp = 20
    net = P.Network(C.to_spec(layers), 2, fp, seed=5)
    net.save_snapshot(str(tmp_path))
    names = [n for n, _ in net.params(0)]
    assert sorted(os.listdir(tmp_path)) == sorted(names)
    fresh = P.Network(C.to_spec(layers), 2, fp, seed=77)
    fresh.load_snapshot(str(tmp_path))
    for name, got in fresh.params(0):
        vals = np.array([float(t) for t in (tmp_path / name).read_text().split()])
        assert np.array_equal(got, O.set_public(vals, fp)), name
