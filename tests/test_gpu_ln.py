"""LayerNorm (LNLayer.cu:66-284) on the GPU against the oracle, bit for bit:
the layer's forward and backward with its gamma / beta update in the fused and
op-by-op forms, on realistic and on arbitrary shares; the reference's own
known answers (Layer.cu:838-1000); networks from the reference's LN model
files, forward and training; hipGraph capture of a forward and of a training
step (the revealed inverse square root runs as a host node); snapshots with
gamma{i} / beta{i} files."""
import os

import numpy as np
import pytest

import oracle as O
import piranha_amd as P
from conftest import rand_u64
from piranha_amd import configs as C

pytestmark = pytest.mark.gpu

U = np.uint64
MODES = [P.MODE_FUSED, P.MODE_PLAIN]
HERE = os.path.dirname(os.path.abspath(__file__))
MODELS = os.path.join(HERE, "golden", "models")


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(HERE, "golden", "fixtures.npz"))


def rec(a, fp):
    return O.reconstruct(a)[0].view(np.int64) / float(1 << fp)


def run_layer(rng, mode, batch, dim, fp, lr, x, delta, gamma_beta=None):
    net = P.Network([(P.LAYER_LN, dim)], batch, fp, seed=1, mode=mode)
    net.set_log_learning_rate(lr)
    if gamma_beta is not None:
        net.set_weight_shares(0, *gamma_beta)
    g, b = net.weights(0)
    dx = P.to_dev(x)
    net.forward_shares(dx.ptr)
    act, xhat, inv = O.ln_forward(x, g, b, batch, fp)
    assert np.array_equal(net.activation(0), act), "activations differ"
    dd = P.to_dev(delta)
    net.layer_backward(0, dd.ptr, dx.ptr)
    want_d, _, want_g, want_b = O.ln_backward(delta, xhat, inv, g, b, batch, fp, lr)
    assert np.array_equal(net.delta(0), want_d), "deltas differ"
    g2, b2 = net.weights(0)
    assert np.array_equal(g2, want_g) and np.array_equal(b2, want_b), "gamma / beta updates differ"
    return net


@pytest.mark.parametrize("mode", MODES, ids=["fused", "plain"])
@pytest.mark.parametrize("batch,dim,fp,lr", [(1, 8, 26, 5), (2, 48, 26, 3), (16, 100, 13, 5), (128, 784, 20, 6)])
def test_layer_realistic(rng, mode, batch, dim, fp, lr):
    x = O.set_public(rng.normal(0, 1, batch * dim), fp)
    delta = O.set_public(rng.normal(0, 0.05, batch * dim), fp)
    gb = (O.set_public(rng.uniform(0.5, 1.5, dim), fp), O.set_public(rng.normal(0, 0.1, dim), fp))
    run_layer(rng, mode, batch, dim, fp, lr, x, delta, gb)


@pytest.mark.parametrize("mode", MODES, ids=["fused", "plain"])
def test_layer_arbitrary_shares(rng, mode):
    # arbitrary 64-bit shares: every division, reveal and host inverse square
    # root meets values of any magnitude (including negative variances)
    batch, dim, fp = 4, 33, 13
    run_layer(rng, mode, batch, dim, fp, 5, rand_u64(rng, 3, 2, batch * dim), rand_u64(rng, 3, 2, batch * dim),
              (rand_u64(rng, 3, 2, dim), rand_u64(rng, 3, 2, dim)))


@pytest.mark.parametrize("mode", MODES, ids=["fused", "plain"])
def test_reference_kats(gold, mode):
    fp = 26
    net = P.Network([(P.LAYER_LN, 8)], 2, fp, seed=1, mode=mode)
    x = O.set_public(np.array([1, 2, 3, 4, -1, -2, 0, 1, 4, 2, 3, 1, -1, 1, 0, -2], float), fp)
    dx = P.to_dev(x)
    net.forward_shares(dx.ptr)
    want = [0, 0.534522, 1.069044, 1.603567, -1.069044, -1.603567, -0.534522, 0,
            1.603567, 0.534522, 1.069044, 0, -1.069044, 0, -0.534522, -1.603567]
    assert np.allclose(rec(net.activation(0), fp), want, atol=1e-2)  # Layer.cu:861-885
    # Layer.cu:943-1000: forward, backward at log_learning_rate 3, gamma and beta
    net = P.Network([(P.LAYER_LN, 48)], 2, fp, seed=1, mode=mode)
    net.set_log_learning_rate(3)
    dx = P.to_dev(O.set_public(gold["lnbwd_input"], fp))
    net.forward_shares(dx.ptr)
    dd = P.to_dev(O.set_public(gold["lnbwd_delta"], fp))
    net.layer_backward(0, dd.ptr, dx.ptr)
    assert np.allclose(rec(net.delta(0), fp), gold["lnbwd_expected"], atol=1e-2)
    g, b = net.weights(0)
    assert np.allclose(rec(g, fp), gold["lnbwd_gamma"], atol=1e-3)
    assert np.allclose(rec(b, fp), gold["lnbwd_beta"], atol=1e-3)


def one_hot(rng, batch, n_classes):
    lab = np.zeros((batch, n_classes))
    lab[np.arange(batch), rng.integers(0, n_classes, batch)] = 1.0
    return lab.ravel()


@pytest.mark.parametrize("mode", MODES, ids=["fused", "plain"])
@pytest.mark.parametrize("fname,batch,fp", [("lenet-ln-norelu.json", 4, 20), ("lenet-test.json", 2, 20)])
def test_ln_networks(fname, batch, fp, mode):
    layers = C.from_model_json(os.path.join(MODELS, fname))
    assert any(l[0] == "ln" for l in layers)
    rng = np.random.default_rng(3)
    net = P.Network(C.to_spec(layers), batch, fp, seed=6, mode=mode)
    params = [net.weights(i) if l[0] in ("fc", "cnn", "ln") else None for i, l in enumerate(layers)]
    x = rng.uniform(0.0, 1.0, batch * C.input_size(layers))
    net.forward(x)
    fwd = O.net_forward(layers, batch, fp, {i: p for i, p in enumerate(params) if p is not None},
                        O.set_public(x, fp))
    for i, (act, der) in enumerate(fwd):
        assert np.array_equal(net.activation(i), act), f"layer {i} ({layers[i][0]}) activation differs"
        if der is not None:
            assert np.array_equal(net.derivative(i), der), f"layer {i} derivative differs"
    labels = one_hot(rng, batch, C.output_size(layers))
    net.set_log_learning_rate(5)
    net.backward(labels)
    deltas, new_params = O.net_backward(layers, batch, fp, params, O.set_public(x, fp), fwd, labels, 5)
    for i in range(len(layers) - 1, -1, -1):
        assert np.array_equal(net.delta(i), deltas[i]), f"layer {i} ({layers[i][0]}) delta differs"
        if params[i] is not None:
            W, b = net.weights(i)
            assert np.array_equal(W, new_params[i][0]), f"layer {i} weights differ"
            if b is not None:
                assert np.array_equal(b, new_params[i][1]), f"layer {i} biases differ"


def test_graph_capture_with_ln():
    layers = C.from_model_json(os.path.join(MODELS, "lenet-ln-norelu.json"))
    batch, fp, lr = 4, 20, 5
    rng = np.random.default_rng(8)
    x = rng.uniform(0.0, 1.0, batch * C.input_size(layers))
    labels = one_hot(rng, batch, C.output_size(layers))
    # forward graph: replays equal the eager forward
    net = P.Network(C.to_spec(layers), batch, fp, seed=2)
    net.forward(x)
    eager = net.activation(len(layers) - 1)
    net.capture()
    for _ in range(3):
        net.replay()
        net.synchronize()
        assert np.array_equal(net.activation(len(layers) - 1), eager)
    # training graph: the captured step follows the eager trajectory
    a, g = (P.Network(C.to_spec(layers), batch, fp, seed=2) for _ in range(2))
    for n in (a, g):
        n.set_log_learning_rate(lr)
        n.forward(x)
        n.set_train_labels(labels)
    traj = []
    for _ in range(3):
        a.forward(x)
        a.backward(labels)
        traj.append([a.weights(i) for i, l in enumerate(layers) if l[0] in ("fc", "cnn", "ln")])
    g.capture_train()
    for step in range(3):
        if step:
            g.replay_train()
            g.synchronize()
        got = [g.weights(i) for i, l in enumerate(layers) if l[0] in ("fc", "cnn", "ln")]
        for (w1, b1), (w2, b2) in zip(got, traj[step]):
            assert np.array_equal(w1, w2) and (b1 is None or np.array_equal(b1, b2)), f"step {step}"


def test_snapshot_gamma_beta(tmp_path, rng):
    fp = 20
    net = P.Network([(P.LAYER_LN, 12)], 3, fp, seed=1)
    net.set_weights(0, rng.uniform(0.5, 1.5, 12), rng.normal(0, 0.1, 12))
    net.save_snapshot(str(tmp_path))
    assert sorted(os.listdir(tmp_path)) == ["beta0", "gamma0"]
    fresh = P.Network([(P.LAYER_LN, 12)], 3, fp, seed=1)
    fresh.load_snapshot(str(tmp_path))
    for k, name in enumerate(("gamma0", "beta0")):
        vals = np.array([float(t) for t in (tmp_path / name).read_text().split()])
        assert np.array_equal(fresh.weights(0)[k], O.set_public(vals, fp))


def test_needs_eight_fractional_bits():
    # epsilon = 1 << (FLOAT_PRECISION - 8) (LNLayer.cu:119) needs FP >= 8
    net = P.Network([(P.LAYER_LN, 8)], 1, 5, seed=1)
    with pytest.raises(P.PirError, match="8 fractional bits"):
        net.forward(np.zeros(8))
