"""Average pooling on the GPU against the oracle, bit for bit: the
gpu::averagepool_im2row / averagepool_expand_delta entry points
(convolution.cuh:194-207, 339-371, 412-437), AveragepoolLayer forward and
backward (AveragepoolLayer.cu:49-125) in the fused and op-by-op forms on
arbitrary 64-bit shares (so the division by f*f meets every operand
magnitude), the reference's layer known-answer tests, and whole networks from
the reference's model files that use average pooling, forward and training."""
import json
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
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models")

GEOMS = [(2, 2, 2, 2, 2, 1), (3, 9, 9, 2, 3, 2), (2, 8, 6, 3, 2, 2), (2, 16, 16, 8, 2, 2), (1, 7, 5, 4, 3, 1),
         (2, 5, 4, 2, 2, 2), (1, 4, 4, 3, 1, 1)]


@pytest.mark.parametrize("batch,H,W,Cc,f,s", GEOMS)
@pytest.mark.parametrize("padding", [0, 1])
def test_averagepool_im2row(rng, batch, H, W, Cc, f, s, padding):
    im = rand_u64(rng, batch * H * W * Cc)
    n = P.lib().pir_averagepool_im2row_size(W, H, f, Cc, batch, s, padding)
    want = O.averagepool_im2row(im, W, H, f, Cc, batch, s, padding)
    assert n == want.size
    dim, out = P.to_dev(im), P.Buffer((n,), U)
    P.check(P.lib().pir_averagepool_im2row(dim.ptr, out.ptr, W, H, f, Cc, batch, s, padding, None))
    assert np.array_equal(out.numpy(), want)


def test_expand_delta_kat():
    din = P.to_dev(np.array([1, 0, 3, 4], U))
    out = P.to_dev(np.zeros(16, U))
    P.check(P.lib().pir_averagepool_expand_delta(din.ptr, 4, out.ptr, 16, 2, 4, None))
    assert out.numpy().tolist() == [1, 0, 1, 0, 1, 0, 1, 0, 3, 4, 3, 4, 3, 4, 3, 4]  # GPU.cu:516-535


@pytest.mark.parametrize("n_in,Din,P_,extra", [(4, 2, 4, 0), (96, 3, 9, 17), (1 << 14, 64, 4, 0)])
def test_expand_delta_random(rng, n_in, Din, P_, extra):
    x = rand_u64(rng, n_in)
    n_out = n_in * P_ + extra
    dx, out = P.to_dev(x), P.to_dev(np.zeros(n_out, U))
    P.check(P.lib().pir_averagepool_expand_delta(dx.ptr, n_in, out.ptr, n_out, Din, P_, None))
    assert np.array_equal(out.numpy(), O.averagepool_expand_delta(x, n_out, Din, P_))


def test_expand_delta_rejects_short_output():
    dx, out = P.to_dev(np.zeros(8, U)), P.to_dev(np.zeros(31, U))
    assert P.lib().pir_averagepool_expand_delta(dx.ptr, 8, out.ptr, 31, 2, 4, None) != 0


@pytest.mark.parametrize("net_mode", MODES, ids=["fused", "plain"])
@pytest.mark.parametrize("batch,H,W,Cc,f,s", GEOMS)
def test_layer_forward_backward(rng, net_mode, batch, H, W, Cc, f, s):
    fp = 13
    net = P.Network([(P.LAYER_AVGPOOL, H, W, Cc, f, s)], batch, fp, seed=1, mode=net_mode)
    x = rand_u64(rng, 3, 2, batch * H * W * Cc)
    dxx = P.to_dev(x)
    net.forward_shares(dxx.ptr)
    assert np.array_equal(net.activation(0), O.avgpool_forward(x, batch, H, W, Cc, f, s))
    OH, OW = (H - f) // s + 1, (W - f) // s + 1
    delta = rand_u64(rng, 3, 2, batch * OH * OW * Cc)
    dd = P.to_dev(delta)
    if not expand_fits(("avgpool", H, W, Cc, f, s)):
        with pytest.raises(Exception):  # the reference would write past the deltas buffer
            net.layer_backward(0, dd.ptr, dxx.ptr)
        return
    net.layer_backward(0, dd.ptr, dxx.ptr)
    assert np.array_equal(net.delta(0), O.avgpool_backward(delta, batch, H, W, Cc, f))


@pytest.mark.parametrize("net_mode", MODES, ids=["fused", "plain"])
def test_layer_kats(net_mode):
    # Layer.cu:419-491 (2x2 images, 2 features, batch 2, 2x2 pool, stride 1)
    fp = 13
    net = P.Network([(P.LAYER_AVGPOOL, 2, 2, 2, 2, 1)], 2, fp, seed=1, mode=net_mode)
    x = O.set_public(np.array([1, -1, -3, 2, 3, -2, 0, 1, 2, -1, 0, 0, -2, -3, -1, 3], float), fp)
    dxx = P.to_dev(x)
    net.forward_shares(dxx.ptr)
    got = O.reconstruct(net.activation(0))[0].view(np.int64) / float(1 << fp)
    assert got.tolist() == [0.25, 0.0, -0.25, -0.25]
    dd = P.to_dev(O.set_public(np.array([1, 0, -1, -2], float), fp))
    net.layer_backward(0, dd.ptr, dxx.ptr)
    got = O.reconstruct(net.delta(0))[0].view(np.int64) / float(1 << fp)
    assert got.tolist() == [0.25, 0.0] * 4 + [-0.25, -0.5] * 4


def expand_fits(l):
    """Whether an average-pool layer's expanded delta fits its deltas buffer."""
    if l[0] != "avgpool":
        return True
    H, W, _, f, s = l[1:]
    return ((H - f) // s + 1) * ((W - f) // s + 1) * f * f <= H * W


def model(name):
    return C.from_model_json(os.path.join(GOLDEN, name))


def one_hot(rng, batch, n_classes):
    lab = np.zeros((batch, n_classes))
    lab[np.arange(batch), rng.integers(0, n_classes, batch)] = 1.0
    return lab.ravel()


NETS = [("lenet-norelu-avgpool.json", 4, 13), ("alexnet-cifar10.json", 2, 20), ("vgg16-cifar10.json", 1, 26)]


@pytest.mark.parametrize("mode", MODES, ids=["fused", "plain"])
@pytest.mark.parametrize("fname,batch,fp", NETS, ids=[n[0] for n in NETS])
def test_network_forward_and_training(fname, batch, fp, mode):
    layers = model(fname)
    assert any(l[0] == "avgpool" for l in layers)
    rng = np.random.default_rng(5)
    net = P.Network(C.to_spec(layers), batch, fp, seed=9, mode=mode)
    params = [net.weights(i) if l[0] in ("fc", "cnn") else None for i, l in enumerate(layers)]
    x = rng.uniform(0.0, 1.0, batch * C.input_size(layers))
    net.forward(x)
    fwd = O.net_forward(layers, batch, fp, {i: p for i, p in enumerate(params) if p is not None},
                        O.set_public(x, fp))
    for i, (act, der) in enumerate(fwd):
        assert np.array_equal(net.activation(i), act), f"layer {i} ({layers[i][0]}) activation differs"
        if der is not None:
            assert np.array_equal(net.derivative(i), der), f"layer {i} derivative differs"
    labels = one_hot(rng, batch, C.output_size(layers))
    lr = 5
    net.set_log_learning_rate(lr)
    if not all(expand_fits(l) for l in layers):
        # the reference's averagepool_expand_delta would write past the layer's
        # deltas (overlapping windows): the backward refuses instead
        with pytest.raises(Exception):
            net.backward(labels)
        return
    net.backward(labels)
    deltas, new_params = O.net_backward(layers, batch, fp, params, O.set_public(x, fp), fwd, labels, lr)
    assert np.array_equal(net.delta(len(layers)), O.reveal_softmax_grad(
        fwd[-1][0], labels, C.output_size(layers), fp, O.log_mini_batch(batch)))
    for i in range(len(layers) - 1, -1, -1):
        assert np.array_equal(net.delta(i), deltas[i]), f"layer {i} ({layers[i][0]}) delta differs"
        if params[i] is not None:
            assert np.array_equal(net.weights(i)[0], new_params[i][0]), f"layer {i} weights differ"


def test_model_files_parse():
    for fname, _, _ in NETS:
        with open(os.path.join(GOLDEN, fname)) as f:
            doc = json.load(f)
        assert len(model(fname)) == len(doc["model"])
