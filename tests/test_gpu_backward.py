"""Backward (training) path on the GPU against the oracle, bit for bit:
gpu::conv_dgrad / conv_wgrad (plain VALU and int8-limb MFMA), reduceSum,
vectorExpand, maxpool_row2im (conv.cuh:405-490, matrix.cuh, convolution.cuh),
RSS convolution(kDgrad / kWgrad) (RSS.inl:641-803), every layer's
Layer::backward with its weight update (FCLayer.cu:133-224, CNNLayer.cu:118-195,
ReLULayer.cu:91-123, MaxpoolLayer.cu:133-180) and whole training steps
(NeuralNetwork.cu:136-149, 207-267, 435-467), in the fused and op-by-op forms.

The reveal-softmax runs on the host in both (the opened logits are copied to
host memory and the C library's exp is applied, as the reference does), so the
softmax delta is compared exactly too.
"""
import os

import numpy as np
import pytest

import oracle as O
import piranha_amd as P
from conftest import rand_u64
from piranha_amd import configs as C

pytestmark = pytest.mark.gpu

U = np.uint64
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures.npz")


def general_shares(rng, n, bits=40):
    v = rng.integers(-(1 << bits), 1 << bits, size=(3, 2, n)).astype(np.int64).view(U)
    return np.ascontiguousarray(v)


def fixed_shares(rng, n, fp, scale=4.0):
    return O.init_list(rng.uniform(-scale, scale, n), fp)


# ---------------------------------------------------------------------------
# gpu:: level

def run_dgrad(dy, w, b, OH, OW, Dout, f, Din, pad, stride, H, W, algo):
    ddy, dw = P.to_dev(dy), P.to_dev(w)
    dx = P.Buffer((b * H * W * Din,), U)
    P.check(P.lib().pir_conv_dgrad(ddy.ptr, dw.ptr, dx.ptr, b, OH, OW, Dout, f, f, Din, pad, pad, stride, 1, H, W,
                                   algo, None))
    return dx.numpy()


def run_wgrad(dy, x, b, OH, OW, Dout, H, W, Din, f, pad, stride, algo):
    ddy, dxx = P.to_dev(dy), P.to_dev(x)
    dw = P.Buffer((Dout * f * f * Din,), U)
    P.check(P.lib().pir_conv_wgrad(ddy.ptr, dxx.ptr, dw.ptr, b, OH, OW, Dout, H, W, Din, f, f, pad, pad, stride, 1,
                                   algo, None))
    return dw.numpy()


ALGOS = [P.ALGO_PLAIN, P.ALGO_LIMB]


@pytest.mark.parametrize("algo", ALGOS, ids=["plain", "limb"])
def test_conv_grad_kats(algo):
    g = np.load(GOLD)
    ones = np.ones(9, U)
    f = np.array([1, 0, 1, 0, 1, 0, 1, 0, 1], U)
    assert run_dgrad(ones, f, 1, 3, 3, 1, 3, 1, 1, 1, 3, 3, algo).tolist() == [2, 3, 2, 3, 5, 3, 2, 3, 2]
    got = run_dgrad(g["dgrad2_delta"].astype(U), g["dgrad2_filters"].astype(U), 2, 3, 3, 10, 3, 4, 1, 2, 5, 5, algo)
    assert got.tolist() == g["dgrad2_expected"].astype(U).tolist()  # GPU.cu:131-172
    im = np.array([1, 2, 1, 2, 3, 2, 1, 2, 1], U)
    assert run_wgrad(ones, im, 1, 3, 3, 1, 3, 3, 1, 3, 1, 1, algo).tolist() == [8, 11, 8, 11, 15, 11, 8, 11, 8]
    got = run_wgrad(g["wgrad2_delta"].astype(U), g["wgrad2_input"].astype(U), 2, 3, 3, 10, 5, 5, 4, 3, 1, 2, algo)
    assert got.tolist() == g["wgrad2_expected"].astype(U).tolist()  # GPU.cu:204-245


GRAD_SHAPES = [  # (b, H, W, Din, Dout, f, pad, stride)
    (2, 9, 9, 2, 3, 3, 1, 2),     # the MNIST layer fixture's geometry
    (2, 8, 8, 64, 32, 3, 1, 1),   # dgrad gather with 32-channel K-steps
    (3, 6, 7, 33, 64, 3, 2, 1),   # pad 2, odd wgrad rows
    (2, 8, 8, 3, 64, 3, 1, 1),    # first-layer style: 27 wgrad rows
    (1, 10, 10, 16, 16, 5, 2, 3), # f 5, stride 3
    (9, 32, 32, 8, 8, 3, 1, 1),   # wgrad K = 9216 crosses the 8192-k flush
    (2, 4, 4, 512, 64, 3, 1, 1),  # deep channels, K = 4608 for dgrad... and wgrad rows 4608
]


@pytest.mark.parametrize("algo", ALGOS, ids=["plain", "limb"])
@pytest.mark.parametrize("b,H,W,Din,Dout,f,pad,s", GRAD_SHAPES)
def test_conv_grads_random(rng, algo, b, H, W, Din, Dout, f, pad, s):
    OH, OW = (H + 2 * pad - f) // s + 1, (W + 2 * pad - f) // s + 1
    dy, w, x = rand_u64(rng, b * OH * OW * Dout), rand_u64(rng, Dout * f * f * Din), rand_u64(rng, b * H * W * Din)
    want = O.conv_dgrad(dy, w, b, OH, OW, Dout, f, f, Din, pad, pad, s, 1, H, W)
    assert np.array_equal(run_dgrad(dy, w, b, OH, OW, Dout, f, Din, pad, s, H, W, algo), want)
    want = O.conv_wgrad(dy, x, b, OH, OW, Dout, H, W, Din, f, f, pad, pad, s, 1)
    assert np.array_equal(run_wgrad(dy, x, b, OH, OW, Dout, H, W, Din, f, pad, s, algo), want)


def test_reduce_sum_vector_expand_row2im(rng):
    L = P.lib()
    # GPU.cu:285-298 and the layer's reduceRows = false form
    a = P.to_dev(np.array([1, 1, 2, 1, 2, 2, 1, 2], U))
    b = P.to_dev(np.zeros(4, U))
    P.check(L.pir_reduce_sum(a.ptr, b.ptr, 1, 2, 4, None))
    assert b.numpy().tolist() == [3, 3, 3, 3]
    rows, cols = 10, 129
    av, bv = rand_u64(rng, rows * cols), rand_u64(rng, rows)
    da, db = P.to_dev(av), P.to_dev(bv)
    P.check(L.pir_reduce_sum(da.ptr, db.ptr, 0, rows, cols, None))
    assert np.array_equal(db.numpy(), O.reduce_sum(av, bv, False, rows, cols))
    v = rand_u64(rng, 1000)
    dv, out = P.to_dev(v), P.Buffer((9000,), U)
    P.check(L.pir_vector_expand(dv.ptr, 1000, out.ptr, 9, None))
    assert np.array_equal(out.numpy(), O.vector_expand(v, 9))
    for (batch, H, W, Cc, f, s) in [(2, 2, 2, 2, 2, 1), (3, 9, 9, 2, 3, 2), (2, 8, 6, 3, 2, 2), (1, 7, 10, 4, 3, 2)]:
        OH, OW = (H - f) // s + 1, (W - f) // s + 1
        r = rand_u64(rng, batch * OH * OW * Cc * f * f)
        base = rand_u64(rng, batch * H * W * Cc)
        dr, do = P.to_dev(r), P.to_dev(base)
        P.check(L.pir_maxpool_row2im(dr.ptr, do.ptr, W, H, f, Cc, batch, s, None))  # the layer's argument order
        assert np.array_equal(do.numpy(), O.maxpool_row2im(r, base, W, H, f, Cc, batch, s))


# ---------------------------------------------------------------------------
# RSS convolution(kDgrad / kWgrad)

@pytest.fixture(params=[P.MODE_FUSED, P.MODE_PLAIN], ids=["fused", "plain"])
def mode(request):
    P.set_mode(request.param)
    yield request.param
    P.set_mode(P.MODE_FUSED)


@pytest.mark.parametrize("op", [P.CONV_DGRAD, P.CONV_WGRAD], ids=["dgrad", "wgrad"])
@pytest.mark.parametrize("algo", ALGOS, ids=["plain", "limb"])
@pytest.mark.parametrize("b,H,W,f,Din,Dout,stride,pad", [(2, 9, 9, 3, 2, 3, 2, 1), (2, 8, 8, 3, 32, 64, 1, 1),
                                                         (3, 5, 6, 3, 5, 7, 1, 0)])
def test_rss_convolution_grad(rng, mode, op, algo, b, H, W, f, Din, Dout, stride, pad):
    fp = 13
    OH, OW = (H + 2 * pad - f) // stride + 1, (W + 2 * pad - f) // stride + 1
    A = general_shares(rng, b * OH * OW * Dout)
    Bv = general_shares(rng, Dout * f * f * Din if op == P.CONV_DGRAD else b * H * W * Din)
    n = b * H * W * Din if op == P.CONV_DGRAD else Dout * f * f * Din
    dA, dB, dC = P.to_dev(A), P.to_dev(Bv), P.Buffer((3, 2, n), U)
    P.check(P.lib().pir_rss_convolution_op(op, dA.ptr, dB.ptr, dC.ptr, b, H, W, f, Din, Dout, stride, pad, fp, algo,
                                           None))
    assert np.array_equal(dC.numpy(), O.convolution_grad(op, A, Bv, b, H, W, f, Din, Dout, stride, pad, fp))


# ---------------------------------------------------------------------------
# Layer::backward of single layers (random shares everywhere)

MODES = [P.MODE_FUSED, P.MODE_PLAIN]


@pytest.mark.parametrize("net_mode", MODES, ids=["fused", "plain"])
@pytest.mark.parametrize("batch,d_in,d_out", [(128, 16, 10), (64, 512, 512), (5, 33, 7)])
def test_fc_backward(rng, net_mode, batch, d_in, d_out):
    fp, lr = 13, 6
    net = P.Network([(P.LAYER_FC, d_in, d_out)], batch, fp, seed=1, mode=net_mode)
    net.set_log_learning_rate(lr)
    W, bias = general_shares(rng, d_in * d_out), general_shares(rng, d_out)
    x, delta = general_shares(rng, batch * d_in), general_shares(rng, batch * d_out)
    net.set_weight_shares(0, W, bias)
    dx, dd = P.to_dev(x), P.to_dev(delta)
    net.layer_backward(0, dd.ptr, dx.ptr)
    want_d, want_W, want_b = O.fc_backward(delta, x, W, bias, batch, d_in, d_out, fp, lr)
    got_W, got_b = net.weights(0)
    assert np.array_equal(net.delta(0), want_d)
    assert np.array_equal(got_W, want_W)
    assert np.array_equal(got_b, want_b)


@pytest.mark.parametrize("net_mode", MODES, ids=["fused", "plain"])
@pytest.mark.parametrize("batch,H,W,Din,Dout,f,s,pad", [(4, 9, 9, 2, 3, 3, 2, 1), (2, 16, 16, 64, 64, 3, 1, 1),
                                                        (2, 8, 8, 3, 64, 3, 1, 1), (3, 6, 6, 32, 16, 3, 1, 0)])
def test_cnn_backward(rng, net_mode, batch, H, W, Din, Dout, f, s, pad):
    fp, lr = 13, 5
    net = P.Network([(P.LAYER_CNN, H, W, Din, Dout, f, s, pad)], batch, fp, seed=1, mode=net_mode)
    net.set_log_learning_rate(lr)
    OH, OW = (H + 2 * pad - f) // s + 1, (W + 2 * pad - f) // s + 1
    Wt = general_shares(rng, Dout * f * f * Din)
    x, delta = general_shares(rng, batch * H * W * Din), general_shares(rng, batch * OH * OW * Dout)
    net.set_weight_shares(0, Wt)
    dx, dd = P.to_dev(x), P.to_dev(delta)
    net.layer_backward(0, dd.ptr, dx.ptr)
    want_d, want_W = O.cnn_backward(delta, x, Wt, batch, H, W, Din, Dout, f, s, pad, fp, lr)
    assert np.array_equal(net.delta(0), want_d)
    assert np.array_equal(net.weights(0)[0], want_W)


@pytest.mark.parametrize("net_mode", MODES, ids=["fused", "plain"])
def test_relu_backward(rng, net_mode):
    fp, n = 13, 4099
    net = P.Network([(P.LAYER_RELU, n)], 1, fp, seed=1, mode=net_mode)
    x = fixed_shares(rng, n, fp)
    x[1] = general_shares(rng, n)[1]  # party B's shares arbitrary too
    dxx = P.to_dev(x)
    net.forward_shares(dxx.ptr)
    _, rp = O.relu_forward(x)
    assert np.array_equal(net.derivative(0), rp)
    delta = general_shares(rng, n)
    dd = P.to_dev(delta)
    net.layer_backward(0, dd.ptr, dxx.ptr)
    assert np.array_equal(net.delta(0), O.relu_backward(delta, rp))


@pytest.mark.parametrize("net_mode", MODES, ids=["fused", "plain"])
@pytest.mark.parametrize("batch,H,W,Cc,pool,s", [(2, 2, 2, 2, 2, 1), (3, 9, 9, 2, 3, 2), (2, 8, 6, 3, 2, 2),
                                                 (2, 16, 16, 8, 2, 2), (2, 8, 8, 2, 4, 4), (1, 9, 9, 3, 3, 3)])
def test_maxpool_backward(rng, net_mode, batch, H, W, Cc, pool, s):
    fp = 13
    net = P.Network([(P.LAYER_MAXPOOL, H, W, Cc, pool, s)], batch, fp, seed=1, mode=net_mode)
    x = fixed_shares(rng, batch * H * W * Cc, fp)
    dxx = P.to_dev(x)
    net.forward_shares(dxx.ptr)
    _, mp = O.maxpool_forward(x, batch, H, W, Cc, pool, s, fp)
    assert np.array_equal(net.derivative(0), mp)
    OH, OW = (H - pool) // s + 1, (W - pool) // s + 1
    delta = general_shares(rng, batch * OH * OW * Cc)
    dd = P.to_dev(delta)
    net.layer_backward(0, dd.ptr, dxx.ptr)
    assert np.array_equal(net.delta(0), O.maxpool_backward(delta, mp, batch, H, W, Cc, pool, s))


# ---------------------------------------------------------------------------
# Whole training steps

def one_hot(rng, batch, n_classes):
    lab = np.zeros((batch, n_classes))
    lab[np.arange(batch), rng.integers(0, n_classes, batch)] = 1.0
    return lab.ravel()


def train_step(layers, batch, fp, lr, mode, seed=11):
    rng = np.random.default_rng(seed)
    net = P.Network(C.to_spec(layers), batch, fp, seed=seed, mode=mode)
    net.set_log_learning_rate(lr)
    params = [net.weights(i) if l[0] in ("fc", "cnn") else None for i, l in enumerate(layers)]
    x = rng.uniform(0.0, 1.0, batch * C.input_size(layers))
    labels = one_hot(rng, batch, C.output_size(layers))
    net.forward(x)
    net.backward(labels)
    return net, params, x, labels


def check_softmax_delta(got, want):
    assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} softmax delta shares differ"


TRAIN_CASES = [
    ("secureml", C.SECUREML, 16, 13, 6),
    ("lenet", C.LENET, 2, 13, 5),
    ("alexnet", C.ALEXNET_CIFAR, 2, 20, 5),
    ("vgg16", C.VGG16_CIFAR, 1, 26, 5),
]


@pytest.mark.parametrize("mode", MODES, ids=["fused", "plain"])
@pytest.mark.parametrize("name,layers,batch,fp,lr", TRAIN_CASES, ids=[c[0] for c in TRAIN_CASES])
def test_training_step(name, layers, batch, fp, lr, mode):
    if mode == P.MODE_PLAIN and name == "vgg16":
        pytest.skip("op-by-op VGG16 training parity is covered through the fused/plain equality test")
    net, params, x, labels = train_step(layers, batch, fp, lr, mode)
    L = len(layers)
    fwd = O.net_forward(layers, batch, fp, {i: p for i, p in enumerate(params) if p is not None},
                        O.set_public(x, fp))
    nc = fwd[-1][0].shape[-1] // batch
    out_delta = net.delta(L)
    want = O.reveal_softmax_grad(fwd[-1][0], labels, nc, fp, O.log_mini_batch(batch))
    check_softmax_delta(out_delta, want)
    deltas, new_params = O.net_backward(layers, batch, fp, params, O.set_public(x, fp), fwd, labels, lr,
                                        out_delta=out_delta)
    for i in range(L - 1, -1, -1):
        assert np.array_equal(net.delta(i), deltas[i]), f"layer {i} ({layers[i][0]}) delta differs"
        if params[i] is not None:
            W, b = net.weights(i)
            assert np.array_equal(W, new_params[i][0]), f"layer {i} weights differ"
            if b is not None:
                assert np.array_equal(b, new_params[i][1]), f"layer {i} biases differ"


def test_training_fused_equals_plain_vgg16():
    layers, batch, fp, lr = C.VGG16_CIFAR, 2, 26, 5
    outs = []
    for mode in (P.MODE_FUSED, P.MODE_PLAIN):
        net, params, _, _ = train_step(layers, batch, fp, lr, mode, seed=4)
        outs.append(([net.delta(i) for i in range(len(layers) + 1)],
                     [net.weights(i) for i, l in enumerate(layers) if l[0] in ("fc", "cnn")]))
    for i, (a, b) in enumerate(zip(outs[0][0], outs[1][0])):
        assert np.array_equal(a, b), f"delta {i} differs between fused and plain forms"
    for j, (a, b) in enumerate(zip(outs[0][1], outs[1][1])):
        assert np.array_equal(a[0], b[0]) and (a[1] is None or np.array_equal(a[1], b[1])), f"weights {j} differ"


def test_backward_drops_captured_graph():
    layers, batch, fp = C.LENET, 2, 13
    rng = np.random.default_rng(5)
    net = P.Network(C.to_spec(layers), batch, fp, seed=3)
    x = rng.uniform(0, 1, batch * C.input_size(layers))
    net.forward(x)
    net.capture()
    net.replay()
    net.backward(one_hot(rng, batch, 10))
    with pytest.raises(P.PirError, match="capture again"):
        net.replay()
    # the next forward (eager or captured) uses the updated weights
    net.forward(x)
    eager = net.activation(len(layers) - 1)
    net.capture()
    net.replay()
    net.synchronize()
    assert np.array_equal(net.activation(len(layers) - 1), eager)


def test_training_reduces_loss():
    """A few SGD steps on one fixed batch lower the softmax cross-entropy of the
    reconstructed logits (the delta chain and the updates point downhill)."""
    layers, batch, fp, lr = C.SECUREML, 32, 20, 3
    rng = np.random.default_rng(2)
    net = P.Network(C.to_spec(layers), batch, fp, seed=9)
    net.set_log_learning_rate(lr)
    x = rng.uniform(0.0, 1.0, batch * C.input_size(layers))
    labels = one_hot(rng, batch, 10)

    def loss():
        net.forward(x)
        z = O.reconstruct(net.activation(len(layers) - 1))[0].view(np.int64) / float(1 << fp)
        z = z.reshape(batch, 10)
        z = z - z.max(1, keepdims=True)
        logp = z - np.log(np.exp(z).sum(1, keepdims=True))
        return float(-(logp * labels.reshape(batch, 10)).sum(1).mean())

    first = loss()
    for _ in range(8):
        net.backward(labels)
        last = loss()
    assert last < first, (first, last)


def test_training_step_api_and_graph_replay():
    """pir_net_train_step equals forward + backward; a captured training step
    (capture_train runs one eager step, each replay one more) follows the same
    trajectory as eager steps, replays interleaved with the host reading
    weights between them."""
    layers, batch, fp, lr = C.LENET, 4, 13, 5
    rng = np.random.default_rng(21)
    x = rng.uniform(0.0, 1.0, batch * C.input_size(layers))
    labels = one_hot(rng, batch, 10)
    nets = [P.Network(C.to_spec(layers), batch, fp, seed=8) for _ in range(3)]
    for net in nets:
        net.set_log_learning_rate(lr)
        net.forward(x)  # fills the input buffer
        net.set_train_labels(labels)
    eager, api, graph = nets
    traj = []
    for step in range(3):
        eager.forward(x)
        eager.backward(labels)
        traj.append([eager.weights(i)[0] for i, l in enumerate(layers) if l[0] in ("fc", "cnn")])
    for step in range(3):
        api.train_step()
        got = [api.weights(i)[0] for i, l in enumerate(layers) if l[0] in ("fc", "cnn")]
        assert all(np.array_equal(a, b) for a, b in zip(got, traj[step])), f"train_step {step}"
    graph.capture_train()
    got = [graph.weights(i)[0] for i, l in enumerate(layers) if l[0] in ("fc", "cnn")]
    assert all(np.array_equal(a, b) for a, b in zip(got, traj[0]))
    for step in (1, 2):
        graph.replay_train()
        graph.synchronize()
        got = [graph.weights(i)[0] for i, l in enumerate(layers) if l[0] in ("fc", "cnn")]
        assert all(np.array_equal(a, b) for a, b in zip(got, traj[step])), f"replay {step}"
    assert np.array_equal(graph.delta(0), eager.delta(0))


def test_capture_train_first_step_allocations():
    """capture_train as the first training call: the output delta (batch x 10
    shares) has the size of the last FC product's split-K workspace (6 splits
    at K = 256), so the persistent backward buffers must exist before the
    eager warm-up step for the captured step to find every temporary cached."""
    layers = [("fc", 32, 256), ("relu", 256), ("fc", 256, 10), ("relu", 10)]
    batch, fp = 8, 13
    rng = np.random.default_rng(4)
    x = rng.uniform(0.0, 1.0, batch * 32)
    labels = one_hot(rng, batch, 10)
    nets = [P.Network(C.to_spec(layers), batch, fp, seed=2) for _ in range(2)]
    for net in nets:
        net.forward(x)
        net.set_train_labels(labels)
    eager, graph = nets
    graph.capture_train()  # one eager step + capture
    graph.replay_train()
    graph.synchronize()
    for _ in range(2):
        eager.train_step()
    for i in (0, 2):
        assert np.array_equal(graph.weights(i)[0], eager.weights(i)[0])
        assert np.array_equal(graph.weights(i)[1], eager.weights(i)[1])
