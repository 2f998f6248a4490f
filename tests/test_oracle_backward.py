"""Pins the backward (training) half of the CPU oracle before it is trusted as
the parity reference for the HIP backward path.

Checks, by source:
  * the reference's known-answer tests: GPU.cu CutlassConvDgrad / Dgrad2 /
    Wgrad / Wgrad2 / ReduceSum, Layer.cu ReLUBackwardBasic /
    MaxpoolBackwardBasic / CNNBackwardBasic (inputs and expected values);
  * the reference's plaintext gradient fixtures (files/test/, converted into
    tests/golden/fixtures.npz), compared as the reference compares them
    (reconstructed doubles within ASSERT_EPSILON);
  * closed forms: dgrad and wgrad are the adjoints of fprop (exact mod-2^64
    inner-product identities), row2im is the scatter of the NHWC pooling
    windows, reduceSum / vectorExpand are numpy sums / repeats, the softmax
    gradient is numpy softmax in fixed point followed by the dividePublic
    truncation.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rand_u64
from piranha_amd import configs as C
from test_oracle import EPS, FP, GOLD, _conv_np, consistent_shares, rec, rec_int

U = np.uint64


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def dot(a, b):
    """Inner product mod 2^64."""
    return int((np.asarray(a, U) * np.asarray(b, U)).sum(dtype=U))


# ---------------------------------------------------------------------------
# gpu:: level

def test_conv_dgrad_kats(gold):
    ones = np.ones(9, U)
    f = np.array([1, 0, 1, 0, 1, 0, 1, 0, 1], U)
    got = O.conv_dgrad(ones, f, 1, 3, 3, 1, 3, 3, 1, 1, 1, 1, 1, 3, 3)
    assert got.tolist() == [2, 3, 2, 3, 5, 3, 2, 3, 2]  # GPU.cu:101-129
    got = O.conv_dgrad(gold["dgrad2_delta"].astype(U), gold["dgrad2_filters"].astype(U),
                       2, 3, 3, 10, 3, 3, 4, 1, 1, 2, 1, 5, 5)
    assert got.tolist() == gold["dgrad2_expected"].astype(U).tolist()  # GPU.cu:131-172


def test_conv_wgrad_kats(gold):
    ones = np.ones(9, U)
    im = np.array([1, 2, 1, 2, 3, 2, 1, 2, 1], U)
    got = O.conv_wgrad(ones, im, 1, 3, 3, 1, 3, 3, 1, 3, 3, 1, 1, 1, 1)
    assert got.tolist() == [8, 11, 8, 11, 15, 11, 8, 11, 8]  # GPU.cu:174-202
    got = O.conv_wgrad(gold["wgrad2_delta"].astype(U), gold["wgrad2_input"].astype(U),
                       2, 3, 3, 10, 5, 5, 4, 3, 3, 1, 1, 2, 1)
    assert got.tolist() == gold["wgrad2_expected"].astype(U).tolist()  # GPU.cu:204-245


GRAD_SHAPES = [  # (b, H, W, Din, Dout, f, pad, stride)
    (2, 5, 5, 4, 10, 3, 1, 2), (1, 7, 6, 3, 5, 3, 0, 1), (2, 6, 7, 2, 3, 3, 2, 2), (1, 8, 9, 3, 4, 5, 2, 3),
    (2, 4, 4, 33, 40, 3, 1, 1), (1, 9, 9, 2, 3, 3, 1, 2),
]


@pytest.mark.parametrize("b,H,W,Din,Dout,f,pad,s", GRAD_SHAPES)
def test_conv_grads_are_adjoints_of_fprop(rng, b, H, W, Din, Dout, f, pad, s):
    """<fprop(x, w), dy> = <x, dgrad(dy, w)> = <w, wgrad(dy, x)> mod 2^64: any
    wrong tap, flip, stride or padding index breaks one of the identities."""
    OH, OW = (H + 2 * pad - f) // s + 1, (W + 2 * pad - f) // s + 1
    x, w = rand_u64(rng, b * H * W * Din), rand_u64(rng, Dout * f * f * Din)
    dy = rand_u64(rng, b * OH * OW * Dout)
    y = O.conv_fprop(x, w, b, H, W, Din, Dout, f, f, pad, pad, s)
    assert np.array_equal(y, _conv_np(x, w, b, H, W, Din, Dout, f, pad, s))
    dx = O.conv_dgrad(dy, w, b, OH, OW, Dout, f, f, Din, pad, pad, s, 1, H, W)
    dw = O.conv_wgrad(dy, x, b, OH, OW, Dout, H, W, Din, f, f, pad, pad, s, 1)
    assert dot(y, dy) == dot(x, dx) == dot(w, dw)
    # and elementwise: dx / dw against one-hot probes of the same identities
    for _ in range(3):
        i = int(rng.integers(0, x.size))
        e = np.zeros(x.size, U)
        e[i] = 1
        assert int(dx[i]) == dot(O.conv_fprop(e, w, b, H, W, Din, Dout, f, f, pad, pad, s), dy)
        j = int(rng.integers(0, w.size))
        e = np.zeros(w.size, U)
        e[j] = 1
        assert int(dw[j]) == dot(O.conv_fprop(x, e, b, H, W, Din, Dout, f, f, pad, pad, s), dy)


def test_reduce_sum_kat_and_closed_form(rng):
    a = np.array([1, 1, 2, 1, 2, 2, 1, 2], U)
    assert O.reduce_sum(a, np.zeros(4, U), True, 2, 4).tolist() == [3, 3, 3, 3]  # GPU.cu:285-298
    rows, cols = 10, 37
    a, b0 = rand_u64(rng, rows * cols), rand_u64(rng, rows)
    want = (b0 + a.reshape(rows, cols).sum(1, dtype=U)).astype(U)
    assert np.array_equal(O.reduce_sum(a, b0, False, rows, cols), want)  # accumulates
    b1 = rand_u64(rng, cols)
    assert np.array_equal(O.reduce_sum(a, b1, True, rows, cols), (b1 + a.reshape(rows, cols).sum(0, dtype=U)).astype(U))


def _row2im_np(rows, batch, H, W, C, f, s):
    OH, OW = (H - f) // s + 1, (W - f) // s + 1
    R = rows.reshape(batch, OH, OW, C, f, f)
    out = np.zeros((batch, H, W, C), U)
    for dh in range(f):
        for dw in range(f):
            out[:, dh:dh + s * OH:s, dw:dw + s * OW:s, :] += R[..., dh, dw]
    return out.ravel()


@pytest.mark.parametrize("batch,H,W,C,f,s", [(2, 2, 2, 2, 2, 1), (3, 9, 9, 2, 3, 2), (2, 8, 6, 3, 2, 2),
                                             (1, 7, 10, 4, 3, 2), (2, 6, 5, 3, 3, 1)])
def test_vector_expand_and_row2im(rng, batch, H, W, C, f, s):
    OH, OW = (H - f) // s + 1, (W - f) // s + 1
    d = rand_u64(rng, batch * OH * OW * C)
    assert np.array_equal(O.vector_expand(d, f * f), np.repeat(d, f * f))
    rows = rand_u64(rng, d.size * f * f)
    # MaxpoolLayer.cu:169-171 passes (imageWidth, imageHeight) into (imageHeight, imageWidth)
    got = O.maxpool_row2im(rows, np.zeros(batch * H * W * C, U), W, H, f, C, batch, s)
    assert np.array_equal(got, _row2im_np(rows, batch, H, W, C, f, s))


# ---------------------------------------------------------------------------
# RSS convolution(kDgrad / kWgrad): local products + dividePublic3out3

@pytest.mark.parametrize("op", [O.CONV_DGRAD, O.CONV_WGRAD])
def test_convolution_grad_closed_form(rng, op):
    batch, H, W, Din, Dout, f, s, pad, fp = 2, 6, 5, 3, 4, 3, 2, 1, 13
    OH, OW = (H + 2 * pad - f) // s + 1, (W + 2 * pad - f) // s + 1
    dy = rng.integers(-(1 << 16), 1 << 16, batch * OH * OW * Dout).astype(np.int64).view(U)
    other = rng.integers(-(1 << 16), 1 << 16, (Dout * f * f * Din) if op == O.CONV_DGRAD
                         else batch * H * W * Din).astype(np.int64).view(U)
    A, B = consistent_shares(rng, dy), consistent_shares(rng, other)
    out = O.convolution_grad(op, A, B, batch, H, W, f, Din, Dout, s, pad, fp)
    if op == O.CONV_DGRAD:
        raw = O.conv_dgrad(dy, other, batch, OH, OW, Dout, f, f, Din, pad, pad, s, 1, H, W)
    else:
        raw = O.conv_wgrad(dy, other, batch, OH, OW, Dout, H, W, Din, f, f, pad, pad, s, 1)
    t = (raw - U(1 << fp)).astype(U).view(np.int64)
    want = np.trunc(t.astype(np.float64) / float(1 << fp)).astype(np.int64) + 1
    assert np.array_equal(rec_int(out), want)
    # placeholder layout: A (v, 0), B (0, 0), C (0, v)
    assert not out[1].any() and not out[0, 1].any() and not out[2, 0].any()


# ---------------------------------------------------------------------------
# Layer-level KATs (Layer.cu)

def test_relu_backward_kat():
    x = O.init_list([-2, -3, 4, 3, 3.5, 1, -1.5, -1, 2, 1 << 12, -17, (1 << 18) - 1], FP)
    _, rp = O.relu_forward(x)
    delta = O.init_list([-7.30e-1, -6.90e-2, -1.60e+0, -1.50e-1, -7.20e-1, 1.60e+0, 1.50e-1, 1.90e+0, -3.00e-1,
                         -9.80e-1, 4.10e-1, -9.70e-1], FP)
    got = O.relu_backward(delta, rp)
    want = [0, 0, -1.60e+0, -1.50e-1, -7.20e-1, 1.60e+0, 0, 0, -3.00e-1, -9.80e-1, 0, -9.70e-1]
    assert np.allclose(rec(got), want, atol=EPS)  # Layer.cu:238-265


def test_maxpool_backward_kat():
    x = O.init_list([1, -1, -3, 2, 3, -2, 0, 1, 2, -1, 0, 0, -2, -3, -1, 3], FP)
    _, mp = O.maxpool_forward(x, 2, 2, 2, 2, 2, 1, FP)
    got = O.maxpool_backward(O.init_list([1, 2, -1, -2], FP), mp, 2, 2, 2, 2, 2, 1)
    want = [0, 0, 0, 2, 1, 0, 0, 0, -1, 0, 0, 0, 0, 0, 0, -2]
    assert np.allclose(rec(got), want, atol=EPS)  # Layer.cu:341-377


CNN_BASIC_W = [
    -0.0017646551, 0.06236978, -0.1734625, -0.2251778, -0.004670009, 0.008731261, 0.12644096, -0.071232304,
    -0.09078175, -0.15610139, 0.18688585, 0.09318139, -0.1939936, -0.046330914, 0.06320529, -0.09716192,
    -0.020917177, 0.14142673, -0.15979224, -0.218654, 0.19572432, -0.09187673, -0.037991285, -0.108501494,
    -0.10263957, -0.14838351, -0.04850757, 0.20364694, 0.024940625, -0.16467114, 0.08561109, -0.059671625,
    0.17637877, -0.1527774, 0.21342279, -0.22074954, -0.13758895, -0.121511586, 0.114238426, -0.10453278,
    0.03987722, 0.23431943, 0.2026092, 0.14871348, 0.012395948, -0.008504704, -0.22007395, 0.09354602,
    0.105174676, 0.13819723, -0.12084066, 0.15074591, -0.1703105, 0.03184168]
CNN_BASIC_X = [1.0, -2.0, 2.0, 1.0, 0.0, 3.0, 0.0, -1.0, 0.0, 2.0, 1.0, 4.0, 1.0, 0.0, -1.0, 1.0, 2.0, -1.0, 3.0,
               1.0, -2.0, 0.0, -3.0, -1.0, -1.0, 1.0, 0.0, 2.0, -1.0, -4.0, 1.0, 0.0, 0.0, 3.0, -4.0, 4.0]
CNN_BASIC_ACT = [1.0683895, -0.859313, -0.40425906, -0.70475876, -0.29679593, 1.2869551, 0.058154106, -0.3632456,
                 0.89918756, -1.1725695, -0.95549667, -0.56751496, -0.6797321, -0.7624123, 0.804101, 0.40834528,
                 0.4625425, -0.63965005, 0.15450963, -1.047125, 0.54290164, 0.7239378, 0.29794204, 0.42340738]
CNN_BASIC_GRADIN = [3.0, 0.0, -1.0, 3.0, -1.0, 3.0, -4.0, 3.0, -2.0, -2.0, 2.0, 0.0, -2.0, 1.0, 0.0, -4.0, 2.0, 3.0,
                    -4.0, 1.0, -3.0, 0.0, -4.0, 0.0]
CNN_BASIC_GRADOUT = [
    -0.2847412, -0.45979947, 1.8705214, 0.5668252, -0.18664983, -0.69746524, 1.3630025, 0.3919149, -0.7356124,
    -0.6634069, 0.38908872, 0.58013153, 0.19281238, 1.2523558, -0.69073474, -1.2081335, 0.084548354, 0.71949667,
    0.13305593, 0.5158497, -0.4520465, 0.08326872, 0.3032997, 1.0061853, 0.5968272, 1.163979, 2.018204,
    0.0052001923, -1.0454829, 0.9028375, 0.27743158, 0.85356665, 0.34817737, -0.22450069, 0.19403028, -0.8145878]


def test_cnn_backward_kat():
    """Layer.cu:715-770: 3x3x2 -> 3 channels, f 3, stride 2, pad 1, batch 2."""
    W = O.init_list(CNN_BASIC_W, FP)
    x = O.init_list(CNN_BASIC_X, FP)
    act = O.cnn_forward(x, W, 2, 3, 3, 2, 3, 3, 2, 1, FP)
    assert np.allclose(rec(act), CNN_BASIC_ACT, atol=EPS)
    deltas, _ = O.cnn_backward(O.init_list(CNN_BASIC_GRADIN, FP), x, W, 2, 3, 3, 2, 3, 3, 2, 1, FP, 5)
    assert np.allclose(rec(deltas), CNN_BASIC_GRADOUT, atol=EPS)


# ---------------------------------------------------------------------------
# Plaintext gradient fixtures (files/test)

def test_fc_backward_fixture(gold):
    """Layer.cu:118-180 (log_learning_rate = 6)."""
    P = lambda k: O.set_public(gold[k], FP)
    deltas, W2, b2 = O.fc_backward(P("fc_gradin"), P("fc_input"), P("fc_weight"), P("fc_bias"), 128, 16, 10, FP, 6)
    assert np.allclose(rec(deltas), gold["fc_gradout"], atol=EPS)
    assert np.allclose(rec(W2), gold["fc_newweight"], atol=EPS)
    assert np.allclose(rec(b2), gold["fc_newbias"], atol=EPS)


def test_fc_backward_closed_form(rng):
    """Exact plaintext restatement of the three updates on small integers."""
    batch, d_in, d_out, fp, lr = 8, 5, 4, 13, 3
    vals = lambda n: rng.integers(-(1 << 15), 1 << 15, n).astype(np.int64)
    dl, x, w, b = vals(batch * d_out), vals(batch * d_in), vals(d_in * d_out), vals(d_out)
    S = lambda v: consistent_shares(rng, v.view(U))
    deltas, W2, b2 = O.fc_backward(S(dl), S(x), S(w), S(b), batch, d_in, d_out, fp, lr)

    def tr(v, k):
        t = (v.astype(U) - U(1 << k)).view(np.int64)
        return np.trunc(t.astype(np.float64) / float(1 << k)).astype(np.int64) + 1

    D, X, Wm = dl.reshape(batch, d_out), x.reshape(batch, d_in), w.reshape(d_in, d_out)
    assert np.array_equal(rec_int(deltas), tr((D @ Wm.T).ravel(), fp))
    assert np.array_equal(rec_int(W2), w - tr((X.T @ D).ravel(), fp + lr))
    # reduceSum(delta, db, false, d_out, batch) reads delta as [d_out][batch]
    db = dl.reshape(d_out, batch).sum(1)
    assert np.array_equal(rec_int(b2), b - tr(db, lr))


def test_relu_backward_fixture(gold):
    _, rp = O.relu_forward(O.set_public(gold["relu_input"], FP))
    got = O.relu_backward(O.set_public(gold["relu_gradin"], FP), rp)
    assert np.allclose(rec(got), gold["relu_gradout"], atol=EPS)


def test_maxpool_backward_fixture(gold):
    """Layer.cu:379-420: 9x9x2, pool 3, stride 2, batch 128."""
    _, mp = O.maxpool_forward(O.set_public(gold["maxpool_input"], FP), 128, 9, 9, 2, 3, 2, FP)
    got = O.maxpool_backward(O.set_public(gold["maxpool_gradin"], FP), mp, 128, 9, 9, 2, 3, 2)
    assert np.allclose(rec(got), gold["maxpool_gradout"], atol=EPS)


def test_cnn_backward_fixture(gold):
    """Layer.cu:772-840: 9x9x2 -> 3 channels, f 3, stride 2, pad 1, batch 128."""
    P = lambda k: O.set_public(gold[k], FP)
    deltas, W2 = O.cnn_backward(P("conv_gradin"), P("conv_input"), P("conv_weight"), 128, 9, 9, 2, 3, 3, 2, 1, FP, 6)
    assert np.allclose(rec(deltas), gold["conv_gradout"], atol=EPS)
    assert np.allclose(rec(W2), gold["conv_newweight"], atol=EPS)


def test_secureml_backward_chain_fixture(gold):
    """One SecureML training step: the reveal-softmax delta and every layer's
    outgoing delta against files/test/secureml/gradin_layer* (dL/d output of
    layer k), all computed with the weights before the update."""
    P = lambda k: O.set_public(gold[k], FP)
    params = {0: (P("secureml_weight0"), P("secureml_bias0")), 2: (P("secureml_weight2"), P("secureml_bias2")),
              4: (P("secureml_weight4"), P("secureml_bias4"))}
    x = P("secureml_input")
    fwd = O.net_forward(C.SECUREML, 128, FP, params, x)
    labels = gold["secureml_labels"]
    d5 = O.reveal_softmax_grad(fwd[-1][0], labels, 10, FP, O.log_mini_batch(128))
    assert np.allclose(rec(d5), gold["secureml_gradin5"], atol=1e-5)
    deltas, _ = O.net_backward(C.SECUREML, 128, FP, [params.get(i) for i in range(6)], x, fwd, labels, 6)
    assert np.array_equal(deltas[5], O.relu_backward(d5, fwd[5][1]))
    for k in range(5):
        assert np.allclose(rec(deltas[k + 1]), gold[f"secureml_gradin{k}"], atol=1e-5), k


# ---------------------------------------------------------------------------
# Reveal-softmax gradient closed form

@pytest.mark.parametrize("batch,n_classes", [(128, 10), (2, 10), (16, 7)])
def test_reveal_softmax_grad_closed_form(rng, batch, n_classes):
    fp = FP
    logits = rng.uniform(-6, 6, batch * n_classes)
    act = consistent_shares(rng, (np.trunc(logits * (1 << fp)).astype(np.int64)).view(U))
    lab = np.zeros((batch, n_classes))
    lab[np.arange(batch), rng.integers(0, n_classes, batch)] = 1
    lmb = O.log_mini_batch(batch)
    got = O.reveal_softmax_grad(act, lab.ravel(), n_classes, fp, lmb)
    z = rec_int(act).astype(np.float64) / (1 << fp)
    e = np.exp(z.reshape(batch, n_classes))
    sm = np.trunc(e / e.sum(1, keepdims=True) * (1 << fp)).astype(np.int64).ravel()
    t = sm - (lab.ravel() * (1 << fp)).astype(np.int64)
    if lmb > 1:
        d = 1 << lmb
        t = np.trunc((t - d).astype(np.float64) / d).astype(np.int64) + 1
    # numpy's exp may differ from the C library's in the last place, which can
    # move a fixed-point value by one unit
    assert np.abs(rec_int(got) - t).max() <= 1
    assert (rec_int(got) == t).mean() > 0.99
