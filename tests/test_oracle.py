"""Pins the CPU oracle (oracle/pir_ref.cpp) before it is trusted as the parity
reference for the HIP path.

Checks, by source:
  * known-answer tests of the reference's own suites (src/test/GPU.cu,
    Func.cu, Layer.cu), with the same inputs and expected values;
  * the reference's plaintext fixtures (files/test/, converted to
    tests/golden/fixtures.npz by tests/golden/make_fixtures.py);
  * closed forms computed independently here with numpy / exact Python
    integers: mod-2^64 products, IEEE-754 truncation, [v >= 0], selects,
    window maxima and a plaintext fixed-point twin of whole networks;
  * share invariants of the placeholder precompute (values live in x0; party B
    keeps the setPublic fall-through value in shareB).
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import rand_u64
from piranha_amd import configs as C

FP = 26  # FLOAT_PRECISION default, globals.h:19-21
EPS = 1e-3  # ASSERT_EPSILON, util.cuh:19
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures.npz")


def rec(x, fp=FP):
    """Party A's reconstruction as doubles (copyToHost, util.cuh:55-79)."""
    return O.reconstruct(x)[0].view(np.int64) / float(1 << fp)


def rec_int(x):
    return O.reconstruct(x)[0].view(np.int64)


def consistent_shares(rng, v):
    """A real replicated sharing of v: x0 + x1 + x2 = v, party p holds (x_p, x_{p+1})."""
    v = np.asarray(v, np.uint64)
    x1, x2 = rand_u64(rng, v.size), rand_u64(rng, v.size)
    x0 = (v - x1 - x2).astype(np.uint64)
    xs = [x0, x1, x2]
    out = np.zeros((3, 2, v.size), np.uint64)
    for p in range(3):
        out[p, 0], out[p, 1] = xs[p], xs[(p + 1) % 3]
    return out


# ---------------------------------------------------------------------------
# gpu:: level KATs (GPU.cu)

def test_gemm_kats_and_closed_form(rng):
    a = np.array([1, 2, 2, 1, 1, 2], np.uint64)
    b = np.array([2, 2, 2, 1, 1, 1, 2, 2, 2, 1, 1, 1], np.uint64)
    assert O.gemm(2, 4, 3, a, 0, b, 0, 0).tolist() == [8, 10, 4, 5, 8, 10, 4, 5]
    assert O.gemm(2, 2, 3, a, 0, a, 1, 0).tolist() == [6, 6, 6, 9]
    assert O.gemm(3, 3, 2, a, 1, a, 0, 0).tolist() == [5, 4, 5, 4, 5, 4, 5, 4, 5]
    assert O.gemm(3, 1, 2, a, 1, np.array([4, 2], np.uint64), 1, 0).tolist() == [8, 10, 8]
    a2 = np.array([1, 2, 2, 1, 0, 1], np.uint64)
    b2 = np.array([1, 3, 2, 1, 0, 1, 1, 0], np.uint64)
    assert O.gemm(3, 4, 2, a2, 1, b2, 1, 1).tolist() == [1, 5, 4, 1, 2, 7, 5, 2, 0, 1, 1, 0]
    # closed form: numpy's uint64 matmul wraps mod 2^64; layouts via transposes
    M, N, K = 7, 9, 13
    A, B = rand_u64(rng, M, K), rand_u64(rng, K, N)
    want = (A @ B).astype(np.uint64)
    for ta in (0, 1):
        for tb in (0, 1):
            for tc in (0, 1):
                Aa = A.ravel() if ta else A.T.ravel()
                Bb = B.ravel() if tb else B.T.ravel()
                got = O.gemm(M, N, K, Aa, ta, Bb, tb, tc)
                assert np.array_equal(got, want.ravel() if tc else want.T.ravel())


def _conv_np(x, w, b, H, W, Din, Dout, f, pad, stride):
    x = x.reshape(b, H, W, Din)
    w = w.reshape(Dout, f, f, Din)
    OH, OW = (H + 2 * pad - f) // stride + 1, (W + 2 * pad - f) // stride + 1
    xp = np.zeros((b, H + 2 * pad, W + 2 * pad, Din), np.uint64)
    xp[:, pad:pad + H, pad:pad + W] = x
    y = np.zeros((b, OH, OW, Dout), np.uint64)
    for oh in range(OH):
        for ow in range(OW):
            patch = xp[:, oh * stride:oh * stride + f, ow * stride:ow * stride + f, :].reshape(b, -1)
            y[:, oh, ow, :] = (patch @ w.reshape(Dout, -1).T).astype(np.uint64)
    return y.ravel()


def test_conv_kat_and_closed_form(rng):
    im = np.array([1, 2, 1, 2, 3, 2, 1, 2, 1], np.uint64)
    f = np.array([1, 0, 1, 0, 1, 0, 1, 0, 1], np.uint64)
    assert O.conv_fprop(im, f, 1, 3, 3, 1, 1, 3, 3, 1, 1, 1).tolist() == [4, 6, 4, 6, 7, 6, 4, 6, 4]
    for (b, H, W, Din, Dout, k, pad, s) in [(2, 5, 5, 4, 10, 3, 1, 2), (1, 7, 6, 3, 5, 3, 0, 1)]:
        x, w = rand_u64(rng, b * H * W * Din), rand_u64(rng, Dout * k * k * Din)
        assert np.array_equal(O.conv_fprop(x, w, b, H, W, Din, Dout, k, k, pad, pad, s),
                              _conv_np(x, w, b, H, W, Din, Dout, k, pad, s))


def test_bitexpand_kat():
    a = np.array([2, 3, 1] + [0] * 29 + [0x2001], np.uint64)
    bits = O.bitexpand(a).reshape(-1, 64)
    assert bits[0, :2].tolist() == [0, 1] and bits[1, :2].tolist() == [1, 1] and bits[2, 0] == 1
    assert np.flatnonzero(bits[32]).tolist() == [0, 13]
    assert bits.sum() == 1 + 2 + 1 + 2


def test_expand_compare_kat():
    out = O.expand_compare(np.array([0, 1, 1, 0]), np.array([1, 0, 0, 1]), 8)
    assert out.tolist() == [0, 1, 1, 0, 1, 0, 0, 1]  # GPU.cu:443-455


def test_element_vector_add_kat():
    a = np.array([1, 3, 2, 2, 3, 1], np.uint64)
    a = O.element_vector_add(a, np.array([1, 1, 2], np.uint64), True, 2, 3)
    assert a.tolist() == [2, 4, 3, 3, 5, 3]
    a = O.element_vector_add(a, np.array([2, 3], np.uint64), False, 2, 3)
    assert a.tolist() == [4, 7, 5, 6, 7, 6]  # GPU.cu:262-283


def test_maxpool_im2row_and_stride_pad_kats():
    im = np.array([1, 1, 2, 2, 1, 1, 2, 2, 2, 2, 1, 1, 2, 2, 0, 0, 1, 1, 2, 2, 3, 4, 1, 1] + [0] * 8, np.uint64)
    pad = np.int64(np.iinfo(np.int64).min // 3).view(np.uint64)
    out = O.maxpool_im2row(im, 4, 4, 3, 2, 1, 1, 0, pad).reshape(-1, 16)
    p = int(pad)
    exp = [
        [1, 2, 1, 2, 1, 2, 1, 2, 3], [1, 2, 1, 2, 1, 2, 1, 2, 4],
        [2, 1, 2, 1, 2, 0, 2, 3, 1], [2, 1, 2, 1, 2, 0, 2, 4, 1],
        [2, 1, 2, 1, 2, 3, 0, 0, 0], [2, 1, 2, 1, 2, 4, 0, 0, 0],
        [1, 2, 0, 2, 3, 1, 0, 0, 0], [1, 2, 0, 2, 4, 1, 0, 0, 0],
    ]  # GPU.cu:457-512
    for r, e in zip(out, exp):
        assert r[:9].tolist() == e and all(int(v) == p for v in r[9:])
    sp = O.stride_pad(np.array([1, 1, 1, 2, 2, 2, 3, 3, 3], np.uint64), 3, 2, 9)
    assert sp.tolist() == [1, 1, 1, 9, 9, 2, 2, 2, 9, 9, 3, 3, 3, 9, 9]  # GPU.cu:536-556
    assert O.expanded_pool_size(2) == 4 and O.expanded_pool_size(3) == 16


def test_truncate_cols():
    a = np.arange(32, dtype=np.uint8)
    assert O.truncate_cols(a, 16, 9).reshape(2, 9).tolist() == [list(range(9)), list(range(16, 25))]


# ---------------------------------------------------------------------------
# Protocol KATs (Func.cu) and closed forms

def test_reconstruct_kat():
    a = O.init_list([1, 1, 2, 8, 12.2, 1], FP)
    assert np.allclose(rec(a), [1, 1, 2, 8, 12.2, 1], atol=EPS)


def test_matmul_kats():
    a = O.init_list([1] * 6, FP)
    b = O.init_list([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], FP)
    c = O.matmul(a, b, 2, 4, 3, 0, 0, 0, FP)
    assert np.allclose(rec(c), [1, 1, 1, 1, 0, 0, 1, 1], atol=EPS)  # Func.cu:33-54
    a = O.init_list(np.eye(4).ravel(), FP)
    vals = [-2.786461, -1.280988, -2.209210, 0.049379, 0.241369, 1.617007, -0.572261, 0.705014, -1.176370,
            -0.814461, 0.992866, 0.856274]
    c = O.matmul(a, O.init_list(vals, FP), 4, 3, 4, 1, 1, 1, FP)
    assert np.allclose(rec(c), vals, atol=EPS)  # Func.cu:56-76


def test_mult_kat():
    a = O.init_list([12, 24, 3, 5, -2, -3], FP, convert=False)
    b = O.init_list([1, 0, 11, 3, -1, 11], FP, convert=False)
    assert rec_int(O.mult(b, a)).tolist() == [12, 0, 33, 15, 2, -33]  # Func.cu:78-97


def test_truncate_kats():
    a = O.divide_public(O.init_list([8, 16, 24, -24], FP), 8)
    assert np.allclose(rec(a), [1, 2, 3, -3], atol=EPS)
    a = O.divide_public(O.init_list([1 << 22], FP), 1 << 21)
    assert np.allclose(rec(a), [2], atol=EPS)
    a = O.init_list([8, 4, 3, -96, 2515014.0], FP)
    d = np.tile((np.uint64(1) << np.array([3, 1, 0, 5, 2], np.uint64)).astype(np.uint64), (3, 1))
    assert np.allclose(rec(O.divide_public_vec(a, d)), [1, 2, 3, -3, 628753.5], atol=EPS)  # Func.cu:99-154


def _ieee_trunc_div(t, d):
    # (S)((double)(S)t / (double)(S)d) with Python floats (IEEE double, RNE)
    st = t - (1 << 64) if t >= 1 << 63 else t
    q = float(st) / float(d)
    return int(q) & ((1 << 64) - 1)


def test_scalar_divide_matches_ieee(rng):
    vals = [0, 1, 2, (1 << 53) - 1, 1 << 53, (1 << 53) + 1, (1 << 54) + 3, (1 << 63) - 1, 1 << 63, (1 << 64) - 1]
    vals += [int(v) for v in rand_u64(rng, 3000)]
    vals += [int(v) >> int(s) for v, s in zip(rand_u64(rng, 3000), rng.integers(0, 64, 3000))]
    for d in (2, 8, 1 << 13, 1 << 20, 1 << 26, 1 << 40, 3, 10):
        for v in vals:
            assert O.scalar_divide(v, d) == _ieee_trunc_div(v, d), (v, d)


def test_divide_3out3_closed_form(rng):
    n = 4000
    a = rand_u64(rng, 3, n)
    for k in (1, 13, 26):
        out = O.divide_public_3out3(a, 1 << k)
        t = (a.sum(axis=0, dtype=np.uint64) - np.uint64(1 << k)).astype(np.uint64)
        q = np.array([_ieee_trunc_div(int(v), 1 << k) for v in t], np.uint64)
        base = (q + np.uint64(1)).astype(np.uint64)
        assert np.array_equal(out[0, 0], base) and np.array_equal(out[2, 1], base)
        assert not out[0, 1].any() and not out[1].any() and not out[2, 0].any()


def test_drelu_kats_and_closed_form(rng):
    d = O.drelu(O.init_list([-1] * 33 + [1], FP), np.uint64)
    assert rec_int(d).tolist() == [0] * 33 + [1]  # Func.cu:287-318
    v = [0.326922, 0.156987, 0.461417, -0.221444, 39.846086, 0, 0, 0] * 3
    assert rec_int(O.drelu(O.init_list(v, FP), np.uint64)).tolist() == [1, 1, 1, 0, 1, 1, 1, 1] * 3
    x = O.init_list(v[:8], FP)
    neg = np.uint64(np.int64(-10 * (1 << FP)).view(np.uint64))
    add = np.array([0, 0, 0, 0, 0, neg, neg, neg], np.uint64)
    for p in range(3):
        for s in range(2):
            x[p, s] += add
    assert rec_int(O.drelu(x, np.uint64)).tolist() == [1, 1, 1, 0, 1, 0, 0, 0]  # Func.cu:351-378 (DRELU3)
    # closed form [v >= 0] on a real (random) replicated sharing
    vals = rand_u64(rng, 2000)
    vals[:4] = np.array([0, 1, 1 << 63, (1 << 63) - 1], np.uint64)
    b = O.drelu(consistent_shares(rng, vals), np.uint8)
    assert np.array_equal(O.reconstruct(b)[0], (vals.view(np.int64) >= 0).astype(np.uint8))


def test_relu_kat_and_closed_form(rng):
    y, d = O.relu(O.init_list([-2, -3, 4, 3, 3.5, 1, -1.5, -1], FP), np.uint64)
    assert np.allclose(rec(y), [0, 0, 4, 3, 3.5, 1, 0, 0], atol=EPS)
    assert rec_int(d).tolist() == [0, 0, 1, 1, 1, 1, 0, 0]  # Func.cu:380-413
    vals = rng.integers(-(1 << 40), 1 << 40, 3000).astype(np.int64)
    y, d = O.relu(consistent_shares(rng, vals.view(np.uint64)))
    assert np.array_equal(O.reconstruct(y)[0].view(np.int64), np.maximum(vals, 0))


def test_select_share_kat_and_closed_form(rng):
    z = O.select_share(O.init_list([1, 2, 10, 1], FP), O.init_list([4, 5, 1, 6], FP),
                       O.init_list([1, 1, 0, 1], FP, convert=False))
    assert np.allclose(rec(z), [4, 5, 10, 6], atol=EPS)  # Func.cu:156-175
    x, y = rand_u64(rng, 500), rand_u64(rng, 500)
    bit = rng.integers(0, 2, 500).astype(np.uint64)
    bs = np.zeros((3, 2, 500), np.uint8)
    bs[0, 0] = bs[2, 1] = bit  # bits as the dReLU output carries them
    z = O.select_share(consistent_shares(rng, x), consistent_shares(rng, y), bs)
    assert np.array_equal(O.reconstruct(z)[0], np.where(bit == 1, y, x))


def test_maxpool_kats():
    r, d, _ = O.maxpool(O.init_list([1, 3, 4, 3, 7, 1, 2, 10], FP), 4, np.uint64)
    assert np.allclose(rec(r), [4, 10], atol=EPS)
    assert rec_int(d).tolist() == [0, 0, 1, 0, 0, 0, 0, 1]  # Func.cu:415-445
    vals = [-0.032290, -0.142006, -0.031253, 0.130512, -0.301328, -0.105484, 0.002150, 0.055205, 0.234268]
    x = O.init_list(vals, FP)
    pools = np.stack([np.stack([O.maxpool_im2row(x[p, s], 3, 3, 3, 1, 1, 1, 0, np.uint64(-10 & (2**64 - 1)))
                                for s in range(2)]) for p in range(3)])
    r, d, _ = O.maxpool(pools, 16, np.uint64)
    dres = np.stack([np.stack([O.truncate_cols(d[p, s].astype(np.uint8), 16, 9) for s in range(2)])
                     for p in range(3)])
    assert np.allclose(rec(r), [0.234268], atol=EPS)
    assert O.reconstruct(dres)[0].tolist() == [0] * 8 + [1]  # Func.cu:447-497


def test_maxpool_closed_form(rng):
    k, w = 16, 300
    vals = rng.integers(-(1 << 45), 1 << 45, (w, k)).astype(np.int64)
    r, d, _ = O.maxpool(consistent_shares(rng, vals.ravel().view(np.uint64)), k, np.uint8)
    assert np.array_equal(O.reconstruct(r)[0].view(np.int64), vals.max(axis=1))
    onehot = O.reconstruct(d)[0].reshape(w, k)
    assert np.array_equal(onehot.sum(axis=1), np.ones(w))
    assert np.array_equal(vals[np.arange(w), onehot.argmax(axis=1)], vals.max(axis=1))


# ---------------------------------------------------------------------------
# Layer fixtures (Layer.cu) from files/test

@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def test_fc_layer_fixture(gold):
    x = O.set_public(gold["fc_input"], FP)
    W, b = O.set_public(gold["fc_weight"], FP), O.set_public(gold["fc_bias"], FP)
    act = O.fc_forward(x, W, b, 128, 16, 10, FP)
    assert np.allclose(rec(act), gold["fc_output"], atol=EPS)  # Layer.cu:52-84


def test_relu_layer_fixture(gold):
    act, _ = O.relu_forward(O.set_public(gold["relu_input"], FP))
    assert np.allclose(rec(act), gold["relu_output"], atol=EPS)


def test_maxpool_layer_fixture(gold):
    act, _ = O.maxpool_forward(O.set_public(gold["maxpool_input"], FP), 128, 9, 9, 2, 3, 2, FP)
    assert np.allclose(rec(act), gold["maxpool_output"], atol=EPS)  # Layer.cu:298-339


def test_conv_layer_fixture(gold):
    act = O.cnn_forward(O.set_public(gold["conv_input"], FP), O.set_public(gold["conv_weight"], FP),
                        128, 9, 9, 2, 3, 3, 2, 1, FP)
    assert np.allclose(rec(act), gold["conv_output"], atol=EPS)  # Layer.cu:670-713


def test_secureml_fixture_chain(gold):
    """FC784-128-ReLU-128-ReLU-10-ReLU at batch 128 against outputlayer0..5."""
    params = {0: (O.set_public(gold["secureml_weight0"], FP), O.set_public(gold["secureml_bias0"], FP)),
              2: (O.set_public(gold["secureml_weight2"], FP), O.set_public(gold["secureml_bias2"], FP)),
              4: (O.set_public(gold["secureml_weight4"], FP), O.set_public(gold["secureml_bias4"], FP))}
    outs = O.net_forward(C.SECUREML, 128, FP, params, O.set_public(gold["secureml_input"], FP))
    for i, (act, _) in enumerate(outs):
        assert np.allclose(rec(act), gold[f"secureml_out{i}"], atol=EPS), f"layer {i}"


# ---------------------------------------------------------------------------
# Plaintext fixed-point twin and share invariants

def _twin_trunc(v, fp):
    t = (v - np.uint64(1 << fp)).astype(np.uint64).view(np.int64)
    q = np.trunc(t.astype(np.float64) / float(1 << fp)).astype(np.int64)  # int64 -> double is RNE
    return (q + 1).view(np.uint64)


def twin_forward(layers, batch, fp, params, x):
    """Exact plaintext restatement: x @ W mod 2^64, the reference truncation
    1 + trunc((v - d) / d), ReLU by sign, max-pool keeping the earlier element on ties."""
    h = (np.trunc(x * float(1 << fp)).astype(np.int64)).view(np.uint64)
    outs = []
    for i, l in enumerate(layers):
        if l[0] == "fc":
            W, b = params[i]
            v = (h.reshape(batch, l[1]) @ W.reshape(l[1], l[2])).astype(np.uint64)
            h = (_twin_trunc(v, fp) + b[None, :]).astype(np.uint64).ravel()
        elif l[0] == "cnn":
            W, _ = params[i]
            H, Wd, Cin, Cout, f, s, p = l[1:]
            h = _twin_trunc(_conv_np(h, W, batch, H, Wd, Cin, Cout, f, p, s), fp)
        elif l[0] == "relu":
            h = np.where(h.view(np.int64) >= 0, h, np.uint64(0)).astype(np.uint64)
        else:
            H, Wd, Cc, pool, s = l[1:]
            OH, OW = (H - pool) // s + 1, (Wd - pool) // s + 1
            img = h.reshape(batch, H, Wd, Cc)
            win = np.stack([img[:, dh:dh + s * OH:s, dw:dw + s * OW:s, :] for dh in range(pool) for dw in range(pool)], -1)
            k = 4 if pool == 2 else 16
            pad = np.uint64(np.int64(-30 * (1 << fp)).view(np.uint64))  # -10 * 2^fp in all three components
            win = np.concatenate([win, np.full(win.shape[:-1] + (k - pool * pool,), pad, np.uint64)], -1)
            while win.shape[-1] > 1:
                e, o = win[..., 0::2], win[..., 1::2]
                keep = (e - o).astype(np.uint64).view(np.int64) >= 0
                win = np.where(keep, e, o)
            h = win[..., 0].ravel()
        outs.append(h.copy())
    return outs


@pytest.mark.parametrize("name,layers,batch,fp", [("secureml", C.SECUREML, 16, 13), ("lenet", C.LENET, 2, 13),
                                                  ("alexnet", C.ALEXNET_CIFAR, 2, 20)])
def test_network_matches_plaintext_twin(rng, name, layers, batch, fp):
    params, plain = {}, {}
    for i, l in enumerate(layers):
        if l[0] == "fc":
            w, b = rng.normal(0, 0.1, l[1] * l[2]), np.full(l[2], 0.01)
            params[i] = (O.set_public(w, fp), O.set_public(b, fp))
            plain[i] = (params[i][0][0, 0], params[i][1][0, 0])
        elif l[0] == "cnn":
            w = rng.normal(0, 0.1, l[3] * l[4] * l[5] * l[5])
            params[i] = (O.set_public(w, fp), None)
            plain[i] = (params[i][0][0, 0], None)
    x = rng.uniform(0, 1, batch * C.input_size(layers))
    outs = O.net_forward(layers, batch, fp, params, O.set_public(x, fp))
    twin = twin_forward(layers, batch, fp, plain, x)
    for i, ((act, _), t) in enumerate(zip(outs, twin)):
        assert np.array_equal(O.reconstruct(act)[0], t), f"{name} layer {i} ({layers[i][0]})"
        if name == "alexnet":
            continue  # 3x3 pools pad with -10*2^fp in every share, so x1, x2 become nonzero
        # placeholder invariants: A.sA == C.sB, x1 == 0; B.sB carries the FC bias, else 0
        assert np.array_equal(act[0, 0], act[2, 1]) and not act[0, 1].any() and not act[1, 0].any()
        assert not act[2, 0].any()
        if layers[i][0] != "fc":
            assert not act[1, 1].any()


# ---------------------------------------------------------------------------
# TPC (2 parties) and FPC (4 parties): matmul, convolution, truncation
# (TPC.inl / FPC.inl).  The reference runs its Func.cu suite typed over
# RSS / TPC / FPC / OPC, so the same KATs apply to every scheme.

SCHEMES = {
    "tpc": (O.tpc_init_list, O.tpc_reconstruct, O.tpc_matmul, O.tpc_convolution, O.tpc_divide_public),
    "fpc": (O.fpc_init_list, O.fpc_reconstruct, O.fpc_matmul, O.fpc_convolution, O.fpc_divide_public),
}


def _every_party(recon, fp=FP):
    return recon.view(np.int64) / float(1 << fp)


@pytest.mark.parametrize("kind", sorted(SCHEMES))
def test_scheme_kats(kind):
    init, recon, mm, conv, div = SCHEMES[kind]
    for r in _every_party(recon(init([1, 1, 2, 8, 12.2, 1], FP))):  # Func.cu:14-31
        assert np.allclose(r, [1, 1, 2, 8, 12.2, 1], atol=EPS)
    c = mm(init([1] * 6, FP), init([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], FP), 2, 4, 3, 0, 0, 0, FP)
    for r in _every_party(recon(c)):  # Func.cu:33-54
        assert np.allclose(r, [1, 1, 1, 1, 0, 0, 1, 1], atol=EPS)
    vals = [-2.786461, -1.280988, -2.209210, 0.049379, 0.241369, 1.617007, -0.572261, 0.705014, -1.176370,
            -0.814461, 0.992866, 0.856274]
    c = mm(init(np.eye(4).ravel(), FP), init(vals, FP), 4, 3, 4, 1, 1, 1, FP)
    for r in _every_party(recon(c)):  # Func.cu:56-76
        assert np.allclose(r, vals, atol=EPS)
    im = [1, 2, 1, 2, 3, 2, 1, 2, 1] * 2
    c = conv(init(im, FP), init([1, 0, 1, 0, 1, 0, 1, 0, 1], FP), 2, 3, 3, 3, 1, 1, 1, 1, FP)
    for r in _every_party(recon(c)):  # Func.cu:177-218
        assert np.allclose(r, [4, 6, 4, 6, 7, 6, 4, 6, 4] * 2, atol=EPS)
    for r in _every_party(recon(div(init([8, 16, 24, -24], FP), 8))):  # Func.cu:99-114
        assert np.allclose(r, [1, 2, 3, -3], atol=EPS)
    for r in _every_party(recon(div(init([1 << 22], FP), 1 << 21))):  # Func.cu:116-131
        assert np.allclose(r, [2], atol=EPS)


def test_fpc_reshare_kat():
    """unitTests.cu:44-88: z, z' known to every party; reshareFPC(i, j) puts z
    (or z' when i, j are opposite parties) into the sharing."""
    z = (np.array([1, 2, 5, -3], np.int64) << FP).view(np.uint64)
    zp = (np.array([7, -3, -1, 4], np.int64) << FP).view(np.uint64)
    Z, ZP = np.tile(z, (4, 1)), np.tile(zp, (4, 1))
    for i, j in [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]:
        out = O.fpc_reconstruct(O.fpc_reshare(Z, ZP, i, j))
        want = zp if (j - i) == 2 else z
        for q in range(4):
            assert np.array_equal(out[q], want), (i, j, q)


def _wrapped_product(X, W, M, N, K, ta, tb):
    """X.W mod 2^64 with the gemm layout flags (transpose = row-major storage)."""
    A = X.reshape(M, K) if ta else X.reshape(K, M).T
    B = W.reshape(K, N) if tb else W.reshape(N, K).T
    return (A.astype(object) @ B.astype(object)) % (1 << 64)


def _trunc_closed_form(raw, k):
    return np.array([(1 + _ieee_trunc_div((int(v) - (1 << k)) % (1 << 64), 1 << k)) % (1 << 64) for v in raw],
                    np.uint64)


@pytest.mark.parametrize("ta,tb,tc", [(1, 1, 1), (0, 0, 0), (1, 0, 1), (0, 1, 0)])
def test_tpc_matmul_closed_form(rng, ta, tb, tc):
    """TPC's Beaver product is exact for any sharing: party A ends with
    1 + trunc((XW - 2^k) / 2^k), party B with 0 (ring identity of
    TPC.inl:1059-1090 with x = y = 1, z = K, then dividePublic)."""
    M, N, K, k = 5, 7, 9, 13
    a, b = rand_u64(rng, 2, M * K), rand_u64(rng, 2, K * N)
    X, W = a[0] + a[1], b[0] + b[1]
    c = O.tpc_matmul(a, b, M, N, K, ta, tb, tc, k)
    P = _wrapped_product(X, W, M, N, K, ta, tb)
    raw = (P if tc else P.T).ravel()
    assert np.array_equal(c[0], _trunc_closed_form(raw, k))
    assert not c[1].any()


def test_tpc_convolution_closed_form(rng):
    b_, H, W, f, Din, Dout, s, p, k = 2, 6, 5, 3, 2, 3, 2, 1, 20
    A, B = rand_u64(rng, 2, b_ * H * W * Din), rand_u64(rng, 2, Dout * f * f * Din)
    c = O.tpc_convolution(A, B, b_, H, W, f, Din, Dout, s, p, k)
    raw = O.conv_fprop(A[0] + A[1], B[0] + B[1], b_, H, W, Din, Dout, f, f, p, p, s)
    assert np.array_equal(c[0], _trunc_closed_form(raw, k))
    assert not c[1].any()


def test_fpc_matmul_public_inputs_closed_form(rng):
    """Values entered as in the reference's tests (FPC initializer list: only
    component x0 is nonzero): every cross term vanishes and each of parties
    B, C, D keeps 1 + trunc((XW - 2^k) / 2^k) in its x0 slot; A holds zeros."""
    M, N, K, k = 4, 6, 8, 26
    X, W = rng.integers(-(1 << 30), 1 << 30, M * K), rng.integers(-(1 << 30), 1 << 30, K * N)
    a, b = O.fpc_init_list(X, 0, convert=False), O.fpc_init_list(W, 0, convert=False)
    c = O.fpc_matmul(a, b, M, N, K, 1, 1, 1, k)
    want = _trunc_closed_form(_wrapped_product(X.astype(np.uint64), W.astype(np.uint64), M, N, K, 1, 1).ravel(), k)
    assert not c[0].any()
    for q, slot in ((1, 2), (2, 1), (3, 0)):
        assert np.array_equal(c[q, slot], want)
        assert not np.delete(c[q], slot, axis=0).any()
