"""The oracle's ResLayer (ResLayer.cu:23-307, a composition of its pinned CNN,
LN and ReLU layers) against a float64 twin of the ResNet basic block as the
architecture defines it (He et al., "Deep Residual Learning for Image
Recognition", 2016, with LayerNorm in place of batch norm, as ResLayer.cu
builds it):

    y = ReLU(LN2(conv3x3(ReLU(LN1(conv3x3_s(x))))) + S(x)),
    S(x) = LN_s(conv1x1_s(x)) when the stride or channel count changes, else x,

block after block.  The reference has no Res test; this pins the block wiring
(which layers, in which order, where the shortcut joins, which stride each
convolution takes) independently of the oracle's own restatement.  The inner
layers are pinned by their own known answers (tests/test_oracle_ln.py and the
CNN / ReLU tests); the twin's LN is the layer's definition in float: each
image's feature vector less its mean, over sqrt(population variance + 2^-8)
(the epsilon 1 << (FP - 8), LNLayer.cu:110-111), then gamma and beta expanded
over the flattened batch by vectorExpand with a factor of the batch count, so
flat element e takes entry e // batch (LNLayer.cu:141-142).

Mistakes planted in copies of the oracle's block (residual joined after the
final ReLU, shortcut without its LN, a missing inner ReLU) must all be caught.
"""
import numpy as np
import pytest

import oracle as O

TOL = 2e-2  # the reference's LN tests allow 1e-2 per layer; two LNs per block


def conv_f(x, w, b, H, W, cin, cout, f, stride, pad):
    x = x.reshape(b, H, W, cin)
    w = w.reshape(cout, f, f, cin)
    OH, OW = (H + 2 * pad - f) // stride + 1, (W + 2 * pad - f) // stride + 1
    xp = np.zeros((b, H + 2 * pad, W + 2 * pad, cin))
    xp[:, pad:pad + H, pad:pad + W] = x
    y = np.zeros((b, OH, OW, cout))
    for oh in range(OH):
        for ow in range(OW):
            patch = xp[:, oh * stride:oh * stride + f, ow * stride:ow * stride + f, :].reshape(b, -1)
            y[:, oh, ow, :] = patch @ w.reshape(cout, -1).T
    return y.reshape(b, -1), OH, OW


def ln_f(h, gamma, beta):
    b = h.shape[0]
    mu = h.mean(axis=1, keepdims=True)
    var = ((h - mu) ** 2).mean(axis=1, keepdims=True)
    xhat = ((h - mu) / np.sqrt(var + 2.0 ** -8)).ravel()
    return (np.repeat(gamma, b) * xhat + np.repeat(beta, b)).reshape(h.shape)


def relu_f(h):
    return np.maximum(h, 0.0)


def twin_res(x, batch, spec, weights):
    """Float64 ResNet basic blocks; `weights` per block: (W1, g1, b1, W2, g2,
    b2) and, for a shortcut block, (Ws, gs, bs)."""
    H, W, cin, planes, nblocks, stride, expansion = spec
    h = x.reshape(batch, -1)
    for blk in range(nblocks):
        s = stride if blk == 0 else 1
        wts = weights[blk]
        c1, OH, OW = conv_f(h, wts[0], batch, H, W, cin, planes, 3, s, 1)
        a = relu_f(ln_f(c1, wts[1], wts[2]))
        c2, _, _ = conv_f(a, wts[3], batch, OH, OW, planes, planes, 3, 1, 1)
        main = ln_f(c2, wts[4], wts[5])
        if s != 1 or cin != expansion * planes:
            cs, _, _ = conv_f(h, wts[6], batch, H, W, cin, expansion * planes, 1, s, 0)
            short = ln_f(cs, wts[7], wts[8])
        else:
            short = h
        h = relu_f(main + short)
        H, W, cin = OH, OW, expansion * planes
    return h.ravel()


def make_weights(rng, spec):
    """Per block float weights and the oracle's snapshot-order share list."""
    structure = O.res_structure(*spec)
    blocks, flat = [], []
    for main, sc in structure:
        vals = []
        for l in main + sc:
            if l[0] == "cnn":
                cin, cout, f = l[3], l[4], l[5]
                vals.append(rng.normal(0, np.sqrt(2.0 / (f * f * cin)), cout * f * f * cin))
            elif l[0] == "ln":
                vals.append(rng.uniform(0.5, 1.5, l[1]))
                vals.append(rng.uniform(-0.2, 0.2, l[1]))
        blocks.append(vals)
    return blocks


def encode(blocks, fp):
    return [O.set_public(v, fp) for vals in blocks for v in vals]


def decode(a, fp):
    return O.reconstruct(a)[0].view(np.int64) / float(1 << fp)


SPECS = [((8, 8, 16, 16, 2, 1, 1), 2),   # two blocks without shortcuts
         ((8, 8, 8, 16, 2, 2, 1), 2),    # stride-2 shortcut block, then a plain one
         ((6, 5, 4, 12, 1, 1, 1), 3)]    # channel change at stride 1 (shortcut, no stride)


@pytest.mark.parametrize("fp", [20, 26])
@pytest.mark.parametrize("spec,batch", SPECS)
def test_res_forward_matches_float_twin(rng, spec, batch, fp):
    H, W, cin = spec[:3]
    blocks = make_weights(rng, spec)
    x = rng.uniform(-1, 1, batch * H * W * cin)
    # the twin sees exactly the fixed-point values the oracle encodes
    xq = decode(O.set_public(x, fp), fp)
    bq = [[decode(O.set_public(v, fp), fp) for v in vals] for vals in blocks]
    got, _ = O.res_forward(O.set_public(x, fp), batch, fp, spec, encode(blocks, fp))
    want = twin_res(xq, batch, spec, bq)
    assert np.allclose(decode(got, fp), want, atol=TOL), np.abs(decode(got, fp) - want).max()


def _variant_forward(x, batch, fp, spec, flat, mistake):
    """The oracle's res_forward with one wiring mistake planted."""
    structure = O.res_structure(*spec)
    params = O._split_params(structure, flat)
    h = O._u(x)
    for (main, sc), (pm, ps) in zip(structure, params):
        if mistake == "no_inner_relu":
            main, pm = main[:2] + main[3:], pm[:2] + pm[3:]
        if mistake == "no_shortcut_ln" and sc:
            residual = O._sub_forward(sc[0], h, ps[0], batch, fp)[0]  # the 1x1 conv alone
        else:
            residual = O._block_forward(sc, ps, h, None, batch, fp)[0][-1] if sc else h
        if mistake == "join_after_relu":
            acts, _, _ = O._block_forward(main, pm, h, None, batch, fp)
            h = acts[-1] + residual
        else:
            acts, _, _ = O._block_forward(main, pm, h, residual, batch, fp)
            h = acts[-1]
    return h


@pytest.mark.parametrize("mistake,spec", [("join_after_relu", (8, 8, 16, 16, 1, 1, 1)),
                                          ("no_shortcut_ln", (8, 8, 8, 16, 1, 2, 1)),
                                          ("no_inner_relu", (8, 8, 16, 16, 1, 1, 1))])
def test_twin_catches_planted_wiring_mistakes(rng, mistake, spec):
    fp, batch = 20, 2
    H, W, cin = spec[:3]
    blocks = make_weights(rng, spec)
    x = rng.uniform(-1, 1, batch * H * W * cin)
    xq = decode(O.set_public(x, fp), fp)
    bq = [[decode(O.set_public(v, fp), fp) for v in vals] for vals in blocks]
    want = twin_res(xq, batch, spec, bq)
    flat = encode(blocks, fp)
    right, _ = O.res_forward(O.set_public(x, fp), batch, fp, spec, flat)
    assert np.allclose(decode(right, fp), want, atol=TOL)
    wrong = _variant_forward(O.set_public(x, fp), batch, fp, spec, flat, mistake)
    assert not np.allclose(decode(wrong, fp), want, atol=TOL), mistake
