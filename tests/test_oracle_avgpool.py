"""The oracle's average-pool path (AveragepoolLayer.cu, convolution.cuh
averagepool_im2row / averagepool_expand_delta) pinned by the reference's own
known-answer tests (GPU.cu AvgpoolDeltaExpand, Layer.cu AveragepoolForwardBasic
/ AveragepoolBackwardBasic) and by closed forms computed here independently:
window sums taken from the NHWC tensor directly, and the placeholder
dividePublic's opened result 1 + trunc((S - d) / d)."""
import numpy as np
import pytest

import oracle as O


def test_expand_delta_kat():
    # GPU.cu:516-535
    out = O.averagepool_expand_delta(np.array([1, 0, 3, 4], np.uint64), 16, 2, 4)
    assert out.tolist() == [1, 0, 1, 0, 1, 0, 1, 0, 3, 4, 3, 4, 3, 4, 3, 4]


LAYER_INPUT = [1, -1, -3, 2, 3, -2, 0, 1, 2, -1, 0, 0, -2, -3, -1, 3]  # Layer.cu:429-432, NHWC 2x2x2, batch 2


@pytest.mark.parametrize("fp", [13, 26])
def test_layer_forward_kat(fp):
    # Layer.cu:419-448: 2x2 pool, stride 1 over 2x2 images
    x = O.set_public(np.array(LAYER_INPUT, float), fp)
    act = O.avgpool_forward(x, 2, 2, 2, 2, 2, 1)
    got = O.reconstruct(act)[0].view(np.int64) / float(1 << fp)
    assert got.tolist() == [0.25, 0.0, -0.25, -0.25]


@pytest.mark.parametrize("fp", [13, 26])
def test_layer_backward_kat(fp):
    # Layer.cu:450-491
    g = O.set_public(np.array([1, 0, -1, -2], float), fp)
    d = O.avgpool_backward(g, 2, 2, 2, 2, 2)
    got = O.reconstruct(d)[0].view(np.int64) / float(1 << fp)
    assert got.tolist() == [0.25, 0.0] * 4 + [-0.25, -0.5] * 4


def opened_divide(S, d):
    """Opened value of the reference's dividePublic with the placeholder
    randomness r = 1, r' = d: 1 + trunc((S - d) / d) (exact integers here)."""
    t = S - d
    q = np.abs(t) // d * np.sign(t)
    return 1 + q


@pytest.mark.parametrize("batch,H,W,C,pool,s", [(2, 6, 4, 3, 2, 2), (1, 7, 5, 2, 3, 1), (3, 8, 8, 4, 2, 2),
                                                (2, 5, 9, 1, 3, 2), (1, 4, 4, 2, 1, 1)])
def test_forward_closed_form(rng, batch, H, W, C, pool, s):
    fp = 13
    vals = rng.integers(-(1 << 20), 1 << 20, size=batch * H * W * C)
    x = O.set_public(vals / float(1 << fp), fp)  # exact fixed-point inputs
    act = O.avgpool_forward(x, batch, H, W, C, pool, s)
    OH, OW = (H - pool) // s + 1, (W - pool) // s + 1
    img = vals.reshape(batch, H, W, C)
    S = np.zeros((batch, OH, OW, C), np.int64)
    for dh in range(pool):
        for dw in range(pool):
            S += img[:, dh:dh + s * (OH - 1) + 1:s, dw:dw + s * (OW - 1) + 1:s, :]
    want = opened_divide(S.ravel(), pool * pool)
    assert np.array_equal(O.reconstruct(act)[0].view(np.int64), want)


def test_im2row_window_order(rng):
    # non-square image: each pool holds its f x f window, row offset outer
    H, W, C, f, s = 5, 3, 2, 2, 1
    im = np.arange(H * W * C, dtype=np.uint64)
    rows = O.averagepool_im2row(im, W, H, f, C, 1, s)
    OH, OW = (H - f) // s + 1, (W - f) // s + 1
    img = im.reshape(H, W, C)
    want = []
    for oh in range(OH):
        for ow in range(OW):
            for c in range(C):
                want += [int(img[oh * s + dh, ow * s + dw, c]) for dh in range(f) for dw in range(f)]
    assert rows.tolist() == want


@pytest.mark.parametrize("batch,H,W,C,pool,s", [(1, 5, 4, 2, 2, 2), (2, 4, 4, 3, 2, 2), (1, 6, 6, 1, 3, 3)])
def test_backward_closed_form(rng, batch, H, W, C, pool, s):
    # deltas entry b*C*P + c + j*C (b over image-window pairs) opens to the
    # divided delta[b*C + c]; entries past n_in * P (5x4 with 2x2 pools covers
    # 16 of 20 per channel group) open to the division of zero, 0
    fp, P = 13, pool * pool
    OH, OW = (H - pool) // s + 1, (W - pool) // s + 1
    n_in, n_out = batch * OH * OW * C, batch * H * W * C
    g = rng.integers(-(1 << 22), 1 << 22, size=n_in)
    d = O.avgpool_backward(O.set_public(g / float(1 << fp), fp), batch, H, W, C, pool)
    got = O.reconstruct(d)[0].view(np.int64)
    want = np.zeros(n_out, np.int64)
    for i in range(n_in * P):
        want[i] = opened_divide(g[(i // (C * P)) * C + i % C], P)
    assert np.array_equal(got, want)
