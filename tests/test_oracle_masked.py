"""The oracle's masked precompute mode (no reference counterpart: the
reference deals placeholder constants only, src/mpc/Precompute.h:73-134, and
never instantiates its AES PRG, src/main.cu:96-104), pinned independently of
its own restatement:

* Philox4x32-10 against the published known-answer vectors of its authors'
  Random123 distribution (Salmon et al., SC'11), and the word stream's layout;
* every masked protocol against closed forms on the values: reshare, product
  and AND reconstruct exactly; dReLU to [v >= 0]; selectShare to b ? y : x;
  ReLU to max(v, 0); maxpool to the window maximum and the placeholder form's
  derivative; truncation to floor or ceil of the exact quotient t / 2^k;
* a whole masked forward layer by layer against the placeholder layers on
  the values the masked layer received (exact for ReLU / maxpool, one unit in
  the last place for the truncated products);
* the shares themselves are not the placeholder's (x1 and x2 are not zero).
"""
import numpy as np
import pytest

import oracle as O
from piranha_amd import configs as C

SEED = 0x0123456789ABCDEF

# Random123 kat_vectors, philox4x32 with 10 rounds: counter, key, output
KATS = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


def reveal(x):
    """Party A's opening: its two shares plus party C's first share."""
    return x[0, 0] + x[0, 1] + x[2, 0]


def reveal_views(x):
    return [x[p, 0] + x[p, 1] + x[(p + 2) % 3, 0] for p in range(3)]


def reveal_xor(x):
    return x[0, 0] ^ x[0, 1] ^ x[2, 0]


def public(v):
    """The placeholder sharing of ring values v (x0 = v: A.sA and C.sB)."""
    v = np.asarray(v, np.uint64)
    s = np.zeros((3, 2, v.size), np.uint64)
    s[0, 0] = v
    s[2, 1] = v
    return s


def signed(v):
    return np.asarray(v, np.uint64).view(np.int64)


def test_philox_kats():
    ctr = np.array([k[0] for k in KATS], np.uint32)
    key = np.array([k[1] for k in KATS], np.uint32)
    assert O.philox(ctr, key).tolist() == [k[2] for k in KATS]


def test_word_stream_layout():
    # word j of element e of a call: half (j & 1) of Philox(key = seed, ctr = (e_lo, e_hi, call, j >> 1))
    seed, call = 0xFFFFFFFFFFFFFFFF, 0xFFFFFFFF
    w = O.prg_fill(seed, call, 6)
    for b in range(3):
        o = O.philox(np.array([[b, 0, call, 0]], np.uint32), np.array([[0xFFFFFFFF, 0xFFFFFFFF]], np.uint32))[0]
        assert int(w[2 * b]) == int(o[0]) | (int(o[1]) << 32)
        assert int(w[2 * b + 1]) == int(o[2]) | (int(o[3]) << 32)
    z = O.prg_fill(0, 0, 2)  # the first KAT block
    assert int(z[0]) == 0xE169C58D6627E8D5 and int(z[1]) == 0x9B00DBD8BC57AC4C


def test_deal_is_a_random_sharing(rng):
    v = rng.integers(0, 2**64, 1000, dtype=np.uint64)
    s = O.deal(v, SEED, 3)
    for view in reveal_views(s):
        assert np.array_equal(view, v)
    assert np.array_equal(s[0, 1], s[1, 0]) and np.array_equal(s[1, 1], s[2, 0]) and np.array_equal(s[2, 1], s[0, 0])
    assert np.count_nonzero(s[1, 1]) > 990  # x2 is not the placeholder's zero
    assert not np.array_equal(s, O.deal(v, SEED, 4))  # another call, other masks


def test_reshare_mult_and(rng):
    n = 500
    c = rng.integers(0, 2**64, (3, n), dtype=np.uint64)
    r = O.reshare_m(c, SEED, 1)
    for view in reveal_views(r):
        assert np.array_equal(view, c[0] + c[1] + c[2])
    assert not np.array_equal(r[:, 0], c)  # masked
    x, y = rng.integers(0, 2**64, (2, n), dtype=np.uint64)
    xs, ys = O.deal(x, SEED, 1), O.deal(y, SEED, 2)
    assert np.array_equal(reveal(O.mult_m(xs, ys, SEED, 5)), x * y)
    for dt in (np.uint8, np.uint64):
        a = rng.integers(0, np.iinfo(dt).max, (2, n), dtype=dt, endpoint=True)
        sh = []
        for k, v in enumerate(a):
            m1, m2 = rng.integers(0, np.iinfo(dt).max, (2, n), dtype=dt, endpoint=True)
            x0 = v ^ m1 ^ m2
            sh.append(np.stack([np.stack([x0, m1]), np.stack([m1, m2]), np.stack([m2, x0])]))
        assert np.array_equal(reveal_xor(O.rss_and_m(sh[0], sh[1], SEED, 9)), a[0] & a[1])


def quotient_bounds(t, k):
    t = signed(t).astype(object)
    lo = [v // (1 << k) for v in t]  # floor
    hi = [-((-v) // (1 << k)) for v in t]  # ceil
    return lo, hi


@pytest.mark.parametrize("k", [13, 26, 40])
def test_truncation_floor_or_ceil(rng, k):
    n = 4000
    t = (rng.integers(-(2**60), 2**60, n)).astype(np.int64)
    t[:6] = [0, 1, -1, (1 << k) - 1, -(1 << k), 1 << k]
    a = rng.integers(0, 2**64, (3, n), dtype=np.uint64)
    a[2] = t.view(np.uint64) - a[0] - a[1]
    out = O.divide_public_3out3_m(a, k, SEED, 7)
    got = [int(v) for v in signed(reveal(out))]
    lo, hi = quotient_bounds(t.view(np.uint64), k)
    assert all(g in (l, h) for g, l, h in zip(got, lo, hi))
    assert sum(g != l for g, l in zip(got, lo)) > n // 10  # both roundings occur (random r)
    # against the placeholder form's 1 + trunc((t - d) / d): at most one unit apart below 2^53
    ph = signed(reveal(O.divide_public_3out3(a.copy(), 1 << k)))
    small = np.abs(t) < 2**53
    assert np.all(np.abs(np.array(got, np.int64)[small] - ph[small]) <= 1)
    # dividePublic of an RSS value: the same bound
    s = O.deal(t.view(np.uint64), SEED, 11)
    got2 = [int(v) for v in signed(reveal(O.divide_public_m(s, k, SEED, 12)))]
    assert all(g in (l, h) for g, l, h in zip(got2, lo, hi))


def test_matmul_and_conv_products(rng):
    M, N, K, fp = 7, 9, 33, 13
    A = (rng.integers(-3000, 3000, M * K)).astype(np.int64).view(np.uint64)
    B = (rng.integers(-3000, 3000, K * N)).astype(np.int64).view(np.uint64)
    a, b = O.deal(A, SEED, 1), O.deal(B, SEED, 2)
    c = O.matmul_m(a, b, M, N, K, True, True, True, fp, SEED, 3)
    exact = O.gemm(M, N, K, A, 1, B, 1, 1)
    lo, hi = quotient_bounds(exact, fp)
    got = [int(v) for v in signed(reveal(c))]
    assert all(g in (l, h) for g, l, h in zip(got, lo, hi))
    # convolution: 2 images 6x5x3, filter 3x3, 4 outputs, pad 1
    b_, H, W, Din, Dout, f = 2, 6, 5, 3, 4, 3
    X = (rng.integers(-3000, 3000, b_ * H * W * Din)).astype(np.int64).view(np.uint64)
    Wt = (rng.integers(-3000, 3000, Dout * f * f * Din)).astype(np.int64).view(np.uint64)
    y = O.convolution_m(O.deal(X, SEED, 4), O.deal(Wt, SEED, 5), b_, H, W, f, Din, Dout, 1, 1, fp, SEED, 6)
    exact = O.conv_fprop(X, Wt, b_, H, W, Din, Dout, f, f, 1, 1, 1)
    lo, hi = quotient_bounds(exact, fp)
    assert all(g in (l, h) for g, l, h in zip([int(v) for v in signed(reveal(y))], lo, hi))


EDGES = np.array([0, 1, 2**64 - 1, 2**63 - 1, 2**63, 2**62, 2**64 - 2**62, 5, 2**64 - 5], np.uint64)


@pytest.mark.parametrize("dt", [np.uint8, np.uint64])
def test_drelu_relu_select(rng, dt):
    n = 300
    v = np.concatenate([EDGES, rng.integers(0, 2**64, n - EDGES.size, dtype=np.uint64)])
    x = O.deal(v, SEED, 1)
    d = O.drelu_m(x, SEED, 2, dt)
    want = (signed(v) >= 0).astype(dt)
    assert np.array_equal(reveal_xor(d), want)
    assert np.count_nonzero(d[1]) > 0  # not confined to the x0 slot
    y, dr = O.relu_m(x, SEED, 3, dt)
    assert np.array_equal(reveal_xor(dr), want)
    assert np.array_equal(reveal(y), np.where(signed(v) >= 0, v, np.uint64(0)))
    # selectShare with a random XOR sharing of bits b
    xv, yv = rng.integers(0, 2**64, (2, n), dtype=np.uint64)
    bits = rng.integers(0, 2, n).astype(dt)
    m1, m2 = rng.integers(0, 2, (2, n)).astype(dt)
    b0 = bits ^ m1 ^ m2
    bs = np.stack([np.stack([b0, m1]), np.stack([m1, m2]), np.stack([m2, b0])])
    z = O.select_share_m(O.deal(xv, SEED, 4), O.deal(yv, SEED, 5), bs, SEED, 6)
    assert np.array_equal(reveal(z), np.where(bits == 1, yv, xv))


@pytest.mark.parametrize("k", [2, 4, 16])
def test_maxpool(rng, k):
    windows = 60
    v = (rng.integers(-50, 50, windows * k)).astype(np.int64).view(np.uint64)  # ties occur
    res, d, _ = O.maxpool_m(O.deal(v, SEED, 1), k, SEED, 2)
    w = signed(v).reshape(windows, k)
    assert np.array_equal(signed(reveal(res)), w.max(axis=1))
    _, d_ph, _ = O.maxpool(public(v), k)
    assert np.array_equal(reveal_xor(d), reveal(d_ph))  # the placeholder's derivative bits
    assert np.count_nonzero(d[1]) > 0


def test_masked_forward_layer_by_layer():
    """LeNet b2 FP13 with N(0, 0.1) weights: each masked layer against the
    placeholder layer on the values it received."""
    layers, batch, fp = C.LENET, 2, 13
    w = C.normal_weights(layers, 0.1, seed=3)
    params = {i: (O.set_public(a, fp), None if b is None else O.set_public(b, fp)) for i, (a, b) in w.items()}
    x = O.set_public(np.random.default_rng(5).uniform(0, 1, batch * C.input_size(layers)), fp)
    xs = O.deal(reveal(x), SEED, 100)
    # the masked mode needs consistent sharings of the weights too: setPublic's
    # party-B quirk (RSS.inl:134-140) only cancels while x1 = x2 = 0
    dealt = {i: (O.deal(reveal(a), SEED, 300 + 2 * i), None if b is None else O.deal(reveal(b), SEED, 301 + 2 * i))
             for i, (a, b) in params.items()}
    outs = O.net_forward_m(layers, batch, fp, dealt, xs, SEED, 200)
    h = xs
    for i, (l, (act, der)) in enumerate(zip(layers, outs)):
        ph = O.net_forward([l], batch, fp, {0: params[i]} if i in params else {}, public(reveal(h)))[0]
        if l[0] in ("relu", "maxpool"):
            assert np.array_equal(reveal(act), reveal(ph[0])), i
            assert np.array_equal(reveal_xor(der), reveal(ph[1])), i
        else:
            diff = signed(reveal(act)) - signed(reveal(ph[0]))
            assert np.all(np.abs(diff) <= 1), (i, np.abs(diff).max())
        assert np.count_nonzero(act[1, 1]) > act.shape[-1] // 2  # random shares, not the placeholder's
        h = act
