"""The oracle's TPC comparison protocols (TPC.inl: Beaver *= and &= :242-312,
selectShare :585-613, dReLU :828-868, ReLU :870-884, maxpool :886-1030,
carryOut :1092-1200), pinned by the reference's typed known answers
(Func.cu:78-98 Mult, :156-175 SelectShare, :287-497 DRELU, DRELU2, DRELU3,
RELU, Maxpool, Maxpool2, all instantiated for TPC<uint64_t>) and by closed
forms on random additive sharings: the product mod 2^64, [v >= 0] on the signed
value, b ? y : x, max(v, 0), and the window maximum with a one-hot derivative
(the even element wins ties)."""
import numpy as np
import pytest

import oracle as O

FP = 26
DT = [np.uint8, np.uint64]


def rec(x):
    return O.tpc_reconstruct(x)[0]


def rec_bits(d):
    return (d[0] + d[1]).astype(np.uint64)  # the sum wraps in the bit type, as reconstruct does


def fixed(v):
    return rec(v).view(np.int64) / float(1 << FP)


def split(rng, v):
    """A random additive sharing (A: v - r, B: r) of the uint64 values v."""
    r = rng.integers(0, 1 << 63, v.size, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, v.size, dtype=np.uint64)
    return np.stack([v - r, r])


def test_mult_kat():
    a = O.tpc_init_list([12, 24, 3, 5, -2, -3], FP, convert=False)
    b = O.tpc_init_list([1, 0, 11, 3, -1, 11], FP, convert=False)
    assert rec(O.tpc_mult(b, a)).view(np.int64).tolist() == [12, 0, 33, 15, 2, -33]


@pytest.mark.parametrize("dt", DT)
def test_select_share_kat(dt):
    x = O.tpc_init_list([1, 2, 10, 1], FP)
    y = O.tpc_init_list([4, 5, 1, 6], FP)
    b = O.tpc_init_list([1, 1, 0, 1], FP, convert=False).astype(dt)
    assert fixed(O.tpc_select_share(x, y, b)).tolist() == [4, 5, 10, 6]


@pytest.mark.parametrize("dt", DT)
def test_drelu_kats(dt):
    x = O.tpc_init_list([-1] * 33 + [1], FP)
    assert rec_bits(O.tpc_drelu(x, dt)).tolist() == [0] * 33 + [1]
    v2 = [0.326922, 0.156987, 0.461417, -0.221444, 39.846086, 0.0, 0.0, 0.0] * 3
    assert rec_bits(O.tpc_drelu(O.tpc_init_list(v2, FP), dt)).tolist() == [1, 1, 1, 0, 1, 1, 1, 1] * 3
    # DRELU3: every party adds (T)(-10 * 2^FP) to its share of the last three values
    x3 = O.tpc_init_list(v2[:8], FP)
    neg = np.uint64((-10 * (1 << FP)) & ((1 << 64) - 1))
    x3[:, 5:] += neg
    assert rec_bits(O.tpc_drelu(x3, dt)).tolist() == [1, 1, 1, 0, 1, 0, 0, 0]


@pytest.mark.parametrize("dt", DT)
def test_relu_kat(dt):
    y, d = O.tpc_relu(O.tpc_init_list([-2, -3, 4, 3, 3.5, 1, -1.5, -1], FP), dt)
    assert fixed(y).tolist() == [0, 0, 4, 3, 3.5, 1, 0, 0]
    assert rec_bits(d).tolist() == [0, 0, 1, 1, 1, 1, 0, 0]


@pytest.mark.parametrize("dt", DT)
def test_maxpool_kats(dt):
    r, d, _ = O.tpc_maxpool(O.tpc_init_list([1, 3, 4, 3, 7, 1, 2, 10], FP), 4, dt)
    assert fixed(r).tolist() == [4, 10]
    assert rec_bits(d).tolist() == [0, 0, 1, 0, 0, 0, 0, 1]
    # Maxpool2: a 3x3 window padded to 16 with -10 in every share, then truncate_cols to 9
    v = [-0.032290, -0.142006, -0.031253, 0.130512, -0.301328, -0.105484, 0.002150, 0.055205, 0.234268]
    x = O.tpc_init_list(v, FP)
    pv = np.uint64((-10) & ((1 << 64) - 1))
    pools = np.stack([O.maxpool_im2row(x[q], 3, 3, 3, 1, 1, 1, 0, pv) for q in range(2)])
    assert pools.shape == (2, 16)
    r, d, _ = O.tpc_maxpool(pools, 16, dt)
    assert np.isclose(fixed(r)[0], 0.234268, atol=1e-6)
    assert rec_bits(d)[:9].tolist() == [0, 0, 0, 0, 0, 0, 0, 0, 1]


def special_values(rng, n):
    v = rng.integers(-(1 << 62), 1 << 62, n).astype(np.int64)
    v[:6] = [0, 1, -1, np.iinfo(np.int64).max, np.iinfo(np.int64).min, -2]
    return v.view(np.uint64)


def test_mult_closed_form(rng):
    x, y = special_values(rng, 300), special_values(rng, 300)
    z = O.tpc_mult(split(rng, x), split(rng, y))
    assert np.array_equal(rec(z), x * y)
    assert not z[1].any()  # party B ends with the triple's z share, 0


@pytest.mark.parametrize("dt", DT)
def test_drelu_closed_form(rng, dt):
    v = special_values(rng, 400)
    d = O.tpc_drelu(split(rng, v), dt)
    assert d[0].tolist() == (v.view(np.int64) >= 0).astype(int).tolist()
    assert not d[1].any()


@pytest.mark.parametrize("dt", DT)
def test_select_relu_closed_form(rng, dt):
    x, y = special_values(rng, 300), special_values(rng, 300)
    bits = rng.integers(0, 2, 300).astype(dt)
    r = rng.integers(0, 255, 300).astype(dt)
    b = np.stack([bits - r, r])  # additive in the bit type
    z = O.tpc_select_share(split(rng, x), split(rng, y), b)
    assert np.array_equal(rec(z), np.where(bits == 1, y, x))
    yv, d = O.tpc_relu(split(rng, x), dt)
    assert np.array_equal(rec(yv).view(np.int64), np.maximum(x.view(np.int64), 0))


@pytest.mark.parametrize("k", [2, 4, 8, 16])
@pytest.mark.parametrize("dt", DT)
def test_maxpool_closed_form(rng, k, dt):
    w = 60
    v = rng.integers(-50, 50, (w, k)).astype(np.int64)  # small range: ties occur
    r, d, _ = O.tpc_maxpool(split(rng, v.ravel().view(np.uint64)), k, dt)
    assert np.array_equal(rec(r).view(np.int64), v.max(axis=1))
    # the tournament keeps the even (earlier) element on ties: the first maximum
    onehot = np.zeros((w, k), int)
    onehot[np.arange(w), v.argmax(axis=1)] = 1
    assert np.array_equal(rec_bits(d).reshape(w, k).astype(int), onehot)
