"""The oracle's FPC comparison protocols (FPC.inl: *= :321-362, &= :374-398,
selectShare :671-712, dReLU :929-989, ReLU :991-1002, maxpool :1004-1181,
carryOut :1297-1416, reshareFPC :1183-1260), pinned by the reference's typed
known answers (Func.cu:78-98 Mult, :156-175 SelectShare, :287-497 DRELU,
DRELU2, DRELU3, RELU, Maxpool, Maxpool2, all instantiated for FPC<uint64_t>)
and by closed forms on random values held in the public component x0 (the
initializer-list sharing; with the placeholder's zero PRGs reshareFPC is not a
correct protocol for general sharings, whose parties then open different
values): the product mod 2^64, [v >= 0] held in the public slots, b ? y : x,
max(v, 0), the window maximum with a one-hot derivative on the first maximum,
and the AND of bits.  dReLU alone is exact on any consistent sharing, and is
checked on those too."""
import numpy as np
import pytest

import oracle as O

FP = 26
DT = [np.uint8, np.uint64]
PUBLIC_SLOT = {1: 2, 2: 1, 3: 0}  # party B: shareC, C: shareB, D: shareA (component x0)


def rec(x):
    r = O.fpc_reconstruct(x)
    assert all(np.array_equal(r[0], r[q]) for q in range(4))  # every party opens the same value
    return r[0]


def rec_bits(d):
    return (d[1, 2] + d[1, 0] + d[1, 1] + d[0, 0]).astype(np.uint64)  # party B's view, in the bit type


def fixed(v):
    return rec(v).view(np.int64) / float(1 << FP)


def share(rng, v, dtype=np.uint64):
    """A random consistent sharing of v: components x1..x3 random, x0 = v - x1 - x2 - x3."""
    v = np.asarray(v).astype(dtype)
    info = np.iinfo(dtype)
    comp = [None] + [rng.integers(0, info.max, v.size, dtype=dtype, endpoint=True) for _ in range(3)]
    comp[0] = (v - comp[1] - comp[2] - comp[3]).astype(dtype)
    return np.stack([np.stack([comp[(q + s + 1) % 4] for s in range(3)]) for q in range(4)])


def public(v, dtype=np.uint64):
    """v in the public component x0 (parties B, C, D's public slots)."""
    v = np.asarray(v).astype(dtype)
    x = np.zeros((4, 3, v.size), dtype)
    for q, s in PUBLIC_SLOT.items():
        x[q, s] = v
    return x


def test_share_helpers(rng):
    v = rng.integers(0, 1 << 63, 50, dtype=np.uint64)
    assert np.array_equal(rec(share(rng, v)), v)
    assert np.array_equal(rec(public(v)), v)


def test_mult_kat():
    a = O.fpc_init_list([12, 24, 3, 5, -2, -3], FP, convert=False)
    b = O.fpc_init_list([1, 0, 11, 3, -1, 11], FP, convert=False)
    assert rec(O.fpc_mult(b, a)).view(np.int64).tolist() == [12, 0, 33, 15, 2, -33]


@pytest.mark.parametrize("dt", DT)
def test_select_share_kat(dt):
    x = O.fpc_init_list([1, 2, 10, 1], FP)
    y = O.fpc_init_list([4, 5, 1, 6], FP)
    b = O.fpc_init_list([1, 1, 0, 1], FP, convert=False).astype(dt)
    assert fixed(O.fpc_select_share(x, y, b)).tolist() == [4, 5, 10, 6]


@pytest.mark.parametrize("dt", DT)
def test_drelu_kats(dt):
    x = O.fpc_init_list([-1] * 33 + [1], FP)
    assert rec_bits(O.fpc_drelu(x, dt)).tolist() == [0] * 33 + [1]
    v2 = [0.326922, 0.156987, 0.461417, -0.221444, 39.846086, 0.0, 0.0, 0.0] * 3
    assert rec_bits(O.fpc_drelu(O.fpc_init_list(v2, FP), dt)).tolist() == [1, 1, 1, 0, 1, 1, 1, 1] * 3
    # DRELU3: every party adds (T)(-10 * 2^FP) to each of its shares of the last three values
    x3 = O.fpc_init_list(v2[:8], FP)
    x3[:, :, 5:] += np.uint64((-10 * (1 << FP)) & ((1 << 64) - 1))
    assert rec_bits(O.fpc_drelu(x3, dt)).tolist() == [1, 1, 1, 0, 1, 0, 0, 0]


@pytest.mark.parametrize("dt", DT)
def test_relu_kat(dt):
    y, d = O.fpc_relu(O.fpc_init_list([-2, -3, 4, 3, 3.5, 1, -1.5, -1], FP), dt)
    assert fixed(y).tolist() == [0, 0, 4, 3, 3.5, 1, 0, 0]
    assert rec_bits(d).tolist() == [0, 0, 1, 1, 1, 1, 0, 0]


@pytest.mark.parametrize("dt", DT)
def test_maxpool_kats(dt):
    r, d, _ = O.fpc_maxpool(O.fpc_init_list([1, 3, 4, 3, 7, 1, 2, 10], FP), 4, dt)
    assert fixed(r).tolist() == [4, 10]
    assert rec_bits(d).tolist() == [0, 0, 1, 0, 0, 0, 0, 1]
    v = [-0.032290, -0.142006, -0.031253, 0.130512, -0.301328, -0.105484, 0.002150, 0.055205, 0.234268]
    x = O.fpc_init_list(v, FP)
    pv = np.uint64((-10) & ((1 << 64) - 1))
    pools = np.stack([np.stack([O.maxpool_im2row(x[q, s], 3, 3, 3, 1, 1, 1, 0, pv) for s in range(3)])
                      for q in range(4)])
    r, d, _ = O.fpc_maxpool(pools, 16, dt)
    # the -10 padding sits in every share, so each party's opened maximum differs in the last units
    views = O.fpc_reconstruct(r).view(np.int64)[:, 0] / float(1 << FP)
    assert np.allclose(views, 0.234268, atol=1e-6)
    assert rec_bits(d)[:9].tolist() == [0, 0, 0, 0, 0, 0, 0, 0, 1]


def special_values(rng, n):
    v = rng.integers(-(1 << 62), 1 << 62, n).astype(np.int64)
    v[:6] = [0, 1, -1, np.iinfo(np.int64).max, np.iinfo(np.int64).min, -2]
    return v.view(np.uint64)


def test_mult_closed_form(rng):
    x, y = special_values(rng, 300), special_values(rng, 300)
    assert np.array_equal(rec(O.fpc_mult(public(x), public(y))), x * y)


@pytest.mark.parametrize("dt", DT)
def test_drelu_closed_form(rng, dt):
    v = special_values(rng, 300)
    d = O.fpc_drelu(share(rng, v), dt)
    want = (v.view(np.int64) >= 0).astype(int)
    for q in range(4):
        for s in range(3):
            assert d[q, s].tolist() == (want.tolist() if PUBLIC_SLOT.get(q) == s else [0] * v.size)


@pytest.mark.parametrize("dt", DT)
def test_select_relu_closed_form(rng, dt):
    x, y = special_values(rng, 300), special_values(rng, 300)
    bits = rng.integers(0, 2, 300).astype(dt)
    z = O.fpc_select_share(public(x), public(y), public(bits, dt))
    assert np.array_equal(rec(z), np.where(bits == 1, y, x))
    yv, _ = O.fpc_relu(public(x), dt)
    assert np.array_equal(rec(yv).view(np.int64), np.maximum(x.view(np.int64), 0))


@pytest.mark.parametrize("dt", DT)
def test_and_public_component(rng, dt):
    a, b = rng.integers(0, 2, 200).astype(dt), rng.integers(0, 2, 200).astype(dt)
    x, y = np.zeros((4, 3, 200), dt), np.zeros((4, 3, 200), dt)
    for q, s in PUBLIC_SLOT.items():
        x[q, s], y[q, s] = a, b
    z = O.fpc_and(x, y)
    for q in range(4):
        for s in range(3):
            assert np.array_equal(z[q, s], (a & b) if PUBLIC_SLOT.get(q) == s else np.zeros(200, dt))


@pytest.mark.parametrize("k", [2, 4, 8, 16])
@pytest.mark.parametrize("dt", DT)
def test_maxpool_closed_form(rng, k, dt):
    w = 40
    v = rng.integers(-50, 50, (w, k)).astype(np.int64)
    r, d, _ = O.fpc_maxpool(public(v.ravel().view(np.uint64)), k, dt)
    assert np.array_equal(rec(r).view(np.int64), v.max(axis=1))
    onehot = np.zeros((w, k), int)
    onehot[np.arange(w), v.argmax(axis=1)] = 1
    assert np.array_equal(rec_bits(d).reshape(w, k).astype(int), onehot)
