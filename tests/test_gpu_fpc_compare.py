"""FPC element-wise protocols on the GPU (FPC.inl: *= and &= with
reshareFPC, dReLU, ReLU, selectShare, maxpool) against the oracle, bit for bit
on all four parties' twelve share arrays, for initializer-list inputs and
arbitrary share values, plus the typed Func.cu known answers through the C ABI."""
import numpy as np
import pytest

import oracle as O
import piranha_amd as P

pytestmark = pytest.mark.gpu

FP = 20
DT = [np.uint8, np.uint64]
KINDS = ["fixed", "general"]


def eb(dt):
    return np.dtype(dt).itemsize


def shares(rng, n, kind):
    if kind == "fixed":
        v = rng.uniform(-8, 8, n)
        v[:4] = [0.0, -0.0, 1.0, -1.0]
        return O.fpc_init_list(v, FP)
    return (rng.integers(0, 1 << 63, (4, 3, n), dtype=np.uint64) * np.uint64(2)
            + rng.integers(0, 2, (4, 3, n), dtype=np.uint64))


def bits(rng, n, dt):
    return rng.choice(np.array([0, 1, 1, 0, 2, 255], np.uint64), size=(4, 3, n)).astype(dt)


@pytest.mark.parametrize("kind", KINDS)
def test_mult(rng, kind):
    n = 555
    x, y = shares(rng, n, kind), shares(rng, n, kind)
    dx, dy = P.to_dev(x), P.to_dev(y)
    P.check(P.lib().pir_fpc_mult(dx.ptr, dy.ptr, n, None))
    assert np.array_equal(dx.numpy(), O.fpc_mult(x, y))


@pytest.mark.parametrize("dt", DT)
def test_and(rng, dt):
    n = 700
    x, y = bits(rng, n, dt), bits(rng, n, dt)
    dx, dy = P.to_dev(x), P.to_dev(y)
    P.check(P.lib().pir_fpc_and(dx.ptr, dy.ptr, n, eb(dt), None))
    assert np.array_equal(dx.numpy(), O.fpc_and(x, y))


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("dt", DT)
def test_drelu_relu(rng, kind, dt):
    n = 800
    x = shares(rng, n, kind)
    dx, dd = P.to_dev(x), P.Buffer((4, 3, n), dt)
    P.check(P.lib().pir_fpc_drelu(dx.ptr, n, dd.ptr, eb(dt), None))
    assert np.array_equal(dd.numpy(), O.fpc_drelu(x, dt))
    dy, dd2 = P.Buffer((4, 3, n), np.uint64), P.Buffer((4, 3, n), dt)
    P.check(P.lib().pir_fpc_relu(dx.ptr, n, dy.ptr, dd2.ptr, eb(dt), None))
    y, d = O.fpc_relu(x, dt)
    assert np.array_equal(dy.numpy(), y)
    assert np.array_equal(dd2.numpy(), d)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("dt", DT)
def test_select_share(rng, kind, dt):
    n = 500
    x, y, b = shares(rng, n, kind), shares(rng, n, kind), bits(rng, n, dt)
    dx, dy, db, dz = P.to_dev(x), P.to_dev(y), P.to_dev(b), P.Buffer((4, 3, n), np.uint64)
    P.check(P.lib().pir_fpc_select_share(dx.ptr, dy.ptr, db.ptr, dz.ptr, n, eb(dt), None))
    assert np.array_equal(dz.numpy(), O.fpc_select_share(x, y, b))


@pytest.mark.parametrize("k", [2, 4, 8, 16])
@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("dt", DT)
def test_maxpool(rng, k, kind, dt):
    n = k * 40
    pools = shares(rng, n, kind)
    if kind == "fixed":  # ties
        pools[:, :, : n // 4] = pools[:, :, :1]
    dp, dr, dd = P.to_dev(pools), P.Buffer((4, 3, n // k), np.uint64), P.Buffer((4, 3, n), dt)
    P.check(P.lib().pir_fpc_maxpool(dp.ptr, n, dr.ptr, dd.ptr, k, eb(dt), None))
    r, d, after = O.fpc_maxpool(pools, k, dt)
    assert np.array_equal(dr.numpy(), r)
    assert np.array_equal(dd.numpy(), d)
    assert np.array_equal(dp.numpy(), after)


def test_func_kats():
    fp = 26

    def fixed(a):
        return O.fpc_reconstruct(a)[1].view(np.int64) / float(1 << fp)

    def opened(d):
        return O.fpc_reconstruct(d)[1].tolist()

    x = O.fpc_init_list([-2, -3, 4, 3, 3.5, 1, -1.5, -1], fp)  # Func.cu:380-413
    dx, dy, dd = P.to_dev(x), P.Buffer((4, 3, 8), np.uint64), P.Buffer((4, 3, 8), np.uint64)
    P.check(P.lib().pir_fpc_relu(dx.ptr, 8, dy.ptr, dd.ptr, 8, None))
    assert fixed(dy.numpy()).tolist() == [0, 0, 4, 3, 3.5, 1, 0, 0]
    assert opened(dd.numpy()) == [0, 0, 1, 1, 1, 1, 0, 0]
    p = O.fpc_init_list([1, 3, 4, 3, 7, 1, 2, 10], fp)  # Func.cu:415-445
    dp, dr, dd = P.to_dev(p), P.Buffer((4, 3, 2), np.uint64), P.Buffer((4, 3, 8), np.uint64)
    P.check(P.lib().pir_fpc_maxpool(dp.ptr, 8, dr.ptr, dd.ptr, 4, 8, None))
    assert fixed(dr.numpy()).tolist() == [4, 10]
    assert opened(dd.numpy()) == [0, 0, 1, 0, 0, 0, 0, 1]
    a = O.fpc_init_list([12, 24, 3, 5, -2, -3], fp, convert=False)  # Func.cu:78-98
    b = O.fpc_init_list([1, 0, 11, 3, -1, 11], fp, convert=False)
    da, db = P.to_dev(a), P.to_dev(b)
    P.check(P.lib().pir_fpc_mult(db.ptr, da.ptr, 6, None))
    assert O.fpc_reconstruct(db.numpy())[1].view(np.int64).tolist() == [12, 0, 33, 15, 2, -33]
    xs, ys = O.fpc_init_list([1, 2, 10, 1], fp), O.fpc_init_list([4, 5, 1, 6], fp)  # Func.cu:156-175
    bs = O.fpc_init_list([1, 1, 0, 1], fp, convert=False)
    dx, dy, dbs, dz = P.to_dev(xs), P.to_dev(ys), P.to_dev(bs), P.Buffer((4, 3, 4), np.uint64)
    P.check(P.lib().pir_fpc_select_share(dx.ptr, dy.ptr, dbs.ptr, dz.ptr, 4, 8, None))
    assert fixed(dz.numpy()).tolist() == [4, 5, 10, 6]


def test_argument_errors(rng):
    x = shares(rng, 32, "fixed")
    dx, dd = P.to_dev(x), P.Buffer((4, 3, 32), np.uint64)
    with pytest.raises(P.PirError, match="element size"):
        P.check(P.lib().pir_fpc_drelu(dx.ptr, 32, dd.ptr, 4, None))
    with pytest.raises(P.PirError, match="k must be"):
        P.check(P.lib().pir_fpc_maxpool(dx.ptr, 32, dd.ptr, dd.ptr, 32, 8, None))
    with pytest.raises(P.PirError, match="multiple of k"):
        P.check(P.lib().pir_fpc_maxpool(dx.ptr, 30, dd.ptr, dd.ptr, 4, 8, None))
    with pytest.raises(P.PirError, match="overlap"):
        P.check(P.lib().pir_fpc_relu(dx.ptr, 32, dx.ptr, dd.ptr, 8, None))
    P.check(P.lib().pir_fpc_drelu(dx.ptr, 0, dd.ptr, 1, None))  # empty


def test_relu_full_size_closed_form(rng):
    # 4 M elements in the public component: [v >= 0] in the public slots and
    # max(v, 0) opened alike by every party
    n = 1 << 22
    v = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    v[:3] = [0, -1, 1]
    x = np.zeros((4, 3, n), np.uint64)
    for q, s in ((1, 2), (2, 1), (3, 0)):
        x[q, s] = v.view(np.uint64)
    dx, dy, dd = P.to_dev(x), P.Buffer((4, 3, n), np.uint64), P.Buffer((4, 3, n), np.uint8)
    P.check(P.lib().pir_fpc_relu(dx.ptr, n, dy.ptr, dd.ptr, 1, None))
    d, y = dd.numpy(), dy.numpy()
    bit = (v >= 0).astype(np.uint8)
    for q in range(4):
        for s in range(3):
            assert np.array_equal(d[q, s], bit if (q, s) in ((1, 2), (2, 1), (3, 0)) else np.zeros(n, np.uint8))
    opened = O.fpc_reconstruct(y)
    assert all(np.array_equal(opened[q].view(np.int64), np.maximum(v, 0)) for q in range(4))
