"""TPC element-wise protocols on the GPU (TPC.inl: Beaver *= and &=, dReLU,
ReLU, selectShare, maxpool) against the oracle, bit for bit on both parties'
shares, for realistic fixed-point inputs and arbitrary share values, plus the
typed Func.cu known answers run through the C ABI."""
import numpy as np
import pytest

import oracle as O
import piranha_amd as P

pytestmark = pytest.mark.gpu

FP = 20
DT = [np.uint8, np.uint64]


def eb(dt):
    return np.dtype(dt).itemsize


def shares(rng, n, kind):
    if kind == "fixed":
        v = rng.uniform(-8, 8, n)
        v[:4] = [0.0, -0.0, 1.0, -1.0]
        return O.tpc_init_list(v, FP)
    return rng.integers(0, 1 << 63, (2, n), dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, (2, n), dtype=np.uint64)


def bits(rng, n, dt):
    return rng.choice(np.array([0, 1, 1, 0, 2, 255], np.uint64), size=(2, n)).astype(dt)


KINDS = ["fixed", "general"]


@pytest.mark.parametrize("kind", KINDS)
def test_mult(rng, kind):
    n = 777
    x, y = shares(rng, n, kind), shares(rng, n, kind)
    dx, dy = P.to_dev(x), P.to_dev(y)
    P.check(P.lib().pir_tpc_mult(dx.ptr, dy.ptr, n, None))
    assert np.array_equal(dx.numpy(), O.tpc_mult(x, y))


@pytest.mark.parametrize("dt", DT)
def test_and(rng, dt):
    n = 901
    x, y = bits(rng, n, dt), bits(rng, n, dt)
    dx, dy = P.to_dev(x), P.to_dev(y)
    P.check(P.lib().pir_tpc_and(dx.ptr, dy.ptr, n, eb(dt), None))
    assert np.array_equal(dx.numpy(), O.tpc_and(x, y))


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("dt", DT)
def test_drelu_relu(rng, kind, dt):
    n = 1000
    x = shares(rng, n, kind)
    dx, dd = P.to_dev(x), P.Buffer((2, n), dt)
    P.check(P.lib().pir_tpc_drelu(dx.ptr, n, dd.ptr, eb(dt), None))
    assert np.array_equal(dd.numpy(), O.tpc_drelu(x, dt))
    dy, dd2 = P.Buffer((2, n), np.uint64), P.Buffer((2, n), dt)
    P.check(P.lib().pir_tpc_relu(dx.ptr, n, dy.ptr, dd2.ptr, eb(dt), None))
    y, d = O.tpc_relu(x, dt)
    assert np.array_equal(dy.numpy(), y)
    assert np.array_equal(dd2.numpy(), d)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("dt", DT)
def test_select_share(rng, kind, dt):
    n = 640
    x, y, b = shares(rng, n, kind), shares(rng, n, kind), bits(rng, n, dt)
    dx, dy, db, dz = P.to_dev(x), P.to_dev(y), P.to_dev(b), P.Buffer((2, n), np.uint64)
    P.check(P.lib().pir_tpc_select_share(dx.ptr, dy.ptr, db.ptr, dz.ptr, n, eb(dt), None))
    assert np.array_equal(dz.numpy(), O.tpc_select_share(x, y, b))


@pytest.mark.parametrize("k", [2, 4, 8, 16])
@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("dt", DT)
def test_maxpool(rng, k, kind, dt):
    n = k * 50
    pools = shares(rng, n, kind)
    if kind == "fixed":  # ties
        pools[:, : n // 4] = pools[:, :1]
    dp, dr, dd = P.to_dev(pools), P.Buffer((2, n // k), np.uint64), P.Buffer((2, n), dt)
    P.check(P.lib().pir_tpc_maxpool(dp.ptr, n, dr.ptr, dd.ptr, k, eb(dt), None))
    r, d, after = O.tpc_maxpool(pools, k, dt)
    assert np.array_equal(dr.numpy(), r)
    assert np.array_equal(dd.numpy(), d)
    assert np.array_equal(dp.numpy(), after)


def test_func_kats():
    fp = 26

    def fixed(a):
        return O.tpc_reconstruct(a)[0].view(np.int64) / float(1 << fp)

    # Func.cu:380-413 (RELU)
    x = O.tpc_init_list([-2, -3, 4, 3, 3.5, 1, -1.5, -1], fp)
    dx, dy, dd = P.to_dev(x), P.Buffer((2, 8), np.uint64), P.Buffer((2, 8), np.uint64)
    P.check(P.lib().pir_tpc_relu(dx.ptr, 8, dy.ptr, dd.ptr, 8, None))
    assert fixed(dy.numpy()).tolist() == [0, 0, 4, 3, 3.5, 1, 0, 0]
    assert O.tpc_reconstruct(dd.numpy())[0].tolist() == [0, 0, 1, 1, 1, 1, 0, 0]
    # Func.cu:415-445 (Maxpool)
    p = O.tpc_init_list([1, 3, 4, 3, 7, 1, 2, 10], fp)
    dp, dr, dd = P.to_dev(p), P.Buffer((2, 2), np.uint64), P.Buffer((2, 8), np.uint64)
    P.check(P.lib().pir_tpc_maxpool(dp.ptr, 8, dr.ptr, dd.ptr, 4, 8, None))
    assert fixed(dr.numpy()).tolist() == [4, 10]
    assert O.tpc_reconstruct(dd.numpy())[0].tolist() == [0, 0, 1, 0, 0, 0, 0, 1]
    # Func.cu:78-98 (Mult) and :156-175 (SelectShare)
    a = O.tpc_init_list([12, 24, 3, 5, -2, -3], fp, convert=False)
    b = O.tpc_init_list([1, 0, 11, 3, -1, 11], fp, convert=False)
    da, db = P.to_dev(a), P.to_dev(b)
    P.check(P.lib().pir_tpc_mult(db.ptr, da.ptr, 6, None))
    assert O.tpc_reconstruct(db.numpy())[0].view(np.int64).tolist() == [12, 0, 33, 15, 2, -33]
    xs, ys = O.tpc_init_list([1, 2, 10, 1], fp), O.tpc_init_list([4, 5, 1, 6], fp)
    bs = O.tpc_init_list([1, 1, 0, 1], fp, convert=False)
    dx, dy, dbs, dz = P.to_dev(xs), P.to_dev(ys), P.to_dev(bs), P.Buffer((2, 4), np.uint64)
    P.check(P.lib().pir_tpc_select_share(dx.ptr, dy.ptr, dbs.ptr, dz.ptr, 4, 8, None))
    assert fixed(dz.numpy()).tolist() == [4, 5, 10, 6]


def test_argument_errors(rng):
    x = shares(rng, 8, "fixed")
    dx, dd = P.to_dev(x), P.Buffer((2, 8), np.uint64)
    with pytest.raises(P.PirError, match="element size"):
        P.check(P.lib().pir_tpc_drelu(dx.ptr, 8, dd.ptr, 2, None))
    big, bd = P.to_dev(shares(rng, 32, "fixed")), P.Buffer((2, 32), np.uint64)
    with pytest.raises(P.PirError, match="k must be"):
        P.check(P.lib().pir_tpc_maxpool(big.ptr, 32, bd.ptr, bd.ptr, 32, 8, None))
    with pytest.raises(P.PirError, match="multiple of k"):
        P.check(P.lib().pir_tpc_maxpool(dx.ptr, 8, dd.ptr, dd.ptr, 3, 8, None))
    with pytest.raises(P.PirError, match="overlap"):
        P.check(P.lib().pir_tpc_relu(dx.ptr, 8, dx.ptr, dd.ptr, 8, None))
    P.check(P.lib().pir_tpc_drelu(dx.ptr, 0, dd.ptr, 1, None))  # empty


def test_relu_full_size_closed_form(rng):
    # 4 M elements (a VGG16 b64 conv1 activation): [v >= 0] and max(v, 0) on
    # every element, with party B's shares zero
    n = 1 << 22
    v = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    v[:3] = [0, -1, 1]
    x = np.zeros((2, n), np.uint64)
    r = rng.integers(0, 1 << 62, n).astype(np.uint64)
    x[0], x[1] = v.view(np.uint64) - r, r
    dx, dy, dd = P.to_dev(x), P.Buffer((2, n), np.uint64), P.Buffer((2, n), np.uint8)
    P.check(P.lib().pir_tpc_relu(dx.ptr, n, dy.ptr, dd.ptr, 1, None))
    d, y = dd.numpy(), dy.numpy()
    assert np.array_equal(d[0], (v >= 0).astype(np.uint8)) and not d[1].any()
    assert np.array_equal((y[0] + y[1]).view(np.int64), np.maximum(v, 0))
