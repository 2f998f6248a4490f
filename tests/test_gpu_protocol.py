"""RSS protocol ops on the GPU, fused and op-by-op forms, against the oracle
(src/mpc/RSS.inl semantics), bit for bit on every party's shares.

Inputs are both placeholder-style sharings (values in x0 only, as the
reference's precompute produces) and arbitrary share tensors (all six arrays
random), so the per-party restatement is exercised beyond the x1 = x2 = 0 case.
"""
import numpy as np
import pytest

import oracle as O
import piranha_amd as P
from conftest import rand_u64

pytestmark = pytest.mark.gpu

FP = 13
MODES = [P.MODE_FUSED, P.MODE_PLAIN]


@pytest.fixture(params=MODES, ids=["fused", "plain"])
def mode(request):
    P.set_mode(request.param)
    yield request.param
    P.set_mode(P.MODE_FUSED)


def fixed_shares(rng, n, scale=8.0, fp=FP):
    return O.init_list(rng.uniform(-scale, scale, n), fp)


def general_shares(rng, n, bits=40):
    v = rng.integers(-(1 << bits), 1 << bits, size=(3, 2, n)).astype(np.int64).view(np.uint64)
    return np.ascontiguousarray(v)


def bit_shares(rng, n, dtype=np.uint8):
    return rng.choice(np.array([0, 1, 1, 0, 2, 255], dtype=np.uint64), size=(3, 2, n)).astype(dtype)


@pytest.fixture(params=["placeholder", "general"])
def kind(request):
    return request.param


def make(rng, kind, n):
    return fixed_shares(rng, n) if kind == "placeholder" else general_shares(rng, n)


def test_reconstruct(rng, mode, kind):
    x = make(rng, kind, 1000)
    dx, out = P.to_dev(x), P.Buffer((3, 1000), np.uint64)
    P.check(P.lib().pir_rss_reconstruct(dx.ptr, 1000, out.ptr, 8, None))
    assert np.array_equal(out.numpy(), O.reconstruct(x))
    b = bit_shares(rng, 777)
    db, ob = P.to_dev(b), P.Buffer((3, 777), np.uint8)
    P.check(P.lib().pir_rss_reconstruct(db.ptr, 777, ob.ptr, 1, None))
    assert np.array_equal(ob.numpy(), O.reconstruct(b))


def test_reshare(rng, mode):
    c = rand_u64(rng, 3, 999)
    dc, out = P.to_dev(c), P.Buffer((3, 2, 999), np.uint64)
    P.check(P.lib().pir_rss_reshare(dc.ptr, 999, out.ptr, None))
    assert np.array_equal(out.numpy(), O.reshare(c))


def test_mult_and(rng, mode, kind):
    x, y = make(rng, kind, 513), make(rng, kind, 513)
    dx, dy = P.to_dev(x), P.to_dev(y)
    P.check(P.lib().pir_rss_mult(dx.ptr, dy.ptr, 513, None))
    assert np.array_equal(dx.numpy(), O.mult(x, y))
    bx, by = bit_shares(rng, 300), bit_shares(rng, 300)
    dbx, dby = P.to_dev(bx), P.to_dev(by)
    P.check(P.lib().pir_rss_and(dbx.ptr, dby.ptr, 300, 1, None))
    assert np.array_equal(dbx.numpy(), O.rss_and(bx, by))
    ux, uy = rand_u64(rng, 3, 2, 300), rand_u64(rng, 3, 2, 300)
    dux, duy = P.to_dev(ux), P.to_dev(uy)
    P.check(P.lib().pir_rss_and(dux.ptr, duy.ptr, 300, 8, None))
    assert np.array_equal(dux.numpy(), O.rss_and(ux, uy))


@pytest.mark.parametrize("k", [1, 3, 13, 21, 26])
def test_divide_public(rng, mode, kind, k):
    n = 2000
    x = make(rng, kind, n)
    if kind == "general":  # include magnitudes beyond 2^53 (double rounding path)
        x[0, 0, :500] = rand_u64(rng, 500)
    dx = P.to_dev(x)
    P.check(P.lib().pir_rss_divide_public(dx.ptr, n, 1 << k, None, None))
    assert np.array_equal(dx.numpy(), O.divide_public(x, 1 << k))


def test_divide_public_vec(rng, mode, kind):
    n = 1000
    x = make(rng, kind, n)
    d = (np.uint64(1) << rng.integers(0, 30, size=(3, n)).astype(np.uint64)).astype(np.uint64)
    dx, dd = P.to_dev(x), P.to_dev(d)
    P.check(P.lib().pir_rss_divide_public(dx.ptr, n, 0, dd.ptr, None))
    assert np.array_equal(dx.numpy(), O.divide_public_vec(x, d))


@pytest.mark.parametrize("d", [3, 9, 7, 1000, 1, 123456789, (1 << 40) + 3])
def test_divide_public_general_denominator(rng, mode, kind, d):
    """Non-power-of-two denominators take the IEEE double division of the
    reference's functor (AveragepoolLayer.cu:81 divides by pool^2, e.g. 9)."""
    n = 2000
    x = make(rng, kind, n)
    x[0, 0, :500] = rand_u64(rng, 500)  # magnitudes beyond 2^53: the int64 -> double rounding
    edge = np.array([0, 1, (1 << 53) - 1, 1 << 53, (1 << 53) + 1, (1 << 63) - 1, 1 << 63, (1 << 64) - 1], np.uint64)
    x[0, 0, 500:508] = edge
    dx = P.to_dev(x)
    P.check(P.lib().pir_rss_divide_public(dx.ptr, n, d, None, None))
    assert np.array_equal(dx.numpy(), O.divide_public(x, d))


def test_divide_public_vec_general(rng, mode, kind):
    n = 1000
    x = make(rng, kind, n)
    d = rng.integers(1, 1 << 20, size=(3, n)).astype(np.uint64)
    d[:, :5] = np.array([1, 2, 3, 9, 0], np.uint64)[:, None].T  # includes d = 0 (the clamped x / 0)
    dx, dd = P.to_dev(x), P.to_dev(d)
    P.check(P.lib().pir_rss_divide_public(dx.ptr, n, 0, dd.ptr, None))
    assert np.array_equal(dx.numpy(), O.divide_public_vec(x, d))


@pytest.mark.parametrize("d", [3, 9, 10, 1, 2, 1 << 30, (1 << 62) + 1])
def test_dd_op_divide(rng, d):
    n = 3000
    x = rand_u64(rng, n)
    x[:8] = np.array([0, 1, (1 << 53) - 1, 1 << 53, (1 << 53) + 1, (1 << 63) - 1, 1 << 63, (1 << 64) - 1], np.uint64)
    dx = P.to_dev(x)
    P.check(P.lib().pir_dd_op(P.OP_DIV, 8, dx.ptr, 1, None, 0, d, n, None))
    assert dx.numpy().tolist() == [O.scalar_divide(v, d) for v in x.tolist()]


@pytest.mark.parametrize("k", [1, 13, 20, 26, 40])
def test_divide_public_3out3(rng, mode, k):
    n = 4096
    a = rand_u64(rng, 3, n)
    a[:, :8] = 0
    # values around +-2^53 and the int64 extremes
    edge = np.array([0, 1, (1 << 53) - 1, 1 << 53, (1 << 53) + 1, (1 << 63) - 1, 1 << 63, (1 << 64) - 1], np.uint64)
    a[0, :8] = edge
    da, out = P.to_dev(a), P.Buffer((3, 2, n), np.uint64)
    P.check(P.lib().pir_rss_divide_public_3out3(da.ptr, n, k, out.ptr, None))
    assert np.array_equal(out.numpy(), O.divide_public_3out3(a, 1 << k))


@pytest.mark.parametrize("algo", [P.ALGO_PLAIN, P.ALGO_LIMB])
@pytest.mark.parametrize("ta,tb,tc", [(1, 1, 1), (0, 0, 0), (1, 0, 1), (0, 1, 0)])
def test_matmul(rng, mode, kind, algo, ta, tb, tc):
    M, N, K = 33, 45, 70
    a, b = make(rng, kind, M * K), make(rng, kind, K * N)
    da, db, dc = P.to_dev(a), P.to_dev(b), P.Buffer((3, 2, M * N), np.uint64)
    P.check(P.lib().pir_rss_matmul(da.ptr, db.ptr, dc.ptr, M, N, K, ta, tb, tc, FP, algo, None))
    assert np.array_equal(dc.numpy(), O.matmul(a, b, M, N, K, ta, tb, tc, FP))


@pytest.mark.parametrize("algo", [P.ALGO_PLAIN, P.ALGO_LIMB])
@pytest.mark.parametrize("b,H,W,f,Din,Dout,stride,pad", [(2, 3, 3, 3, 1, 1, 1, 1), (2, 9, 9, 3, 2, 3, 2, 1),
                                                         (3, 8, 8, 3, 16, 20, 1, 1),
                                                         # Cin = 3 with paired 16-byte stores: 1024 rows x 64
                                                         # columns; 90 rows (a partial tile) x 96 columns
                                                         (4, 16, 16, 3, 3, 64, 1, 1), (3, 11, 9, 3, 3, 96, 2, 1)])
def test_convolution(rng, mode, kind, algo, b, H, W, f, Din, Dout, stride, pad):
    a, w = make(rng, kind, b * H * W * Din), make(rng, kind, Dout * f * f * Din)
    OH = (H + 2 * pad - f) // stride + 1
    OW = (W + 2 * pad - f) // stride + 1
    n = b * OH * OW * Dout
    da, dw, dc = P.to_dev(a), P.to_dev(w), P.Buffer((3, 2, n), np.uint64)
    P.check(P.lib().pir_rss_convolution(da.ptr, dw.ptr, dc.ptr, b, H, W, f, Din, Dout, stride, pad, FP, algo, None))
    assert np.array_equal(dc.numpy(), O.convolution(a, w, b, H, W, f, Din, Dout, stride, pad, FP))


@pytest.mark.parametrize("n", [700, 701])  # the 4-element-per-thread form, and the per-element form
@pytest.mark.parametrize("dt", [np.uint8, np.uint64])
def test_drelu_relu(rng, mode, kind, dt, n):
    x = make(rng, kind, n)
    x[:, :, :4] = 0
    eb = np.dtype(dt).itemsize
    dx, dr = P.to_dev(x), P.Buffer((3, 2, n), dt)
    P.check(P.lib().pir_rss_drelu(dx.ptr, n, dr.ptr, eb, None))
    assert np.array_equal(dr.numpy(), O.drelu(x, dt))
    dy, dd = P.Buffer((3, 2, n), np.uint64), P.Buffer((3, 2, n), dt)
    P.check(P.lib().pir_rss_relu(dx.ptr, n, dy.ptr, dd.ptr, eb, None))
    y, d = O.relu(x, dt)
    assert np.array_equal(dy.numpy(), y)
    assert np.array_equal(dd.numpy(), d)


def test_select_share(rng, mode, kind):
    n = 600
    x, y, b = make(rng, kind, n), make(rng, kind, n), bit_shares(rng, n)
    dx, dy, db, dz = P.to_dev(x), P.to_dev(y), P.to_dev(b), P.Buffer((3, 2, n), np.uint64)
    P.check(P.lib().pir_rss_select_share(dx.ptr, dy.ptr, db.ptr, dz.ptr, n, 1, None))
    assert np.array_equal(dz.numpy(), O.select_share(x, y, b))


@pytest.mark.parametrize("k", [2, 4, 16])
@pytest.mark.parametrize("dt", [np.uint8, np.uint64])
def test_maxpool(rng, mode, kind, k, dt):
    n = k * 40
    pools = make(rng, kind, n)
    dp, dr, dd = P.to_dev(pools), P.Buffer((3, 2, n // k), np.uint64), P.Buffer((3, 2, n), dt)
    P.check(P.lib().pir_rss_maxpool(dp.ptr, n, dr.ptr, dd.ptr, k, np.dtype(dt).itemsize, None))
    r, d, p_after = O.maxpool(pools, k, dt)
    assert np.array_equal(dr.numpy(), r)
    assert np.array_equal(dd.numpy(), d)
    assert np.array_equal(dp.numpy(), p_after)


# gpu::transpose (matrix.cuh:43-59) and getMSBs / setMSBs (bitwise.cuh:23-39)

def test_transpose_kat():
    a = P.to_dev(np.array([1, 4, 2, 5, 3, 6], np.uint64))
    b = P.to_dev(np.zeros(6, np.uint64))
    P.check(P.lib().pir_transpose(a.ptr, b.ptr, 2, 3, 8, None))
    assert b.numpy().tolist() == [1, 2, 3, 4, 5, 6]  # GPU.cu:247-261


@pytest.mark.parametrize("rows,cols", [(1, 7), (5, 9), (300, 257)])
@pytest.mark.parametrize("eb", [8, 1])
def test_transpose(rng, rows, cols, eb):
    a = rand_u64(rng, rows * cols)
    if eb == 1:
        a = (a & np.uint64(0xFF)).astype(np.uint8)
    da, db = P.to_dev(a), P.to_dev(np.zeros_like(a))
    P.check(P.lib().pir_transpose(da.ptr, db.ptr, rows, cols, eb, None))
    want = O.transpose(a.astype(np.uint64), rows, cols)
    assert np.array_equal(db.numpy().astype(np.uint64), want)


@pytest.mark.parametrize("bw", [1, 8, 64])
@pytest.mark.parametrize("eb", [8, 1])
def test_get_set_msbs(rng, bw, eb):
    n = 1001
    bits = rand_u64(rng, n * bw + 3)
    if eb == 1:
        bits = (bits & np.uint64(1)).astype(np.uint8)
    dt = np.uint64 if eb == 8 else np.uint8
    db, dm = P.to_dev(bits), P.to_dev(np.zeros(n, dt))
    P.check(P.lib().pir_get_msbs(db.ptr, dm.ptr, n, bw, eb, None))
    assert np.array_equal(dm.numpy().astype(np.uint64), O.get_msbs(bits.astype(np.uint64), n, bw))
    P.check(P.lib().pir_set_msbs(db.ptr, bits.size, 1, bw, eb, None))
    assert np.array_equal(db.numpy().astype(np.uint64), O.set_msbs(bits.astype(np.uint64), 1, bw))


@pytest.mark.parametrize("op", [P.OP_SHL, P.OP_SHR])
def test_dd_op_shift_clamps(rng, op):
    # amounts >= 64 clamp as on the reference's GPU (the oracle's rule; pinned there)
    amounts = np.array([0, 1, 7, 31, 32, 63, 64, 65, 127, 200, 1 << 32, 1 << 63, (1 << 64) - 1], np.uint64)
    n = amounts.size * 50
    x = rand_u64(rng, n)
    y = np.tile(amounts, 50)
    dx, dy = P.to_dev(x), P.to_dev(y)
    P.check(P.lib().pir_dd_op(op, 8, dx.ptr, 1, dy.ptr, 1, 0, n, None))
    want = O.op_local(np.broadcast_to(x, (3, 2, n)), np.broadcast_to(y, (3, n)), 6 if op == P.OP_SHL else 5)[0, 0]
    assert np.array_equal(dx.numpy(), want)
