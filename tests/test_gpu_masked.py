"""The masked precompute mode on the GPU, bit for bit against the oracle's
restatement at the same seed and call numbers: the Philox4x32-10 blocks and
word stream, the dealer, every masked RSS protocol (fused form) on random
consistent sharings, and whole masked forwards (eager and hipGraph replay)
after the network's weights and input are re-dealt as random sharings.  The
values themselves are pinned by tests/test_oracle_masked.py (closed forms);
here the shares must match exactly.  Functions without a masked form refuse
the mode."""
import ctypes

import numpy as np
import pytest

import oracle as O
import piranha_amd as P
from conftest import rand_u64
from piranha_amd import configs as C

pytestmark = pytest.mark.gpu

SEED = 0x5EED0000DEADBEEF


@pytest.fixture(autouse=True)
def placeholder_after():
    yield
    P.set_mode(P.MODE_FUSED)
    P.set_precompute(P.PRECOMPUTE_PLACEHOLDER)


def masked(call):
    P.set_mode(P.MODE_FUSED)
    P.set_precompute(P.PRECOMPUTE_MASKED, SEED, call)


def test_philox_and_stream():
    L = P.lib()
    ctr = np.array([[0, 0, 0, 0], [0xFFFFFFFF] * 4, [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344]], np.uint32)
    key = np.array([[0, 0], [0xFFFFFFFF] * 2, [0xA4093822, 0x299F31D0]], np.uint32)
    rng = np.random.default_rng(1)
    ctr = np.concatenate([ctr, rng.integers(0, 2**32, (1000, 4), dtype=np.uint32)])
    key = np.concatenate([key, rng.integers(0, 2**32, (1000, 2), dtype=np.uint32)])
    dc, dk, do = P.to_dev(ctr), P.to_dev(key), P.Buffer(ctr.shape, np.uint32)
    P.check(L.pir_prg_philox(dc.ptr, dk.ptr, ctr.shape[0], do.ptr, None))
    assert np.array_equal(do.numpy(), O.philox(ctr, key))
    for n in (1, 2, 1001):
        out = P.Buffer((n,), np.uint64)
        P.check(L.pir_prg_fill(SEED, 77, n, out.ptr, None))
        assert np.array_equal(out.numpy(), O.prg_fill(SEED, 77, n))


def test_deal_and_rerandomize(rng):
    L = P.lib()
    n = 1237
    v = rand_u64(rng, n)
    dv, sh = P.to_dev(v), P.Buffer((3, 2, n), np.uint64)
    P.check(L.pir_rss_deal(dv.ptr, n, SEED, 5, sh.ptr, None))
    want = O.deal(v, SEED, 5)
    assert np.array_equal(sh.numpy(), want)
    P.check(L.pir_rss_rerandomize(sh.ptr, n, SEED, 6, sh.ptr, None))  # in place
    assert np.array_equal(sh.numpy(), O.deal(v, SEED, 6))


def run(fn, *args):
    P.check(getattr(P.lib(), fn)(*args))


def test_reshare_mult_and(rng):
    n = 999
    c = rand_u64(rng, 3, n)
    masked(10)
    dc, out = P.to_dev(c), P.Buffer((3, 2, n), np.uint64)
    run("pir_rss_reshare", dc.ptr, n, out.ptr, None)
    assert np.array_equal(out.numpy(), O.reshare_m(c, SEED, 10))
    x, y = O.deal(rand_u64(rng, n), SEED, 1), O.deal(rand_u64(rng, n), SEED, 2)
    dx, dy = P.to_dev(x), P.to_dev(y)
    run("pir_rss_mult", dx.ptr, dy.ptr, n, None)  # call 11
    assert np.array_equal(dx.numpy(), O.mult_m(x, y, SEED, 11))
    for eb, dt in ((1, np.uint8), (8, np.uint64)):
        a = rng.integers(0, np.iinfo(dt).max, (3, 2, n), dtype=dt, endpoint=True)
        b = rng.integers(0, np.iinfo(dt).max, (3, 2, n), dtype=dt, endpoint=True)
        da, db = P.to_dev(a), P.to_dev(b)
        call = P.prg_call()
        run("pir_rss_and", da.ptr, db.ptr, n, eb, None)
        assert np.array_equal(da.numpy(), O.rss_and_m(a, b, SEED, call))


@pytest.mark.parametrize("k", [13, 26])
def test_truncations(rng, k):
    n = 777
    a = rand_u64(rng, 3, n)
    masked(20)
    da, out = P.to_dev(a), P.Buffer((3, 2, n), np.uint64)
    run("pir_rss_divide_public_3out3", da.ptr, n, k, out.ptr, None)
    assert np.array_equal(out.numpy(), O.divide_public_3out3_m(a, k, SEED, 20))
    s = O.deal(rand_u64(rng, n), SEED, 3)
    ds = P.to_dev(s)
    run("pir_rss_divide_public", ds.ptr, n, 1 << k, None, None)  # call 21
    assert np.array_equal(ds.numpy(), O.divide_public_m(s, k, SEED, 21))


@pytest.mark.parametrize("algo", [P.ALGO_LIMB, P.ALGO_PLAIN])
@pytest.mark.parametrize("ta,tb,tc", [(1, 1, 1), (0, 1, 0), (1, 0, 1)])
def test_matmul(rng, algo, ta, tb, tc):
    M, N, K, fp = 37, 70, 300, 13
    a, b = O.deal(rand_u64(rng, M * K), SEED, 1), O.deal(rand_u64(rng, K * N), SEED, 2)
    masked(30)
    da, db, dc = P.to_dev(a), P.to_dev(b), P.Buffer((3, 2, M * N), np.uint64)
    run("pir_rss_matmul", da.ptr, db.ptr, dc.ptr, M, N, K, ta, tb, tc, fp, algo, None)
    assert np.array_equal(dc.numpy(), O.matmul_m(a, b, M, N, K, ta, tb, tc, fp, SEED, 30))


def test_matmul_split_k(rng):
    # K long enough for several K splits: the masked epilogue reduces the partials
    M, N, K, fp = 64, 64, 9000, 26
    a, b = O.deal(rand_u64(rng, M * K), SEED, 1), O.deal(rand_u64(rng, K * N), SEED, 2)
    masked(31)
    da, db, dc = P.to_dev(a), P.to_dev(b), P.Buffer((3, 2, M * N), np.uint64)
    run("pir_rss_matmul", da.ptr, db.ptr, dc.ptr, M, N, K, 1, 1, 1, fp, P.ALGO_LIMB, None)
    assert np.array_equal(dc.numpy(), O.matmul_m(a, b, M, N, K, 1, 1, 1, fp, SEED, 31))


@pytest.mark.parametrize("geom", [(2, 9, 7, 3, 3, 16, 1, 1), (1, 8, 8, 32, 3, 64, 1, 1), (2, 12, 12, 3, 5, 8, 2, 0)])
def test_convolution(rng, geom):
    b, H, W, Din, f, Dout, stride, pad = geom
    fp = 13
    x = O.deal(rand_u64(rng, b * H * W * Din), SEED, 1)
    w = O.deal(rand_u64(rng, Dout * f * f * Din), SEED, 2)
    OH, OW = (H + 2 * pad - f) // stride + 1, (W + 2 * pad - f) // stride + 1
    masked(40)
    dx, dw, dy = P.to_dev(x), P.to_dev(w), P.Buffer((3, 2, b * OH * OW * Dout), np.uint64)
    run("pir_rss_convolution", dx.ptr, dw.ptr, dy.ptr, b, H, W, f, Din, Dout, stride, pad, fp, P.ALGO_LIMB, None)
    assert np.array_equal(dy.numpy(), O.convolution_m(x, w, b, H, W, f, Din, Dout, stride, pad, fp, SEED, 40))


@pytest.mark.parametrize("n", [700, 701])
@pytest.mark.parametrize("dt", [np.uint8, np.uint64])
def test_drelu_relu_select(rng, dt, n):
    eb = np.dtype(dt).itemsize
    v = rand_u64(rng, n)
    v[:4] = [0, 1, 2**64 - 1, 2**63]
    x = O.deal(v, SEED, 1)
    dx = P.to_dev(x)
    masked(50)
    d = P.Buffer((3, 2, n), dt)
    run("pir_rss_drelu", dx.ptr, n, d.ptr, eb, None)
    assert np.array_equal(d.numpy(), O.drelu_m(x, SEED, 50, dt))
    y, dd = P.Buffer((3, 2, n), np.uint64), P.Buffer((3, 2, n), dt)
    run("pir_rss_relu", dx.ptr, n, y.ptr, dd.ptr, eb, None)  # call 51
    wy, wd = O.relu_m(x, SEED, 51, dt)
    assert np.array_equal(y.numpy(), wy) and np.array_equal(dd.numpy(), wd)
    xv, yv = O.deal(rand_u64(rng, n), SEED, 2), O.deal(rand_u64(rng, n), SEED, 3)
    bits = rng.integers(0, 2, (3, 2, n)).astype(dt)
    z = P.Buffer((3, 2, n), np.uint64)
    dxv, dyv, dbits = P.to_dev(xv), P.to_dev(yv), P.to_dev(bits)  # kept alive until the call returns
    run("pir_rss_select_share", dxv.ptr, dyv.ptr, dbits.ptr, z.ptr, n, eb, None)
    assert np.array_equal(z.numpy(), O.select_share_m(xv, yv, bits, SEED, 52))


@pytest.mark.parametrize("k", [2, 4, 16])
@pytest.mark.parametrize("dt", [np.uint8, np.uint64])
def test_maxpool(rng, k, dt):
    n = k * 53
    v = (rng.integers(-40, 40, n)).astype(np.int64).view(np.uint64)
    pools = O.deal(v, SEED, 1)
    masked(60)
    dp, dr, dd = P.to_dev(pools), P.Buffer((3, 2, n // k), np.uint64), P.Buffer((3, 2, n), dt)
    run("pir_rss_maxpool", dp.ptr, n, dr.ptr, dd.ptr, k, np.dtype(dt).itemsize, None)
    r, d, after = O.maxpool_m(pools, k, SEED, 60, dt)
    assert np.array_equal(dr.numpy(), r)
    assert np.array_equal(dd.numpy(), d)
    assert np.array_equal(dp.numpy(), after)


NETS = [
    ("secureml", C.SECUREML, 16, 13),
    ("lenet", C.LENET, 2, 13),
    ("alexnet", C.ALEXNET_CIFAR, 2, 20),
]


@pytest.mark.parametrize("name,layers,batch,fp", NETS, ids=[c[0] for c in NETS])
def test_masked_network(name, layers, batch, fp):
    net = P.Network(C.to_spec(layers), batch, fp, seed=4, mode=P.MODE_FUSED)
    for i, (w, b) in C.normal_weights(layers, 0.1, seed=4).items():
        net.set_weights(i, w, b)
    x = np.random.default_rng(8).uniform(0.0, 1.0, batch * C.input_size(layers))
    net.forward(x)  # setPublic of the input into the network's buffer
    net.rerandomize(SEED ^ 1)
    net.set_precompute(P.PRECOMPUTE_MASKED, SEED, 1000)
    net.forward_shares(None)
    net.synchronize()  # the network's stream does not order against the copies below
    ptr, n = net.input_ptr()
    xin = np.empty((3, 2, n), np.uint64)
    P.check(P.lib().pir_memcpy_d2h(xin.ctypes.data, ptr, xin.nbytes, None))
    params = {i: net.weights(i) for i, l in enumerate(layers) if l[0] in ("fc", "cnn")}
    print(f"masked network {name} b{batch}: oracle forward", flush=True)
    want = O.net_forward_m(layers, batch, fp, params, xin, SEED, 1000)
    for i, (act, der) in enumerate(want):
        assert np.array_equal(net.activation(i), act), f"layer {i} ({layers[i][0]}) activation shares differ"
        if der is not None:
            assert np.array_equal(net.derivative(i), der), f"layer {i} derivative shares differ"
    # the captured forward replays the call numbers of its eager pass
    net.set_precompute(P.PRECOMPUTE_MASKED, SEED, 1000)
    net.capture()
    for _ in range(2):
        net.replay()
        net.synchronize()
        assert np.array_equal(net.activation(len(layers) - 1), want[-1][0])


def test_masked_mode_refusals(rng):
    n = 64
    x = P.to_dev(O.deal(rand_u64(rng, n), SEED, 1))
    y, d = P.Buffer((3, 2, n), np.uint64), P.Buffer((3, 2, n), np.uint8)
    L = P.lib()
    P.set_mode(P.MODE_PLAIN)
    P.set_precompute(P.PRECOMPUTE_MASKED, SEED, 0)
    assert L.pir_rss_relu(x.ptr, n, y.ptr, d.ptr, 1, None) != 0  # no op-by-op masked form
    assert "fused" in L.pir_last_error().decode()
    P.set_mode(P.MODE_FUSED)
    assert L.pir_rss_get_powers(x.ptr, n, y.ptr, None) != 0
    t = P.to_dev(rand_u64(rng, 2, n))
    assert L.pir_tpc_mult(t.ptr, t.ptr, n, None) != 0
    assert L.pir_rss_divide_public(x.ptr, n, 3, None, None) != 0  # not a power of two
    P.set_precompute(P.PRECOMPUTE_PLACEHOLDER)
    assert L.pir_rss_relu(x.ptr, n, y.ptr, d.ptr, 1, None) == 0
    # a network in the masked mode refuses a training step
    layers = C.SECUREML
    net = P.Network(C.to_spec(layers), 2, 13, seed=1, mode=P.MODE_FUSED)
    net.forward(np.zeros(2 * 784))
    net.set_precompute(P.PRECOMPUTE_MASKED, SEED, 0)
    labels = np.zeros((2, 10))
    labels[:, 1] = 1
    with pytest.raises(P.PirError):
        net.backward(labels.ravel())
