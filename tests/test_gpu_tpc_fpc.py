"""2-party (TPC) and 4-party (FPC) matmul / convolution / truncation on the GPU
against the oracle (src/mpc/TPC.inl, FPC.inl semantics), bit for bit on every
party's shares, in the fused limb form and the reference's op-by-op form.

Inputs are both the reference's initializer-list sharings (the value on one
component only) and arbitrary share tensors (every array random): the FPC
placeholder reshare is not a correct protocol for general shares, so the
second kind checks the per-party restatement itself."""
import numpy as np
import pytest

import oracle as O
import piranha_amd as P
from conftest import rand_u64

pytestmark = pytest.mark.gpu

FP = 13
SHAPE = {"tpc": (2,), "fpc": (4, 3)}
INIT = {"tpc": O.tpc_init_list, "fpc": O.fpc_init_list}


@pytest.fixture(params=[P.MODE_FUSED, P.MODE_PLAIN], ids=["fused", "plain"])
def mode(request):
    P.set_mode(request.param)
    yield request.param
    P.set_mode(P.MODE_FUSED)


@pytest.fixture(params=["tpc", "fpc"])
def scheme(request):
    return request.param


@pytest.fixture(params=["public", "general"])
def kind(request):
    return request.param


def make(rng, scheme, kind, n):
    if kind == "public":
        return INIT[scheme](rng.uniform(-8.0, 8.0, n), FP)
    v = rng.integers(-(1 << 40), 1 << 40, size=SHAPE[scheme] + (n,)).astype(np.int64).view(np.uint64)
    return np.ascontiguousarray(v)


def lib_fn(scheme, name):
    return getattr(P.lib(), f"pir_{scheme}_{name}")


def test_reconstruct(rng, scheme):
    x = rand_u64(rng, *SHAPE[scheme], 1000)
    nparty = SHAPE[scheme][0]
    dx, out = P.to_dev(x), P.Buffer((nparty, 1000), np.uint64)
    P.check(lib_fn(scheme, "reconstruct")(dx.ptr, 1000, out.ptr, None))
    want = O.tpc_reconstruct(x) if scheme == "tpc" else O.fpc_reconstruct(x)
    assert np.array_equal(out.numpy(), want)


@pytest.mark.parametrize("k", [1, 13, 26])
def test_divide_public(rng, scheme, kind, k):
    a = make(rng, scheme, kind, 3000)
    da = P.to_dev(a)
    P.check(lib_fn(scheme, "divide_public")(da.ptr, 3000, 1 << k, None))
    want = O.tpc_divide_public(a, 1 << k) if scheme == "tpc" else O.fpc_divide_public(a, 1 << k)
    assert np.array_equal(da.numpy(), want)


@pytest.mark.parametrize("i,j", [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)])
def test_fpc_reshare(rng, i, j):
    z, zp, base = rand_u64(rng, 4, 500), rand_u64(rng, 4, 500), rand_u64(rng, 4, 3, 500)
    dz, dzp, dout = P.to_dev(z), P.to_dev(zp), P.to_dev(base)
    P.check(P.lib().pir_fpc_reshare(dz.ptr, dzp.ptr, 500, i, j, dout.ptr, None))
    assert np.array_equal(dout.numpy(), O.fpc_reshare(z, zp, i, j, base))


@pytest.mark.parametrize("algo", [P.ALGO_PLAIN, P.ALGO_LIMB])
@pytest.mark.parametrize("ta,tb,tc", [(1, 1, 1), (0, 0, 0), (1, 0, 1), (0, 1, 0)])
def test_matmul(rng, scheme, mode, kind, algo, ta, tb, tc):
    M, N, K = 33, 45, 70
    a, b = make(rng, scheme, kind, M * K), make(rng, scheme, kind, K * N)
    da, db, dc = P.to_dev(a), P.to_dev(b), P.Buffer(SHAPE[scheme] + (M * N,), np.uint64)
    P.check(lib_fn(scheme, "matmul")(da.ptr, db.ptr, dc.ptr, M, N, K, ta, tb, tc, FP, algo, None))
    ref = O.tpc_matmul if scheme == "tpc" else O.fpc_matmul
    assert np.array_equal(dc.numpy(), ref(a, b, M, N, K, ta, tb, tc, FP))


def test_matmul_long_k(rng, scheme):
    """K beyond one workgroup's 8192-k accumulation chunk (split-K path)."""
    M, N, K = 20, 24, 9000
    a, b = make(rng, scheme, "general", M * K), make(rng, scheme, "general", K * N)
    da, db, dc = P.to_dev(a), P.to_dev(b), P.Buffer(SHAPE[scheme] + (M * N,), np.uint64)
    P.check(lib_fn(scheme, "matmul")(da.ptr, db.ptr, dc.ptr, M, N, K, 1, 1, 1, FP, P.ALGO_LIMB, None))
    ref = O.tpc_matmul if scheme == "tpc" else O.fpc_matmul
    assert np.array_equal(dc.numpy(), ref(a, b, M, N, K, 1, 1, 1, FP))


@pytest.mark.parametrize("algo", [P.ALGO_PLAIN, P.ALGO_LIMB])
@pytest.mark.parametrize("b,H,W,f,Din,Dout,stride,pad", [(2, 3, 3, 3, 1, 1, 1, 1), (2, 9, 9, 3, 2, 3, 2, 1),
                                                         (3, 8, 8, 3, 32, 20, 1, 1)])
def test_convolution(rng, scheme, mode, kind, algo, b, H, W, f, Din, Dout, stride, pad):
    a, w = make(rng, scheme, kind, b * H * W * Din), make(rng, scheme, kind, Dout * f * f * Din)
    OH = (H + 2 * pad - f) // stride + 1
    OW = (W + 2 * pad - f) // stride + 1
    n = b * OH * OW * Dout
    da, dw, dc = P.to_dev(a), P.to_dev(w), P.Buffer(SHAPE[scheme] + (n,), np.uint64)
    P.check(lib_fn(scheme, "convolution")(da.ptr, dw.ptr, dc.ptr, b, H, W, f, Din, Dout, stride, pad, FP, algo, None))
    ref = O.tpc_convolution if scheme == "tpc" else O.fpc_convolution
    assert np.array_equal(dc.numpy(), ref(a, w, b, H, W, f, Din, Dout, stride, pad, FP))


def test_invalid_arguments_are_rejected(scheme):
    with pytest.raises(P.PirError):
        P.check(lib_fn(scheme, "matmul")(None, None, None, 0, 4, 4, 1, 1, 1, FP, P.ALGO_LIMB, None))
    if scheme == "fpc":
        with pytest.raises(P.PirError):
            P.check(P.lib().pir_fpc_reshare(None, None, 4, 2, 1, None, None))
