"""Ring GEMM / convolution on the GPU (plain VALU and int8-limb MFMA) against
the oracle's mod-2^64 products (gemm.cuh:64-112, conv.cuh:359-401 semantics)."""
import itertools

import numpy as np
import pytest

import oracle as O
import piranha_amd as P
from conftest import rand_u64

pytestmark = pytest.mark.gpu


def run_gemm(M, N, K, A, ta, B, tb, tc, algo):
    dA, dB = P.to_dev(A), P.to_dev(B)
    dC = P.Buffer((M * N,), np.uint64)
    P.check(P.lib().pir_gemm(M, N, K, dA.ptr, int(ta), dB.ptr, int(tb), dC.ptr, int(tc), algo, None))
    return dC.numpy()


@pytest.mark.parametrize("algo", [P.ALGO_PLAIN, P.ALGO_LIMB])
def test_gemm_kats(algo):
    # GPU.cu:12-78 (CutlassGemm, CutlassGemmTranspose, CutlassGemmTranspose2)
    a = np.array([1, 2, 2, 1, 1, 2], np.uint64)
    b = np.array([2, 2, 2, 1, 1, 1, 2, 2, 2, 1, 1, 1], np.uint64)
    assert run_gemm(2, 4, 3, a, 0, b, 0, 0, algo).tolist() == [8, 10, 4, 5, 8, 10, 4, 5]
    assert run_gemm(2, 2, 3, a, 0, a, 1, 0, algo).tolist() == [6, 6, 6, 9]
    assert run_gemm(3, 3, 2, a, 1, a, 0, 0, algo).tolist() == [5, 4, 5, 4, 5, 4, 5, 4, 5]
    assert run_gemm(3, 1, 2, a, 1, np.array([4, 2], np.uint64), 1, 0, algo).tolist() == [8, 10, 8]
    a2 = np.array([1, 2, 2, 1, 0, 1], np.uint64)
    b2 = np.array([1, 3, 2, 1, 0, 1, 1, 0], np.uint64)
    assert run_gemm(3, 4, 2, a2, 1, b2, 1, 1, algo).tolist() == [1, 5, 4, 1, 2, 7, 5, 2, 0, 1, 1, 0]


@pytest.mark.parametrize("algo", [P.ALGO_PLAIN, P.ALGO_LIMB])
@pytest.mark.parametrize("ta,tb,tc", list(itertools.product([0, 1], repeat=3)))
def test_gemm_random_layouts(rng, algo, ta, tb, tc):
    M, N, K = 37, 70, 83
    A, B = rand_u64(rng, M * K), rand_u64(rng, K * N)
    got = run_gemm(M, N, K, A, ta, B, tb, tc, algo)
    assert np.array_equal(got, O.gemm(M, N, K, A, ta, B, tb, tc))


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (64, 64, 32), (130, 65, 9000), (257, 10, 784)])
def test_limb_gemm_shapes(rng, M, N, K):
    # K = 9000 crosses the 8192-k int32 flush; N = 10 / M = 1 exercise padding
    A, B = rand_u64(rng, M * K), rand_u64(rng, K * N)
    got = run_gemm(M, N, K, A, 1, B, 1, 1, P.ALGO_LIMB)
    assert np.array_equal(got, O.gemm(M, N, K, A, 1, B, 1, 1))


@pytest.mark.parametrize("split", [1, 2, 5, 16])
def test_limb_split_k(rng, split):
    M, N, K = 96, 72, 2000
    A0, A1 = rand_u64(rng, M * K), rand_u64(rng, M * K)
    B0, B1 = rand_u64(rng, K * N), rand_u64(rng, K * N)
    L = P.lib()
    dA0, dA1, dB0, dB1 = (P.to_dev(x) for x in (A0, A1, B0, B1))
    nbytes = L.pir_limb_packed_bytes(K, N, 2)
    packed = P.Buffer((nbytes,), np.uint8)
    b0, _k0 = P._ptr_array([dB0.ptr, dB0.ptr])
    b1, _k1 = P._ptr_array([dB1.ptr, None])
    P.check(L.pir_limb_pack(b0, b1, 2, K, N, 1, packed.ptr, None))
    a, _k2 = P._ptr_array([dA0.ptr, dA1.ptr])
    dC = P.Buffer((M * N,), np.uint64)
    P.check(L.pir_limb_gemm(a, 2, M, K, 1, packed.ptr, N, dC.ptr, 1, split, None))
    want = O.gemm(M, N, K, A0, 1, (B0 + B1).astype(np.uint64), 1, 1) + O.gemm(M, N, K, A1, 1, B0, 1, 1)
    assert np.array_equal(dC.numpy(), want)


@pytest.mark.parametrize("algo", [P.ALGO_PLAIN, P.ALGO_LIMB])
def test_ring_gemm3(rng, algo):
    M, N, K = 65, 66, 200
    A0, A1 = rand_u64(rng, M * K), rand_u64(rng, M * K)
    B0, B1 = rand_u64(rng, K * N), rand_u64(rng, K * N)
    d = [P.to_dev(x) for x in (A0, A1, B0, B1)]
    dC = P.Buffer((M * N,), np.uint64)
    P.check(P.lib().pir_ring_gemm3(algo, M, N, K, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, dC.ptr, None))
    want = O.gemm(M, N, K, A0, 1, B0, 1, 1) + O.gemm(M, N, K, A0, 1, B1, 1, 1) + O.gemm(M, N, K, A1, 1, B0, 1, 1)
    assert np.array_equal(dC.numpy(), want)


def run_conv(x, w, b, H, W, Din, Dout, f, pad, stride, algo):
    OH = (H + 2 * pad - f) // stride + 1
    OW = (W + 2 * pad - f) // stride + 1
    dx, dw = P.to_dev(x), P.to_dev(w)
    dy = P.Buffer((b * OH * OW * Dout,), np.uint64)
    P.check(P.lib().pir_conv_fprop(dx.ptr, dw.ptr, dy.ptr, b, H, W, Din, Dout, f, f, pad, pad, stride, 1, algo, None))
    return dy.numpy()


@pytest.mark.parametrize("algo", [P.ALGO_PLAIN, P.ALGO_LIMB])
def test_conv_kat(algo):
    # GPU.cu:80-99 (CutlassConvFprop)
    im = np.array([1, 2, 1, 2, 3, 2, 1, 2, 1], np.uint64)
    f = np.array([1, 0, 1, 0, 1, 0, 1, 0, 1], np.uint64)
    assert run_conv(im, f, 1, 3, 3, 1, 1, 3, 1, 1, algo).tolist() == [4, 6, 4, 6, 7, 6, 4, 6, 4]


@pytest.mark.parametrize("algo", [P.ALGO_PLAIN, P.ALGO_LIMB])
@pytest.mark.parametrize("b,H,W,Din,Dout,f,pad,stride", [
    (2, 5, 5, 4, 10, 3, 1, 2),     # GPU.cu:131-172 geometry
    (3, 9, 9, 3, 64, 3, 1, 1),     # VGG conv1 channel count
    (2, 12, 12, 20, 50, 5, 0, 1),  # LeNet conv2
    (1, 8, 8, 64, 70, 3, 1, 1),    # Cin % 8 == 0 fast path, N padding
    (2, 32, 32, 3, 96, 11, 9, 4),  # AlexNet conv1
])
def test_conv_random(rng, algo, b, H, W, Din, Dout, f, pad, stride):
    x, w = rand_u64(rng, b * H * W * Din), rand_u64(rng, Dout * f * f * Din)
    got = run_conv(x, w, b, H, W, Din, Dout, f, pad, stride, algo)
    want = O.conv_fprop(x, w, b, H, W, Din, Dout, f, f, pad, pad, stride)
    assert np.array_equal(got, want)
