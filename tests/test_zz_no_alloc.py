"""pir_ring_gemm3_noalloc, the allocation-free limb form of the RSS local
product (localMatMul, src/mpc/RSS.inl:1096-1110): bit-exact against the
allocating entry point pir_ring_gemm3 (limb and plain forms) and the oracle's
mod-2^64 product, on two calls in a row (the second smaller, so stale scratch
contents from the first would show), including a shape whose K is split
across workgroups (the split-K partials live in the scratch array too), and
the clear error for a shape beyond the scratch array."""
import numpy as np
import pytest

import oracle as O
import piranha_amd as P
from conftest import rand_u64

pytestmark = pytest.mark.gpu


def operands(rng, M, N, K):
    return rand_u64(rng, M * K), rand_u64(rng, M * K), rand_u64(rng, K * N), rand_u64(rng, K * N)


def oracle_gemm3(M, N, K, A0, A1, B0, B1):
    g = lambda A, B: O.gemm(M, N, K, A, 1, B, 1, 1)  # noqa: E731
    return g(A0, B0) + g(A0, B1) + g(A1, B0)


def run(fn_name, M, N, K, ops, algo=None):
    dev = [P.to_dev(x) for x in ops]
    dC = P.Buffer((M * N,), np.uint64)
    fn = getattr(P.lib(), fn_name)
    args = [M, N, K] + [d.ptr for d in dev] + [dC.ptr, None]
    print(f"{fn_name} algo={algo} M={M} N={N} K={K}", flush=True)
    P.check(fn(algo, *args) if algo is not None else fn(*args))
    out = dC.numpy()
    return out


def test_noalloc_matches_ring_gemm3_twice(rng):
    # the first shape splits K (2 * 7040 / 32 = 440 K-steps > 256 per workgroup),
    # the second is smaller in every dimension and not a multiple of the tiles
    for (M, N, K) in [(192, 256, 7040), (37, 70, 83)]:
        ops = operands(rng, M, N, K)
        got = run("pir_ring_gemm3_noalloc", M, N, K, ops)
        want = run("pir_ring_gemm3", M, N, K, ops, algo=P.ALGO_LIMB)
        assert np.array_equal(got, want), (M, N, K)
        assert np.array_equal(got, run("pir_ring_gemm3", M, N, K, ops, algo=P.ALGO_PLAIN)), (M, N, K)
        assert np.array_equal(got, oracle_gemm3(M, N, K, *ops)), (M, N, K)


def test_noalloc_square_1024_then_512(rng):
    for n in (1024, 512):
        ops = operands(rng, n, n, n)
        got = run("pir_ring_gemm3_noalloc", n, n, n, ops)
        assert np.array_equal(got, run("pir_ring_gemm3", n, n, n, ops, algo=P.ALGO_LIMB)), n


def test_noalloc_rejects_oversize():
    L = P.lib()
    # 16 * 8192 * 8192 bytes of packed operand alone exceed the scratch array;
    # the size check comes before any device work, so null operands are never read
    rc = L.pir_ring_gemm3_noalloc(8192, 8192, 8192, 8, 8, 8, 8, 8, None)
    assert rc != 0
    assert "maximum" in L.pir_last_error().decode()
