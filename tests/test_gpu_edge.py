"""Empty inputs through the C ABI: every entry point accepts n = 0 (or an
empty GEMM side) and returns success without touching memory, in both
execution forms (the reference's DeviceData / thrust calls are no-ops on
empty ranges)."""
import numpy as np
import pytest

import piranha_amd as P

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[P.MODE_FUSED, P.MODE_PLAIN], ids=["fused", "plain"])
def mode(request):
    P.set_mode(request.param)
    yield request.param
    P.set_mode(P.MODE_FUSED)


def test_empty_protocol_calls(mode):
    L = P.lib()
    a, b, c = (P.Buffer((3, 2, 4), np.uint64) for _ in range(3))
    d = P.Buffer((3, 2, 4), np.uint8)
    P.check(L.pir_rss_reconstruct(a.ptr, 0, c.ptr, 8, None))
    P.check(L.pir_rss_reshare(a.ptr, 0, c.ptr, None))
    P.check(L.pir_rss_mult(a.ptr, b.ptr, 0, None))
    P.check(L.pir_rss_and(a.ptr, b.ptr, 0, 8, None))
    P.check(L.pir_rss_divide_public(a.ptr, 0, 1 << 13, None, None))
    P.check(L.pir_rss_divide_public_3out3(a.ptr, 0, 13, c.ptr, None))
    P.check(L.pir_rss_drelu(a.ptr, 0, d.ptr, 1, None))
    P.check(L.pir_rss_relu(a.ptr, 0, c.ptr, d.ptr, 1, None))
    P.check(L.pir_rss_select_share(a.ptr, b.ptr, d.ptr, c.ptr, 0, 1, None))
    P.check(L.pir_rss_maxpool(a.ptr, 0, c.ptr, d.ptr, 4, 1, None))
    P.check(L.pir_tpc_reconstruct(a.ptr, 0, c.ptr, None))
    P.check(L.pir_tpc_divide_public(a.ptr, 0, 1 << 13, None))
    P.check(L.pir_fpc_reconstruct(a.ptr, 0, c.ptr, None))
    P.check(L.pir_fpc_divide_public(a.ptr, 0, 1 << 13, None))
    P.check(L.pir_dd_op(P.OP_ADD, 8, a.ptr, 1, b.ptr, 1, 0, 0, None))
    P.check(L.pir_bitexpand(a.ptr, 0, d.ptr, 1, None))
    P.synchronize()


@pytest.mark.parametrize("algo", [P.ALGO_PLAIN, P.ALGO_LIMB])
@pytest.mark.parametrize("M,N", [(0, 5), (5, 0)])
def test_empty_gemm(algo, M, N):
    a, b, c = (P.Buffer((64,), np.uint64) for _ in range(3))
    P.check(P.lib().pir_gemm(M, N, 4, a.ptr, 1, b.ptr, 1, c.ptr, 1, algo, None))
    P.synchronize()
