"""The limb-GEMM autotuner: tuned plans (tile shape, split-K) change only the
launch configuration, never a share."""
import numpy as np
import pytest

import oracle as O
import piranha_amd as P
from piranha_amd import configs as C
from conftest import rand_u64

pytestmark = pytest.mark.gpu


def test_autotuned_network_matches_oracle(tmp_path):
    layers, batch, fp = C.LENET, 8, 13
    P.autotune_clear()
    net = P.Network(C.to_spec(layers), batch, fp, seed=6, mode=P.MODE_FUSED)
    params = {i: net.weights(i) for i, l in enumerate(layers) if l[0] in ("fc", "cnn")}
    x = np.random.default_rng(5).uniform(0, 1, batch * C.input_size(layers))
    net.forward(x)
    net.autotune()
    assert P.autotune_entries() >= 4  # 2 conv + 2 FC shapes
    net.forward(x)
    net.capture()
    net.replay()
    net.synchronize()
    want = O.net_forward(layers, batch, fp, params, O.set_public(x, fp))
    for i, (act, _) in enumerate(want):
        assert np.array_equal(net.activation(i), act), f"layer {i} differs after autotuning"
    path = tmp_path / "plans.txt"
    P.autotune_save(path)
    n = P.autotune_entries()
    P.autotune_clear()
    P.autotune_load(path)
    assert P.autotune_entries() == n
    P.autotune_clear()


@pytest.mark.parametrize("n", [300, 1024])
def test_autotuned_ring_gemm3(rng, n):
    A0, A1, B0, B1 = (rand_u64(rng, n * n) for _ in range(4))
    d = [P.to_dev(v) for v in (A0, A1, B0, B1)]
    want = P.Buffer((n * n,), np.uint64)
    P.check(P.lib().pir_ring_gemm3(P.ALGO_PLAIN, n, n, n, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, want.ptr, None))
    got = P.Buffer((n * n,), np.uint64)
    P.autotune_clear()
    P.check(P.lib().pir_autotune_set(1))
    try:
        for _ in range(2):  # first call tunes, second uses the table
            P.check(P.lib().pir_ring_gemm3(P.ALGO_LIMB, n, n, n, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, got.ptr,
                                           None))
            assert np.array_equal(got.numpy(), want.numpy())
    finally:
        P.check(P.lib().pir_autotune_set(0))
    assert P.autotune_entries() == 1
    P.autotune_clear()
