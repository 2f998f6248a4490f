"""The measured workloads at their configured sizes, against the oracle:

* the uneven-batch workload: 64 concurrent AlexNet-CIFAR queries (batch per
  query drawn from 1..32, 986 images, FP20) run as one forward over the
  concatenated batch and replayed from its hipGraph; every query's slice of
  every layer equals a separate forward of that query alone, and three
  queries (batch 1, batch 32 and a mid-size one) equal the oracle;
* the ring-GEMM microbenchmark's 3-term RSS local product
  (localMatMul, src/mpc/RSS.inl:1096-1110) at M = N = K = 512 and 1024 in full,
  and at 2048 .. 8192 through sampled rows plus a random projection C.v with
  odd v (any wrong element changes C.v), in both GEMM forms;
* SecureML and LeNet at batch 128 with N(0, 0.1) weights, VGG16-CIFAR at
  batch 4 with its He-init weights, in both execution forms.
"""
import numpy as np
import pytest

import oracle as O
import piranha_amd as P
from conftest import rand_u64
from piranha_amd import configs as C

pytestmark = pytest.mark.gpu


def per_image(layers):
    """Per-image element count of each layer's activation."""
    out = []
    for i in range(len(layers)):
        out.append(C.output_size(layers[: i + 1]))
    return out


def oracle_check(net, layers, batch, fp, x, params=None):
    if params is None:
        params = {i: net.weights(i) for i, l in enumerate(layers) if l[0] in ("fc", "cnn")}
    want = O.net_forward(layers, batch, fp, params, O.set_public(x, fp))
    for i, (act, der) in enumerate(want):
        assert np.array_equal(net.activation(i), act), f"layer {i} ({layers[i][0]}) activation shares differ"
        if der is not None:
            assert np.array_equal(net.derivative(i), der), f"layer {i} derivative shares differ"


# ---- uneven batch: 64 AlexNet-CIFAR queries ---------------------------------

def test_uneven_64_queries_alexnet():
    layers, fp, seed = C.ALEXNET_CIFAR, 20, 1
    qb = C.uneven_query_batches()
    total = sum(qb)
    assert len(qb) == 64 and total == 986
    per = per_image(layers)
    x = np.random.default_rng(seed).uniform(0.0, 1.0, total * C.input_size(layers))

    print(f"uneven: concatenated forward of {total} images", flush=True)
    big = P.Network(C.to_spec(layers), total, fp, seed=seed, mode=P.MODE_FUSED)
    big.forward(x)
    big.capture()
    big.replay()
    big.synchronize()
    acts = [big.activation(i) for i in range(len(layers))]
    ders = {i: big.derivative(i) for i, l in enumerate(layers) if l[0] in ("relu", "maxpool")}
    weights = {i: big.weights(i) for i, l in enumerate(layers) if l[0] in ("fc", "cnn")}
    del big

    nets = {}
    off = 0
    checked_oracle = set()
    mid = sorted(qb)[len(qb) // 2]
    for q, b in enumerate(qb):
        if b not in nets:
            nets[b] = P.Network(C.to_spec(layers), b, fp, seed=seed, mode=P.MODE_FUSED)
            for i, w in weights.items():  # the same seed deals the same weights at every batch size
                got = nets[b].weights(i)
                assert np.array_equal(got[0], w[0]) and (w[1] is None or np.array_equal(got[1], w[1]))
        net = nets[b]
        xq = x[off * C.input_size(layers):(off + b) * C.input_size(layers)]
        print(f"uneven: query {q} batch {b} images {off}..{off + b - 1}", flush=True)
        net.forward(xq)
        for i in range(len(layers)):
            sl = slice(off * per[i], (off + b) * per[i])
            assert np.array_equal(net.activation(i), acts[i][:, :, sl]), f"query {q} layer {i} differs"
            if i in ders:  # a maxpool layer keeps one derivative entry per window element
                dp = ders[i].shape[2] // total
                dsl = slice(off * dp, (off + b) * dp)
                assert np.array_equal(net.derivative(i), ders[i][:, :, dsl]), f"query {q} layer {i} derivative"
        if b in (1, 32, mid) and b not in checked_oracle:
            checked_oracle.add(b)
            print(f"uneven: query {q} batch {b} against the oracle", flush=True)
            oracle_check(net, layers, b, fp, xq, weights)
        off += b
    assert off == total
    assert checked_oracle == {1, 32, mid}


# ---- ring GEMM microbenchmark sizes -----------------------------------------

def gemm3_dev(algo, n, d):
    Cd = P.Buffer((n * n,), np.uint64)
    print(f"ring gemm3 algo={algo} n={n}", flush=True)
    P.check(P.lib().pir_ring_gemm3(algo, n, n, n, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, Cd.ptr, None))
    return Cd.numpy()


def oracle_gemm3(M, N, K, A0, A1, B0, B1):
    """A0.B0 + A0.B1 + A1.B0 mod 2^64 by the oracle's GEMM (A rows M x K)."""
    g = lambda A, B: O.gemm(M, N, K, A, 1, B, 1, 1)  # noqa: E731
    return g(A0, B0) + g(A0, B1) + g(A1, B0)


@pytest.mark.parametrize("n", [512, 1024])
def test_ring_gemm3_full(n):
    rng = np.random.default_rng(n)
    ops = [rand_u64(rng, n * n) for _ in range(4)]
    d = [P.to_dev(a) for a in ops]
    want = oracle_gemm3(n, n, n, *ops)
    for algo in (P.ALGO_LIMB, P.ALGO_PLAIN):
        assert np.array_equal(gemm3_dev(algo, n, d), want), algo


@pytest.mark.parametrize("n", [2048, 4096, 8192])
def test_ring_gemm3_sampled(n):
    rng = np.random.default_rng(n)
    A0, A1, B0, B1 = (rand_u64(rng, n * n) for _ in range(4))
    d = [P.to_dev(a) for a in (A0, A1, B0, B1)]
    rows = np.sort(rng.choice(n, 8, replace=False))
    pick = lambda A: np.ascontiguousarray(A.reshape(n, n)[rows]).ravel()  # noqa: E731
    want_rows = oracle_gemm3(len(rows), n, n, pick(A0), pick(A1), B0, B1).reshape(len(rows), n)
    # C.v = A0.((B0 + B1).v) + A1.(B0.v) mod 2^64, with odd v
    v = rand_u64(rng, n) | np.uint64(1)
    mv = lambda A, x: O.gemm(n, 1, n, A, 1, x, 1, 1)  # noqa: E731
    want_proj = mv(A0, mv(B0, v) + mv(B1, v)) + mv(A1, mv(B0, v))
    for algo in (P.ALGO_LIMB, P.ALGO_PLAIN):
        got = gemm3_dev(algo, n, d)
        assert np.array_equal(got.reshape(n, n)[rows], want_rows), (algo, "rows")
        assert np.array_equal(mv(got, v), want_proj), (algo, "projection")
        del got


# ---- networks at the configured batch sizes ----------------------------------

NET_CASES = [
    ("secureml_b128", C.SECUREML, 128, 13, 0.1),
    ("lenet_b128", C.LENET, 128, 13, 0.1),
    ("vgg16_b4", C.VGG16_CIFAR, 4, 26, None),  # He init, the layers' default
]


@pytest.mark.parametrize("mode", [P.MODE_FUSED, P.MODE_PLAIN], ids=["fused", "plain"])
@pytest.mark.parametrize("name,layers,batch,fp,sigma", NET_CASES, ids=[c[0] for c in NET_CASES])
def test_configured_network(name, layers, batch, fp, sigma, mode):
    print(f"network {name} mode {mode}", flush=True)
    net = P.Network(C.to_spec(layers), batch, fp, seed=1, mode=mode)
    if sigma is not None:
        for i, (w, b) in C.normal_weights(layers, sigma, seed=1).items():
            net.set_weights(i, w, b)
    x = np.random.default_rng(1).uniform(0.0, 1.0, batch * C.input_size(layers))
    net.forward(x)
    oracle_check(net, layers, batch, fp, x)
