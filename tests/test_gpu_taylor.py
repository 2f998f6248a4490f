"""getPowers and the Taylor-series functions (RSS.inl:570-612, 1211-1320) on
the GPU, fused and op-by-op forms, against the oracle bit for bit on every
party's shares, and against the reference's own known answers
(Func.cu:499-615: Pow exactly, Sqrt within 5e-2, Inverse within 1e-1).

Inputs cover the series' working range, zero, negative values, values whose
power overflows the coefficient functors' 32-bit `1 << x` (large inputs, where
the reference's results collapse), and arbitrary share tensors whose parties
open different powers."""
import numpy as np
import pytest

import oracle as O
import piranha_amd as P

pytestmark = pytest.mark.gpu

MODES = [P.MODE_FUSED, P.MODE_PLAIN]
FNS = [P.TAYLOR_SQRT, P.TAYLOR_INVERSE, P.TAYLOR_SIGMOID]
ENTRY = {P.TAYLOR_SQRT: "pir_rss_sqrt", P.TAYLOR_INVERSE: "pir_rss_inverse", P.TAYLOR_SIGMOID: "pir_rss_sigmoid"}


@pytest.fixture(params=MODES, ids=["fused", "plain"])
def mode(request):
    P.set_mode(request.param)
    yield request.param
    P.set_mode(P.MODE_FUSED)


def values(rng, n):
    """Log-uniform positives over [1e-4, 1e5] with both signs and specials."""
    v = np.exp(rng.uniform(np.log(1e-4), np.log(1e5), n))
    v[: n // 8] *= -1
    v[n // 8: n // 8 + 6] = [0.0, 1.0, 2.0, 0.5, 3.0e7, -1.0]
    return v


def make(rng, kind, n, fp):
    if kind == "placeholder":
        return O.init_list(values(rng, n), fp)
    return np.ascontiguousarray(rng.integers(-(1 << 40), 1 << 40, size=(3, 2, n)).astype(np.int64).view(np.uint64))


def rec(a, fp):
    return O.reconstruct(a)[0].view(np.int64) / float(1 << fp)


def gpu_powers(x):
    n = x.shape[-1]
    dx, dp = P.to_dev(x), P.Buffer((3, n), np.uint64)
    P.check(P.lib().pir_rss_get_powers(dx.ptr, n, dp.ptr, None))
    return dp.numpy()


def gpu_named(x, fn, fp):
    n = x.shape[-1]
    dx, dy = P.to_dev(x), P.Buffer((3, 2, n), np.uint64)
    P.check(getattr(P.lib(), ENTRY[fn])(dx.ptr, n, dy.ptr, fp, None))
    return dy.numpy()


@pytest.mark.parametrize("kind", ["placeholder", "general"])
def test_get_powers(rng, mode, kind):
    x = make(rng, kind, 300, 20)
    assert np.array_equal(gpu_powers(x), O.get_powers(x))


def test_get_powers_closed_form(rng):
    v = np.concatenate([rng.integers(1, 1 << 62, 500), [1, 2, 3, (1 << 62) + 5, (1 << 63) - 1]]).astype(np.uint64)
    x = np.zeros((3, 2, v.size), np.uint64)
    x[0, 0] = x[2, 1] = v
    pw = gpu_powers(x)
    assert all(pw[q].tolist() == [int(t).bit_length() - 1 for t in v.tolist()] for q in range(3))


def test_pow_kat(mode):
    # Func.cu:499-521: initializer list, then >>= FLOAT_PRECISION
    fp = 26
    x = O.op_scalar(O.init_list(np.array([1, 2, 4, 8, 9, 15, 16], float), fp), fp, 5)
    assert all(gpu_powers(x)[q].tolist() == [0, 1, 2, 3, 3, 3, 4] for q in range(3))


@pytest.mark.parametrize("fn", FNS, ids=["sqrt", "inverse", "sigmoid"])
@pytest.mark.parametrize("fp", [26, 13])
@pytest.mark.parametrize("kind", ["placeholder", "general"])
def test_named(rng, mode, fn, fp, kind):
    x = make(rng, kind, 400, fp)
    assert np.array_equal(gpu_named(x, fn, fp), O.taylor(x, fn, fp))


@pytest.mark.parametrize("coeffs", [(1.0, 0.0, 0.0), (0.5, 0.0, 0.25), (0.0, 0.3, 0.0), (-1.5, 2.0, -0.75)])
@pytest.mark.parametrize("fn", FNS, ids=["sqrt", "inverse", "sigmoid"])
def test_taylor_series_coefficients(rng, mode, coeffs, fn):
    # each combination of zero / non-zero a1 and a2 takes a different set of steps
    fp = 20
    x = make(rng, "placeholder", 250, fp)
    n = x.shape[-1]
    dx, dy = P.to_dev(x), P.Buffer((3, 2, n), np.uint64)
    P.check(P.lib().pir_rss_taylor_series(dx.ptr, n, dy.ptr, *coeffs, fn, fp, None))
    assert np.array_equal(dy.numpy(), O.taylor_series(x, *coeffs, fn, fp))


def test_sqrt_kat(mode):
    # Func.cu:523-572
    fp = 26
    xs = [4e-2, 8e-2, 4e-1, 8.98e-1, 1.5, 1.75, 2.0, 2.25, 2.5, 2.75, 3.0, 3.25, 3.5]
    r = gpu_named(O.init_list(np.array(xs), fp), P.TAYLOR_SQRT, fp)
    assert np.allclose(rec(r, fp), np.sqrt(xs), atol=5e-2)


def test_inverse_kat(mode):
    # Func.cu:574-615
    fp = 26
    xs = [0.25, 0.3, 0.375, 0.49, 0.5, 0.51, 0.75, 1, 1.5]
    r = gpu_named(O.init_list(np.array(xs), fp), P.TAYLOR_INVERSE, fp)
    assert np.allclose(rec(r, fp), 1 / np.array(xs), atol=1e-1)


def test_fused_matches_plain_large(rng):
    # a size the oracle would take minutes on: the two execution forms agree
    fp, n = 26, 1 << 16
    x = make(rng, "placeholder", n, fp)
    outs = []
    for m in MODES:
        P.set_mode(m)
        try:
            outs.append(gpu_named(x, P.TAYLOR_INVERSE, fp))
        finally:
            P.set_mode(P.MODE_FUSED)
    assert np.array_equal(outs[0], outs[1])
    # spot-check a slice against the oracle
    assert np.array_equal(outs[0][:, :, :512], O.taylor(np.ascontiguousarray(x[:, :, :512]), P.TAYLOR_INVERSE, fp))


def test_empty(mode):
    x = np.zeros((3, 2, 0), np.uint64)
    d = P.Buffer((8,), np.uint64)
    P.check(P.lib().pir_rss_get_powers(d.ptr, 0, d.ptr, None))
    P.check(P.lib().pir_rss_sqrt(d.ptr, 0, d.ptr, 26, None))
    assert x.size == 0


def test_argument_errors(rng):
    x = make(rng, "placeholder", 16, 20)
    dx, dy = P.to_dev(x), P.Buffer((3, 2, 16), np.uint64)
    with pytest.raises(P.PirError, match="overlap"):
        P.check(P.lib().pir_rss_sqrt(dx.ptr, 16, dx.ptr, 20, None))
    with pytest.raises(P.PirError, match="overlap"):
        P.check(P.lib().pir_rss_get_powers(dx.ptr, 16, dx.ptr + 8 * 16, None))
    with pytest.raises(P.PirError, match="unknown function"):
        P.check(P.lib().pir_rss_taylor_series(dx.ptr, 16, dy.ptr, 1.0, 0.0, 0.0, 3, 20, None))
    with pytest.raises(P.PirError, match="precision"):
        P.check(P.lib().pir_rss_inverse(dx.ptr, 16, dy.ptr, 63, None))
