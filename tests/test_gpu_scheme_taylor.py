"""getPowers and the Taylor-series functions of the TPC and FPC schemes on the
GPU (TPC.inl:615-642,1202-1290, FPC.inl:714-741,1418-1526) against the oracle,
bit for bit on every party's shares: initializer-list inputs over the series'
working range, zero, negatives and overflowing powers, and arbitrary share
values; plus the typed Func.cu Pow / Sqrt / Inverse known answers."""
import numpy as np
import pytest

import oracle as O
import piranha_amd as P

pytestmark = pytest.mark.gpu

SHAPE = {"tpc": (2,), "fpc": (4, 3)}
INIT = {"tpc": O.tpc_init_list, "fpc": O.fpc_init_list}
REC = {"tpc": O.tpc_reconstruct, "fpc": O.fpc_reconstruct}
NAMES = {P.TAYLOR_SQRT: "sqrt", P.TAYLOR_INVERSE: "inverse", P.TAYLOR_SIGMOID: "sigmoid"}
KINDS = ["tpc", "fpc"]


def values(rng, n):
    v = np.exp(rng.uniform(np.log(1e-4), np.log(1e5), n))
    v[: n // 8] *= -1
    v[n // 8: n // 8 + 6] = [0.0, 1.0, 2.0, 0.5, 3.0e7, -1.0]
    return v


def make(rng, kind, n, fp, general):
    if not general:
        return INIT[kind](values(rng, n), fp)
    return np.ascontiguousarray(rng.integers(-(1 << 40), 1 << 40, size=SHAPE[kind] + (n,)).astype(np.int64)
                                .view(np.uint64))


def gpu_powers(kind, x):
    n = x.shape[-1]
    dx, dp = P.to_dev(x), P.Buffer((SHAPE[kind][0], n), np.uint64)
    P.check(getattr(P.lib(), f"pir_{kind}_get_powers")(dx.ptr, n, dp.ptr, None))
    return dp.numpy()


def gpu_named(kind, x, fn, fp):
    n = x.shape[-1]
    dx, dy = P.to_dev(x), P.Buffer(x.shape, np.uint64)
    P.check(getattr(P.lib(), f"pir_{kind}_{NAMES[fn]}")(dx.ptr, n, dy.ptr, fp, None))
    return dy.numpy()


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("general", [False, True], ids=["init", "general"])
def test_get_powers(rng, kind, general):
    x = make(rng, kind, 300, 20, general)
    assert np.array_equal(gpu_powers(kind, x), O.scheme_get_powers(kind, x))


@pytest.mark.parametrize("kind", KINDS)
def test_pow_kat(kind):
    x = INIT[kind]([1, 2, 4, 8, 9, 15, 16], 26, convert=False)  # Func.cu:499-521
    pw = gpu_powers(kind, x)
    assert all(pw[q].tolist() == [0, 1, 2, 3, 3, 3, 4] for q in range(pw.shape[0]))


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("fn", [P.TAYLOR_SQRT, P.TAYLOR_INVERSE, P.TAYLOR_SIGMOID], ids=["sqrt", "inverse", "sigmoid"])
@pytest.mark.parametrize("fp", [26, 13])
@pytest.mark.parametrize("general", [False, True], ids=["init", "general"])
def test_named(rng, kind, fn, fp, general):
    x = make(rng, kind, 300, fp, general)
    assert np.array_equal(gpu_named(kind, x, fn, fp), O.scheme_taylor(kind, x, fn, fp))


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("coeffs", [(1.0, 0.0, 0.0), (0.5, 0.0, 0.25), (0.0, 0.3, 0.0), (-1.5, 2.0, -0.75)])
def test_taylor_series_coefficients(rng, kind, coeffs):
    fp = 20
    x = make(rng, kind, 200, fp, False)
    n = x.shape[-1]
    dx, dy = P.to_dev(x), P.Buffer(x.shape, np.uint64)
    P.check(getattr(P.lib(), f"pir_{kind}_taylor_series")(dx.ptr, n, dy.ptr, *coeffs, P.TAYLOR_INVERSE, fp, None))
    assert np.array_equal(dy.numpy(), O.scheme_taylor_series(kind, x, *coeffs, P.TAYLOR_INVERSE, fp))


@pytest.mark.parametrize("kind", KINDS)
def test_sqrt_inverse_kats(kind):
    fp = 26
    xs = [4e-2, 8e-2, 4e-1, 8.98e-1, 1.5, 1.75, 2.0, 2.25, 2.5, 2.75, 3.0, 3.25, 3.5]  # Func.cu:523-572
    r = REC[kind](gpu_named(kind, INIT[kind](xs, fp), P.TAYLOR_SQRT, fp))
    assert all(np.allclose(r[q].view(np.int64) / 2.0**fp, np.sqrt(xs), atol=5e-2) for q in range(r.shape[0]))
    xi = [0.25, 0.3, 0.375, 0.49, 0.5, 0.51, 0.75, 1, 1.5]  # Func.cu:574-615
    r = REC[kind](gpu_named(kind, INIT[kind](xi, fp), P.TAYLOR_INVERSE, fp))
    assert all(np.allclose(r[q].view(np.int64) / 2.0**fp, 1 / np.array(xi), atol=1e-1) for q in range(r.shape[0]))


@pytest.mark.parametrize("kind", KINDS)
def test_argument_errors(rng, kind):
    x = make(rng, kind, 16, 20, False)
    dx, dy = P.to_dev(x), P.Buffer(x.shape, np.uint64)
    lib = P.lib()
    with pytest.raises(P.PirError, match="overlap"):
        P.check(getattr(lib, f"pir_{kind}_sqrt")(dx.ptr, 16, dx.ptr, 20, None))
    with pytest.raises(P.PirError, match="overlap"):
        P.check(getattr(lib, f"pir_{kind}_get_powers")(dx.ptr, 16, dx.ptr, None))
    with pytest.raises(P.PirError, match="unknown function"):
        P.check(getattr(lib, f"pir_{kind}_taylor_series")(dx.ptr, 16, dy.ptr, 1.0, 0.0, 0.0, 3, 20, None))
    with pytest.raises(P.PirError, match="precision"):
        P.check(getattr(lib, f"pir_{kind}_inverse")(dx.ptr, 16, dy.ptr, 63, None))
    P.check(getattr(lib, f"pir_{kind}_sigmoid")(dx.ptr, 0, dy.ptr, 20, None))  # empty


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("general", [False, True], ids=["init", "general"])
def test_divide_public_local(rng, kind, general):
    n = 400
    a = make(rng, kind, n, 20, general)
    d = rng.integers(1, 1 << 20, (SHAPE[kind][0], n)).astype(np.uint64)
    d[:, :8] = [1, 2, 3, 0, 7, 1 << 20, 1 << 40, (1 << 64) - 1]  # including 0 and a "negative" denominator
    da, dd = P.to_dev(a), P.to_dev(d)
    P.check(getattr(P.lib(), f"pir_{kind}_divide_public_local")(da.ptr, n, dd.ptr, None))
    assert np.array_equal(da.numpy(), O.scheme_divide_public_vec(kind, a, d))
