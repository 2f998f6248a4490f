"""The oracle's getPowers and taylorSeries for the TPC and FPC schemes
(TPC.inl:1202-1290, FPC.inl:1418-1526), pinned by the typed Func.cu known
answers (Pow :499-521 exactly, Sqrt :523-572 within 5e-2, Inverse :574-615
within 1e-1, all instantiated for TPC<uint64_t> and FPC<uint64_t>), by the
closed form floor(log2 v) of getPowers, and by the RSS restatement: on
initializer-list inputs every scheme opens the same value, since the series is
a fixed sequence of public-denominator divisions of the same products."""
import numpy as np
import pytest

import oracle as O

FP = 26
INIT = {"tpc": O.tpc_init_list, "fpc": O.fpc_init_list}
REC = {"tpc": O.tpc_reconstruct, "fpc": O.fpc_reconstruct}


def opened(kind, x):
    r = REC[kind](x)
    assert all(np.array_equal(r[0], r[q]) for q in range(r.shape[0]))
    return r[0]


@pytest.mark.parametrize("kind", ["tpc", "fpc"])
def test_pow_kat(kind):
    x = INIT[kind]([1, 2, 4, 8, 9, 15, 16], FP, convert=False)  # the reference's >>= FP of fixed inputs
    pw = O.scheme_get_powers(kind, x)
    assert all(pw[q].tolist() == [0, 1, 2, 3, 3, 3, 4] for q in range(pw.shape[0]))


@pytest.mark.parametrize("kind", ["tpc", "fpc"])
def test_pow_closed_form(rng, kind):
    v = np.concatenate([rng.integers(1, 1 << 62, 100), [1, 2, 3, (1 << 63) - 1]]).astype(np.int64)
    x = INIT[kind](np.zeros(v.size), FP, convert=False)
    slots = [(0,)] if kind == "tpc" else [(1, 2), (2, 1), (3, 0)]
    for sl in slots:
        x[sl] = v.view(np.uint64)
    pw = O.scheme_get_powers(kind, x)
    assert all(pw[q].tolist() == [int(t).bit_length() - 1 for t in v.tolist()] for q in range(pw.shape[0]))


@pytest.mark.parametrize("kind", ["tpc", "fpc"])
def test_sqrt_inverse_kats(kind):
    xs = [4e-2, 8e-2, 4e-1, 8.98e-1, 1.5, 1.75, 2.0, 2.25, 2.5, 2.75, 3.0, 3.25, 3.5]
    r = O.scheme_taylor(kind, INIT[kind](xs, FP), O.TAYLOR_SQRT, FP)
    assert np.allclose(opened(kind, r).view(np.int64) / 2.0**FP, np.sqrt(xs), atol=5e-2)
    xi = [0.25, 0.3, 0.375, 0.49, 0.5, 0.51, 0.75, 1, 1.5]
    r = O.scheme_taylor(kind, INIT[kind](xi, FP), O.TAYLOR_INVERSE, FP)
    assert np.allclose(opened(kind, r).view(np.int64) / 2.0**FP, 1 / np.array(xi), atol=1e-1)


@pytest.mark.parametrize("kind", ["tpc", "fpc"])
@pytest.mark.parametrize("fn", [O.TAYLOR_SQRT, O.TAYLOR_INVERSE, O.TAYLOR_SIGMOID])
@pytest.mark.parametrize("fp", [26, 13])
def test_matches_rss(rng, kind, fn, fp):
    v = np.exp(rng.uniform(np.log(1e-3), np.log(1e4), 120))
    v[:3] = [0.0, -1.5, 1.0]
    r = O.scheme_taylor(kind, INIT[kind](v, fp), fn, fp)
    want = O.reconstruct(O.taylor(O.init_list(v, fp), fn, fp))[0]
    assert np.array_equal(opened(kind, r), want)


@pytest.mark.parametrize("kind", ["tpc", "fpc"])
def test_truncate_vec_kat(kind):
    # Func.cu:133-154 (TruncateVec)
    a = INIT[kind]([1 << 3, 2 << 1, 3 << 0, -3 << 5, 2515014.0], FP)
    d = np.array([1 << 3, 1 << 1, 1, 1 << 5, 1 << 2], np.uint64)
    r = O.scheme_divide_public_vec(kind, a, np.tile(d, (a.shape[0], 1)))
    got = opened(kind, r).view(np.int64) / 2.0**FP
    assert np.allclose(got, [1, 2, 3, -3, 628753.5], atol=1e-3)


@pytest.mark.parametrize("kind", ["tpc", "fpc"])
def test_divide_public_vec_matches_scalar(rng, kind):
    a = INIT[kind](rng.uniform(-100, 100, 64), FP)
    for d in (1, 3, 1 << 13, 1000003):
        dv = np.full((a.shape[0], 64), d, np.uint64)
        scalar = O.tpc_divide_public(a, d) if kind == "tpc" else O.fpc_divide_public(a, d)
        assert np.array_equal(O.scheme_divide_public_vec(kind, a, dv), scalar)
