"""The oracle's getPowers and taylorSeries-based sqrt / inverse / sigmoid
(RSS.inl:570-612, 1211-1320) pinned by the reference's own known answers
(Func.cu:499-615: Pow exactly, Sqrt within 5e-2, Inverse within 1e-1), by the
closed form floor(log2 x) of getPowers on random positive values, and by the
relative error of the range-reduced series over decades of input.  sigmoid has
no reference test; its coefficient functor uses the C library's float exp where
the reference uses its device's, so its parity is pinned only up to that."""
import numpy as np
import pytest

import oracle as O

FP = 26


def rec(a, fp=FP):
    return O.reconstruct(a)[0].view(np.int64) / float(1 << fp)


def test_pow_kat():
    # Func.cu:499-521: initializer list, then >>= FLOAT_PRECISION
    x = O.op_scalar(O.init_list(np.array([1, 2, 4, 8, 9, 15, 16], float), FP), FP, 5)
    pw = O.get_powers(x)
    assert all(pw[q].tolist() == [0, 1, 2, 3, 3, 3, 4] for q in range(3))


def test_pow_closed_form(rng):
    v = np.concatenate([rng.integers(1, 1 << 62, 200), [1, 2, 3, (1 << 62) + 5, (1 << 63) - 1]]).astype(np.uint64)
    x = np.zeros((3, 2, v.size), np.uint64)
    x[0, 0] = x[2, 1] = v  # the initializer-list sharing A(v, 0), C(0, v)
    pw = O.get_powers(x)
    assert all(pw[q].tolist() == [int(t).bit_length() - 1 for t in v.tolist()] for q in range(3))


def test_sqrt_kat():
    xs = [4e-2, 8e-2, 4e-1, 8.98e-1, 1.5, 1.75, 2.0, 2.25, 2.5, 2.75, 3.0, 3.25, 3.5]
    r = O.taylor(O.init_list(np.array(xs), FP), O.TAYLOR_SQRT, FP)
    assert np.allclose(rec(r), np.sqrt(xs), atol=5e-2)  # Func.cu:523-572


def test_inverse_kat():
    xs = [0.25, 0.3, 0.375, 0.49, 0.5, 0.51, 0.75, 1, 1.5]
    r = O.taylor(O.init_list(np.array(xs), FP), O.TAYLOR_INVERSE, FP)
    assert np.allclose(rec(r), 1 / np.array(xs), atol=1e-1)  # Func.cu:574-615


@pytest.mark.parametrize("fn,f,lo,hi,tol", [(O.TAYLOR_SQRT, np.sqrt, 0.05, 15.0, 0.03),
                                           (O.TAYLOR_INVERSE, lambda v: 1 / v, 0.1, 15.0, 0.03)])
def test_series_tracks_function(rng, fn, f, lo, hi, tol):
    # the reference's range reduction: with p = floor(log2 x) + 1 the series is
    # evaluated on x / 2^p in [0.5, 1) and rescaled by fn(2^(p - FP)); the
    # relative error stays bounded over decades of input, up to the range the
    # 32-bit `1 << p` of the coefficient functors allows (x * 2^FP < 2^30)
    v = np.exp(rng.uniform(np.log(lo), np.log(hi), 200))
    r = rec(O.taylor(O.init_list(v, FP), fn, FP))
    assert np.all(np.abs(r - f(v)) <= tol * np.abs(f(v)) + 1e-3), np.max(np.abs(r - f(v)) / np.abs(f(v)))
