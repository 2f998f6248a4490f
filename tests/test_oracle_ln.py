"""The oracle's LayerNorm layer (LNLayer.cu:66-284) pinned by the reference's
own known-answer tests (Layer.cu:838-1000: LNForwardBasic, LNForwardBasic2,
LNForward, LNBackwardBasic, LNBackward with its updated gamma and beta), at the
tolerances those tests use, and the revealed inverse square root
(inv_sqrt_fixed_point_functor, functors.h:92-103) against numpy's IEEE sqrt
and division."""
import os

import numpy as np
import pytest

import oracle as O

FP = 26  # FLOAT_PRECISION of the reference tests (globals.h:19-21)
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def rec(a, fp=FP):
    return O.reconstruct(a)[0].view(np.int64) / float(1 << fp)


def params(n):
    return O.set_public(np.ones(n), FP), O.set_public(np.zeros(n), FP)  # LNLayer.cu:28-35


def test_forward_basic():
    x = O.set_public(np.array([1, 2, 3, 4, -1, -2, 0, 1], float), FP)
    act, _, _ = O.ln_forward(x, *params(8), 1, FP)
    want = [0, 0.534522, 1.069044, 1.603567, -1.069044, -1.603567, -0.534522, 0]
    assert np.allclose(rec(act), want, atol=1e-2)  # Layer.cu:838-859


def test_forward_basic2():
    x = O.set_public(np.array([1, 2, 3, 4, -1, -2, 0, 1, 4, 2, 3, 1, -1, 1, 0, -2], float), FP)
    act, _, _ = O.ln_forward(x, *params(8), 2, FP)
    want = [0, 0.534522, 1.069044, 1.603567, -1.069044, -1.603567, -0.534522, 0,
            1.603567, 0.534522, 1.069044, 0, -1.069044, 0, -0.534522, -1.603567]
    assert np.allclose(rec(act), want, atol=1e-2)  # Layer.cu:861-885


def test_forward_fixture(gold):
    x = O.set_public(gold["lnfwd_input"], FP)
    act, _, _ = O.ln_forward(x, *params(48), 2, FP)
    assert np.allclose(rec(act), gold["lnfwd_expected"], atol=1e-2)  # Layer.cu:887-912


def test_backward_basic():
    x = O.set_public(np.array([1, 2, 3, 4, -1, -2, 0, 1], float), FP)
    g, b = params(8)
    _, xhat, inv = O.ln_forward(x, g, b, 1, FP)
    d = O.set_public(np.array([-0.1901, 0.9232, 0.0208, -0.5175, -0.4921, 0.7767, -0.0701, 0.5335]), FP)
    deltas, _, _, _ = O.ln_backward(d, xhat, inv, g, b, 1, FP, 5)
    want = [-0.16738572, 0.46327254, 0.01649384, -0.23566524, -0.39996027, 0.24266748, -0.1388174, 0.21939475]
    assert np.allclose(rec(deltas), want, atol=1e-2)  # Layer.cu:914-941


def test_backward_fixture(gold):
    # Layer.cu:943-1000 with log_learning_rate = 3
    x = O.set_public(gold["lnbwd_input"], FP)
    g, b = params(48)
    _, xhat, inv = O.ln_forward(x, g, b, 2, FP)
    d = O.set_public(gold["lnbwd_delta"], FP)
    deltas, _, g2, b2 = O.ln_backward(d, xhat, inv, g, b, 2, FP, 3)
    assert np.allclose(rec(deltas), gold["lnbwd_expected"], atol=1e-2)
    assert np.allclose(rec(g2), gold["lnbwd_gamma"], atol=1e-3)  # ASSERT_EPSILON, util.cuh:19
    assert np.allclose(rec(b2), gold["lnbwd_beta"], atol=1e-3)


def test_backward_leaves_inputs_alone(gold):
    x = O.set_public(gold["lnbwd_input"], FP)
    g, b = params(48)
    _, xhat, inv = O.ln_forward(x, g, b, 2, FP)
    keep = [a.copy() for a in (xhat, g, b)]
    O.ln_backward(O.set_public(gold["lnbwd_delta"], FP), xhat, inv, g, b, 2, FP, 3)
    for a, k in zip((xhat, g, b), keep):
        assert np.array_equal(a, k)


@pytest.mark.parametrize("fp", [13, 20, 26])
def test_inv_sqrt_fixed(fp):
    rng = np.random.default_rng(fp)
    xs = np.concatenate([rng.integers(1, 1 << 40, 2000), [1, 2, 3, 1 << fp, (1 << fp) - 1, (1 << 62) + 12345]])
    for x in xs.tolist():
        v = np.float64(np.int64(x)) / np.float64(1 << fp)
        f = (np.float64(1.0) / np.sqrt(v)) * np.float64(1 << fp)
        want = 2**64 - 1 if f >= 2.0**64 else int(f)
        assert O.inv_sqrt_fixed(x, fp) == want, x
    assert O.inv_sqrt_fixed(0, fp) == 2**64 - 1  # 1 / sqrt(0) = +inf clamps
    assert O.inv_sqrt_fixed((1 << 64) - 5, fp) == 0  # sqrt of a negative: NaN -> 0
