"""Small-Cin convolutions (Cin % 32 != 0: the per-element im2col loader of the
limb GEMM) over geometries beyond the networks' own, against the oracle
(gpu::conv_fprop, conv.cuh:359-401; the RSS convolution, RSS.inl:705-803),
bit for bit: non-square filters, padding on one axis only, strides 1-4, Cin
1..40, tiles that span two or more images, and every tile shape / split-K
plan, including splits whose workgroups start in the middle of a segment.
Output channel counts of 65..96 take the 32x128 tile's helper form (three
column blocks, the fourth wave summing their s = 7 limb products), for both
im2col loaders (Cin % 32 != 0 and Cin % 32 == 0); counts up to 32 leave the
64x64 tile's second column of waves idle, for both loaders too."""
import os

import numpy as np
import pytest

import oracle as O
import piranha_amd as P
from conftest import rand_u64

pytestmark = pytest.mark.gpu

FP = 13

# b, H, W, Din, Dout, fh, fw, ph, pw, stride
GEOMS = [
    (3, 32, 32, 3, 64, 3, 3, 1, 1, 1),     # VGG conv1
    (2, 32, 32, 3, 96, 11, 11, 9, 9, 4),   # AlexNet-CIFAR conv1
    (3, 28, 28, 1, 20, 5, 5, 0, 0, 1),     # LeNet conv1
    (2, 12, 12, 20, 50, 5, 5, 0, 0, 1),    # LeNet conv2
    (4, 10, 13, 5, 33, 4, 3, 2, 1, 3),
    (2, 16, 16, 16, 70, 3, 3, 1, 1, 2),
    (4, 19, 20, 20, 130, 4, 5, 1, 0, 1),
    (1, 13, 4, 40, 7, 5, 1, 1, 0, 1),
    (3, 13, 7, 2, 9, 3, 5, 2, 2, 2),
    (2, 5, 5, 4, 10, 3, 3, 1, 1, 2),       # 3x3 output: a tile spans more than two images
    (2, 8, 8, 64, 96, 3, 3, 1, 1, 1),      # Cin % 32 == 0, three column blocks
    (3, 6, 7, 32, 72, 3, 2, 1, 0, 1),      # Cin % 32 == 0, N = 72: the third block partly past N
    (2, 8, 8, 32, 16, 3, 3, 1, 1, 1),      # Cin % 32 == 0, one column block (N <= 32)
]


@pytest.mark.parametrize("g", GEOMS)
def test_conv_fprop_geometry(rng, g):
    b, H, W, Din, Dout, fh, fw, ph, pw, stride = g
    OH, OW = (H + 2 * ph - fh) // stride + 1, (W + 2 * pw - fw) // stride + 1
    x, w = rand_u64(rng, b * H * W * Din), rand_u64(rng, Dout * fh * fw * Din)
    want = O.conv_fprop(x, w, b, H, W, Din, Dout, fh, fw, ph, pw, stride)
    dx, dw = P.to_dev(x), P.to_dev(w)
    for algo in (P.ALGO_LIMB, P.ALGO_PLAIN):
        dy = P.Buffer((b * OH * OW * Dout,), np.uint64)
        P.check(P.lib().pir_conv_fprop(dx.ptr, dw.ptr, dy.ptr, b, H, W, Din, Dout, fh, fw, ph, pw, stride, 1, algo,
                                       None))
        assert np.array_equal(dy.numpy(), want), f"algo {algo}"


@pytest.fixture
def plans():
    yield
    P.autotune_clear()
    P.lib().pir_autotune_set(0)


def general_shares(rng, n, bits=40):
    v = rng.integers(-(1 << bits), 1 << bits, size=(3, 2, n)).astype(np.int64).view(np.uint64)
    return np.ascontiguousarray(v)


def force_plan(tmp_path, wide, split):
    """Rewrites every entry of the tuning table to (wide, split); the 32x128
    tile only where the packed 64-column tiles pair up."""
    path = os.path.join(tmp_path, "t.tune")
    P.autotune_save(path)
    lines = []
    for line in open(path):
        key = line.split(":")[0].strip()
        n = int(key.split()[1])
        lines.append(f"{key} : {wide if ((n + 63) // 64) % 2 == 0 else 0} {split}\n")
    with open(path, "w") as f:
        f.writelines(lines)
    P.autotune_clear()
    P.autotune_load(path)


# b, H, W, f, Din, Dout, stride, pad (square, as pir_rss_convolution takes them)
RSS_GEOMS = [(3, 32, 32, 3, 3, 64, 1, 1), (2, 32, 32, 11, 3, 96, 4, 9), (2, 12, 12, 5, 20, 50, 1, 0),
             (2, 10, 10, 3, 5, 40, 2, 2), (2, 8, 8, 3, 32, 80, 1, 1), (2, 6, 6, 3, 32, 24, 1, 1),
             (3, 12, 12, 5, 1, 20, 1, 0)]


@pytest.mark.parametrize("g", RSS_GEOMS)
@pytest.mark.parametrize("wide,split", [(1, 1), (0, 1), (1, 3), (0, 5)])
def test_rss_convolution_forced_plans(rng, plans, tmp_path, g, wide, split):
    b, H, W, f, Din, Dout, stride, pad = g
    OH = (H + 2 * pad - f) // stride + 1
    n = b * OH * OH * Dout
    a, w = general_shares(rng, b * H * W * Din), general_shares(rng, Dout * f * f * Din)
    want = O.convolution(a, w, b, H, W, f, Din, Dout, stride, pad, FP)
    da, dw = P.to_dev(a), P.to_dev(w)
    P.set_mode(P.MODE_FUSED)
    P.autotune_clear()
    P.lib().pir_autotune_set(1)
    dc = P.Buffer((3, 2, n), np.uint64)
    P.check(P.lib().pir_rss_convolution(da.ptr, dw.ptr, dc.ptr, b, H, W, f, Din, Dout, stride, pad, FP,
                                        P.ALGO_LIMB, None))
    assert np.array_equal(dc.numpy(), want), "autotuned plan"
    P.lib().pir_autotune_set(0)
    assert P.autotune_entries() >= 1
    force_plan(str(tmp_path), wide, split)
    dc2 = P.Buffer((3, 2, n), np.uint64)
    P.check(P.lib().pir_rss_convolution(da.ptr, dw.ptr, dc2.ptr, b, H, W, f, Din, Dout, stride, pad, FP,
                                        P.ALGO_LIMB, None))
    assert np.array_equal(dc2.numpy(), want), f"plan wide={wide} split={split}"
