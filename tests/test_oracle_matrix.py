"""The oracle's gpu::transpose (matrix.cuh:43-59) and getMSBs / setMSBs
(bitwise.cuh:23-39): the reference's Transpose known answer (GPU.cu:247-261)
and numpy closed forms (a transpose of the column-major view; every
bit_width-th element)."""
import numpy as np
import pytest

import oracle as O
from conftest import rand_u64


def test_transpose_kat():
    assert O.transpose(np.array([1, 4, 2, 5, 3, 6], np.uint64), 2, 3).tolist() == [1, 2, 3, 4, 5, 6]


@pytest.mark.parametrize("rows,cols", [(1, 7), (7, 1), (5, 9), (64, 33)])
def test_transpose_closed_form(rng, rows, cols):
    a = rand_u64(rng, rows * cols)
    assert np.array_equal(O.transpose(a, rows, cols), a.reshape(cols, rows).T.ravel())


@pytest.mark.parametrize("bw", [1, 8, 64])
def test_msbs_closed_form(rng, bw):
    n = 37
    bits = rand_u64(rng, n * bw + 5)  # a ragged tail is ignored
    assert np.array_equal(O.get_msbs(bits, n, bw), bits[bw - 1:n * bw:bw])
    out = O.set_msbs(bits, 1, bw)
    want = bits.copy()
    want[bw - 1:(bits.size // bw) * bw:bw] = 1
    assert np.array_equal(out, want)


SHIFTS = np.array([0, 1, 7, 31, 32, 63, 64, 65, 127, 200, 1 << 32, 1 << 63, (1 << 64) - 1], np.uint64)


def clamped_shifts(x, y):
    """The reference's shift functors on its GPU (functors.cuh:46-59, 86-92):
    amounts >= 64 clamp, so << gives 0 and >> fills with the sign."""
    big = y >= 64
    ys = np.where(big, 0, y).astype(np.uint64)
    shl = np.where(big, np.uint64(0), x << ys)
    sar = (x.view(np.int64) >> ys.astype(np.int64)).view(np.uint64)
    fill = np.where(x.view(np.int64) < 0, np.uint64((1 << 64) - 1), np.uint64(0))
    return shl, np.where(big, fill, sar)


def test_shift_amounts_clamp(rng):
    n = SHIFTS.size * 4
    x = rand_u64(rng, 3, 2, n)
    y = np.tile(SHIFTS, (3, 4))
    shl, sar = clamped_shifts(x, y[:, None, :])
    assert np.array_equal(O.op_local(x, y, 6), shl)  # SHL on both shares of every party
    assert np.array_equal(O.op_local(x, y, 5), sar)  # SHR (arithmetic)
