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
