"""CPU oracle for the 3-party RSS secure-inference path (test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  It wraps ``oracle/build/libpir_ref.so``,
a literal per-party C++ restatement of the reference's RSS protocol
(src/mpc/RSS.inl, src/gpu/*.cuh, src/nn/*Layer.cu), and is the thing the HIP
path is checked against -- never the thing measured or shipped.

Array conventions: an RSS value for all three parties is a numpy array of
shape (3, 2, n) (party, share, element); a party-local value is (3, n).  A
TPC value (two parties, src/mpc/TPC.inl) is (2, n); an FPC value (four
parties, three shares each, src/mpc/FPC.inl) is (4, 3, n).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libpir_ref.so")

ADD, SUB, MUL, XOR, AND, SHR, SHL, DIV = range(8)

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_dp = ctypes.POINTER(ctypes.c_double)
_sz = ctypes.c_size_t
_i = ctypes.c_int
_u64 = ctypes.c_uint64

_SIGS = {
    "ref_gemm": [_i, _i, _i, _u64p, _i, _u64p, _i, _u64p, _i],
    "ref_conv_fprop": [_u64p, _u64p, _u64p] + [_i] * 11,
    "ref_bitexpand_u8": [_u64p, _sz, _u8p],
    "ref_bitexpand_u64": [_u64p, _sz, _u64p],
    "ref_set_carry_out_msb_u8": [_u8p, _u8p, _u8p, _sz, _i, _i],
    "ref_expand_compare_u8": [_u8p, _u8p, _u8p, _sz, _sz],
    "ref_element_vector_add": [_u64p, _u64p, _i, _i, _i],
    "ref_truncate_cols_u8": [_u8p, _sz, _u8p, _i, _i],
    "ref_stride_pad": [_u64p, _u64p, _sz, _i, _i, _u64],
    "ref_expanded_pool_size": [_i],
    "ref_maxpool_im2row": [_u64p, _u64p, _i, _i, _i, _i, _i, _i, _i, _u64],
    "ref_rss_set_public": [_u64p, _sz, _dp, _i],
    "ref_rss_init_list": [_u64p, _sz, _dp, _i, _i],
    "ref_rss_fill_u64": [_u64p, _sz, _u64],
    "ref_rss_fill_u8": [_u8p, _sz, ctypes.c_uint8],
    "ref_reconstruct_u64": [_u64p, _sz, _u64p],
    "ref_reconstruct_u8": [_u8p, _sz, _u8p],
    "ref_reshare_u64": [_u64p, _sz, _u64p],
    "ref_reconstruct3out3": [_u64p, _sz, _u64p],
    "ref_rss_mult": [_u64p, _u64p, _sz],
    "ref_rss_and_u8": [_u8p, _u8p, _sz],
    "ref_rss_and_u64": [_u64p, _u64p, _sz],
    "ref_rss_op_rss": [_u64p, _u64p, _sz, _i],
    "ref_rss_op_local": [_u64p, _u64p, _sz, _i],
    "ref_rss_op_scalar": [_u64p, _sz, _u64, _i],
    "ref_divide_public": [_u64p, _sz, _u64],
    "ref_divide_public_vec": [_u64p, _sz, _u64p],
    "ref_divide_public_3out3": [_u64p, _sz, _u64, _u64p],
    "ref_scalar_divide": [_u64, _u64],
    "ref_matmul": [_u64p, _sz, _u64p, _sz, _u64p] + [_i] * 7,
    "ref_local_matmul": [_u64p, _sz, _u64p, _sz, _u64p] + [_i] * 6,
    "ref_convolution": [_u64p, _sz, _u64p, _sz, _u64p, _sz] + [_i] * 9,
    "ref_drelu_u8": [_u64p, _sz, _u8p],
    "ref_drelu_u64": [_u64p, _sz, _u64p],
    "ref_relu_u8": [_u64p, _sz, _u64p, _u8p],
    "ref_relu_u64": [_u64p, _sz, _u64p, _u64p],
    "ref_select_share_u8": [_u64p, _u64p, _u8p, _u64p, _sz],
    "ref_select_share_u64": [_u64p, _u64p, _u64p, _u64p, _sz],
    "ref_maxpool_u8": [_u64p, _sz, _u64p, _u8p, _i],
    "ref_maxpool_u64": [_u64p, _sz, _u64p, _u64p, _i],
    "ref_fc_forward": [_u64p, _u64p, _u64p, _u64p, _i, _i, _i, _i],
    "ref_cnn_forward": [_u64p, _u64p, _u64p] + [_i] * 9,
    "ref_relu_forward": [_u64p, _sz, _u64p, _u8p],
    "ref_maxpool_forward": [_u64p, _u64p, _u8p] + [_i] * 7,
    "ref_tpc_init_list": [_u64p, _sz, _dp, _i, _i],
    "ref_tpc_reconstruct": [_u64p, _sz, _u64p],
    "ref_tpc_divide_public": [_u64p, _sz, _u64],
    "ref_tpc_matmul": [_u64p, _sz, _u64p, _sz, _u64p] + [_i] * 7,
    "ref_tpc_convolution": [_u64p, _sz, _u64p, _sz, _u64p, _sz] + [_i] * 9,
    "ref_fpc_init_list": [_u64p, _sz, _dp, _i, _i],
    "ref_fpc_reconstruct": [_u64p, _sz, _u64p],
    "ref_fpc_divide_public": [_u64p, _sz, _u64],
    "ref_fpc_reshare": [_u64p, _u64p, _sz, _i, _i, _u64p],
    "ref_fpc_matmul": [_u64p, _sz, _u64p, _sz, _u64p] + [_i] * 7,
    "ref_fpc_convolution": [_u64p, _sz, _u64p, _sz, _u64p, _sz] + [_i] * 9,
    "ref_conv_dgrad": [_u64p, _u64p, _u64p] + [_i] * 13,
    "ref_conv_wgrad": [_u64p, _u64p, _u64p] + [_i] * 13,
    "ref_reduce_sum": [_u64p, _u64p, _i, _i, _i],
    "ref_vector_expand": [_u64p, _sz, _u64p, _sz],
    "ref_maxpool_row2im": [_u64p, _sz, _u64p, _sz] + [_i] * 6,
    "ref_convolution_grad": [_i, _u64p, _sz, _u64p, _sz, _u64p, _sz] + [_i] * 9,
    "ref_fc_backward": [_u64p] * 5 + [_i] * 5,
    "ref_cnn_backward": [_u64p] * 4 + [_i] * 10,
    "ref_relu_backward": [_u64p, _u8p, _sz, _u64p],
    "ref_maxpool_backward": [_u64p, _u8p, _u64p] + [_i] * 6,
    "ref_reveal_softmax_grad": [_u64p, _dp, _sz, _u64p, _i, _i, _i],
    "ref_averagepool_im2row": [_u64p, _u64p] + [_i] * 7,
    "ref_averagepool_expand_delta": [_u64p, _sz, _u64p, _sz, _i, _i],
    "ref_avgpool_forward": [_u64p, _u64p] + [_i] * 6,
    "ref_avgpool_backward": [_u64p, _sz, _u64p, _sz, _i, _i],
    "ref_ln_forward": [_u64p] * 6 + [_sz, _i, _i],
    "ref_ln_backward": [_u64p] * 6 + [_sz, _i, _i, _i],
    "ref_inv_sqrt_fixed": [_u64, _i],
    "ref_calc_fn": [_i, _u64, _i],
    "ref_transpose_u64": [_u64p, _u64p, _i, _i],
    "ref_get_powers": [_u64p, _sz, _u64p],
    "ref_taylor": [_u64p, _sz, _u64p, _i, _i],
    "ref_tpc_mult": [_u64p, _u64p, _sz],
    "ref_tpc_divide_public_vec": [_u64p, _sz, _u64p],
    "ref_fpc_divide_public_vec": [_u64p, _sz, _u64p],
    "ref_tpc_get_powers": [_u64p, _sz, _u64p],
    "ref_fpc_get_powers": [_u64p, _sz, _u64p],
    "ref_tpc_taylor_series": [_u64p, _sz, _u64p, ctypes.c_double, ctypes.c_double, ctypes.c_double, _i, _i],
    "ref_fpc_taylor_series": [_u64p, _sz, _u64p, ctypes.c_double, ctypes.c_double, ctypes.c_double, _i, _i],
    "ref_fpc_mult": [_u64p, _u64p, _sz],
    "ref_fpc_and": [ctypes.c_void_p, ctypes.c_void_p, _sz, _i],
    "ref_fpc_drelu": [_u64p, _sz, ctypes.c_void_p, _i],
    "ref_fpc_select_share": [_u64p, _u64p, ctypes.c_void_p, _u64p, _sz, _i],
    "ref_fpc_relu": [_u64p, _sz, _u64p, ctypes.c_void_p, _i],
    "ref_fpc_maxpool": [_u64p, _sz, _u64p, ctypes.c_void_p, _i, _i],
    "ref_tpc_and": [ctypes.c_void_p, ctypes.c_void_p, _sz, _i],
    "ref_tpc_drelu": [_u64p, _sz, ctypes.c_void_p, _i],
    "ref_tpc_select_share": [_u64p, _u64p, ctypes.c_void_p, _u64p, _sz, _i],
    "ref_tpc_relu": [_u64p, _sz, _u64p, ctypes.c_void_p, _i],
    "ref_tpc_maxpool": [_u64p, _sz, _u64p, ctypes.c_void_p, _i, _i],
    "ref_taylor_series": [_u64p, _sz, _u64p, ctypes.c_double, ctypes.c_double, ctypes.c_double, _i, _i],
    "ref_get_msbs_u64": [_u64p, _u64p, _sz, _sz],
    "ref_set_msbs_u64": [_u64p, _sz, _u64, _sz],
}
_u32 = ctypes.c_uint32
_u32p = ctypes.POINTER(ctypes.c_uint32)
_M = [_u64, _u32]  # seed, call of the masked precompute mode
_SIGS.update({
    "ref_philox": [_u32p, _u32p, _sz, _u32p],
    "ref_prg_fill": [_u64, _u32, _sz, _u64p],
    "ref_deal": [_u64p, _sz, _u64, _u32, _u64p],
    "ref_reshare_m": [_u64p, _sz, _u64p] + _M,
    "ref_rss_mult_m": [_u64p, _u64p, _sz] + _M,
    "ref_rss_and_m_u8": [_u8p, _u8p, _sz] + _M,
    "ref_rss_and_m_u64": [_u64p, _u64p, _sz] + _M,
    "ref_divide_public_m": [_u64p, _sz, _i] + _M,
    "ref_divide_public_3out3_m": [_u64p, _sz, _i, _u64p] + _M,
    "ref_matmul_m": [_u64p, _sz, _u64p, _sz, _u64p] + [_i] * 7 + _M,
    "ref_convolution_m": [_u64p, _sz, _u64p, _sz, _u64p, _sz] + [_i] * 9 + _M,
    "ref_drelu_m_u8": [_u64p, _sz, _u8p] + _M,
    "ref_drelu_m_u64": [_u64p, _sz, _u64p] + _M,
    "ref_select_share_m_u8": [_u64p, _u64p, _u8p, _u64p, _sz] + _M,
    "ref_select_share_m_u64": [_u64p, _u64p, _u64p, _u64p, _sz] + _M,
    "ref_relu_m_u8": [_u64p, _sz, _u64p, _u8p] + _M,
    "ref_relu_m_u64": [_u64p, _sz, _u64p, _u64p] + _M,
    "ref_maxpool_m_u8": [_u64p, _sz, _u64p, _u8p, _i] + _M,
    "ref_maxpool_m_u64": [_u64p, _sz, _u64p, _u64p, _i] + _M,
    "ref_fc_forward_m": [_u64p, _u64p, _u64p, _u64p, _i, _i, _i, _i] + _M,
    "ref_cnn_forward_m": [_u64p, _u64p, _u64p] + [_i] * 9 + _M,
    "ref_maxpool_forward_m": [_u64p, _u64p, _u8p] + [_i] * 7 + _M,
})

_RET = {"ref_expanded_pool_size": _i, "ref_maxpool_im2row": _sz, "ref_averagepool_im2row": _sz,
        "ref_inv_sqrt_fixed": _u64, "ref_calc_fn": _u64,
        "ref_scalar_divide": _u64, "ref_version": _i}

_lib = None


def build():
    """Compile the oracle library with its Makefile (g++, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RET.get(name, None)
        L.ref_version.restype = _i
        _lib = L
    return _lib


def _p(a):
    """ctypes pointer to a C-contiguous numpy array (uint64, uint32, uint8 or float64)."""
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be contiguous"
    if a.dtype == np.uint64:
        return a.ctypes.data_as(_u64p)
    if a.dtype == np.uint32:
        return a.ctypes.data_as(_u32p)
    if a.dtype == np.uint8:
        return a.ctypes.data_as(_u8p)
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    raise TypeError(a.dtype)


def _u(a, dtype=np.uint64):
    return np.ascontiguousarray(a, dtype=dtype)


# --- gpu:: level ------------------------------------------------------------

def gemm(M, N, K, A, ta, B, tb, tc):
    A, B = _u(A), _u(B)
    C = np.zeros(M * N, np.uint64)
    lib().ref_gemm(M, N, K, _p(A), int(ta), _p(B), int(tb), _p(C), int(tc))
    return C


def conv_fprop(x, w, b, H, W, Din, Dout, fh, fw, ph, pw, stride, dil=1):
    x, w = _u(x), _u(w)
    OH = (H + 2 * ph - ((fh - 1) * dil + 1)) // stride + 1
    OW = (W + 2 * pw - ((fw - 1) * dil + 1)) // stride + 1
    y = np.zeros(b * OH * OW * Dout, np.uint64)
    lib().ref_conv_fprop(_p(x), _p(w), _p(y), b, H, W, Din, Dout, fh, fw, ph, pw, stride, dil)
    return y


def bitexpand(a, out_dtype=np.uint8):
    a = _u(a)
    b = np.zeros(a.size * 64, out_dtype)
    if out_dtype == np.uint8:
        lib().ref_bitexpand_u8(_p(a), a.size, _p(b))
    else:
        lib().ref_bitexpand_u64(_p(a), a.size, _p(b))
    return b


def expanded_pool_size(f):
    return lib().ref_expanded_pool_size(f)


def maxpool_im2row(im, H, W, f, Din, batch, stride, padding, pad_val):
    im = _u(im)
    n = lib().ref_maxpool_im2row(_p(im), ctypes.cast(None, _u64p), H, W, f, Din, batch, stride, padding,
                                 int(pad_val) & 0xFFFFFFFFFFFFFFFF)
    out = np.zeros(n, np.uint64)
    lib().ref_maxpool_im2row(_p(im), _p(out), H, W, f, Din, batch, stride, padding,
                             int(pad_val) & 0xFFFFFFFFFFFFFFFF)
    return out


def stride_pad(a, stride, pad, pad_val):
    a = _u(a)
    out = np.zeros(a.size // stride * (stride + pad), np.uint64)
    lib().ref_stride_pad(_p(a), _p(out), a.size, stride, pad, int(pad_val) & 0xFFFFFFFFFFFFFFFF)
    return out


def truncate_cols(a, start_cols, end_cols):
    a = _u(a, np.uint8)
    b = np.zeros(a.size // start_cols * end_cols, np.uint8)
    lib().ref_truncate_cols_u8(_p(a), a.size, _p(b), start_cols, end_cols)
    return b


def element_vector_add(a, b, rowwise, rows, cols):
    a = _u(a).copy()
    b = _u(b)
    lib().ref_element_vector_add(_p(a), _p(b), int(rowwise), rows, cols)
    return a


def expand_compare(b, inv, out_n):
    b, inv = _u(b, np.uint8), _u(inv, np.uint8)
    out = np.zeros(out_n, np.uint8)
    lib().ref_expand_compare_u8(_p(b), _p(inv), _p(out), b.size, out_n)
    return out


def set_carry_out_msb(rbits, abits, n, bit_width, xor_msb):
    rbits, abits = _u(rbits, np.uint8).copy(), _u(abits, np.uint8).copy()
    msb = np.zeros(n, np.uint8)
    lib().ref_set_carry_out_msb_u8(_p(rbits), _p(abits), _p(msb), n, bit_width, int(xor_msb))
    return rbits, abits, msb


# --- shares ---------------------------------------------------------------

def set_public(vals, fp):
    """RSS setPublic for all parties (RSS.inl:122-142), returns (3, 2, n) uint64."""
    vals = np.ascontiguousarray(vals, np.float64).ravel()
    x = np.zeros((3, 2, vals.size), np.uint64)
    lib().ref_rss_set_public(_p(x), vals.size, _p(vals), fp)
    return x


def init_list(vals, fp, convert=True):
    """RSS initializer-list constructor (RSS.inl:388-414)."""
    vals = np.ascontiguousarray(vals, np.float64).ravel()
    x = np.zeros((3, 2, vals.size), np.uint64)
    lib().ref_rss_init_list(_p(x), vals.size, _p(vals), int(convert), fp)
    return x


def fill(n, v, dtype=np.uint64):
    x = np.zeros((3, 2, n), dtype)
    if dtype == np.uint8:
        lib().ref_rss_fill_u8(_p(x), n, v)
    else:
        lib().ref_rss_fill_u64(_p(x), n, int(v) & 0xFFFFFFFFFFFFFFFF)
    return x


def reconstruct(x):
    x = np.ascontiguousarray(x)
    n = x.shape[-1]
    out = np.zeros((3, n), x.dtype)
    if x.dtype == np.uint8:
        lib().ref_reconstruct_u8(_p(x), n, _p(out))
    else:
        lib().ref_reconstruct_u64(_p(x), n, _p(out))
    return out


def reshare(c):
    c = _u(c)
    n = c.shape[-1]
    out = np.zeros((3, 2, n), np.uint64)
    lib().ref_reshare_u64(_p(c), n, _p(out))
    return out


def reconstruct3out3(a):
    a = _u(a)
    out = np.zeros_like(a)
    lib().ref_reconstruct3out3(_p(a), a.shape[-1], _p(out))
    return out


def mult(x, y):
    x, y = _u(x).copy(), _u(y)
    lib().ref_rss_mult(_p(x), _p(y), x.shape[-1])
    return x


def rss_and(x, y):
    x = np.ascontiguousarray(x).copy()
    y = np.ascontiguousarray(y, x.dtype)
    if x.dtype == np.uint8:
        lib().ref_rss_and_u8(_p(x), _p(y), x.shape[-1])
    else:
        lib().ref_rss_and_u64(_p(x), _p(y), x.shape[-1])
    return x


def op_rss(x, y, op):
    x, y = _u(x).copy(), _u(y)
    lib().ref_rss_op_rss(_p(x), _p(y), x.shape[-1], op)
    return x


def op_local(x, y, op):
    x, y = _u(x).copy(), _u(y)
    lib().ref_rss_op_local(_p(x), _p(y), x.shape[-1], op)
    return x


def op_scalar(x, v, op):
    x = _u(x).copy()
    lib().ref_rss_op_scalar(_p(x), x.shape[-1], int(v) & 0xFFFFFFFFFFFFFFFF, op)
    return x


def divide_public(a, d):
    a = _u(a).copy()
    lib().ref_divide_public(_p(a), a.shape[-1], int(d))
    return a


def divide_public_vec(a, d):
    a, d = _u(a).copy(), _u(d)
    lib().ref_divide_public_vec(_p(a), a.shape[-1], _p(d))
    return a


def divide_public_3out3(a, d):
    a = _u(a).copy()
    out = np.zeros((3, 2, a.shape[-1]), np.uint64)
    lib().ref_divide_public_3out3(_p(a), a.shape[-1], int(d), _p(out))
    return out


def scalar_divide(x, d):
    return lib().ref_scalar_divide(int(x) & 0xFFFFFFFFFFFFFFFF, int(d) & 0xFFFFFFFFFFFFFFFF)


def matmul(a, b, M, N, K, ta, tb, tc, trunc):
    a, b = _u(a), _u(b)
    c = np.zeros((3, 2, M * N), np.uint64)
    lib().ref_matmul(_p(a), a.shape[-1], _p(b), b.shape[-1], _p(c), M, N, K, int(ta), int(tb), int(tc), trunc)
    return c


def local_matmul(a, b, M, N, K, ta, tb, tc):
    a, b = _u(a), _u(b)
    c = np.zeros((3, M * N), np.uint64)
    lib().ref_local_matmul(_p(a), a.shape[-1], _p(b), b.shape[-1], _p(c), M, N, K, int(ta), int(tb), int(tc))
    return c


def convolution(A, B, batch, H, W, f, Din, Dout, stride, pad, trunc):
    A, B = _u(A), _u(B)
    OH = (H + 2 * pad - f) // stride + 1
    OW = (W + 2 * pad - f) // stride + 1
    n = batch * OH * OW * Dout
    C = np.zeros((3, 2, n), np.uint64)
    lib().ref_convolution(_p(A), A.shape[-1], _p(B), B.shape[-1], _p(C), n, batch, H, W, f, Din, Dout, stride, pad,
                          trunc)
    return C


def drelu(x, out_dtype=np.uint8):
    x = _u(x)
    n = x.shape[-1]
    out = np.zeros((3, 2, n), out_dtype)
    (lib().ref_drelu_u8 if out_dtype == np.uint8 else lib().ref_drelu_u64)(_p(x), n, _p(out))
    return out


def relu(x, out_dtype=np.uint8):
    x = _u(x)
    n = x.shape[-1]
    out = np.zeros((3, 2, n), np.uint64)
    dout = np.zeros((3, 2, n), out_dtype)
    (lib().ref_relu_u8 if out_dtype == np.uint8 else lib().ref_relu_u64)(_p(x), n, _p(out), _p(dout))
    return out, dout


def select_share(x, y, b):
    x, y = _u(x), _u(y)
    b = np.ascontiguousarray(b)
    n = x.shape[-1]
    z = np.zeros((3, 2, n), np.uint64)
    (lib().ref_select_share_u8 if b.dtype == np.uint8 else lib().ref_select_share_u64)(_p(x), _p(y), _p(b), _p(z), n)
    return z


def maxpool(pools, k, out_dtype=np.uint8):
    """maxpool over gathered pools (RSS.inl:867-1017); returns (result, dresult, pools_after)."""
    pools = _u(pools).copy()
    n = pools.shape[-1]
    result = np.zeros((3, 2, n // k), np.uint64)
    dres = np.zeros((3, 2, n), out_dtype)
    (lib().ref_maxpool_u8 if out_dtype == np.uint8 else lib().ref_maxpool_u64)(_p(pools), n, _p(result), _p(dres), k)
    return result, dres, pools


# --- layers -----------------------------------------------------------------

def fc_forward(x, W, b, batch, d_in, d_out, fp):
    x, W, b = _u(x), _u(W), _u(b)
    act = np.zeros((3, 2, batch * d_out), np.uint64)
    lib().ref_fc_forward(_p(x), _p(W), _p(b), _p(act), batch, d_in, d_out, fp)
    return act


def cnn_forward(x, W, batch, H, Wd, Din, Dout, f, stride, pad, fp):
    x, W = _u(x), _u(W)
    OH = (H - f + 2 * pad) // stride + 1
    OW = (Wd - f + 2 * pad) // stride + 1
    act = np.zeros((3, 2, batch * OH * OW * Dout), np.uint64)
    lib().ref_cnn_forward(_p(x), _p(W), _p(act), batch, H, Wd, Din, Dout, f, stride, pad, fp)
    return act


def relu_forward(x):
    x = _u(x)
    n = x.shape[-1]
    act = np.zeros((3, 2, n), np.uint64)
    rp = np.zeros((3, 2, n), np.uint8)
    lib().ref_relu_forward(_p(x), n, _p(act), _p(rp))
    return act, rp


def maxpool_forward(x, batch, H, W, C, pool, stride, fp):
    x = _u(x)
    OH = (H - pool) // stride + 1
    OW = (W - pool) // stride + 1
    n = batch * C * OH * OW
    act = np.zeros((3, 2, n), np.uint64)
    mp = np.zeros((3, 2, n * pool * pool), np.uint8)
    lib().ref_maxpool_forward(_p(x), _p(act), _p(mp), batch, H, W, C, pool, stride, fp)
    return act, mp


def averagepool_im2row(im, imageWidth, imageHeight, f, Din, batch, stride, padding=0):
    """gpu::averagepool_im2row (convolution.cuh:339-371) of one share array."""
    im = np.ascontiguousarray(im, np.uint64)
    n = lib().ref_averagepool_im2row(_p(im), ctypes.cast(None, _u64p), imageWidth, imageHeight, f, Din, batch, stride, padding)
    out = np.zeros(n, np.uint64)
    lib().ref_averagepool_im2row(_p(im), _p(out), imageWidth, imageHeight, f, Din, batch, stride, padding)
    return out


def averagepool_expand_delta(x, out_n, Din, pool_size):
    """gpu::averagepool_expand_delta (convolution.cuh:194-207, 412-437) into a
    zeroed array of out_n entries."""
    x = np.ascontiguousarray(x, np.uint64)
    out = np.zeros(out_n, np.uint64)
    lib().ref_averagepool_expand_delta(_p(x), x.size, _p(out), out_n, Din, pool_size)
    return out


def avgpool_forward(x, batch, H, W, C, pool, stride):
    """AveragepoolLayer::forward (AveragepoolLayer.cu:49-96)."""
    x = _u(x)
    OH = (H - pool) // stride + 1
    OW = (W - pool) // stride + 1
    act = np.zeros((3, 2, batch * C * OH * OW), np.uint64)
    lib().ref_avgpool_forward(_p(x), _p(act), batch, H, W, C, pool, stride)
    return act


def avgpool_backward(delta, batch, H, W, C, pool):
    """AveragepoolLayer::backward (AveragepoolLayer.cu:98-125)."""
    delta = _u(delta)
    deltas = np.zeros((3, 2, batch * H * W * C), np.uint64)
    lib().ref_avgpool_backward(_p(delta), delta.shape[-1], _p(deltas), deltas.shape[-1], C, pool)
    return deltas


def ln_forward(x, gamma, beta, num_batches, fp):
    """LNLayer::forward (LNLayer.cu:66-162).  Returns (activations, xhat, invSigma)."""
    x, gamma, beta = _u(x), _u(gamma), _u(beta)
    n = x.shape[-1]
    act = np.zeros((3, 2, n), np.uint64)
    xhat = np.zeros((3, 2, n), np.uint64)
    inv_sigma = np.zeros((3, 2, num_batches), np.uint64)
    lib().ref_ln_forward(_p(x), _p(gamma), _p(beta), _p(xhat), _p(inv_sigma), _p(act), n, num_batches, fp)
    return act, xhat, inv_sigma


def ln_backward(delta, xhat, inv_sigma, gamma, beta, num_batches, fp, lr):
    """LNLayer::backward (LNLayer.cu:166-284).  Returns (deltas, xhat after the
    pass, updated gamma, updated beta)."""
    delta, inv_sigma = _u(delta), _u(inv_sigma)
    xhat, gamma, beta = _u(xhat).copy(), _u(gamma).copy(), _u(beta).copy()
    n = delta.shape[-1]
    deltas = np.zeros((3, 2, n), np.uint64)
    lib().ref_ln_backward(_p(delta), _p(xhat), _p(inv_sigma), _p(gamma), _p(beta), _p(deltas), n, num_batches, fp, lr)
    return deltas, xhat, gamma, beta


def transpose(a, rows, cols):
    """gpu::transpose (matrix.cuh:43-59) of one array."""
    a = np.ascontiguousarray(a, np.uint64)
    b = np.zeros(rows * cols, np.uint64)
    lib().ref_transpose_u64(_p(a), _p(b), rows, cols)
    return b


def get_msbs(bits, n, bit_width):
    bits = np.ascontiguousarray(bits, np.uint64)
    out = np.zeros(n, np.uint64)
    lib().ref_get_msbs_u64(_p(bits), _p(out), n, bit_width)
    return out


def set_msbs(bits, val, bit_width):
    bits = np.ascontiguousarray(bits, np.uint64).copy()
    lib().ref_set_msbs_u64(_p(bits), bits.size, int(val), bit_width)
    return bits


def get_powers(x):
    """getPowers (RSS.inl:1211-1235): opened [3][n] floor(log2) of each value."""
    x = _u(x)
    n = x.shape[-1]
    pw = np.zeros((3, n), np.uint64)
    lib().ref_get_powers(_p(x), n, _p(pw))
    return pw


TAYLOR_SQRT, TAYLOR_INVERSE, TAYLOR_SIGMOID = 0, 1, 2


def taylor_series(x, a0, a1, a2, fn, fp):
    """taylorSeries (RSS.inl:1237-1320) with explicit coefficients."""
    x = _u(x)
    out = np.zeros_like(x)
    lib().ref_taylor_series(_p(x), x.shape[-1], _p(out), a0, a1, a2, fn, fp)
    return out


def taylor(x, fn, fp):
    """sqrt / inverse / sigmoid of RSS shares (RSS.inl:570-612, taylorSeries :1237-1320)."""
    x = _u(x)
    out = np.zeros_like(x)
    lib().ref_taylor(_p(x), x.shape[-1], _p(out), fn, fp)
    return out


def calc_fn(fn, x, fp):
    """One entry of taylorSeries's coefficient table, calc_fn (functors.h:22-31):
    fn 0 sqrt, 1 inverse, 2 sigmoid of the float 2^(x - fp), times 1 << x."""
    return lib().ref_calc_fn(fn, int(x), fp)


def inv_sqrt_fixed(x, fp):
    """inv_sqrt_fixed_point_functor (functors.h:92-103) of one ring element."""
    return lib().ref_inv_sqrt_fixed(int(x) & 0xFFFFFFFFFFFFFFFF, fp)


# --- ResLayer (ResLayer.cu) as a composition of the oracle's layers ---------

def res_structure(H, W, in_planes, planes, num_blocks, stride, expansion):
    """ResLayer's inner layers (ResLayer.cu:23-103): per block a (main,
    shortcut) pair of symbolic layer lists; the shortcut is empty when the
    block input is added directly."""
    def block(ih, iw, inp, s):
        oh, ow = (ih + 2 - 3) // s + 1, (iw + 2 - 3) // s + 1
        n = oh * ow * planes
        main = [("cnn", ih, iw, inp, planes, 3, s, 1), ("ln", n), ("relu", n),
                ("cnn", oh, ow, planes, planes, 3, 1, 1), ("ln", n), ("relu", n)]
        sc = []
        if s != 1 or inp != expansion * planes:
            sc = [("cnn", ih, iw, inp, expansion * planes, 1, s, 0), ("ln", oh * ow * expansion * planes)]
        return main, sc, oh, ow

    out = []
    m, sc, ih, iw = block(H, W, in_planes, stride)
    out.append((m, sc))
    for _ in range(1, num_blocks):
        m, sc, ih, iw = block(ih, iw, planes * expansion, 1)
        out.append((m, sc))
    return out


def _n_params(l):
    return {"cnn": 1, "ln": 2}.get(l[0], 0)


def _split_params(structure, flat):
    """The flat snapshot-order list (ResLayer.cu:134-167: each block's layers,
    then its shortcut) -> per block (main, shortcut) lists of per-layer lists."""
    flat = list(flat)
    out, k = [], 0
    for main, sc in structure:
        parts = []
        for layers in (main, sc):
            per = []
            for l in layers:
                per.append(flat[k:k + _n_params(l)])
                k += _n_params(l)
            parts.append(per)
        out.append(tuple(parts))
    assert k == len(flat), "parameter count does not match the ResLayer"
    return out


def _sub_forward(l, h, p, batch, fp):
    if l[0] == "cnn":
        return cnn_forward(h, p[0], batch, *l[1:], fp), None
    if l[0] == "ln":
        act, xhat, inv = ln_forward(h, p[0], p[1], batch, fp)
        return act, (xhat, inv)
    return relu_forward(h)


def _sub_backward(l, delta, inp, p, extra, batch, fp, lr):
    if l[0] == "cnn":
        d, W2 = cnn_backward(delta, inp, p[0], batch, *l[1:], fp, lr)
        return d, [W2]
    if l[0] == "ln":
        d, _, g2, b2 = ln_backward(delta, extra[0], extra[1], p[0], p[1], batch, fp, lr)
        return d, [g2, b2]
    return relu_backward(delta, extra), []


def _block_forward(layers, params, inp, residual, batch, fp):
    acts, extras, h = [], [], inp
    for l, p in zip(layers[:-1], params[:-1]):
        a, e = _sub_forward(l, h, p, batch, fp)
        acts.append(a)
        extras.append(e)
        h = a
    lli = acts[-1].copy() if residual is None else acts[-1] + residual  # share-wise, mod 2^64
    a, e = _sub_forward(layers[-1], lli, params[-1], batch, fp)
    acts.append(a)
    extras.append(e)
    return acts, extras, lli


def res_forward(x, batch, fp, spec, flat_params):
    """ResLayer::forward (ResLayer.cu:196-236).  Returns (output, state) with
    the state backward needs."""
    structure = res_structure(*spec)
    params = _split_params(structure, flat_params)
    state, h = [], _u(x)
    for (main, sc), (pm, ps) in zip(structure, params):
        if sc:
            sc_acts, sc_extras, sc_lli = _block_forward(sc, ps, h, None, batch, fp)
            residual = sc_acts[-1]
        else:
            sc_acts = sc_extras = sc_lli = None
            residual = h
        acts, extras, lli = _block_forward(main, pm, h, residual, batch, fp)
        state.append((h, acts, extras, lli, sc_acts, sc_extras, sc_lli))
        h = acts[-1]
    return h, state


def _block_backward(layers, params, acts, extras, lli, delta, forward_input, batch, fp, lr):
    last = len(layers) - 1
    dl, newp = [None] * len(layers), [None] * len(layers)
    dl[last], newp[last] = _sub_backward(layers[last], delta, lli, params[last], extras[last], batch, fp, lr)
    for i in range(last - 1, 0, -1):
        dl[i], newp[i] = _sub_backward(layers[i], dl[i + 1], acts[i - 1], params[i], extras[i], batch, fp, lr)
    dl[0], newp[0] = _sub_backward(layers[0], dl[1], forward_input, params[0], extras[0], batch, fp, lr)
    return dl, newp


def res_backward(delta, x, batch, fp, lr, spec, flat_params):
    """ResLayer::backward (ResLayer.cu:250-307), with every inner weight update.
    Returns (dL/d input, updated parameters in snapshot order).  The block
    without a shortcut adds the scalar 1 to its first layer's delta, as the
    reference does."""
    structure = res_structure(*spec)
    params = _split_params(structure, flat_params)
    _, state = res_forward(x, batch, fp, spec, flat_params)
    new = [None] * len(structure)
    d = _u(delta)
    for b in range(len(structure) - 1, -1, -1):
        main, sc = structure[b]
        pm, ps = params[b]
        inp, acts, extras, lli, sc_acts, sc_extras, sc_lli = state[b]
        dl, npm = _block_backward(main, pm, acts, extras, lli, d, inp, batch, fp, lr)
        d0 = dl[0].copy()
        nps = []
        if sc:
            sdl, nps = _block_backward(sc, ps, sc_acts, sc_extras, sc_lli, dl[-1], inp, batch, fp, lr)
            d0 = d0 + sdl[0]
        else:
            d0[0, 0] += np.uint64(1)  # RSS += 1: A.sA and C.sB
            d0[2, 1] += np.uint64(1)
        new[b] = (npm, nps)
        d = d0
    flat = []
    for npm, nps in new:
        for per in list(npm) + list(nps):
            flat.extend(per)
    return d, flat


def net_forward(layers, batch, fp, params, x_shares):
    """NeuralNetwork::forward (NeuralNetwork.cu:108-133) over the oracle layers.

    layers: symbolic specs (piranha_amd.configs form); params[i] = (W, b) share
    tensors for fc / cnn layers (b None for cnn); x_shares: input share tensor.
    Returns the list of (activation, derivative-or-None) per layer.
    """
    out = []
    h = x_shares
    for i, l in enumerate(layers):
        kind = l[0]
        if kind == "fc":
            W, b = params[i]
            h = fc_forward(h, W, b, batch, l[1], l[2], fp)
            out.append((h, None))
        elif kind == "cnn":
            W, _ = params[i]
            H, Wd, Cin, Cout, f, s, p = l[1:]
            h = cnn_forward(h, W, batch, H, Wd, Cin, Cout, f, s, p, fp)
            out.append((h, None))
        elif kind == "relu":
            h, d = relu_forward(h)
            out.append((h, d))
        elif kind == "maxpool":
            H, Wd, C, pool, s = l[1:]
            h, d = maxpool_forward(h, batch, H, Wd, C, pool, s, fp)
            out.append((h, d))
        elif kind == "avgpool":
            H, Wd, C, pool, s = l[1:]
            h = avgpool_forward(h, batch, H, Wd, C, pool, s)
            out.append((h, None))
        elif kind == "ln":
            gamma, beta = params[i]
            h, _, _ = ln_forward(h, gamma, beta, batch, fp)
            out.append((h, None))
        elif kind == "res":
            h, _ = res_forward(h, batch, fp, l[1:], params[i])
            out.append((h, None))
        else:
            raise ValueError(kind)
    return out


# --- backward (training) path ------------------------------------------------

def conv_dgrad(dy, w, b, OH, OW, Dout, fh, fw, Din, ph, pw, stride, dil, H, W):
    """gpu::conv_dgrad (conv.cuh:405-446): dx [b][H][W][Din]."""
    dy, w = _u(dy), _u(w)
    dx = np.zeros(b * H * W * Din, np.uint64)
    lib().ref_conv_dgrad(_p(dy), _p(w), _p(dx), b, OH, OW, Dout, fh, fw, Din, ph, pw, stride, dil, H, W)
    return dx


def conv_wgrad(dy, x, b, OH, OW, Dout, H, W, Din, fh, fw, ph, pw, stride, dil):
    """gpu::conv_wgrad (conv.cuh:448-490): dw [Dout][fh][fw][Din]."""
    dy, x = _u(dy), _u(x)
    dw = np.zeros(Dout * fh * fw * Din, np.uint64)
    lib().ref_conv_wgrad(_p(dy), _p(x), _p(dw), b, OH, OW, Dout, H, W, Din, fh, fw, ph, pw, stride, dil)
    return dw


def reduce_sum(a, b, reduce_rows, rows, cols):
    """gpu::reduceSum (matrix.cuh:109-126); accumulates into a copy of b."""
    a, b = _u(a), _u(b).copy()
    lib().ref_reduce_sum(_p(a), _p(b), int(reduce_rows), rows, cols)
    return b


def vector_expand(a, factor):
    a = _u(a)
    b = np.zeros(a.size * factor, np.uint64)
    lib().ref_vector_expand(_p(a), a.size, _p(b), factor)
    return b


def maxpool_row2im(rows, out, image_height, image_width, f, Din, batch, stride):
    """gpu::maxpool_row2im (convolution.cuh:110-151); adds into a copy of out."""
    rows, out = _u(rows), _u(out).copy()
    lib().ref_maxpool_row2im(_p(rows), rows.size, _p(out), out.size, image_height, image_width, f, Din, batch, stride)
    return out


CONV_DGRAD, CONV_WGRAD = 1, 2


def convolution_grad(op, A, B, batch, H, W, f, Din, Dout, stride, pad, trunc):
    """RSS convolution(kDgrad / kWgrad) (RSS.inl:641-790): A = delta, B = filter
    (dgrad) or forward input (wgrad)."""
    A, B = _u(A), _u(B)
    n = batch * H * W * Din if op == CONV_DGRAD else Dout * f * f * Din
    C = np.zeros((3, 2, n), np.uint64)
    lib().ref_convolution_grad(op, _p(A), A.shape[-1], _p(B), B.shape[-1], _p(C), n, batch, H, W, f, Din, Dout,
                               stride, pad, trunc)
    return C


def fc_backward(delta, x, W, b, batch, d_in, d_out, fp, lr):
    """FCLayer::backward (FCLayer.cu:133-224) -> (deltas, new W, new bias)."""
    delta, x = _u(delta), _u(x)
    W, b = _u(W).copy(), _u(b).copy()
    deltas = np.zeros((3, 2, batch * d_in), np.uint64)
    lib().ref_fc_backward(_p(delta), _p(x), _p(W), _p(b), _p(deltas), batch, d_in, d_out, fp, lr)
    return deltas, W, b


def cnn_backward(delta, x, W, batch, H, Wd, Din, Dout, f, stride, pad, fp, lr):
    """CNNLayer::backward (CNNLayer.cu:118-195) -> (deltas, new W)."""
    delta, x = _u(delta), _u(x)
    W = _u(W).copy()
    deltas = np.zeros((3, 2, batch * H * Wd * Din), np.uint64)
    lib().ref_cnn_backward(_p(delta), _p(x), _p(W), _p(deltas), batch, H, Wd, Din, Dout, f, stride, pad, fp, lr)
    return deltas, W


def relu_backward(delta, relu_prime):
    """ReLULayer::backward (ReLULayer.cu:91-123)."""
    delta, rp = _u(delta), _u(relu_prime, np.uint8)
    n = delta.shape[-1]
    deltas = np.zeros((3, 2, n), np.uint64)
    lib().ref_relu_backward(_p(delta), _p(rp), n, _p(deltas))
    return deltas


def maxpool_backward(delta, max_prime, batch, H, W, C, pool, stride):
    """MaxpoolLayer::backward (MaxpoolLayer.cu:133-180)."""
    delta, mp = _u(delta), _u(max_prime, np.uint8)
    deltas = np.zeros((3, 2, batch * H * W * C), np.uint64)
    lib().ref_maxpool_backward(_p(delta), _p(mp), _p(deltas), batch, H, W, C, pool, stride)
    return deltas


def reveal_softmax_grad(act, labels, n_classes, fp, log_mini_batch):
    """NeuralNetwork::_reveal_softmax_grad (NeuralNetwork.cu:207-267)."""
    act = _u(act)
    labels = np.ascontiguousarray(labels, np.float64).ravel()
    n = act.shape[-1]
    assert labels.size == n
    deltas = np.zeros((3, 2, n), np.uint64)
    lib().ref_reveal_softmax_grad(_p(act), _p(labels), n, _p(deltas), n_classes, fp, log_mini_batch)
    return deltas


def log_mini_batch(batch):
    """LOG_MINI_BATCH = log2(MINI_BATCH_SIZE) stored in an int (model.cu:39)."""
    return int(np.log2(batch))


def net_backward(layers, batch, fp, params, x_shares, fwd, labels, lr, out_delta=None):
    """NeuralNetwork::backward (NeuralNetwork.cu:136-149, 435-467) over the
    oracle layers: reveal-softmax delta, then every layer's backward in
    reverse order, each with the previous layer's activation (the network
    input for layer 0).  fwd is net_forward's output for x_shares.  With
    out_delta given, the backward pass starts from it instead of the
    reveal-softmax delta of the labels.

    Returns (deltas, new_params): deltas[i] is layer i's outgoing delta
    (layer 0 included), new_params the updated (W, b) per layer.
    """
    L = len(layers)
    n_classes = fwd[-1][0].shape[-1] // batch
    if out_delta is None:
        delta = reveal_softmax_grad(fwd[-1][0], labels, n_classes, fp, log_mini_batch(batch))
    else:
        delta = _u(out_delta)
    deltas = [None] * L
    new_params = list(params)
    for i in range(L - 1, -1, -1):
        kind = layers[i][0]
        inp = fwd[i - 1][0] if i > 0 else x_shares
        if kind == "fc":
            W, b = params[i]
            d, W2, b2 = fc_backward(delta, inp, W, b, batch, layers[i][1], layers[i][2], fp, lr)
            new_params[i] = (W2, b2)
        elif kind == "cnn":
            W, _ = params[i]
            H, Wd, Cin, Cout, f, s, p = layers[i][1:]
            d, W2 = cnn_backward(delta, inp, W, batch, H, Wd, Cin, Cout, f, s, p, fp, lr)
            new_params[i] = (W2, None)
        elif kind == "relu":
            d = relu_backward(delta, fwd[i][1])
        elif kind == "maxpool":
            H, Wd, C, pool, s = layers[i][1:]
            d = maxpool_backward(delta, fwd[i][1], batch, H, Wd, C, pool, s)
        elif kind == "avgpool":
            H, Wd, C, pool, s = layers[i][1:]
            d = avgpool_backward(delta, batch, H, Wd, C, pool)
        elif kind == "ln":
            gamma, beta = params[i]
            _, xhat, inv_sigma = ln_forward(inp, gamma, beta, batch, fp)  # the forward's saved state
            d, _, g2, b2 = ln_backward(delta, xhat, inv_sigma, gamma, beta, batch, fp, lr)
            new_params[i] = (g2, b2)
        elif kind == "res":
            d, new_params[i] = res_backward(delta, inp, batch, fp, lr, layers[i][1:], params[i])
        else:
            raise ValueError(kind)
        deltas[i] = d
        delta = d
    return deltas, new_params


# --- TPC (2 parties x 1 share) and FPC (4 parties x 3 shares) ----------------

_NSH = {"tpc": (2,), "fpc": (4, 3)}


def _scheme_init_list(kind, vals, fp, convert):
    vals = np.ascontiguousarray(vals, np.float64).ravel()
    x = np.zeros(_NSH[kind] + (vals.size,), np.uint64)
    getattr(lib(), f"ref_{kind}_init_list")(_p(x), vals.size, _p(vals), int(convert), fp)
    return x


def tpc_init_list(vals, fp, convert=True):
    """TPC initializer-list constructor (TPC.inl:341-363): party A holds the values."""
    return _scheme_init_list("tpc", vals, fp, convert)


def fpc_init_list(vals, fp, convert=True):
    """FPC initializer-list constructor (FPC.inl:527-561)."""
    return _scheme_init_list("fpc", vals, fp, convert)


def tpc_reconstruct(x):
    x = _u(x)
    out = np.zeros((2, x.shape[-1]), np.uint64)
    lib().ref_tpc_reconstruct(_p(x), x.shape[-1], _p(out))
    return out


def fpc_reconstruct(x):
    x = _u(x)
    out = np.zeros((4, x.shape[-1]), np.uint64)
    lib().ref_fpc_reconstruct(_p(x), x.shape[-1], _p(out))
    return out


def tpc_divide_public(a, d):
    a = _u(a).copy()
    lib().ref_tpc_divide_public(_p(a), a.shape[-1], int(d))
    return a


def fpc_divide_public(a, d):
    a = _u(a).copy()
    lib().ref_fpc_divide_public(_p(a), a.shape[-1], int(d))
    return a


def fpc_reshare(z, zp, i, j, out=None):
    """reshareFPC (FPC.inl:1183-1260) of per-party z, z' (4, n), added into out."""
    z, zp = _u(z), _u(zp)
    n = z.shape[-1]
    out = np.zeros((4, 3, n), np.uint64) if out is None else _u(out).copy()
    lib().ref_fpc_reshare(_p(z), _p(zp), n, i, j, _p(out))
    return out


def _scheme_matmul(kind, a, b, M, N, K, ta, tb, tc, trunc):
    a, b = _u(a), _u(b)
    c = np.zeros(_NSH[kind] + (M * N,), np.uint64)
    getattr(lib(), f"ref_{kind}_matmul")(_p(a), a.shape[-1], _p(b), b.shape[-1], _p(c), M, N, K, int(ta), int(tb),
                                         int(tc), trunc)
    return c


def tpc_matmul(a, b, M, N, K, ta, tb, tc, trunc):
    """matmul (TPC.inl:574-583, localMatMul :1059-1090)."""
    return _scheme_matmul("tpc", a, b, M, N, K, ta, tb, tc, trunc)


def fpc_matmul(a, b, M, N, K, ta, tb, tc, trunc):
    """matmul (FPC.inl:635-666)."""
    return _scheme_matmul("fpc", a, b, M, N, K, ta, tb, tc, trunc)


def _scheme_conv(kind, A, B, batch, H, W, f, Din, Dout, stride, pad, trunc):
    A, B = _u(A), _u(B)
    OH = (H + 2 * pad - f) // stride + 1
    OW = (W + 2 * pad - f) // stride + 1
    n = batch * OH * OW * Dout
    C = np.zeros(_NSH[kind] + (n,), np.uint64)
    getattr(lib(), f"ref_{kind}_convolution")(_p(A), A.shape[-1], _p(B), B.shape[-1], _p(C), n, batch, H, W, f, Din,
                                              Dout, stride, pad, trunc)
    return C


def tpc_convolution(A, B, batch, H, W, f, Din, Dout, stride, pad, trunc):
    """convolution(kFprop) (TPC.inl:644-690, 788-826)."""
    return _scheme_conv("tpc", A, B, batch, H, W, f, Din, Dout, stride, pad, trunc)


def fpc_convolution(A, B, batch, H, W, f, Din, Dout, stride, pad, trunc):
    """convolution(kFprop) (FPC.inl:744-782, 861-927)."""
    return _scheme_conv("fpc", A, B, batch, H, W, f, Din, Dout, stride, pad, trunc)


def _eb(dtype):
    return np.dtype(dtype).itemsize


def tpc_mult(x, y):
    """TPC *= TPC (TPC.inl:242-272, Beaver triples x = y = z = fill(1))."""
    x = _u(x).copy()
    lib().ref_tpc_mult(_p(x), _p(_u(y)), x.shape[-1])
    return x


def tpc_and(x, y):
    """TPC &= TPC (TPC.inl:281-312, boolean triples x = y = z = fill(1))."""
    x = np.ascontiguousarray(x).copy()
    y = np.ascontiguousarray(y, x.dtype)
    lib().ref_tpc_and(x.ctypes.data, y.ctypes.data, x.shape[-1], _eb(x.dtype))
    return x


def tpc_drelu(x, out_dtype=np.uint8):
    """dReLU (TPC.inl:828-868): (2, n) shares of [x >= 0]."""
    x = _u(x)
    out = np.zeros((2, x.shape[-1]), out_dtype)
    lib().ref_tpc_drelu(_p(x), x.shape[-1], out.ctypes.data, _eb(out_dtype))
    return out


def tpc_select_share(x, y, b):
    """selectShare (TPC.inl:585-613): z = b ? y : x."""
    x, y, b = _u(x), _u(y), np.ascontiguousarray(b)
    z = np.zeros_like(x)
    lib().ref_tpc_select_share(_p(x), _p(y), b.ctypes.data, _p(z), x.shape[-1], _eb(b.dtype))
    return z


def tpc_relu(x, out_dtype=np.uint8):
    """ReLU (TPC.inl:870-884): (result, dresult)."""
    x = _u(x)
    y = np.zeros_like(x)
    d = np.zeros((2, x.shape[-1]), out_dtype)
    lib().ref_tpc_relu(_p(x), x.shape[-1], _p(y), d.ctypes.data, _eb(out_dtype))
    return y, d


def tpc_maxpool(pools, k, out_dtype=np.uint8):
    """maxpool over gathered pools (TPC.inl:886-1030); returns (result, dresult, pools_after)."""
    pools = _u(pools).copy()
    n = pools.shape[-1]
    result = np.zeros((2, n // k), np.uint64)
    d = np.zeros((2, n), out_dtype)
    lib().ref_tpc_maxpool(_p(pools), n, _p(result), d.ctypes.data, k, _eb(out_dtype))
    return result, d, pools


def fpc_mult(x, y):
    """FPC *= FPC (FPC.inl:321-362): cross terms, share-wise product, six reshareFPC."""
    x = _u(x).copy()
    lib().ref_fpc_mult(_p(x), _p(_u(y)), x.shape[-1])
    return x


def fpc_and(x, y):
    """FPC &= FPC (FPC.inl:374-398), the reshare adding in the element type."""
    x = np.ascontiguousarray(x).copy()
    y = np.ascontiguousarray(y, x.dtype)
    lib().ref_fpc_and(x.ctypes.data, y.ctypes.data, x.shape[-1], _eb(x.dtype))
    return x


def fpc_drelu(x, out_dtype=np.uint8):
    """dReLU (FPC.inl:929-989): (4, 3, n) shares of [x >= 0]."""
    x = _u(x)
    out = np.zeros((4, 3, x.shape[-1]), out_dtype)
    lib().ref_fpc_drelu(_p(x), x.shape[-1], out.ctypes.data, _eb(out_dtype))
    return out


def fpc_select_share(x, y, b):
    """selectShare (FPC.inl:671-712): z = b ? y : x."""
    x, y, b = _u(x), _u(y), np.ascontiguousarray(b)
    z = np.zeros_like(x)
    lib().ref_fpc_select_share(_p(x), _p(y), b.ctypes.data, _p(z), x.shape[-1], _eb(b.dtype))
    return z


def fpc_relu(x, out_dtype=np.uint8):
    """ReLU (FPC.inl:991-1002): (result, dresult)."""
    x = _u(x)
    y = np.zeros_like(x)
    d = np.zeros((4, 3, x.shape[-1]), out_dtype)
    lib().ref_fpc_relu(_p(x), x.shape[-1], _p(y), d.ctypes.data, _eb(out_dtype))
    return y, d


def fpc_maxpool(pools, k, out_dtype=np.uint8):
    """maxpool over gathered pools (FPC.inl:1004-1181); returns (result, dresult, pools_after)."""
    pools = _u(pools).copy()
    n = pools.shape[-1]
    result = np.zeros((4, 3, n // k), np.uint64)
    d = np.zeros((4, 3, n), out_dtype)
    lib().ref_fpc_maxpool(_p(pools), n, _p(result), d.ctypes.data, k, _eb(out_dtype))
    return result, d, pools


TAYLOR_COEFFS = {TAYLOR_SQRT: (0.424, 0.584, 0.0), TAYLOR_INVERSE: (4.245, -5.857, 2.630),
                 TAYLOR_SIGMOID: (0.494286, 0.275589, -0.038751)}  # RSS.inl / TPC.inl / FPC.inl sqrt, inverse, sigmoid


def scheme_get_powers(kind, x):
    """getPowers of a TPC (2, n) or FPC (4, 3, n) value: each party's opened power."""
    x = _u(x)
    n = x.shape[-1]
    pw = np.zeros((_NSH[kind][0], n), np.uint64)
    getattr(lib(), f"ref_{kind}_get_powers")(_p(x), n, _p(pw))
    return pw


def scheme_taylor_series(kind, x, a0, a1, a2, fn, fp):
    """taylorSeries of a TPC or FPC value (TPC.inl:1228-1290, FPC.inl:1443-1526)."""
    x = _u(x)
    out = np.zeros_like(x)
    getattr(lib(), f"ref_{kind}_taylor_series")(_p(x), x.shape[-1], _p(out), a0, a1, a2, fn, fp)
    return out


def scheme_taylor(kind, x, fn, fp):
    """sqrt / inverse / sigmoid of a TPC or FPC value with the reference's coefficients."""
    return scheme_taylor_series(kind, x, *TAYLOR_COEFFS[fn], fn, fp)


def scheme_divide_public_vec(kind, a, d):
    """dividePublic(Share, DeviceData) of a TPC or FPC value by per-party denominators d (P, n)."""
    a = _u(a).copy()
    getattr(lib(), f"ref_{kind}_divide_public_vec")(_p(a), a.shape[-1], _p(_u(d)))
    return a


# --- masked precompute mode ----------------------------------------------------
# The build's own mode (no reference counterpart: the reference deals only
# placeholder constants, src/mpc/Precompute.h:73-134).  `seed` keys
# Philox4x32-10, `call` is the protocol call's number; see pir_ref.cpp.

def philox(ctr, key):
    """Philox4x32-10 blocks: ctr (n, 4) and key (n, 2) uint32 -> (n, 4) uint32."""
    ctr = np.ascontiguousarray(ctr, np.uint32).reshape(-1, 4)
    key = np.ascontiguousarray(key, np.uint32).reshape(-1, 2)
    out = np.zeros_like(ctr)
    lib().ref_philox(_p(ctr), _p(key), ctr.shape[0], _p(out))
    return out


def prg_fill(seed, call, n):
    out = np.zeros(n, np.uint64)
    lib().ref_prg_fill(seed, call, n, _p(out))
    return out


def deal(values, seed, call):
    values = _u(values)
    out = np.zeros((3, 2, values.size), np.uint64)
    lib().ref_deal(_p(values), values.size, seed, call, _p(out))
    return out


def reshare_m(c, seed, call):
    c = _u(c).copy()
    n = c.shape[-1]
    out = np.zeros((3, 2, n), np.uint64)
    lib().ref_reshare_m(_p(c), n, _p(out), seed, call)
    return out


def mult_m(x, y, seed, call):
    x, y = _u(x).copy(), _u(y)
    lib().ref_rss_mult_m(_p(x), _p(y), x.shape[-1], seed, call)
    return x


def rss_and_m(x, y, seed, call):
    x = np.ascontiguousarray(x).copy()
    y = np.ascontiguousarray(y, x.dtype)
    (lib().ref_rss_and_m_u8 if x.dtype == np.uint8 else lib().ref_rss_and_m_u64)(_p(x), _p(y), x.shape[-1], seed, call)
    return x


def divide_public_m(a, k, seed, call):
    a = _u(a).copy()
    lib().ref_divide_public_m(_p(a), a.shape[-1], k, seed, call)
    return a


def divide_public_3out3_m(a, k, seed, call):
    a = _u(a).copy()
    n = a.shape[-1]
    out = np.zeros((3, 2, n), np.uint64)
    lib().ref_divide_public_3out3_m(_p(a), n, k, _p(out), seed, call)
    return out


def matmul_m(a, b, M, N, K, ta, tb, tc, trunc, seed, call):
    a, b = _u(a), _u(b)
    c = np.zeros((3, 2, M * N), np.uint64)
    lib().ref_matmul_m(_p(a), a.shape[-1], _p(b), b.shape[-1], _p(c), M, N, K, int(ta), int(tb), int(tc), trunc,
                       seed, call)
    return c


def convolution_m(A, B, batch, H, W, f, Din, Dout, stride, pad, trunc, seed, call):
    A, B = _u(A), _u(B)
    OH, OW = (H + 2 * pad - f) // stride + 1, (W + 2 * pad - f) // stride + 1
    C = np.zeros((3, 2, batch * OH * OW * Dout), np.uint64)
    lib().ref_convolution_m(_p(A), A.shape[-1], _p(B), B.shape[-1], _p(C), C.shape[-1], batch, H, W, f, Din, Dout,
                            stride, pad, trunc, seed, call)
    return C


def drelu_m(x, seed, call, out_dtype=np.uint8):
    x = _u(x)
    n = x.shape[-1]
    out = np.zeros((3, 2, n), out_dtype)
    (lib().ref_drelu_m_u8 if out_dtype == np.uint8 else lib().ref_drelu_m_u64)(_p(x), n, _p(out), seed, call)
    return out


def select_share_m(x, y, b, seed, call):
    x, y = _u(x), _u(y)
    b = np.ascontiguousarray(b)
    n = x.shape[-1]
    z = np.zeros((3, 2, n), np.uint64)
    (lib().ref_select_share_m_u8 if b.dtype == np.uint8 else lib().ref_select_share_m_u64)(
        _p(x), _p(y), _p(b), _p(z), n, seed, call)
    return z


def relu_m(x, seed, call, out_dtype=np.uint8):
    x = _u(x)
    n = x.shape[-1]
    out = np.zeros((3, 2, n), np.uint64)
    dout = np.zeros((3, 2, n), out_dtype)
    (lib().ref_relu_m_u8 if out_dtype == np.uint8 else lib().ref_relu_m_u64)(_p(x), n, _p(out), _p(dout), seed, call)
    return out, dout


def maxpool_m(pools, k, seed, call, out_dtype=np.uint8):
    pools = _u(pools).copy()
    n = pools.shape[-1]
    result = np.zeros((3, 2, n // k), np.uint64)
    dres = np.zeros((3, 2, n), out_dtype)
    (lib().ref_maxpool_m_u8 if out_dtype == np.uint8 else lib().ref_maxpool_m_u64)(
        _p(pools), n, _p(result), _p(dres), k, seed, call)
    return result, dres, pools


def net_forward_m(layers, batch, fp, params, x_shares, seed, call):
    """NeuralNetwork::forward in the masked precompute mode: layer i uses call
    number call + i (fc, cnn, relu, maxpool layers)."""
    out = []
    h = _u(x_shares)
    for i, l in enumerate(layers):
        kind, c = l[0], call + i
        if kind == "fc":
            W, b = _u(params[i][0]), _u(params[i][1])
            act = np.zeros((3, 2, batch * l[2]), np.uint64)
            lib().ref_fc_forward_m(_p(h), _p(W), _p(b), _p(act), batch, l[1], l[2], fp, seed, c)
            h = act
            out.append((h, None))
        elif kind == "cnn":
            W = _u(params[i][0])
            H, Wd, Cin, Cout, f, s, p = l[1:]
            OH, OW = (H - f + 2 * p) // s + 1, (Wd - f + 2 * p) // s + 1
            act = np.zeros((3, 2, batch * OH * OW * Cout), np.uint64)
            lib().ref_cnn_forward_m(_p(h), _p(W), _p(act), batch, H, Wd, Cin, Cout, f, s, p, fp, seed, c)
            h = act
            out.append((h, None))
        elif kind == "relu":
            h, d = relu_m(h, seed, c)
            out.append((h, d))
        elif kind == "maxpool":
            H, Wd, C, pool, s = l[1:]
            OH, OW = (H - pool) // s + 1, (Wd - pool) // s + 1
            n = batch * C * OH * OW
            act = np.zeros((3, 2, n), np.uint64)
            mp = np.zeros((3, 2, n * pool * pool), np.uint8)
            lib().ref_maxpool_forward_m(_p(h), _p(act), _p(mp), batch, H, Wd, C, pool, s, fp, seed, c)
            h = act
            out.append((h, mp))
        else:
            raise ValueError(f"{kind}: no masked form")
    return out
