"""MI355X-native 3-party RSS secure-inference path (Piranha's per-party
secret-share compute path) -- Python handle on libpir_hip.so.

The product is the C ABI in include/pir.h (HIP kernels + C++ host runtime);
this module only loads it with ctypes and moves numpy arrays to and from
device memory, for tests and benchmarks.  It fails loudly if the library is
missing: there is no CPU fallback.

Share arrays: an RSS tensor for all parties is (3, 2, n) uint64 (party,
share, element); derivative / bit shares are (3, 2, n) uint8 or uint64.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpir_hip.so")

ALGO_PLAIN, ALGO_LIMB = 0, 1
MODE_FUSED, MODE_PLAIN = 0, 1
PRECOMPUTE_PLACEHOLDER, PRECOMPUTE_MASKED = 0, 1
LAYER_FC, LAYER_CNN, LAYER_RELU, LAYER_MAXPOOL, LAYER_AVGPOOL, LAYER_LN, LAYER_RES = 0, 1, 2, 3, 4, 5, 6
CONV_FPROP, CONV_DGRAD, CONV_WGRAD = 0, 1, 2
LIMB_WIDE, LIMB_VARIANTS = 8, 16  # kernel variant = A-operand form (0..6) + 8 for the 32x128 tile
OP_ADD, OP_SUB, OP_MUL, OP_XOR, OP_AND, OP_SHR, OP_SHL, OP_DIV = range(8)
TAYLOR_SQRT, TAYLOR_INVERSE, TAYLOR_SIGMOID = range(3)

_vp = ctypes.c_void_p
_i = ctypes.c_int
_sz = ctypes.c_size_t
_u64 = ctypes.c_uint64
_ll = ctypes.c_longlong
_d = ctypes.c_double

_SIGS = {
    "pir_version": ([], _i),
    "pir_last_error": ([], ctypes.c_char_p),
    "pir_malloc": ([ctypes.POINTER(_vp), _sz], _i),
    "pir_free": ([_vp], _i),
    "pir_memcpy_h2d": ([_vp, _vp, _sz, _vp], _i),
    "pir_memcpy_d2h": ([_vp, _vp, _sz, _vp], _i),
    "pir_memcpy_d2d": ([_vp, _vp, _sz, _vp], _i),
    "pir_fill_u64": ([_vp, _u64, _sz, _vp], _i),
    "pir_fill_u8": ([_vp, ctypes.c_uint8, _sz, _vp], _i),
    "pir_stream_create": ([ctypes.POINTER(_vp)], _i),
    "pir_stream_destroy": ([_vp], _i),
    "pir_release_cached": ([], _i),
    "pir_cached_bytes": ([], _sz),
    "pir_stream_synchronize": ([_vp], _i),
    "pir_set_precompute": ([_i, _u64, ctypes.c_uint32], _i),
    "pir_get_prg_call": ([ctypes.POINTER(ctypes.c_uint32)], _i),
    "pir_prg_philox": ([_vp, _vp, _sz, _vp, _vp], _i),
    "pir_prg_fill": ([_u64, ctypes.c_uint32, _sz, _vp, _vp], _i),
    "pir_rss_deal": ([_vp, _sz, _u64, ctypes.c_uint32, _vp, _vp], _i),
    "pir_rss_rerandomize": ([_vp, _sz, _u64, ctypes.c_uint32, _vp, _vp], _i),
    "pir_net_set_precompute": ([_vp, _i, _u64, ctypes.c_uint32], _i),
    "pir_net_rerandomize": ([_vp, _u64], _i),
    "pir_measure_peak": ([_i, _sz, _i, ctypes.POINTER(_d), ctypes.POINTER(_d), _vp], _i),
    "pir_device_synchronize": ([], _i),
    "pir_set_mode": ([_i], _i),
    "pir_gemm": ([_i, _i, _i, _vp, _i, _vp, _i, _vp, _i, _i, _vp], _i),
    "pir_conv_fprop": ([_vp, _vp, _vp] + [_i] * 12 + [_vp], _i),
    "pir_conv_dgrad": ([_vp, _vp, _vp] + [_i] * 14 + [_vp], _i),
    "pir_conv_wgrad": ([_vp, _vp, _vp] + [_i] * 14 + [_vp], _i),
    "pir_reduce_sum": ([_vp, _vp, _i, _i, _i, _vp], _i),
    "pir_vector_expand": ([_vp, _sz, _vp, _i, _vp], _i),
    "pir_maxpool_row2im": ([_vp, _vp] + [_i] * 6 + [_vp], _i),
    "pir_ring_gemm3": ([_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "pir_ring_gemm3_noalloc": ([_i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "pir_limb_packed_bytes": ([_i, _i, _i], _sz),
    "pir_limb_pack": ([_vp, _vp, _i, _i, _i, _i, _vp, _vp], _i),
    "pir_limb_gemm": ([_vp, _i, _i, _i, _i, _vp, _i, _vp, _i, _i, _vp], _i),
    "pir_bitexpand": ([_vp, _sz, _vp, _i, _vp], _i),
    "pir_set_carry_out_msb": ([_vp, _vp, _vp, _sz, _i, _i, _i, _vp], _i),
    "pir_expand_compare": ([_vp, _vp, _vp, _sz, _sz, _i, _vp], _i),
    "pir_element_vector_add": ([_vp, _vp, _i, _i, _i, _vp], _i),
    "pir_truncate_cols": ([_vp, _sz, _vp, _i, _i, _i, _vp], _i),
    "pir_stride_pad": ([_vp, _vp, _sz, _i, _i, _u64, _vp], _i),
    "pir_maxpool_im2row_size": ([_i] * 7, _sz),
    "pir_maxpool_im2row": ([_vp, _vp] + [_i] * 7 + [_u64, _vp], _i),
    "pir_averagepool_im2row_size": ([_i] * 7, _sz),
    "pir_transpose": ([_vp, _vp, _i, _i, _i, _vp], _i),
    "pir_get_msbs": ([_vp, _vp, _sz, _sz, _i, _vp], _i),
    "pir_set_msbs": ([_vp, _sz, _u64, _sz, _i, _vp], _i),
    "pir_averagepool_im2row": ([_vp, _vp] + [_i] * 7 + [_vp], _i),
    "pir_averagepool_expand_delta": ([_vp, _sz, _vp, _sz, _i, _i, _vp], _i),
    "pir_dd_op": ([_i, _i, _vp, _ll, _vp, _ll, _u64, _sz, _vp], _i),
    "pir_rss_reconstruct": ([_vp, _sz, _vp, _i, _vp], _i),
    "pir_rss_reshare": ([_vp, _sz, _vp, _vp], _i),
    "pir_rss_mult": ([_vp, _vp, _sz, _vp], _i),
    "pir_rss_and": ([_vp, _vp, _sz, _i, _vp], _i),
    "pir_rss_divide_public": ([_vp, _sz, _u64, _vp, _vp], _i),
    "pir_rss_divide_public_3out3": ([_vp, _sz, _i, _vp, _vp], _i),
    "pir_rss_matmul": ([_vp, _vp, _vp] + [_i] * 8 + [_vp], _i),
    "pir_rss_convolution": ([_vp, _vp, _vp] + [_i] * 10 + [_vp], _i),
    "pir_rss_convolution_op": ([_i, _vp, _vp, _vp] + [_i] * 10 + [_vp], _i),
    "pir_rss_drelu": ([_vp, _sz, _vp, _i, _vp], _i),
    "pir_rss_get_powers": ([_vp, _sz, _vp, _vp], _i),
    "pir_rss_taylor_series": ([_vp, _sz, _vp, _d, _d, _d, _i, _i, _vp], _i),
    "pir_rss_sqrt": ([_vp, _sz, _vp, _i, _vp], _i),
    "pir_rss_inverse": ([_vp, _sz, _vp, _i, _vp], _i),
    "pir_rss_sigmoid": ([_vp, _sz, _vp, _i, _vp], _i),
    "pir_rss_relu": ([_vp, _sz, _vp, _vp, _i, _vp], _i),
    "pir_rss_select_share": ([_vp, _vp, _vp, _vp, _sz, _i, _vp], _i),
    "pir_rss_maxpool": ([_vp, _sz, _vp, _vp, _i, _i, _vp], _i),
    "pir_net_create": ([ctypes.POINTER(_vp), _vp, _i, _i, _i, _i, _i, _vp], _i),
    "pir_net_destroy": ([_vp], _i),
    "pir_net_num_layers": ([_vp], _i),
    "pir_model_read": ([ctypes.c_char_p, _vp, _vp, _vp, _vp, _vp, _i], _i),
    "pir_net_load_model": ([ctypes.POINTER(_vp), ctypes.c_char_p, _i, _i, _i, _i, _vp], _i),
    "pir_net_batch_size": ([_vp], _i),
    "pir_net_layer_spec": ([_vp, _i, _vp], _i),
    "pir_net_save_snapshot": ([_vp, ctypes.c_char_p], _i),
    "pir_net_load_snapshot": ([_vp, ctypes.c_char_p], _i),
    "pir_net_num_params": ([_vp, _i], _i),
    "pir_net_param": ([_vp, _i, _i, ctypes.POINTER(_vp), ctypes.POINTER(_sz)], _i),
    "pir_net_param_name": ([_vp, _i, _i, ctypes.c_char_p, _sz], _i),
    "pir_net_set_param_shares": ([_vp, _i, _i, _vp], _i),
    "pir_net_set_weights": ([_vp, _i, _vp, _sz, _vp, _sz], _i),
    "pir_net_forward": ([_vp, _vp, _sz], _i),
    "pir_net_forward_shares": ([_vp, _vp], _i),
    "pir_net_input": ([_vp, ctypes.POINTER(_vp), ctypes.POINTER(_sz)], _i),
    "pir_net_synchronize": ([_vp], _i),
    "pir_net_activation": ([_vp, _i, ctypes.POINTER(_vp), ctypes.POINTER(_sz)], _i),
    "pir_net_derivative": ([_vp, _i, ctypes.POINTER(_vp), ctypes.POINTER(_sz)], _i),
    "pir_net_weights": ([_vp, _i, ctypes.POINTER(_vp), ctypes.POINTER(_sz), ctypes.POINTER(_vp),
                         ctypes.POINTER(_sz)], _i),
    "pir_net_set_weight_shares": ([_vp, _i, _vp, _vp], _i),
    "pir_net_set_log_learning_rate": ([_vp, _i], _i),
    "pir_net_backward": ([_vp, _vp, _sz], _i),
    "pir_net_backward_shares": ([_vp, _vp], _i),
    "pir_net_layer_backward": ([_vp, _i, _vp, _vp], _i),
    "pir_net_delta": ([_vp, _i, ctypes.POINTER(_vp), ctypes.POINTER(_sz)], _i),
    "pir_net_set_train_labels": ([_vp, _vp, _sz], _i),
    "pir_net_train_labels": ([_vp, ctypes.POINTER(_vp), ctypes.POINTER(_sz)], _i),
    "pir_net_train_step": ([_vp], _i),
    "pir_net_capture_train": ([_vp], _i),
    "pir_net_replay_train": ([_vp], _i),
    "pir_net_autotune_train": ([_vp], _i),
    "pir_net_profile_train_gemm": ([_vp, _vp, _vp, _vp], _i),
    "pir_net_profile_train_layers": ([_vp, _vp, _i], _i),
    "pir_net_capture": ([_vp, _vp], _i),
    "pir_net_replay": ([_vp], _i),
    "pir_net_set_mode": ([_vp, _i], _i),
    "pir_net_profile_layers": ([_vp, _vp, _i], _i),
    "pir_net_profile_gemm": ([_vp, _vp, _vp, _vp], _i),
    "pir_net_autotune": ([_vp], _i),
    "pir_autotune_set": ([_i], _i),
    "pir_autotune_entries": ([], _i),
    "pir_autotune_clear": ([], _i),
    "pir_autotune_save": ([ctypes.c_char_p], _i),
    "pir_autotune_load": ([ctypes.c_char_p], _i),
    "pir_tpc_reconstruct": ([_vp, _sz, _vp, _vp], _i),
    "pir_tpc_divide_public": ([_vp, _sz, _u64, _vp], _i),
    "pir_tpc_mult": ([_vp, _vp, _sz, _vp], _i),
    "pir_fpc_mult": ([_vp, _vp, _sz, _vp], _i),
    "pir_tpc_divide_public_local": ([_vp, _sz, _vp, _vp], _i),
    "pir_fpc_divide_public_local": ([_vp, _sz, _vp, _vp], _i),
    "pir_tpc_get_powers": ([_vp, _sz, _vp, _vp], _i),
    "pir_fpc_get_powers": ([_vp, _sz, _vp, _vp], _i),
    "pir_tpc_taylor_series": ([_vp, _sz, _vp, _d, _d, _d, _i, _i, _vp], _i),
    "pir_fpc_taylor_series": ([_vp, _sz, _vp, _d, _d, _d, _i, _i, _vp], _i),
    "pir_tpc_sqrt": ([_vp, _sz, _vp, _i, _vp], _i),
    "pir_tpc_inverse": ([_vp, _sz, _vp, _i, _vp], _i),
    "pir_tpc_sigmoid": ([_vp, _sz, _vp, _i, _vp], _i),
    "pir_fpc_sqrt": ([_vp, _sz, _vp, _i, _vp], _i),
    "pir_fpc_inverse": ([_vp, _sz, _vp, _i, _vp], _i),
    "pir_fpc_sigmoid": ([_vp, _sz, _vp, _i, _vp], _i),
    "pir_fpc_and": ([_vp, _vp, _sz, _i, _vp], _i),
    "pir_fpc_drelu": ([_vp, _sz, _vp, _i, _vp], _i),
    "pir_fpc_relu": ([_vp, _sz, _vp, _vp, _i, _vp], _i),
    "pir_fpc_select_share": ([_vp, _vp, _vp, _vp, _sz, _i, _vp], _i),
    "pir_fpc_maxpool": ([_vp, _sz, _vp, _vp, _i, _i, _vp], _i),
    "pir_tpc_and": ([_vp, _vp, _sz, _i, _vp], _i),
    "pir_tpc_drelu": ([_vp, _sz, _vp, _i, _vp], _i),
    "pir_tpc_relu": ([_vp, _sz, _vp, _vp, _i, _vp], _i),
    "pir_tpc_select_share": ([_vp, _vp, _vp, _vp, _sz, _i, _vp], _i),
    "pir_tpc_maxpool": ([_vp, _sz, _vp, _vp, _i, _i, _vp], _i),
    "pir_tpc_matmul": ([_vp, _vp, _vp] + [_i] * 8 + [_vp], _i),
    "pir_tpc_convolution": ([_vp, _vp, _vp] + [_i] * 10 + [_vp], _i),
    "pir_fpc_reconstruct": ([_vp, _sz, _vp, _vp], _i),
    "pir_fpc_divide_public": ([_vp, _sz, _u64, _vp], _i),
    "pir_fpc_reshare": ([_vp, _vp, _sz, _i, _i, _vp, _vp], _i),
    "pir_fpc_matmul": ([_vp, _vp, _vp] + [_i] * 8 + [_vp], _i),
    "pir_fpc_convolution": ([_vp, _vp, _vp] + [_i] * 10 + [_vp], _i),
}

_lib = None


class PirError(RuntimeError):
    pass


def lib():
    """Load libpir_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PirError(f"{LIB_PATH} is missing: build it with `python -m piranha_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, ret) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ret
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise PirError(f"status {rc}: {lib().pir_last_error().decode()}")


class Buffer:
    """Device buffer holding a numpy-shaped array (freed on deletion)."""

    def __init__(self, shape, dtype, stream=None):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.stream = stream
        p = _vp()
        check(lib().pir_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        b = cls(a.shape, a.dtype, stream)
        if b.nbytes:
            check(lib().pir_memcpy_h2d(b.ptr, a.ctypes.data, b.nbytes, stream))
        return b

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().pir_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, self.stream))
        return out

    def zero(self):
        if self.dtype.itemsize == 8:
            check(lib().pir_fill_u64(self.ptr, 0, self.nbytes // 8, self.stream))
        else:
            check(lib().pir_fill_u8(self.ptr, 0, self.nbytes, self.stream))
        return self

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.pir_free(self.ptr)
            self.ptr = None


def to_dev(a, stream=None):
    return Buffer.from_numpy(a, stream)


PEAK_HBM_COPY, PEAK_HBM_READ, PEAK_HBM_WRITE, PEAK_MFMA_I8, PEAK_VALU_U64, PEAK_MFMA_I8_RANDOM = range(6)


def measure_peak(which, nbytes=1 << 30, reps=10):
    """(rate, ms per launch) of one of the measured peaks (see include/pir.h)."""
    r, ms = _d(), _d()
    check(lib().pir_measure_peak(which, nbytes, reps, ctypes.byref(r), ctypes.byref(ms), None))
    return r.value, ms.value


def synchronize():
    check(lib().pir_device_synchronize())


def set_mode(mode):
    check(lib().pir_set_mode(mode))


def set_precompute(mode, seed=0, call=0):
    """The calling thread's precompute mode for the free-standing calls (the
    masked mode draws call number `call`, then the next ones, from `seed`)."""
    check(lib().pir_set_precompute(mode, seed, call))


def prg_call():
    c = ctypes.c_uint32()
    check(lib().pir_get_prg_call(ctypes.byref(c)))
    return c.value


def _ptr_array(ptrs):
    arr = (_vp * len(ptrs))(*[p for p in ptrs])
    return ctypes.cast(arr, _vp), arr


def read_model(path):
    """loadModel's reading of a model file (model.cu:24-115) without building
    it: (input_size, num_classes, batch_size, [8-int layer spec, ...])."""
    ins, ncl, bs, nl = _i(), _i(), _i(), _i()
    check(lib().pir_model_read(os.fsencode(path), ctypes.byref(ins), ctypes.byref(ncl), ctypes.byref(bs),
                               ctypes.byref(nl), None, 0))
    spec = np.zeros(8 * max(nl.value, 1), np.int32)
    check(lib().pir_model_read(os.fsencode(path), None, None, None, ctypes.byref(nl), spec.ctypes.data, nl.value))
    return ins.value, ncl.value, bs.value, [tuple(int(v) for v in spec[8 * i:8 * i + 8]) for i in range(nl.value)]


class Network:
    """Handle on a pir_net (NeuralNetwork of the reference, all three parties)."""

    def __init__(self, spec, batch, frac_bits, seed=0, mode=MODE_FUSED, model_file=None):
        h = _vp()
        if model_file is not None:  # loadModel (model.cu:24-115); batch <= 0 keeps the file's batch_size
            check(lib().pir_net_load_model(ctypes.byref(h), os.fsencode(model_file), batch or 0, frac_bits, seed, mode,
                                           None))
        else:
            flat = []
            for layer in spec:
                row = list(layer) + [0] * (8 - len(layer))
                flat.extend(row[:8])
            self._spec = np.asarray(flat, np.int32)
            check(lib().pir_net_create(ctypes.byref(h), self._spec.ctypes.data, len(spec), batch, frac_bits, seed,
                                       mode, None))
        self.h = h.value
        self.batch = lib().pir_net_batch_size(self.h)
        self.frac_bits = frac_bits
        self.spec = [self.layer_spec(i) for i in range(self.num_layers)]

    @classmethod
    def from_model_file(cls, path, frac_bits, batch=0, seed=0, mode=MODE_FUSED):
        return cls(None, batch, frac_bits, seed=seed, mode=mode, model_file=path)

    def layer_spec(self, layer):
        v = np.zeros(8, np.int32)
        check(lib().pir_net_layer_spec(self.h, layer, v.ctypes.data))
        return tuple(int(x) for x in v)

    def params(self, layer):
        """[(snapshot name, shares (3, 2, n) uint64)] of every parameter tensor of a layer."""
        return _params_of(self, layer)

    def set_param_shares(self, layer, index, shares):
        shares = np.ascontiguousarray(shares, np.uint64)
        check(lib().pir_net_set_param_shares(self.h, layer, index, shares.ctypes.data))

    def save_snapshot(self, path):
        check(lib().pir_net_save_snapshot(self.h, os.fsencode(path)))

    def load_snapshot(self, path):
        check(lib().pir_net_load_snapshot(self.h, os.fsencode(path)))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.pir_net_destroy(self.h)
            self.h = None

    @property
    def num_layers(self):
        return lib().pir_net_num_layers(self.h)

    def set_mode(self, mode):
        check(lib().pir_net_set_mode(self.h, mode))

    def set_precompute(self, mode, seed=0, call=0):
        check(lib().pir_net_set_precompute(self.h, mode, seed, call))

    def rerandomize(self, seed):
        """Every parameter and the input buffer re-dealt as random sharings."""
        check(lib().pir_net_rerandomize(self.h, seed))

    def set_weights(self, layer, w, b=None):
        w = np.ascontiguousarray(w, np.float64).ravel()
        bp, bn = (None, 0)
        if b is not None:
            b = np.ascontiguousarray(b, np.float64).ravel()
            bp, bn = b.ctypes.data, b.size
        check(lib().pir_net_set_weights(self.h, layer, w.ctypes.data, w.size, bp, bn))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float64).ravel()
        check(lib().pir_net_forward(self.h, x.ctypes.data, x.size))

    def forward_shares(self, dev_ptr=None):
        check(lib().pir_net_forward_shares(self.h, dev_ptr))

    def input_ptr(self):
        p, n = _vp(), _sz()
        check(lib().pir_net_input(self.h, ctypes.byref(p), ctypes.byref(n)))
        return p.value, n.value

    def synchronize(self):
        check(lib().pir_net_synchronize(self.h))

    def capture(self, dev_ptr=None):
        check(lib().pir_net_capture(self.h, dev_ptr))

    def autotune(self):
        """One eager forward with the limb-GEMM autotuner on (see include/pir.h)."""
        check(lib().pir_net_autotune(self.h))

    def replay(self):
        check(lib().pir_net_replay(self.h))

    def profile_layers(self):
        """Device milliseconds of each layer over one eager forward (HIP events)."""
        ms = np.zeros(self.num_layers, np.float32)
        check(lib().pir_net_profile_layers(self.h, ms.ctypes.data, ms.size))
        return ms

    def profile_gemm(self):
        """One eager forward with HIP events around every limb GEMM launch.

        Returns {variant: (kernel_ms, algorithmic_int8_ops, launches)} for the
        variants launched: A-operand form (0 dense row-major, 1 dense
        column-major, 2 conv, 3 conv with Cin % 32 == 0, 4 dgrad, 5 dgrad with
        Cin % 32 == 0, 6 wgrad), plus LIMB_WIDE for the 32x128 workgroup tile
        (limb_gemm_kernel<form, 1>; else <form, 2>)."""
        ms, ops = np.zeros(LIMB_VARIANTS, np.float64), np.zeros(LIMB_VARIANTS, np.float64)
        n = np.zeros(LIMB_VARIANTS, np.int32)
        check(lib().pir_net_profile_gemm(self.h, ms.ctypes.data, ops.ctypes.data, n.ctypes.data))
        return {v: (float(ms[v]), float(ops[v]), int(n[v])) for v in range(LIMB_VARIANTS) if n[v]}

    # ---- training ---------------------------------------------------------

    def set_log_learning_rate(self, log_lr):
        check(lib().pir_net_set_log_learning_rate(self.h, log_lr))

    def set_weight_shares(self, layer, w, b=None):
        """Weights (and FC biases) from share arrays (3, 2, n) uint64."""
        w = np.ascontiguousarray(w, np.uint64)
        bp = None
        if b is not None:
            b = np.ascontiguousarray(b, np.uint64)
            bp = b.ctypes.data
        check(lib().pir_net_set_weight_shares(self.h, layer, w.ctypes.data, bp))

    def backward(self, labels):
        """NeuralNetwork::backward after a forward: labels (batch x classes)."""
        labels = np.ascontiguousarray(labels, np.float64).ravel()
        check(lib().pir_net_backward(self.h, labels.ctypes.data, labels.size))

    def backward_shares(self, dev_ptr):
        check(lib().pir_net_backward_shares(self.h, dev_ptr))

    def layer_backward(self, layer, delta_ptr, input_ptr):
        """Layer::backward(delta, forwardInput) of one layer (device shares)."""
        check(lib().pir_net_layer_backward(self.h, layer, delta_ptr, input_ptr))

    def set_train_labels(self, labels):
        labels = np.ascontiguousarray(labels, np.float64).ravel()
        check(lib().pir_net_set_train_labels(self.h, labels.ctypes.data, labels.size))

    def train_step(self):
        """Forward from the input buffer + backward from the stored labels."""
        check(lib().pir_net_train_step(self.h))

    def capture_train(self):
        check(lib().pir_net_capture_train(self.h))

    def replay_train(self):
        check(lib().pir_net_replay_train(self.h))

    def autotune_train(self):
        check(lib().pir_net_autotune_train(self.h))

    def profile_train_gemm(self):
        """profile_gemm over one eager training step (forward + backward)."""
        ms, ops = np.zeros(LIMB_VARIANTS, np.float64), np.zeros(LIMB_VARIANTS, np.float64)
        n = np.zeros(LIMB_VARIANTS, np.int32)
        check(lib().pir_net_profile_train_gemm(self.h, ms.ctypes.data, ops.ctypes.data, n.ctypes.data))
        return {v: (float(ms[v]), float(ops[v]), int(n[v])) for v in range(LIMB_VARIANTS) if n[v]}

    def profile_train_layers(self):
        """Device ms of one eager training step: (forward per layer, softmax
        delta, backward per layer)."""
        L = self.num_layers
        ms = np.zeros(2 * L + 1, np.float32)
        check(lib().pir_net_profile_train_layers(self.h, ms.ctypes.data, ms.size))
        return ms[:L], float(ms[L]), ms[L + 1:]

    def delta(self, layer):
        """dL/d(input of layer) after a backward pass (layer == num_layers: the
        reveal-softmax delta at the network output)."""
        return self._read(lib().pir_net_delta, layer, np.uint64)

    def _read(self, fn, layer, dtype):
        p, n = _vp(), _sz()
        check(fn(self.h, layer, ctypes.byref(p), ctypes.byref(n)))
        self.synchronize()
        out = np.empty((3, 2, n.value), dtype)
        check(lib().pir_memcpy_d2h(out.ctypes.data, p.value, out.nbytes, None))
        return out

    def activation(self, layer):
        return self._read(lib().pir_net_activation, layer, np.uint64)

    def derivative(self, layer):
        return self._read(lib().pir_net_derivative, layer, np.uint8)

    def weights(self, layer):
        p, n, bp, bn = _vp(), _sz(), _vp(), _sz()
        check(lib().pir_net_weights(self.h, layer, ctypes.byref(p), ctypes.byref(n), ctypes.byref(bp),
                                    ctypes.byref(bn)))
        self.synchronize()
        w = np.empty((3, 2, n.value), np.uint64)
        check(lib().pir_memcpy_d2h(w.ctypes.data, p.value, w.nbytes, None))
        b = None
        if bp.value:
            b = np.empty((3, 2, bn.value), np.uint64)
            check(lib().pir_memcpy_d2h(b.ctypes.data, bp.value, b.nbytes, None))
        return w, b


def _params_of(net, layer):
    out = []
    for k in range(lib().pir_net_num_params(net.h, layer)):
        p, n = _vp(), _sz()
        check(lib().pir_net_param(net.h, layer, k, ctypes.byref(p), ctypes.byref(n)))
        name = ctypes.create_string_buffer(64)
        check(lib().pir_net_param_name(net.h, layer, k, name, 64))
        net.synchronize()
        a = np.empty((3, 2, n.value), np.uint64)
        check(lib().pir_memcpy_d2h(a.ctypes.data, p.value, a.nbytes, None))
        out.append((name.value.decode(), a))
    return out


def autotune_entries():
    return lib().pir_autotune_entries()


def autotune_clear():
    check(lib().pir_autotune_clear())


def autotune_save(path):
    check(lib().pir_autotune_save(os.fsencode(path)))


def autotune_load(path):
    check(lib().pir_autotune_load(os.fsencode(path)))
