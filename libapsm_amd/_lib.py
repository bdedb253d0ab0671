"""ctypes loader for libapsm_amd.so (the C ABI in include/apsm_capi.h)."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "libapsm_amd.so")

APSM_OK = 0
APSM_ERR_ARG = -1
APSM_ERR_DEVICE = -2
APSM_ERR_NUMERIC = -3
APSM_ERR_UNSUPPORTED = -4


class ApsmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class ApsmParamsC(ctypes.Structure):
    _fields_ = [
        ("gaussian_kernel_weight", ctypes.c_float),
        ("gaussian_kernel_variance", ctypes.c_float),
        ("eB", ctypes.c_float),
        ("norm_constraint", ctypes.c_float),
        ("window_size", ctypes.c_uint32),
        ("sample_step", ctypes.c_uint32),
        ("train_passes", ctypes.c_uint32),
        ("dictionary_size", ctypes.c_uint32),
        ("start_with_full_window", ctypes.c_uint8),
        ("lls_initialization", ctypes.c_uint8),
        ("training_shuffle", ctypes.c_uint8),
        ("reserved", ctypes.c_uint8),
        ("detect_version", ctypes.c_int32),
        ("linear_kernel_weight", ctypes.c_float),
        ("detect_precision", ctypes.c_int32),
    ]


class ServerConfigC(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_uint64), ("frame_arena", ctypes.c_uint64), ("batch_max_frames", ctypes.c_uint32),
                ("detects_in_flight", ctypes.c_uint32), ("batch_max_bytes", ctypes.c_uint64)]


class FrameC(ctypes.Structure):
    _fields_ = [
        ("arrival_s", ctypes.c_double),
        ("link", ctypes.c_uint32),
        ("n_data", ctypes.c_uint32),
        ("n_pilots", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("rx_data", ctypes.c_void_p),
        ("rx_pilots", ctypes.c_void_p),
        ("tx_pilots", ctypes.c_void_p),
        ("est", ctypes.c_void_p),
    ]


class FrameResultC(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("reserved", ctypes.c_uint32),
        ("ticket", ctypes.c_uint64),
        ("arrival_s", ctypes.c_double),
        ("admit_s", ctypes.c_double),
        ("output_s", ctypes.c_double),
        ("trained_s", ctypes.c_double),
        ("detect_round", ctypes.c_int64),
        ("last_round", ctypes.c_int64),
    ]


# Every symbol include/apsm_capi.h declares, with (restype, argtypes).
_P = ctypes.POINTER
_vp = ctypes.c_void_p
_f32p = _P(ctypes.c_float)
_u32 = ctypes.c_uint32
SIGNATURES = {
    "apsm_params_default": (None, [_P(ApsmParamsC)]),
    "apsm_last_error": (ctypes.c_char_p, []),
    "apsm_device_count": (ctypes.c_int, []),
    "apsm_create": (ctypes.c_int, [_P(ApsmParamsC), _u32, _u32, _P(_vp)]),
    "apsm_destroy": (None, [_vp]),
    "apsm_reset": (ctypes.c_int, [_vp]),
    "apsm_train": (ctypes.c_int, [_vp, _vp, _vp, _u32, _f32p]),
    "apsm_detect": (ctypes.c_int, [_vp, _vp, _u32, _vp, _f32p]),
    "apsm_train_dev": (ctypes.c_int, [_vp, _vp, _vp, _u32]),
    "apsm_detect_dev": (ctypes.c_int, [_vp, _vp, _u32, _vp]),
    "apsm_kernel_times": (ctypes.c_int, [_vp, _f32p, _f32p]),
    "apsm_profile_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lls_train_lowlevel_noalloc": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp, _vp, _vp]),
    "apsm_profile_read": (ctypes.c_int, [_vp, _P(ctypes.c_double), _P(ctypes.c_uint64)]),
    "apsm_synchronize": (ctypes.c_int, [_vp]),
    "apsm_stream": (_vp, [_vp]),
    "apsm_get_state": (ctypes.c_int, [_vp, _P(ctypes.c_uint64), _P(ctypes.c_uint32), _f32p, _vp, _vp, _u32]),
    "lls_create": (ctypes.c_int, [_u32, _P(_vp)]),
    "lls_destroy": (None, [_vp]),
    "lls_train": (ctypes.c_int, [_vp, _vp, _vp, _u32, _f32p]),
    "lls_detect": (ctypes.c_int, [_vp, _vp, _u32, _vp, _f32p]),
    "lls_train_dev": (ctypes.c_int, [_vp, _vp, _vp, _u32]),
    "lls_detect_dev": (ctypes.c_int, [_vp, _vp, _u32, _vp]),
    "lls_get_weights": (ctypes.c_int, [_vp, _vp]),
    "lls_synchronize": (ctypes.c_int, [_vp]),
    "apsm_get_dictionary": (ctypes.c_int, [_vp, _vp, _u32, _P(ctypes.c_uint32)]),
    "apsm_norm_ratio": (ctypes.c_int, [_vp, _u32, _f32p]),
    "apsm_set_detect_precision": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "apsm_server_create": (ctypes.c_int, [_P(ServerConfigC), _P(_vp)]),
    "apsm_server_destroy": (None, [_vp]),
    "apsm_server_add_link": (ctypes.c_int, [_vp, _P(ApsmParamsC), _u32, _u32, _u32, _P(ctypes.c_uint32)]),
    "apsm_server_link": (_vp, [_vp, _u32]),
    "apsm_server_submit": (ctypes.c_int, [_vp, _P(FrameC), _P(ctypes.c_uint64)]),
    "apsm_server_poll": (ctypes.c_int, [_vp, ctypes.c_int, _P(FrameResultC), _u32, _P(ctypes.c_uint32)]),
    "apsm_server_replay": (ctypes.c_int, [_vp, _P(FrameC), _u32, _P(FrameResultC)]),
    "apsm_server_stats": (ctypes.c_int, [_vp, _P(ctypes.c_uint64), _P(ctypes.c_uint64), _P(ctypes.c_uint64)]),
    "apsm_host_malloc": (ctypes.c_int, [_P(_vp), ctypes.c_uint64]),
    "apsm_host_free": (ctypes.c_int, [_vp]),
    "apsm_train_history": (ctypes.c_int, [_vp, _vp, _vp, _u32, _vp, _vp, _u32, _u32, ctypes.c_float, _vp, _u32,
                                          _P(ctypes.c_uint32), _f32p]),
    "lls_create_on_stream": (ctypes.c_int, [_u32, _vp, _P(_vp)]),
    "lls_train_lowlevel": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "lls_detect_lowlevel": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_float, ctypes.c_float]),
    "lls_signal_power": (ctypes.c_int, [_vp, _vp, _u32, _f32p]),
    "lls_weights_dev": (_vp, [_vp]),
    "lls_stream": (_vp, [_vp]),
    "apsm_device_malloc": (ctypes.c_int, [_P(_vp), ctypes.c_uint64]),
    "apsm_device_free": (ctypes.c_int, [_vp]),
    "apsm_memcpy_h2d": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64]),
    "apsm_memcpy_d2h": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64]),
    "apsm_device_synchronize": (ctypes.c_int, []),
    "apsm_batch_create": (ctypes.c_int, [_u32, _P(ApsmParamsC), _P(_u32), _P(_u32), _P(_vp)]),
    "apsm_batch_destroy": (None, [_vp]),
    "apsm_batch_plan": (ctypes.c_int, [_vp, _P(_vp), _P(_vp), _P(_u32), _P(_vp), _P(_u32), _P(_vp), ctypes.c_int]),
    "apsm_batch_run": (ctypes.c_int, [_vp]),
    "apsm_batch_synchronize": (ctypes.c_int, [_vp]),
    "apsm_batch_last_ms": (ctypes.c_int, [_vp, _f32p]),
    "apsm_batch_last_split": (ctypes.c_int, [_vp, _f32p, _f32p]),
    "apsm_batch_instance": (_vp, [_vp, _u32]),
    "apsm_batch_tuning": (ctypes.c_int, [_vp, _P(ctypes.c_int32), _u32, _P(_u32)]),
    "apsm_stream_create": (ctypes.c_int, [_P(_vp)]),
    "apsm_stream_destroy": (ctypes.c_int, [_vp]),
    "apsm_stream_synchronize": (ctypes.c_int, [_vp]),
}

_lib = None


def lib():
    """Load libapsm_amd.so; raises if it has not been built (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise ImportError(
                f"{library_path} is missing: build it with `make -C libapsm_amd/csrc` "
                "(or __graft_entry__.build())")
        L = ctypes.CDLL(library_path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != APSM_OK:
        raise ApsmError(rc, lib().apsm_last_error().decode())


class PinnedArray:
    """A numpy array in pinned host memory of the library's HIP runtime (so that the device
    copies of frames overlap the device work). .array is the numpy view."""

    def __init__(self, shape, dtype):
        import numpy as np
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = max(1, int(np.prod(self.shape)) * self.dtype.itemsize)
        self._p = ctypes.c_void_p()
        check(lib().apsm_host_malloc(ctypes.byref(self._p), self.nbytes))
        buf = (ctypes.c_char * self.nbytes).from_address(self._p.value)
        self.array = np.frombuffer(buf, dtype=self.dtype, count=int(np.prod(self.shape))).reshape(self.shape)

    @property
    def ptr(self):
        return self._p.value or 0

    def __del__(self):
        p = getattr(self, "_p", None)
        if p is not None and p.value and _lib is not None:
            self.array = None
            lib().apsm_host_free(p)
            self._p = ctypes.c_void_p()


class DeviceArray:
    """A device allocation of the library's HIP runtime holding one numpy array's bytes.

    DeviceArray(a) uploads a; DeviceArray.empty(shape, dtype) allocates; .ptr is the device
    address; .numpy() downloads (synchronous)."""

    def __init__(self, a=None, shape=None, dtype=None):
        import numpy as np
        if a is not None:
            a = np.ascontiguousarray(a)
            shape, dtype = a.shape, a.dtype
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._p = ctypes.c_void_p()
        check(lib().apsm_device_malloc(ctypes.byref(self._p), self.nbytes))
        if a is not None:
            check(lib().apsm_memcpy_h2d(self._p, a.ctypes.data, self.nbytes))

    @classmethod
    def empty(cls, shape, dtype):
        return cls(shape=shape, dtype=dtype)

    @property
    def ptr(self):
        return self._p.value or 0

    def numpy(self):
        import numpy as np
        out = np.empty(self.shape, self.dtype)
        check(lib().apsm_memcpy_d2h(out.ctypes.data, self._p, self.nbytes))
        return out

    def upload(self, a):
        import numpy as np
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().apsm_memcpy_h2d(self._p, a.ctypes.data, self.nbytes))

    def __del__(self):
        p = getattr(self, "_p", None)
        if p is not None and p.value and _lib is not None:
            _lib.apsm_device_free(p)
            self._p = ctypes.c_void_p()
