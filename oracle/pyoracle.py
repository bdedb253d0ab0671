"""ctypes binding of the CPU oracle (oracle/apsm_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

Array conventions follow the reference's host API (cpp/lib/noma/types.cuh:69-100):
rx is [antenna][sample] complex, passed as float32 [N][S][2]; tx is [sample] complex,
passed as float32 [S][2]. Soft estimates come back as float64 [S][2].
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("gaussian_kernel_weight", ctypes.c_double),
        ("gaussian_kernel_variance", ctypes.c_double),
        ("eB", ctypes.c_double),
        ("norm_constraint", ctypes.c_double),
        ("window_size", ctypes.c_int32),
        ("sample_step", ctypes.c_int32),
        ("train_passes", ctypes.c_int32),
        ("dictionary_size", ctypes.c_int32),
        ("start_with_full_window", ctypes.c_int32),
        ("training_shuffle", ctypes.c_int32),
        ("lls_initialization", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("linear_kernel_weight", ctypes.c_double),
    ]


# Defaults of the reference factory (cpp/lib/apsm/apsm_noma_detector.cu:361-372).
DEFAULTS = dict(
    gaussianKernelWeight=0.5,
    gaussianKernelVariance=0.05,
    eB=0.001,
    normConstraint=0.0,
    windowSize=20,
    sampleStep=2,
    trainPasses=1,
    dictionarySize=685,
    startWithFullWindow=False,
    trainingShuffle=False,
    llsInitialization=False,
)


def make_params(cfg):
    c = dict(DEFAULTS)
    c.update({k: v for k, v in cfg.items() if k not in ("otype", "detectVersion")})
    wL = float(c.get("linearKernelWeight", 1.0 - float(c["gaussianKernelWeight"])))
    return OrcParams(
        float(c["gaussianKernelWeight"]), float(c["gaussianKernelVariance"]), float(c["eB"]),
        float(c["normConstraint"]), int(c["windowSize"]), int(c["sampleStep"]),
        int(c["trainPasses"]), int(c["dictionarySize"]), int(bool(c["startWithFullWindow"])),
        int(bool(c["trainingShuffle"])), int(bool(c["llsInitialization"])), 0, wL)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        P = ctypes.POINTER
        f32p, f64p = P(ctypes.c_float), P(ctypes.c_double)
        i64 = ctypes.c_int64
        L.orc_last_error.restype = ctypes.c_char_p
        for pre in ("orc", "cpu"):
            getattr(L, pre + "_apsm_create").restype = ctypes.c_void_p
            getattr(L, pre + "_apsm_create").argtypes = [P(OrcParams), ctypes.c_int]
            getattr(L, pre + "_apsm_destroy").argtypes = [ctypes.c_void_p]
            getattr(L, pre + "_apsm_detect").argtypes = [ctypes.c_void_p, f32p, i64, f64p]
        L.orc_apsm_train.argtypes = [ctypes.c_void_p, f32p, f32p, i64, ctypes.c_int]
        L.cpu_apsm_train.argtypes = [ctypes.c_void_p, f32p, f32p, i64]
        L.orc_apsm_reset.argtypes = [ctypes.c_void_p]
        L.orc_apsm_state.restype = i64
        L.orc_apsm_state.argtypes = [ctypes.c_void_p, f64p, f64p, i64, P(i64), f64p]
        L.orc_apsm_eval_rows.argtypes = [ctypes.c_void_p, f64p, i64, f64p]
        L.orc_apsm_norm_ratio.restype = ctypes.c_double
        L.orc_apsm_norm_ratio.argtypes = [ctypes.c_void_p]
        L.orc_apsm_set_state.argtypes = [ctypes.c_void_p, f64p, f64p, f64p, i64, ctypes.c_double]
        L.orc_lls_train.argtypes = [f32p, f32p, ctypes.c_int, i64, ctypes.c_int, f64p]
        L.orc_lls_detect.argtypes = [f32p, ctypes.c_int, i64, f64p, f64p]
        L.orc_shuffle_perm.argtypes = [i64, P(i64)]
        L.orc_variance_heuristic.restype = ctypes.c_float
        L.orc_variance_heuristic.argtypes = [f32p, ctypes.c_int, i64]
        L.orc_mod_scale.restype = ctypes.c_double
        L.orc_mod_map.argtypes = [ctypes.c_int, ctypes.c_uint32, f64p, f64p]
        L.orc_mod_demap.restype = ctypes.c_uint32
        L.orc_mod_demap.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_symbol_errors.restype = i64
        L.orc_symbol_errors.argtypes = [f32p, f32p, i64, ctypes.c_int, ctypes.c_int]
        L.orc_bit_errors.restype = i64
        L.orc_bit_errors.argtypes = [f32p, f32p, i64, ctypes.c_int]
        L.orc_evm.restype = ctypes.c_double
        L.orc_evm.argtypes = [f32p, f32p, i64]
        L.orc_multi_create.restype = ctypes.c_void_p
        L.orc_multi_create.argtypes = [P(OrcParams), ctypes.c_int, ctypes.c_int]
        L.orc_multi_destroy.argtypes = [ctypes.c_void_p]
        L.orc_multi_train.argtypes = [ctypes.c_void_p, f32p, f32p, i64]
        L.orc_multi_detect.argtypes = [ctypes.c_void_p, f32p, i64, f64p]
        L.orc_multi_state.restype = i64
        L.orc_multi_state.argtypes = [ctypes.c_void_p, f64p, f64p, i64, P(i64), f64p]
        L.orc_omp_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def as_rx(rx):
    """complex [N][S] -> float32 [N][S][2]"""
    rx = np.asarray(rx)
    if np.iscomplexobj(rx):
        rx = np.stack([rx.real, rx.imag], axis=-1)
    return np.ascontiguousarray(rx, dtype=np.float32)


def as_tx(tx):
    tx = np.asarray(tx)
    if np.iscomplexobj(tx):
        tx = np.stack([tx.real, tx.imag], axis=-1)
    return np.ascontiguousarray(tx, dtype=np.float32)


class Apsm:
    """FP64 (fp32=False) or FP32 (fp32=True) restatement of one APSM detector."""

    def __init__(self, cfg, num_antennas, fp32=False):
        self.L = lib()
        self.pre = "cpu" if fp32 else "orc"
        self.N = num_antennas
        self.params = make_params(cfg)
        self.h = getattr(self.L, self.pre + "_apsm_create")(ctypes.byref(self.params), num_antennas)
        if not self.h:
            raise ValueError(self.L.orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            getattr(self.L, self.pre + "_apsm_destroy")(self.h)
            self.h = None

    def train(self, rx, tx, incremental=False):
        rx, prx = _f32(as_rx(rx))
        tx, ptx = _f32(as_tx(tx))
        assert rx.shape[0] == self.N and rx.shape[1] == tx.shape[0]
        if self.pre == "cpu":
            rc = self.L.cpu_apsm_train(self.h, prx, ptx, rx.shape[1])
        else:
            rc = self.L.orc_apsm_train(self.h, prx, ptx, rx.shape[1], int(incremental))
        if rc != 0:
            raise ValueError(self.L.orc_last_error().decode())

    def detect(self, rx):
        rx, prx = _f32(as_rx(rx))
        est = np.zeros((rx.shape[1], 2), np.float64)
        getattr(self.L, self.pre + "_apsm_detect")(self.h, prx, rx.shape[1],
                                                   est.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return est

    def reset(self):
        assert self.pre == "orc"
        self.L.orc_apsm_reset(self.h)

    def state(self, cap=1 << 20):
        theta = np.zeros(2 * self.N)
        coeffs = np.zeros(cap)
        lo = ctypes.c_int64()
        var = ctypes.c_double()
        used = self.L.orc_apsm_state(self.h, theta.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), cap,
                                     ctypes.byref(lo), ctypes.byref(var))
        n = min(cap, used - lo.value)
        return dict(used=used, lo=lo.value, theta=theta, coeffs=coeffs[:n].copy(), var=var.value)

    def set_state(self, rows, coeffs, theta, var):
        rows, prows = _f64(rows)
        coeffs, pc = _f64(coeffs)
        theta, pt = _f64(theta)
        rc = self.L.orc_apsm_set_state(self.h, prows, pc, pt, rows.shape[0], float(var))
        if rc != 0:
            raise ValueError(self.L.orc_last_error().decode())

    def norm_ratio(self):
        return self.L.orc_apsm_norm_ratio(self.h)

    def eval_rows(self, u):
        u, pu = _f64(u)
        out = np.zeros(u.shape[0])
        self.L.orc_apsm_eval_rows(self.h, pu, u.shape[0], out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out


class ApsmMulti:
    """FP64 streamed form for U detectors sharing one pilot stream (the window Gram matrix is
    kept between steps; kernel values are shared by the users). Each user's result equals
    Apsm(cfg, N).train(rx, tx[u], incremental=True). tx: complex [U][T]; detect returns
    float64 [U][M][2]."""

    def __init__(self, cfg, num_antennas, num_users):
        self.L = lib()
        self.N, self.U = num_antennas, num_users
        self.params = make_params(cfg)
        self.h = self.L.orc_multi_create(ctypes.byref(self.params), num_antennas, num_users)
        if not self.h:
            raise ValueError(self.L.orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_multi_destroy(self.h)
            self.h = None

    def train(self, rx, tx):
        rx, prx = _f32(as_rx(rx))
        tx, ptx = _f32(as_tx(np.atleast_2d(tx)))
        assert rx.shape[0] == self.N and tx.shape[0] == self.U and rx.shape[1] == tx.shape[1]
        if self.L.orc_multi_train(self.h, prx, ptx, rx.shape[1]) != 0:
            raise ValueError(self.L.orc_last_error().decode())

    def detect(self, rx):
        rx, prx = _f32(as_rx(rx))
        est = np.zeros((self.U, rx.shape[1], 2), np.float64)
        self.L.orc_multi_detect(self.h, prx, rx.shape[1], est.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return est

    def state(self, cap=1 << 16):
        theta = np.zeros((self.U, 2 * self.N))
        coeffs = np.zeros((self.U, cap))
        lo = ctypes.c_int64()
        var = ctypes.c_double()
        used = self.L.orc_multi_state(self.h, theta.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), cap,
                                      ctypes.byref(lo), ctypes.byref(var))
        n = min(cap, used - lo.value)
        return dict(used=used, lo=lo.value, theta=theta, coeffs=coeffs[:, :n].copy(), var=var.value)


def lls_train(rx, tx, method="cholesky"):
    rx, prx = _f32(as_rx(rx))
    tx, ptx = _f32(as_tx(tx))
    w = np.zeros(2 * rx.shape[0])
    rc = lib().orc_lls_train(prx, ptx, rx.shape[0], rx.shape[1], 0 if method == "cholesky" else 1,
                             w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if rc != 0:
        raise ValueError("LLS solve failed")
    return w


def lls_detect(rx, w):
    rx, prx = _f32(as_rx(rx))
    w, pw = _f64(w)
    est = np.zeros((rx.shape[1], 2))
    lib().orc_lls_detect(prx, rx.shape[0], rx.shape[1], pw, est.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return est


def shuffle_perm(T):
    p = np.zeros(T, np.int64)
    lib().orc_shuffle_perm(T, p.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return p


def variance_heuristic(rx):
    rx, prx = _f32(as_rx(rx))
    return float(lib().orc_variance_heuristic(prx, rx.shape[0], rx.shape[1]))


def mod_scale(bps):
    return lib().orc_mod_scale(bps)


def mod_map(bps, bits):
    re, im = ctypes.c_double(), ctypes.c_double()
    lib().orc_mod_map(bps, bits, ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def mod_demap(bps, z):
    return lib().orc_mod_demap(bps, float(np.real(z)), float(np.imag(z)))


def symbol_errors(est, ref, bps, cli_compat=False):
    est, pe = _f32(as_tx(est))
    ref, pr = _f32(as_tx(ref))
    return lib().orc_symbol_errors(pe, pr, est.shape[0], bps, int(cli_compat))


def bit_errors(est, ref, bps):
    est, pe = _f32(as_tx(est))
    ref, pr = _f32(as_tx(ref))
    return lib().orc_bit_errors(pe, pr, est.shape[0], bps)


def evm(est, ref):
    est, pe = _f32(as_tx(est))
    ref, pr = _f32(as_tx(ref))
    return lib().orc_evm(pe, pr, est.shape[0])


def set_threads(n):
    lib().orc_set_threads(int(n))


def threads():
    return lib().orc_omp_threads()
