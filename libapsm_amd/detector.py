"""Detector classes mirroring the reference API, over the C ABI of libapsm_amd.so.

Reference interfaces mirrored (paths relative to the reference repository):
  NomaDetector            cpp/lib/noma/noma_detector.cuh:65-75
  NomaDetectorFactory     cpp/lib/noma/noma_detector_factory.cuh:70-74, .cu:57-65
  ApsmNomaDetector        cpp/lib/apsm/apsm_noma_detector.cuh:83-109
  ApsmNomaDetectorFactory cpp/lib/apsm/apsm_noma_detector.cu:348-383 (keys, defaults)
  LlsNomaDetector         cpp/lib/noma/lls_detector.cuh:75-99
  apsm_parameters         cpp/lib/apsm/apsm_parameters.cuh:68-86
  apsm_versions           cpp/lib/apsm/apsm_versions.cuh:75-105

Differences forced by Python: detect() returns the estimate (the reference fills an
output vector and returns the time); the time of the last call is kept in
`last_train_ms` / `last_detect_ms`. Configuration errors raise ValueError where the
reference throws std::invalid_argument.
"""
import ctypes

import numpy as np

from ._lib import ApsmError, ApsmParamsC, check, lib

APSM_VERSIONS = {
    "original": 0x10001,
    "oldfast": 0x10002,
    "shmem": 0x10006,
    "balanced": 0x10007,
}

# detectPrecision values (an extension of the reference configuration: arithmetic of the
# detect cross terms; everything else stays FP32).
APSM_PRECISIONS = {
    "fp32": 0,
    "fp32-vector": 1,
    "fp32-matrix": 2,
    "bf16": 3,
    "fp16": 4,
}

_VERSION_STRINGS = {
    0x10001: "APSM Detect original",
    0x10002: "APSM Detect old fast code",
    0x10006: "APSM Detect Store vectors in shared memory",
    0x10007: "APSM Detect Balance memory and computation",
}


def apsm_get_version_string(version_id):
    return _VERSION_STRINGS.get(version_id, "Unknown")


class apsm_parameters:  # noqa: N801 - reference class name
    """Parameter set with the reference factory defaults."""

    def __init__(self):
        self.gaussianKernelWeight = 0.5
        self.linearKernelWeight = 0.5
        self.gaussianKernelVariance = 0.05
        self.windowSize = 20
        self.sampleStep = 2
        self.trainPasses = 1
        self.startWithFullWindow = False
        self.llsInitialization = False
        self.trainingShuffle = False
        self.dictionarySize = 685
        self.normConstraint = 0.0
        self.eB = 0.001
        self.detectVersion = APSM_VERSIONS["balanced"]
        self.detectPrecision = APSM_PRECISIONS["fp32"]

    def to_c(self):
        return ApsmParamsC(
            float(self.gaussianKernelWeight), float(self.gaussianKernelVariance), float(self.eB),
            float(self.normConstraint), int(self.windowSize), int(self.sampleStep), int(self.trainPasses),
            int(self.dictionarySize), int(bool(self.startWithFullWindow)), int(bool(self.llsInitialization)),
            int(bool(self.trainingShuffle)), 0, int(self.detectVersion), float(self.linearKernelWeight),
            int(self.detectPrecision))


def _rx_array(rx, num_antennas=None):
    a = np.ascontiguousarray(np.asarray(rx), dtype=np.complex64)
    if a.ndim != 2:
        raise ValueError("received signal must be [antenna][sample]")
    if num_antennas is not None and a.shape[0] != num_antennas:
        raise ValueError(f"expected {num_antennas} antennas, got {a.shape[0]}")
    return a


def _tx_array(tx, num_users, T):
    a = np.ascontiguousarray(np.asarray(tx), dtype=np.complex64)
    if a.ndim == 1:
        a = a[None, :]
    if a.shape != (num_users, T):
        raise ValueError(f"pilots must be [{num_users}][{T}], got {list(a.shape)}")
    return a


class NomaTrainHistory:
    """Per-step EVM / SER / BER on the training and validation sets
    (cpp/lib/noma/noma_train_history.cuh:64-99)."""

    def __init__(self, bitPerSymbol, modulationScale):
        self.bitPerSymbol = int(bitPerSymbol)
        self.modulationScale = float(modulationScale)
        self.evmTrain, self.serTrain, self.berTrain = [], [], []
        self.evmVal, self.serVal, self.berVal = [], [], []

    def extend(self, rows):
        for r in np.asarray(rows, np.float32):
            self.evmTrain.append(float(r[0]))
            self.serTrain.append(float(r[1]))
            self.berTrain.append(float(r[2]))
            self.evmVal.append(float(r[3]))
            self.serVal.append(float(r[4]))
            self.berVal.append(float(r[5]))

    def toCsv(self, filename):
        with open(filename, "w") as f:
            f.write("step,train_evm,train_ser,train_ber,val_evm,val_ser,val_ber\n")
            for i in range(len(self.evmTrain)):
                f.write(f"{i},{self.evmTrain[i]:g},{self.serTrain[i]:g},{self.berTrain[i]:g},"
                        f"{self.evmVal[i]:g},{self.serVal[i]:g},{self.berVal[i]:g}\n")


class NomaDetector:
    def train(self, rxSigTraining, txSigTraining, valSet=None):
        raise NotImplementedError

    def detect(self, rxSigData):
        raise NotImplementedError

    def resetState(self):
        raise NotImplementedError


class ApsmNomaDetector(NomaDetector):
    """APSM detector; num_users > 1 trains several users' detectors on the same rx in one
    handle (they share the Gaussian dictionary, which is made of the received pilots)."""

    def __init__(self, par, numAntennas, numUsers=1):
        self.par = par
        self.numAntennas = int(numAntennas)
        self.numUsers = int(numUsers)
        self._h = ctypes.c_void_p()
        L = lib()
        cpar = par.to_c()
        check(L.apsm_create(ctypes.byref(cpar), self.numAntennas, self.numUsers, ctypes.byref(self._h)))
        self.last_train_ms = 0.0
        self.last_detect_ms = 0.0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().apsm_destroy(h)
            self._h = ctypes.c_void_p()

    @property
    def handle(self):
        return self._h

    def train(self, rxSigTraining, txSigTraining, valSet=None):
        """valSet = (NomaTrainHistory, rx_val, tx_val) records per-step metrics
        (cpp/lib/apsm/apsm_train.cu:147-158); with several users the history holds user 0."""
        rx = _rx_array(rxSigTraining, self.numAntennas)
        tx = _tx_array(txSigTraining, self.numUsers, rx.shape[1])
        ms = ctypes.c_float()
        if valSet is None:
            check(lib().apsm_train(self._h, rx.ctypes.data, tx.ctypes.data, rx.shape[1], ctypes.byref(ms)))
        else:
            hist, rxv, txv = valSet
            rxv = _rx_array(rxv, self.numAntennas)
            txv = _tx_array(txv, self.numUsers, rxv.shape[1])
            per_user = self.train_history(rx, tx, rxv, txv, hist.bitPerSymbol, hist.modulationScale, ms)
            hist.extend(per_user[:, 0, :])
        self.last_train_ms = ms.value
        return ms.value

    def train_history(self, rx, tx, rxv, txv, bits_per_symbol, modulation_scale, ms=None):
        """Per-step metrics [steps][users][6] = train EVM, SER, BER, validation EVM, SER, BER."""
        rx = _rx_array(rx, self.numAntennas)
        tx = _tx_array(tx, self.numUsers, rx.shape[1])
        rxv = _rx_array(rxv, self.numAntennas)
        txv = _tx_array(txv, self.numUsers, rxv.shape[1])
        cap = rx.shape[1] + 1
        out = np.zeros((cap, self.numUsers, 6), np.float32)
        n = ctypes.c_uint32()
        ms = ms if ms is not None else ctypes.c_float()
        check(lib().apsm_train_history(self._h, rx.ctypes.data, tx.ctypes.data, rx.shape[1], rxv.ctypes.data,
                                       txv.ctypes.data, rxv.shape[1], int(bits_per_symbol), float(modulation_scale),
                                       out.ctypes.data, cap, ctypes.byref(n), ctypes.byref(ms)))
        return out[:n.value]

    def computeNormRatio(self, user=0):
        r = ctypes.c_float()
        check(lib().apsm_norm_ratio(self._h, int(user), ctypes.byref(r)))
        return r.value

    def getDictionary(self):
        """Complex basis vectors of the live dictionary, [N][entries], oldest first."""
        n = ctypes.c_uint32()
        check(lib().apsm_get_dictionary(self._h, None, 0, ctypes.byref(n)))
        d = np.zeros((self.numAntennas, max(1, n.value)), np.complex64)
        check(lib().apsm_get_dictionary(self._h, d.ctypes.data, max(1, n.value), ctypes.byref(n)))
        return d[:, :n.value]

    def detect(self, rxSigData):
        rx = _rx_array(rxSigData, self.numAntennas)
        M = rx.shape[1]
        est = np.empty((self.numUsers, M), np.complex64)
        ms = ctypes.c_float()
        check(lib().apsm_detect(self._h, rx.ctypes.data, M, est.ctypes.data, ctypes.byref(ms)))
        self.last_detect_ms = ms.value
        return est[0] if self.numUsers == 1 else est

    def setDetectPrecision(self, precision):
        """Detect arithmetic for the following detects ("fp32", "fp32-vector", "fp32-matrix",
        "bf16", "fp16"); the trained state is kept."""
        if precision not in APSM_PRECISIONS:
            raise ValueError("Unknown detect precision: " + str(precision))
        check(lib().apsm_set_detect_precision(self._h, APSM_PRECISIONS[precision]))

    def train_device(self, d_rx, d_tx, T):
        """Device pointers (ints) in the host layouts; enqueued on the handle's stream."""
        check(lib().apsm_train_dev(self._h, ctypes.c_void_p(d_rx), ctypes.c_void_p(d_tx), int(T)))

    def detect_device(self, d_rx, M, d_est):
        check(lib().apsm_detect_dev(self._h, ctypes.c_void_p(d_rx), int(M), ctypes.c_void_p(d_est)))

    def synchronize(self):
        check(lib().apsm_synchronize(self._h))

    def kernel_times(self):
        """(train sequencer ms, detect ms) of the latest launches, from HIP events."""
        t, d = ctypes.c_float(), ctypes.c_float()
        check(lib().apsm_kernel_times(self._h, ctypes.byref(t), ctypes.byref(d)))
        return t.value, d.value

    PROFILE_KINDS = ("sequencer", "block_pre", "detect")

    def profile_enable(self, enable=True):
        """Bracket every sequencer, block-precompute and detect launch with HIP events."""
        check(lib().apsm_profile_enable(self._h, int(bool(enable))))

    def profile_read(self):
        """{kind: (device ms summed, launches)} since the previous read (waits for the events)."""
        ms = (ctypes.c_double * 3)()
        n = (ctypes.c_uint64 * 3)()
        check(lib().apsm_profile_read(self._h, ms, n))
        return {k: (ms[i], n[i]) for i, k in enumerate(self.PROFILE_KINDS)}

    def stream(self):
        return lib().apsm_stream(self._h)

    def resetState(self):
        check(lib().apsm_reset(self._h))

    def getTrainingState(self):
        """dict(used_rows, dict_entries, variance, theta [U][2N], coeffs [U][2*dict])."""
        used = ctypes.c_uint64()
        nd = ctypes.c_uint32()
        var = ctypes.c_float()
        check(lib().apsm_get_state(self._h, ctypes.byref(used), ctypes.byref(nd), ctypes.byref(var), None, None, 0))
        theta = np.zeros((self.numUsers, 2 * self.numAntennas), np.float32)
        coeffs = np.zeros((self.numUsers, max(1, 2 * nd.value)), np.float32)
        check(lib().apsm_get_state(self._h, None, None, None, theta.ctypes.data, coeffs.ctypes.data, coeffs.shape[1]))
        return dict(used_rows=used.value, dict_entries=nd.value, variance=var.value, theta=theta,
                    coeffs=coeffs[:, :2 * nd.value])


class ApsmNomaDetectorFactory:
    VALID_KEYS = ("otype", "llsInitialization", "trainingShuffle", "gaussianKernelWeight",
                  "gaussianKernelVariance", "windowSize", "sampleStep", "trainPasses",
                  "startWithFullWindow", "dictionarySize", "normConstraint", "eB", "detectVersion",
                  "detectPrecision")

    @staticmethod
    def params_from_config(config):
        for key in config:
            if key not in ApsmNomaDetectorFactory.VALID_KEYS:
                raise ValueError("Invalid key: " + key)
        par = apsm_parameters()
        par.llsInitialization = bool(config.get("llsInitialization", False))
        par.trainingShuffle = bool(config.get("trainingShuffle", False))
        par.gaussianKernelWeight = float(config.get("gaussianKernelWeight", 0.5))
        par.linearKernelWeight = 1.0 - par.gaussianKernelWeight
        par.gaussianKernelVariance = float(config.get("gaussianKernelVariance", 0.05))
        par.windowSize = int(config.get("windowSize", 20))
        par.sampleStep = int(config.get("sampleStep", 2))
        par.trainPasses = int(config.get("trainPasses", 1))
        par.startWithFullWindow = bool(config.get("startWithFullWindow", False))
        par.dictionarySize = int(config.get("dictionarySize", 685))
        par.normConstraint = float(config.get("normConstraint", 0.0))
        par.eB = float(config.get("eB", 0.001))
        version = config.get("detectVersion", "balanced")
        if version not in APSM_VERSIONS:
            raise ValueError("Unknown detect version: " + str(version))
        par.detectVersion = APSM_VERSIONS[version]
        precision = config.get("detectPrecision", "fp32")
        if precision not in APSM_PRECISIONS:
            raise ValueError("Unknown detect precision: " + str(precision))
        par.detectPrecision = APSM_PRECISIONS[precision]
        return par

    @staticmethod
    def build(config, numAntennas, numUsers=1):
        return ApsmNomaDetector(ApsmNomaDetectorFactory.params_from_config(config), numAntennas, numUsers)


class LlsNomaDetector(NomaDetector):
    """Least-squares linear detector (FP32 normal equations, FP64 Cholesky, on the GPU)."""

    def __init__(self, numAntennas):
        self.numAntennas = int(numAntennas)
        self._h = ctypes.c_void_p()
        check(lib().lls_create(self.numAntennas, ctypes.byref(self._h)))
        self.last_train_ms = 0.0
        self.last_detect_ms = 0.0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().lls_destroy(h)
            self._h = ctypes.c_void_p()

    def train(self, rxSigTraining, txSigTraining, valSet=None):
        rx = _rx_array(rxSigTraining, self.numAntennas)
        tx = _tx_array(txSigTraining, 1, rx.shape[1])
        ms = ctypes.c_float()
        check(lib().lls_train(self._h, rx.ctypes.data, tx.ctypes.data, rx.shape[1], ctypes.byref(ms)))
        self.last_train_ms = ms.value
        return ms.value

    def detect(self, rxSigData):
        rx = _rx_array(rxSigData, self.numAntennas)
        est = np.empty(rx.shape[1], np.complex64)
        ms = ctypes.c_float()
        check(lib().lls_detect(self._h, rx.ctypes.data, rx.shape[1], est.ctypes.data, ctypes.byref(ms)))
        self.last_detect_ms = ms.value
        return est

    def train_device(self, d_rx, d_tx, T):
        check(lib().lls_train_dev(self._h, ctypes.c_void_p(d_rx), ctypes.c_void_p(d_tx), int(T)))

    def detect_device(self, d_rx, M, d_est):
        check(lib().lls_detect_dev(self._h, ctypes.c_void_p(d_rx), int(M), ctypes.c_void_p(d_est)))

    def synchronize(self):
        check(lib().lls_synchronize(self._h))

    def getWeightMat(self):
        w = np.zeros(128, np.float32)
        check(lib().lls_get_weights(self._h, w.ctypes.data))
        return w[:self._nW]

    _nW_override = None

    @property
    def _nW(self):
        return self._nW_override or 2 * self.numAntennas

    def trainLowLevel(self, rxSigMat, txSigMat, numInputDims, numOutputDims, numSamples):
        """Device pointers: column-major A (numSamples x numInputDims) and y (numSamples)."""
        check(lib().lls_train_lowlevel(self._h, ctypes.c_void_p(rxSigMat), ctypes.c_void_p(txSigMat),
                                       int(numInputDims), int(numOutputDims), int(numSamples)))
        self._nW_override = int(numInputDims)

    def detectLowLevel(self, rxSigMat, estMat, numInputDims, numOutputDims, numSamples, alpha=1.0, beta=0.0):
        check(lib().lls_detect_lowlevel(self._h, ctypes.c_void_p(rxSigMat), ctypes.c_void_p(estMat), int(numInputDims),
                                        int(numOutputDims), int(numSamples), float(alpha), float(beta)))

    @staticmethod
    def trainLowLevelNoAlloc(A, y, w, status, numInputDims, numOutputDims, numSamples, stream=None):
        """trainLowLevel without a handle and without device allocation (fixed per-device
        scratch): device pointers A (column-major numSamples x numInputDims), y (numSamples),
        outputs w (numInputDims floats) and status (one int32: 0, or -1 if not positive
        definite); enqueued on `stream` (hipStream_t or None) without synchronisation."""
        check(lib().lls_train_lowlevel_noalloc(ctypes.c_void_p(A), ctypes.c_void_p(y), int(numInputDims),
                                               int(numOutputDims), int(numSamples), ctypes.c_void_p(w),
                                               ctypes.c_void_p(status), ctypes.c_void_p(stream)))

    def computeSignalPower(self, data, numValues):
        """sqrt(sum d^2 / n) of a device array (lls_detector.cu:236-242)."""
        out = ctypes.c_float()
        check(lib().lls_signal_power(self._h, ctypes.c_void_p(data), int(numValues), ctypes.byref(out)))
        return out.value

    def resetState(self):
        pass  # state is replaced by the next train (lls_detector.cu:220-223)


class LlsNomaDetectorFactory:
    @staticmethod
    def build(config, numAntennas):
        for key in config:
            if key != "otype":
                raise ValueError("Invalid key: " + key)
        return LlsNomaDetector(numAntennas)


class NomaDetectorFactory:
    @staticmethod
    def build(config, numAntennas, numUsers=1):
        otype = config.get("otype")
        if otype == "APSM":
            return ApsmNomaDetectorFactory.build(config, numAntennas, numUsers)
        if otype == "LLS":
            return LlsNomaDetectorFactory.build(config, numAntennas)
        raise ValueError("Unknown algorithm: " + str(otype))


__all__ = ["ApsmError", "ApsmNomaDetector", "ApsmNomaDetectorFactory", "LlsNomaDetector",
           "LlsNomaDetectorFactory", "NomaDetector", "NomaDetectorFactory", "apsm_parameters",
           "APSM_VERSIONS", "apsm_get_version_string"]
