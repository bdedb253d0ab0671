"""Batched detector instances over the C ABI (apsm_batch_*): many independent APSM detectors,
each with its own parameters, antenna count and users, trained and detected as one launch
sequence (captured in a hipGraph). A run gives every instance the result of reset + train +
detect on a handle of its own."""
import ctypes

import numpy as np

from ._lib import ApsmParamsC, DeviceArray, check, lib
from .detector import ApsmNomaDetectorFactory


class ApsmBatch:
    """configs: list of detector JSON configurations (reference keys); antennas, users: per
    instance counts."""

    def __init__(self, configs, antennas, users=None):
        n = len(configs)
        users = users if users is not None else [1] * n
        self.n = n
        self.antennas = [int(a) for a in antennas]
        self.users = [int(u) for u in users]
        params = (ApsmParamsC * n)(*[ApsmNomaDetectorFactory.params_from_config(dict(c, otype="APSM")).to_c()
                                     for c in configs])
        self._b = ctypes.c_void_p()
        check(lib().apsm_batch_create(n, params, (ctypes.c_uint32 * n)(*self.antennas),
                                      (ctypes.c_uint32 * n)(*self.users), ctypes.byref(self._b)))
        self._keep = []

    def __del__(self):
        b = getattr(self, "_b", None)
        if b is not None and b.value:
            lib().apsm_batch_destroy(b)
            self._b = ctypes.c_void_p()

    def plan(self, rx_train, tx_train, rx_data, use_graph=True):
        """rx_train[i] complex [N_i][T_i], tx_train[i] complex [U_i][T_i] (or [T_i]), rx_data[i]
        complex [N_i][M_i]; uploads them, plans, autotunes and leaves one completed run."""
        n = self.n
        self.T = [np.asarray(r).shape[1] for r in rx_train]
        self.M = [np.asarray(r).shape[1] for r in rx_data]
        up = lambda z: DeviceArray(np.ascontiguousarray(z, dtype=np.complex64).view(np.float32))  # noqa: E731
        self.d_rxt = [up(r) for r in rx_train]
        self.d_txt = [up(np.atleast_2d(t)) for t in tx_train]
        self.d_rxd = [up(r) for r in rx_data]
        self.d_est = [DeviceArray.empty((self.users[i], self.M[i], 2), np.float32) for i in range(n)]
        P = ctypes.c_void_p * n
        check(lib().apsm_batch_plan(self._b, P(*[d.ptr for d in self.d_rxt]), P(*[d.ptr for d in self.d_txt]),
                                    (ctypes.c_uint32 * n)(*self.T), P(*[d.ptr for d in self.d_rxd]),
                                    (ctypes.c_uint32 * n)(*self.M), P(*[d.ptr for d in self.d_est]), int(bool(use_graph))))

    def plan_device(self, d_rx_train, d_tx_train, T, d_rx_data, M, d_est, use_graph=True):
        """Device pointers (ints) in the *_dev layouts."""
        n = self.n
        P = ctypes.c_void_p * n
        self.T, self.M = list(T), list(M)
        check(lib().apsm_batch_plan(self._b, P(*d_rx_train), P(*d_tx_train), (ctypes.c_uint32 * n)(*T), P(*d_rx_data),
                                    (ctypes.c_uint32 * n)(*M), P(*d_est), int(bool(use_graph))))

    def run(self):
        check(lib().apsm_batch_run(self._b))

    def synchronize(self):
        check(lib().apsm_batch_synchronize(self._b))

    def last_ms(self):
        ms = ctypes.c_float()
        check(lib().apsm_batch_last_ms(self._b, ctypes.byref(ms)))
        return ms.value

    def last_split(self):
        """(train ms, detect ms) of the most recent run."""
        t, d = ctypes.c_float(), ctypes.c_float()
        check(lib().apsm_batch_last_split(self._b, ctypes.byref(t), ctypes.byref(d)))
        return t.value, d.value

    def estimates(self, i):
        """complex64 [U_i][M_i] of the last run (after synchronize)."""
        e = self.d_est[i].numpy()
        return e[..., 0] + 1j * e[..., 1]

    def tuning(self):
        rows = ctypes.c_uint32()
        check(lib().apsm_batch_tuning(self._b, None, 0, ctypes.byref(rows)))
        t = np.zeros((max(1, rows.value), 8), np.int32)
        check(lib().apsm_batch_tuning(self._b, t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), t.shape[0],
                                      ctypes.byref(rows)))
        kernels = {-1: "vector", 0: "matrix-fp32", 1: "matrix-bf16", 2: "matrix-fp16"}
        return [dict(nmax=int(r[0]), umax=int(r[1]), threads=int(r[2]), blocks_per_cu=int(r[3]), items=int(r[4]),
                     kernel=kernels.get(int(r[5]), str(r[5])), entries_per_tile=int(r[6]), samples_per_workgroup=int(r[7]))
                for r in t[:rows.value]]
