"""Serving engine over the C ABI (apsm_server_*): links whose detector state stays on the device,
and frames of them served by continuous batching. A frame of a link is detected with the link's
current state, then the link trains incrementally on the frame's pilots; frames of different
links share batched launches, and a frame that arrives while others are on the GPU joins the
next round. Each result equals apsm_detect followed by apsm_train on a handle of its own."""
import ctypes

import numpy as np

from ._lib import FrameC, FrameResultC, ServerConfigC, check, lib
from .detector import ApsmNomaDetectorFactory

APSM_ERR_BUSY = -5


class Frame:
    """One request: rx_data complex [N][n], pilots rx complex [N][n_p] and tx complex [U][n_p]
    (n_p may be 0); est (complex64 [U][n]) receives the estimates. Arrays may be numpy arrays
    or views of pinned memory (PinnedArray)."""

    def __init__(self, link, rx_data, rx_pilots=None, tx_pilots=None, arrival_s=0.0, est=None):
        self.link = int(link)
        self.rx_data = _c64(rx_data)
        self.rx_pilots = None if rx_pilots is None else _c64(rx_pilots)
        self.tx_pilots = None if tx_pilots is None else _c64(np.atleast_2d(tx_pilots))
        self.arrival_s = float(arrival_s)
        self.est = est
        self.result = None

    def to_c(self, users):
        n = self.rx_data.shape[1]
        npil = 0 if self.rx_pilots is None else self.rx_pilots.shape[1]
        if self.est is None:
            self.est = np.zeros((users, n), np.complex64)
        if npil and self.tx_pilots.shape != (users, npil):
            raise ValueError("tx pilots must be [users][pilots]")
        return FrameC(self.arrival_s, self.link, n, npil, 0, self.rx_data.ctypes.data,
                      self.rx_pilots.ctypes.data if npil else None, self.tx_pilots.ctypes.data if npil else None,
                      self.est.ctypes.data)


def _c64(a):
    a = np.asarray(a)
    if a.dtype == np.complex64 and a.flags["C_CONTIGUOUS"]:
        return a
    return np.ascontiguousarray(a, dtype=np.complex64)


def _result(r):
    return dict(status=r.status, ticket=r.ticket, arrival_s=r.arrival_s, admit_s=r.admit_s, output_s=r.output_s,
                trained_s=r.trained_s, latency_s=r.output_s - r.arrival_s, detect_round=r.detect_round,
                last_round=r.last_round)


class ApsmServer:
    """hbm_budget: device bytes for link states and the frame arena; frame_arena: device bytes
    for frames in flight (a frame waits until it fits)."""

    def __init__(self, hbm_budget, frame_arena, batch_max_frames=0, detects_in_flight=0, batch_max_bytes=0):
        cfg = ServerConfigC(int(hbm_budget), int(frame_arena), int(batch_max_frames), int(detects_in_flight),
                            int(batch_max_bytes))
        self._s = ctypes.c_void_p()
        check(lib().apsm_server_create(ctypes.byref(cfg), ctypes.byref(self._s)))
        self.links = []  # (antennas, users)
        self._pending = {}

    def __del__(self):
        s = getattr(self, "_s", None)
        if s is not None and s.value:
            lib().apsm_server_destroy(s)
            self._s = ctypes.c_void_p()

    def add_link(self, config, num_antennas, num_users=1, max_pilots=4096):
        """Admits a link with a detector JSON configuration (reference keys). Returns its id;
        raises ApsmError with code APSM_ERR_BUSY when the HBM budget cannot hold it."""
        par = ApsmNomaDetectorFactory.params_from_config(dict(config, otype="APSM"))
        p = par.to_c()
        lid = ctypes.c_uint32()
        check(lib().apsm_server_add_link(self._s, ctypes.byref(p), int(num_antennas), int(num_users), int(max_pilots),
                                         ctypes.byref(lid)))
        self.links.append((int(num_antennas), int(num_users)))
        return lid.value

    def _users(self, f):
        if not 0 <= f.link < len(self.links):
            raise ValueError(f"unknown link {f.link}")
        return self.links[f.link][1]

    def replay(self, frames):
        """Serves frames (arrival_s non-decreasing, relative to the call) and returns when all
        have finished; sets frame.result (dict with latency_s = output_s - arrival_s)."""
        arr = (FrameC * len(frames))(*[f.to_c(self._users(f)) for f in frames])
        res = (FrameResultC * len(frames))()
        check(lib().apsm_server_replay(self._s, arr, len(frames), res))
        for f, r in zip(frames, res):
            f.result = _result(r)
        return [f.result for f in frames]

    def submit(self, frame):
        c = frame.to_c(self._users(frame))
        t = ctypes.c_uint64()
        check(lib().apsm_server_submit(self._s, ctypes.byref(c), ctypes.byref(t)))
        self._pending[t.value] = frame
        return t.value

    def poll(self, wait=False, cap=1024):
        """Makes progress and returns the frames that finished (their .result set)."""
        out = (FrameResultC * cap)()
        n = ctypes.c_uint32()
        done = []
        while True:
            check(lib().apsm_server_poll(self._s, 1 if wait else 0, out, cap, ctypes.byref(n)))
            for r in out[:n.value]:
                f = self._pending.pop(r.ticket)
                f.result = _result(r)
                done.append(f)
            if n.value < cap:
                return done

    def stats(self):
        r, m, b = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().apsm_server_stats(self._s, ctypes.byref(r), ctypes.byref(m), ctypes.byref(b)))
        return dict(rounds=r.value, max_frames_per_round=m.value, link_bytes=b.value)

    def link_state(self, link):
        """(used_rows, dict_entries, variance, theta [U][2N], coeffs [U][2*dict]) of a link."""
        h = lib().apsm_server_link(self._s, int(link))
        N, U = self.links[link]
        used, nd, var = ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_float()
        check(lib().apsm_get_state(h, ctypes.byref(used), ctypes.byref(nd), ctypes.byref(var), None, None, 0))
        theta = np.zeros((U, 2 * N), np.float32)
        coeffs = np.zeros((U, max(1, 2 * nd.value)), np.float32)
        check(lib().apsm_get_state(h, None, None, None, theta.ctypes.data, coeffs.ctypes.data, coeffs.shape[1]))
        return dict(used_rows=used.value, dict_entries=nd.value, variance=var.value, theta=theta,
                    coeffs=coeffs[:, :2 * nd.value])
