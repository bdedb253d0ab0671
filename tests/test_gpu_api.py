"""The rest of the detector API on the GPU: LLS low-level entry points, signal power,
external streams, norm ratio, dictionary readback and the per-step training history.

Tolerances: LLS weights relative 1e-4 against an FP64 least-squares solve; norm ratio
relative 1e-3 against the FP64 oracle; history EVM absolute 2e-3 and SER / BER within two
decisions of the oracle per step (FP32 device against FP64 oracle near decision boundaries).
"""
import ctypes

import numpy as np
import pytest

from helpers import small_problem, to_ri

pytestmark = pytest.mark.gpu


def _dev(a):
    from libapsm_amd._lib import DeviceArray
    return DeviceArray(np.ascontiguousarray(a, np.float32))


def test_lls_lowlevel_matches_fp64_lstsq(require_gpu):
    from libapsm_amd import LlsNomaDetector
    rng = np.random.default_rng(11)
    m, n = 3000, 20
    A = rng.standard_normal((m, n))
    w_true = rng.standard_normal(n)
    y = A @ w_true + 0.1 * rng.standard_normal(m)
    dA = _dev(A.T.reshape(-1))  # column-major: column i is contiguous
    dy = _dev(y)
    lls = LlsNomaDetector(4)
    lls.trainLowLevel(dA.ptr, dy.ptr, n, 1, m)
    w = lls.getWeightMat()
    ref = np.linalg.lstsq(A.astype(np.float32).astype(np.float64), y.astype(np.float32).astype(np.float64), rcond=None)[0]
    rel = np.abs(w - ref).max() / np.abs(ref).max()
    print(f"lowlevel weights rel err {rel:.2e}")
    assert rel < 1e-4
    # inputs are not overwritten (the reference's Sgels overwrites them; ours reads only)
    assert np.array_equal(dA.numpy(), np.asarray(A.T.reshape(-1), np.float32))
    # est = alpha * A w + beta * est
    est0 = rng.standard_normal(m).astype(np.float32)
    dEst = _dev(est0)
    lls.detectLowLevel(dA.ptr, dEst.ptr, n, 1, m, 2.0, 0.5)
    lls.synchronize()
    got = dEst.numpy().astype(np.float64)
    want = 2.0 * (A.astype(np.float32).astype(np.float64) @ w.astype(np.float64)) + 0.5 * est0
    assert np.abs(got - want).max() < 1e-4 * np.abs(want).max()
    # beta = 0 ignores the previous contents, NaN included
    dEst.upload(np.full(m, np.nan, np.float32))
    lls.detectLowLevel(dA.ptr, dEst.ptr, n, 1, m)
    lls.synchronize()
    assert np.isfinite(dEst.numpy()).all()


def test_lls_signal_power(require_gpu):
    from libapsm_amd import LlsNomaDetector
    rng = np.random.default_rng(3)
    x = rng.standard_normal(100_003).astype(np.float32) * 1.7
    d = _dev(x)
    p = LlsNomaDetector(2).computeSignalPower(d.ptr, x.size)
    ref = np.sqrt(np.mean(x.astype(np.float64) ** 2))
    assert abs(p - ref) < 1e-5 * ref


def test_lls_on_external_stream(require_gpu):
    from libapsm_amd._lib import check, lib
    L = lib()
    s = ctypes.c_void_p()
    check(L.apsm_stream_create(ctypes.byref(s)))
    h = ctypes.c_void_p()
    check(L.lls_create_on_stream(4, s, ctypes.byref(h)))
    assert L.lls_stream(h) == s.value
    d = small_problem(N=4, T=200, M=50, seed=5)
    rx = np.ascontiguousarray(to_ri(d["rx_train"]))
    tx = np.ascontiguousarray(to_ri(d["tx_train"][0]))
    check(L.lls_train(h, rx.ctypes.data, tx.ctypes.data, 200, None))
    w = np.zeros(8, np.float32)
    check(L.lls_get_weights(h, w.ctypes.data))
    L.lls_destroy(h)
    # the lent stream survives the handle
    check(L.apsm_stream_synchronize(s))
    check(L.apsm_stream_destroy(s))
    from oracle import pyoracle
    ref = pyoracle.lls_train(d["rx_train"], d["tx_train"][0])
    assert np.abs(w - ref).max() < 1e-4 * np.abs(ref).max()


@pytest.mark.parametrize("dc", [300, 60])
def test_norm_ratio_and_dictionary(require_gpu, oracle, dc):
    from libapsm_amd import ApsmNomaDetectorFactory
    cfg = dict(windowSize=8, dictionarySize=dc, eB=0.01, gaussianKernelVariance=0.1, gaussianKernelWeight=0.4)
    d = small_problem(N=4, T=150, M=10, seed=21)
    det = ApsmNomaDetectorFactory.build(dict(cfg, otype="APSM"), 4)
    det.train(d["rx_train"], d["tx_train"][0])
    o = oracle.Apsm(cfg, 4)
    o.train(d["rx_train"], d["tx_train"][0], incremental=True)
    r_gpu, r_orc = det.computeNormRatio(), o.norm_ratio()
    print(f"norm ratio gpu {r_gpu:.6g} oracle {r_orc:.6g}")
    assert abs(r_gpu - r_orc) <= 1e-3 * abs(r_orc)
    dic = det.getDictionary()
    G = min(150, dc)
    assert dic.shape == (4, G)
    assert np.array_equal(dic, d["rx_train"][:, 150 - G:].astype(np.complex64))


@pytest.mark.parametrize("case", [
    dict(windowSize=5, dictionarySize=200, eB=0.01, gaussianKernelVariance=0.1),
    dict(windowSize=5, dictionarySize=30, eB=0.01, gaussianKernelVariance=0.1, sampleStep=4),
    dict(windowSize=6, dictionarySize=200, eB=0.01, gaussianKernelVariance=0.1, startWithFullWindow=True),
    dict(windowSize=5, dictionarySize=200, eB=0.01, gaussianKernelVariance=0.1, trainingShuffle=True),
], ids=["base", "fifo-q4", "fullwin", "shuffle"])
def test_train_history_matches_oracle(require_gpu, oracle, case):
    from libapsm_amd import ApsmNomaDetectorFactory, NomaTrainHistory, datagen
    N, T, Tv = 4, 60, 80
    d = small_problem(N=N, T=T, M=Tv, seed=17)
    bps = 2
    scale = datagen.mod_scale(bps)
    det = ApsmNomaDetectorFactory.build(dict(case, otype="APSM"), N)
    hist = NomaTrainHistory(bps, scale)
    det.train(d["rx_train"], d["tx_train"][0], (hist, d["rx_data"], d["tx_data"][0]))

    # oracle: the same pilots (in training order) one step per call, detect after each step
    rx, tx = d["rx_train"], d["tx_train"][0]
    if case.get("trainingShuffle"):
        perm = oracle.shuffle_perm(T)
        rx, tx = rx[:, perm], tx[perm]
    ocfg = {k: v for k, v in case.items() if k != "trainingShuffle"}
    o = oracle.Apsm(ocfg, N)
    q2 = case.get("sampleStep", 2) // 2
    first = case["windowSize"] if case.get("startWithFullWindow") else q2
    c0, rows = 0, []
    ytr, yv = to_ri(tx), to_ri(d["tx_data"][0])
    while c0 + (first if c0 == 0 else q2) <= T:
        n = first if c0 == 0 else q2
        o.train(rx[:, c0:c0 + n], tx[c0:c0 + n], incremental=True)
        c0 += n
        et = o.detect(rx).astype(np.float32)
        ev = o.detect(d["rx_data"]).astype(np.float32)
        rows.append([oracle.evm(et, ytr), oracle.symbol_errors(et, ytr, bps) / T, oracle.bit_errors(et, ytr, bps) / (T * bps),
                     oracle.evm(ev, yv), oracle.symbol_errors(ev, yv, bps) / Tv, oracle.bit_errors(ev, yv, bps) / (Tv * bps)])
    ref = np.array(rows)
    got = np.stack([hist.evmTrain, hist.serTrain, hist.berTrain, hist.evmVal, hist.serVal, hist.berVal], axis=1)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    assert np.abs(got[:, 0] - ref[:, 0]).max() < 2e-3
    assert np.abs(got[:, 3] - ref[:, 3]).max() < 2e-3
    assert np.abs(got[:, 1] - ref[:, 1]).max() <= 2.0 / T + 1e-7
    assert np.abs(got[:, 4] - ref[:, 4]).max() <= 2.0 / Tv + 1e-7
    assert np.abs(got[:, 2] - ref[:, 2]).max() <= 2.0 / T + 1e-7
    assert np.abs(got[:, 5] - ref[:, 5]).max() <= 2.0 / Tv + 1e-7
    # the state after the history run equals a plain train on the same pilots
    det2 = ApsmNomaDetectorFactory.build(dict(case, otype="APSM"), N)
    det2.train(d["rx_train"], d["tx_train"][0])
    e1, e2 = det.detect(d["rx_data"]), det2.detect(d["rx_data"])
    assert np.abs(e1 - e2).max() < 1e-4


def test_train_history_rejects_passes(require_gpu):
    from libapsm_amd import ApsmError, ApsmNomaDetectorFactory, NomaTrainHistory
    d = small_problem(N=4, T=40, M=20, seed=1)
    det = ApsmNomaDetectorFactory.build(dict(otype="APSM", windowSize=4, dictionarySize=40, trainPasses=2), 4)
    with pytest.raises(ApsmError):
        det.train(d["rx_train"], d["tx_train"][0], (NomaTrainHistory(2, 0.7071), d["rx_data"], d["tx_data"][0]))


@pytest.mark.parametrize("N,T,shuffle", [(4, 2, False), (4, 301, False), (8, 3000, True), (16, 1200, False)])
def test_variance_heuristic_on_device(require_gpu, oracle, N, T, shuffle):
    """gaussianKernelVariance = 0 selects the pair-distance heuristic; the device's radix select
    returns the oracle's sorted-vector element (within the last-bit differences of |a - b|)."""
    from libapsm_amd import ApsmNomaDetectorFactory
    d = small_problem(N=N, T=T, M=10, seed=T)
    det = ApsmNomaDetectorFactory.build(dict(otype="APSM", windowSize=1, dictionarySize=T, gaussianKernelVariance=0.0,
                                             trainingShuffle=shuffle), N)
    det.train(d["rx_train"], d["tx_train"][0])
    got = det.getTrainingState()["variance"]
    ref = oracle.variance_heuristic(d["rx_train"])
    print(f"variance heuristic gpu {got:.9g} oracle {ref:.9g}")
    assert got == pytest.approx(ref, rel=2e-6)
