"""Parity of the HIP path (through the C ABI) with the FP64 oracle.

Tolerances (FP32 device arithmetic against the FP64 oracle, unit-energy symbols):
  * soft outputs: max |est_gpu - est_oracle| <= SOFT_ATOL (absolute, per real component);
  * hard decisions: identical wherever the oracle's soft output is farther than DELTA from
    a decision boundary;
  * LLS weights: relative error <= 1e-4.
"""
import zlib

import numpy as np
import pytest

from helpers import equidistant, load_offline, small_problem, to_ri

pytestmark = pytest.mark.gpu

SOFT_ATOL = 2e-3
DELTA = 1e-2


def gpu_run(cfg, N, rx_t, tx_t, rx_d, users=1):
    from libapsm_amd import ApsmNomaDetectorFactory
    det = ApsmNomaDetectorFactory.build(dict(cfg, otype="APSM"), N, users)
    det.train(rx_t, tx_t)
    est = det.detect(rx_d)
    return det, est


def orc_run(oracle, cfg, N, rx_t, tx_t, rx_d):
    o = oracle.Apsm(cfg, N)
    o.train(rx_t, tx_t, incremental=True)
    return o, o.detect(rx_d)


def boundary_margin(est_ri, bps):
    """Distance of each soft output to the nearest decision boundary (either dimension)."""
    from libapsm_amd.datagen import mod_scale
    s = mod_scale(bps)
    ns = 1 << max(1, bps // 2)
    bnds = np.array([2 * k * s for k in range(-(ns // 2 - 1), ns // 2)])
    mr = np.abs(est_ri[:, 0, None] - bnds[None, :]).min(axis=1)
    mi = np.abs(est_ri[:, 1, None] - bnds[None, :]).min(axis=1)
    return np.minimum(mr, mi)


def compare(est_gpu, est_orc, bps):
    from libapsm_amd.datagen import hard_decision
    g = to_ri(est_gpu).astype(np.float64)
    err = np.abs(g - est_orc).max()
    print(f"max soft err {err:.3e}  rms {np.sqrt(np.mean((g - est_orc) ** 2)):.3e}")
    assert err <= SOFT_ATOL
    zo = est_orc[:, 0] + 1j * est_orc[:, 1]
    far = boundary_margin(est_orc, bps) > DELTA
    hg = hard_decision(np.asarray(est_gpu), bps)
    ho = hard_decision(zo, bps)
    assert np.array_equal(hg[far], ho[far])


CASES = [
    ("base", dict(windowSize=10, dictionarySize=200, eB=0.01, gaussianKernelVariance=0.1), 4, 150),
    ("fifo", dict(windowSize=10, dictionarySize=40, eB=0.01, gaussianKernelVariance=0.1), 4, 150),
    ("fullwin-q4", dict(windowSize=10, dictionarySize=200, eB=0.01, gaussianKernelVariance=0.1,
                        startWithFullWindow=True, sampleStep=4), 4, 150),
    ("passes3", dict(windowSize=8, dictionarySize=150, eB=0.001, gaussianKernelVariance=0.05, trainPasses=3), 4, 150),
    ("shuffle", dict(windowSize=6, dictionarySize=150, eB=0.01, gaussianKernelVariance=0.2, trainingShuffle=True), 4, 150),
    ("llsinit", dict(windowSize=6, dictionarySize=150, eB=0.01, gaussianKernelVariance=0.2, llsInitialization=True,
                     gaussianKernelWeight=0.3), 4, 150),
    ("n8-w80", dict(windowSize=80, dictionarySize=400, eB=0.01, gaussianKernelVariance=0.25 / 64), 8, 400),
    ("n16-w160-fifo", dict(windowSize=160, dictionarySize=256, eB=0.01, gaussianKernelVariance=0.25 / 256), 16, 500),
    ("n32", dict(windowSize=20, dictionarySize=300, eB=0.01, gaussianKernelVariance=0.25 / 1024), 32, 300),
    ("n64", dict(windowSize=20, dictionarySize=300, eB=0.01, gaussianKernelVariance=0.25 / 4096), 64, 300),
    ("n5-odd", dict(windowSize=7, dictionarySize=100, eB=0.01, gaussianKernelVariance=0.1), 5, 120),
    ("w512", dict(windowSize=512, dictionarySize=600, eB=0.01, gaussianKernelVariance=0.25 / 64), 8, 700),
    ("q4-fifo", dict(windowSize=12, dictionarySize=50, eB=0.01, gaussianKernelVariance=0.1, sampleStep=4), 4, 300),
    ("w64", dict(windowSize=64, dictionarySize=200, eB=0.01, gaussianKernelVariance=0.25 / 64), 8, 500),
    ("w100-fifo", dict(windowSize=100, dictionarySize=150, eB=0.01, gaussianKernelVariance=0.25 / 64), 8, 600),
    ("w128", dict(windowSize=128, dictionarySize=700, eB=0.01, gaussianKernelVariance=0.25 / 256), 16, 600),
    ("w160-evict", dict(windowSize=160, dictionarySize=400, eB=0.01, gaussianKernelVariance=0.25 / 256), 16, 1500),
    ("w40-q4", dict(windowSize=40, dictionarySize=90, eB=0.001, gaussianKernelVariance=0.25 / 64, sampleStep=4), 8, 500),
    ("nc", dict(windowSize=8, dictionarySize=150, eB=0.01, gaussianKernelVariance=0.1, normConstraint=0.05), 4, 150),
    ("nc-fifo-q4", dict(windowSize=8, dictionarySize=30, eB=0.01, gaussianKernelVariance=0.1, normConstraint=0.02,
                        sampleStep=4), 4, 200),
    ("nc-passes2", dict(windowSize=6, dictionarySize=150, eB=0.001, gaussianKernelVariance=0.1, normConstraint=0.05,
                        trainPasses=2), 4, 150),
    ("nc-n16-w80", dict(windowSize=80, dictionarySize=200, eB=0.01, gaussianKernelVariance=0.25 / 256,
                        normConstraint=0.05), 16, 400),
]


@pytest.mark.parametrize("name,cfg,N,T", CASES, ids=[c[0] for c in CASES])
def test_train_detect_matches_oracle(require_gpu, oracle, name, cfg, N, T):
    print(f"case {name} N={N} T={T}", flush=True)
    d = small_problem(N=N, U=3, T=T, M=333, seed=zlib.crc32(name.encode()) % 1000)
    det, est = gpu_run(cfg, N, d["rx_train"], d["tx_train"][1], d["rx_data"])
    o, est_o = orc_run(oracle, cfg, N, d["rx_train"], d["tx_train"][1], d["rx_data"])
    compare(est, est_o, 2)
    st, so = det.getTrainingState(), o.state()
    assert st["used_rows"] == so["used"]
    assert st["dict_entries"] == (so["used"] - so["lo"]) // 2
    np.testing.assert_allclose(st["coeffs"][0], so["coeffs"], rtol=0, atol=2e-4)
    np.testing.assert_allclose(st["theta"][0], so["theta"], rtol=0, atol=2e-3)


def test_multi_user_handle_matches_oracle_and_single_handles(require_gpu, oracle):
    cfg = dict(windowSize=12, dictionarySize=120, eB=0.01, gaussianKernelVariance=0.1)
    d = small_problem(N=6, U=3, T=200, M=300, seed=21)
    det, est = gpu_run(cfg, 6, d["rx_train"], d["tx_train"], d["rx_data"], users=3)
    assert est.shape == (3, 300)
    for u in range(3):
        _, est_o = orc_run(oracle, cfg, 6, d["rx_train"], d["tx_train"][u], d["rx_data"])
        compare(est[u], est_o, 2)
        _, est_1 = gpu_run(cfg, 6, d["rx_train"], d["tx_train"][u], d["rx_data"])
        np.testing.assert_array_equal(est[u], est_1)  # same kernel, same order: bitwise


def test_two_train_calls_continue_the_stream(require_gpu, oracle):
    from libapsm_amd import ApsmNomaDetectorFactory
    cfg = dict(windowSize=10, dictionarySize=60, eB=0.01, gaussianKernelVariance=0.1)
    d = small_problem(N=4, T=160, M=200, seed=9)
    det = ApsmNomaDetectorFactory.build(dict(cfg, otype="APSM"), 4)
    det.train(d["rx_train"][:, :70], d["tx_train"][0, :70])
    det.train(d["rx_train"][:, 70:], d["tx_train"][0, 70:])
    est = det.detect(d["rx_data"])
    _, est_o = orc_run(oracle, cfg, 4, d["rx_train"], d["tx_train"][0], d["rx_data"])
    compare(est, est_o, 2)
    det.resetState()
    det.train(d["rx_train"], d["tx_train"][0])
    compare(det.detect(d["rx_data"]), est_o, 2)


@pytest.mark.parametrize("M", [1, 63, 64, 255, 257, 1000])
def test_ragged_detect_sizes(require_gpu, oracle, M):
    cfg = dict(windowSize=10, dictionarySize=100, eB=0.01, gaussianKernelVariance=0.1)
    d = small_problem(N=4, T=100, M=M, seed=M)
    _, est = gpu_run(cfg, 4, d["rx_train"], d["tx_train"][0], d["rx_data"])
    _, est_o = orc_run(oracle, cfg, 4, d["rx_train"], d["tx_train"][0], d["rx_data"])
    compare(est, est_o, 2)


def test_detect_before_train_is_zero(require_gpu):
    from libapsm_amd import ApsmNomaDetectorFactory
    det = ApsmNomaDetectorFactory.build({"otype": "APSM"}, 4)
    d = small_problem(N=4, T=10, M=50, seed=1)
    assert np.all(det.detect(d["rx_data"]) == 0)


def test_lls_matches_oracle(require_gpu, oracle):
    from libapsm_amd import LlsNomaDetector
    for N, T in [(4, 50), (16, 5000), (64, 3000)]:
        d = small_problem(N=N, T=T, M=400, seed=N)
        lls = LlsNomaDetector(N)
        lls.train(d["rx_train"], d["tx_train"][0])
        w = lls.getWeightMat()
        w_o = oracle.lls_train(d["rx_train"], d["tx_train"][0])
        rel = np.abs(w - w_o).max() / np.abs(w_o).max()
        print(f"LLS N={N} T={T} weight rel err {rel:.2e}")
        assert rel < 1e-4
        est = lls.detect(d["rx_data"])
        est_o = oracle.lls_detect(d["rx_data"], w_o)
        assert np.abs(to_ri(est) - est_o).max() < 1e-4 * (1 + np.abs(est_o).max())


def test_lls_singular_system_is_reported(require_gpu):
    from libapsm_amd import ApsmError, LlsNomaDetector
    rx = np.zeros((4, 20), np.complex64)
    lls = LlsNomaDetector(4)
    with pytest.raises(ApsmError):
        lls.train(rx, np.ones(20, np.complex64))


@pytest.mark.parametrize("mod,domain", [("QPSK", "time"), ("QPSK", "ofdm"), ("QAM16", "ofdm")])
def test_bundled_reference_data(require_gpu, oracle, mod, domain):
    d = load_offline(mod, domain, antennas=equidistant(8))
    cfg = dict(windowSize=80)
    _, est = gpu_run(cfg, 8, d["rx_train"], d["tx_train"][1], d["rx_data"])
    _, est_o = orc_run(oracle, cfg, 8, d["rx_train"], d["tx_train"][1], d["rx_data"])
    compare(est, est_o, d["bps"])
    ref = to_ri(d["tx_data"][1])
    e_g = oracle.symbol_errors(to_ri(est), ref, d["bps"])
    e_o = oracle.symbol_errors(est_o, ref, d["bps"])
    print(f"{mod}/{domain}: SER gpu {e_g} oracle {e_o} of {ref.shape[0]}")
    assert abs(e_g - e_o) <= max(1, 0.01 * e_o)


def test_norm_constraint_two_calls_and_ratio(require_gpu, oracle):
    """The norm-constraint state (explicit theta, Gaussian part of f, c^T K c) carries over to
    the next train call; the device's norm ratio matches the oracle's after the constraint acted."""
    from libapsm_amd import ApsmNomaDetectorFactory
    cfg = dict(windowSize=8, dictionarySize=50, eB=0.01, gaussianKernelVariance=0.1, normConstraint=0.03)
    d = small_problem(N=4, T=160, M=200, seed=12)
    det = ApsmNomaDetectorFactory.build(dict(cfg, otype="APSM"), 4)
    det.train(d["rx_train"][:, :90], d["tx_train"][0, :90])
    det.train(d["rx_train"][:, 90:], d["tx_train"][0, 90:])
    o, est_o = orc_run(oracle, cfg, 4, d["rx_train"], d["tx_train"][0], d["rx_data"])
    compare(det.detect(d["rx_data"]), est_o, 2)
    assert det.computeNormRatio() == pytest.approx(o.norm_ratio(), rel=2e-3)
