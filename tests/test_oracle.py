"""Pins of the CPU oracle (oracle/apsm_oracle.cpp) before it is trusted as the checker.

The reference holds no golden vectors, known-answer tests or expected outputs for this
path (it has no tests at all); what pins the oracle here:
  * two independently organised train algorithms (reference schedule vs incremental
    window-Gram form) that must agree to FP64 rounding;
  * closed forms: the first projection step from the zero state; the purely linear model
    (wG = 0) reduces to (p - jq)^T x; the LS solution equals numpy's complex lstsq;
  * invariants: dictionary capacity above the stream length changes nothing; two train
    calls equal one call on the concatenation; Cholesky and Householder QR agree;
  * the bundled reference data: LLS and APSM reach the error-free regime on QPSK, which a
    wrong embedding pairing or sign cannot;
  * brute force on tiny inputs: the variance heuristic and the SER decision regions.
"""
import json
import os

import numpy as np
import pytest

from helpers import GOLDEN, equidistant, load_offline, small_problem, to_ri

CFGS = [
    dict(windowSize=10, dictionarySize=200, eB=0.01, gaussianKernelVariance=0.1),
    dict(windowSize=10, dictionarySize=40, eB=0.01, gaussianKernelVariance=0.1),  # FIFO wrap
    dict(windowSize=10, dictionarySize=200, eB=0.01, gaussianKernelVariance=0.1, startWithFullWindow=True, sampleStep=4),
    dict(windowSize=8, dictionarySize=150, eB=0.001, gaussianKernelVariance=0.05, trainPasses=3),
    dict(windowSize=6, dictionarySize=150, eB=0.01, gaussianKernelVariance=0.2, trainingShuffle=True),
    dict(windowSize=6, dictionarySize=150, eB=0.01, gaussianKernelVariance=0.2, llsInitialization=True,
         gaussianKernelWeight=0.3),
    dict(windowSize=8, dictionarySize=150, eB=0.01, gaussianKernelVariance=0.1, normConstraint=0.3),
    dict(windowSize=8, dictionarySize=30, eB=0.01, gaussianKernelVariance=0.1, normConstraint=0.3, sampleStep=4),
    dict(windowSize=6, dictionarySize=150, eB=0.001, gaussianKernelVariance=0.1, normConstraint=0.2, trainPasses=2),
]


@pytest.mark.parametrize("cfg", CFGS)
def test_naive_equals_incremental(oracle, cfg):
    d = small_problem(N=4, T=150, seed=3)
    a, b = oracle.Apsm(cfg, 4), oracle.Apsm(cfg, 4)
    a.train(d["rx_train"], d["tx_train"][1], incremental=False)
    b.train(d["rx_train"], d["tx_train"][1], incremental=True)
    sa, sb = a.state(), b.state()
    assert sa["used"] == sb["used"] and sa["lo"] == sb["lo"]
    np.testing.assert_allclose(sa["coeffs"], sb["coeffs"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(sa["theta"], sb["theta"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(a.detect(d["rx_data"]), b.detect(d["rx_data"]), rtol=0, atol=1e-12)


def test_first_step_closed_form(oracle):
    # From the zero state f = 0, so one step on one complex pilot (W = 2 rows) gives
    # beta_t = soft(y_t, eB) / 2 / (wL |x|^2 + wG), c = beta, theta = beta_a a + beta_b b.
    rng = np.random.default_rng(7)
    N, wG, eB, v = 3, 0.4, 0.05, 0.3
    x = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    s = np.complex64(0.7 - 0.2j)
    cfg = dict(windowSize=1, dictionarySize=4, eB=eB, gaussianKernelVariance=v, gaussianKernelWeight=wG)
    o = oracle.Apsm(cfg, N)
    o.train(x[:, None], np.array([s]))
    st = o.state()
    xd = x.astype(np.complex128)
    a = np.concatenate([xd.real, xd.imag])
    b = np.concatenate([xd.imag, -xd.real])
    wL = 1 - wG
    nrm = np.sum(np.abs(xd) ** 2)

    def soft(r):
        return r - eB if r > eB else (r + eB if r < -eB else 0.0)

    ba = soft(float(s.real)) / 2 / (wL * nrm + wG)
    bb = soft(float(s.imag)) / 2 / (wL * nrm + wG)
    np.testing.assert_allclose(st["coeffs"], [ba, bb], rtol=1e-12)
    np.testing.assert_allclose(st["theta"], ba * a + bb * b, rtol=1e-12, atol=1e-15)
    assert st["var"] == pytest.approx(np.float32(v) * N * N, rel=1e-7)
    # f(a) of the trained model on the same sample: K(a,a) = 1, K(b,a) = exp(-2|x|^2/(2 var))
    var = st["var"]
    kab = np.exp(-2 * nrm / (2 * var))
    fa = wL * (st["theta"] @ a) + wG * (ba * 1.0 + bb * kab)
    fb = wL * (st["theta"] @ b) + wG * (ba * kab + bb * 1.0)
    est = o.detect(x[:, None])
    np.testing.assert_allclose(est[0], [fa, fb], rtol=1e-12)


def test_linear_model_identity(oracle):
    # wG = 0: the model is linear and f(a_x) + j f(b_x) = (p - jq)^T x with theta = [p; q]
    d = small_problem(N=5, T=120, seed=11)
    cfg = dict(windowSize=12, dictionarySize=300, eB=0.01, gaussianKernelVariance=0.1, gaussianKernelWeight=0.0)
    o = oracle.Apsm(cfg, 5)
    o.train(d["rx_train"], d["tx_train"][0])
    th = o.state()["theta"]
    v = th[:5] - 1j * th[5:]
    est = o.detect(d["rx_data"])
    ref = v @ d["rx_data"].astype(np.complex128)
    np.testing.assert_allclose(est[:, 0] + 1j * est[:, 1], ref, rtol=1e-9, atol=1e-12)


def test_dictionary_capacity_above_stream_is_neutral(oracle):
    d = small_problem(N=4, T=100, seed=5)
    base = dict(windowSize=10, eB=0.01, gaussianKernelVariance=0.1)
    a = oracle.Apsm(dict(base, dictionarySize=100), 4)
    b = oracle.Apsm(dict(base, dictionarySize=5000), 4)
    for o in (a, b):
        o.train(d["rx_train"], d["tx_train"][2])
    np.testing.assert_array_equal(a.detect(d["rx_data"]), b.detect(d["rx_data"]))
    # and a smaller capacity does change the result (the eviction is live)
    c = oracle.Apsm(dict(base, dictionarySize=30), 4)
    c.train(d["rx_train"], d["tx_train"][2])
    assert np.abs(c.detect(d["rx_data"]) - a.detect(d["rx_data"])).max() > 1e-6


@pytest.mark.parametrize("incremental", [False, True])
def test_two_calls_equal_one(oracle, incremental):
    d = small_problem(N=4, T=160, seed=9)
    cfg = dict(windowSize=10, dictionarySize=60, eB=0.01, gaussianKernelVariance=0.1)
    a, b = oracle.Apsm(cfg, 4), oracle.Apsm(cfg, 4)
    a.train(d["rx_train"], d["tx_train"][0], incremental=incremental)
    b.train(d["rx_train"][:, :70], d["tx_train"][0, :70], incremental=incremental)
    b.train(d["rx_train"][:, 70:], d["tx_train"][0, 70:], incremental=incremental)
    np.testing.assert_allclose(a.detect(d["rx_data"]), b.detect(d["rx_data"]), rtol=0, atol=1e-12)


def test_lls_matches_complex_lstsq(oracle):
    d = small_problem(N=6, T=300, seed=2)
    rx, s = d["rx_train"].astype(np.complex128), d["tx_train"][1].astype(np.complex128)
    v, *_ = np.linalg.lstsq(rx.T, s, rcond=None)  # min sum |x_k^T v - s_k|^2
    w_ref = np.concatenate([v.real, -v.imag])
    for method in ("cholesky", "qr"):
        w = oracle.lls_train(d["rx_train"], d["tx_train"][1], method)
        np.testing.assert_allclose(w, w_ref, rtol=1e-9, atol=1e-11)
    # residual orthogonality A^T (A w - y) = 0 over the real embedding
    w = oracle.lls_train(d["rx_train"], d["tx_train"][1])
    est = oracle.lls_detect(d["rx_train"], w)
    r = (est[:, 0] + 1j * est[:, 1]) - s
    assert np.abs(rx.conj() @ r).max() < 1e-9
    np.testing.assert_allclose(est[:, 0] + 1j * est[:, 1], v @ rx, rtol=1e-9, atol=1e-11)


def test_bundled_qpsk_anchor(oracle):
    # README example: QPSK, time domain, user 1, 8 equidistant antennas, windowSize 80
    d = load_offline("QPSK", "time", antennas=equidistant(8))
    cfg = dict(windowSize=80)
    o = oracle.Apsm(cfg, 8)
    o.train(d["rx_train"], d["tx_train"][1], incremental=True)
    est = o.detect(d["rx_data"])
    ref = to_ri(d["tx_data"][1])
    assert oracle.symbol_errors(est, ref, 2) == 0
    assert 0.05 < oracle.evm(est, ref) < 0.2
    w = oracle.lls_train(d["rx_train"], d["tx_train"][1])
    assert oracle.symbol_errors(oracle.lls_detect(d["rx_data"], w), ref, 2) == 0
    # a wrong pairing of the embedding (b = [Im; +Re]) breaks the detector
    bad_rx = d["rx_data"].conj()
    assert oracle.symbol_errors(o.detect(bad_rx), ref, 2) > 500


def test_variance_heuristic_bruteforce(oracle):
    rng = np.random.default_rng(4)
    rx = (rng.standard_normal((3, 40)) + 1j * rng.standard_normal((3, 40))).astype(np.complex64)
    T = rx.shape[1]
    i, j = np.triu_indices(T, 1)
    dist = np.abs(rx[:, i] - rx[:, j]).astype(np.float64).sum(axis=0)
    nd = np.sort(np.sqrt(dist))
    v = nd[int(np.float32(len(nd)) * np.float32(0.7))] * T ** (-1 / 5)
    assert oracle.variance_heuristic(rx) == pytest.approx(v * v, rel=1e-5)


def test_shuffle_permutation_fixture(oracle):
    with open(os.path.join(GOLDEN, "shuffle_perm.json")) as f:
        fx = json.load(f)
    for T, head in fx.items():
        p = oracle.shuffle_perm(int(T))
        assert sorted(p.tolist()) == list(range(int(T)))
        assert p[:len(head)].tolist() == head


@pytest.mark.parametrize("bps", [1, 2, 4, 6])
def test_modulation_map_demap(oracle, bps):
    pts = [oracle.mod_map(bps, b) for b in range(1 << bps)]
    for b, z in enumerate(pts):
        assert oracle.mod_demap(bps, z) == b
    if bps > 1:
        assert np.mean(np.abs(pts) ** 2) == pytest.approx(1.0, rel=1e-12)
        # Gray property: nearest neighbours differ in exactly one bit
        s = oracle.mod_scale(bps)
        for b, z in enumerate(pts):
            for b2, z2 in enumerate(pts):
                if abs(abs(z - z2) - 2 * s) < 1e-9:
                    assert bin(b ^ b2).count("1") == 1
    expect = {1: 1.0, 2: 1 / np.sqrt(2), 4: 1 / np.sqrt(10), 6: 1 / np.sqrt(42)}[bps]
    assert oracle.mod_scale(bps) == pytest.approx(expect, rel=1e-12)


@pytest.mark.parametrize("bps", [2, 4])
def test_symbol_errors_are_decision_regions(oracle, bps):
    rng = np.random.default_rng(bps)
    pts = np.array([oracle.mod_map(bps, b) for b in range(1 << bps)])
    true = pts[rng.integers(0, len(pts), 4000)]
    est = true + (rng.standard_normal(4000) + 1j * rng.standard_normal(4000)) * 0.25
    # brute force: nearest constellation point by Euclidean distance
    nearest = pts[np.argmin(np.abs(est[:, None] - pts[None, :]), axis=1)]
    s = oracle.mod_scale(bps)
    # skip points within 1e-6 of a region boundary
    ok = (np.abs(((est.real / s) % 2) - 0) > 1e-6) & (np.abs(((est.imag / s) % 2) - 0) > 1e-6)
    e_brute = int(np.sum((nearest != true) & ok))
    e_orc = oracle.symbol_errors(to_ri(est[ok]), to_ri(true[ok]), bps)
    assert e_orc == e_brute
    assert oracle.bit_errors(to_ri(true), to_ri(true), bps) == 0


def _gauss_quad(rows, c, var):
    """c^T K c for real rows [n][L], K = exp(-0.5/var |r_i - r_j|^2) (numpy, FP64)."""
    d = ((rows[:, None, :] - rows[None, :, :]) ** 2).sum(-1)
    return float(c @ np.exp(-0.5 / var * d) @ c)


def _real_rows(rx):
    """Real embedding rows of complex samples [N][S]: row 2k = [Re; Im], row 2k+1 = [Im; -Re]."""
    x = np.asarray(rx, np.complex128).T
    a = np.concatenate([x.real, x.imag], 1)
    b = np.concatenate([x.imag, -x.real], 1)
    return np.stack([a, b], 1).reshape(-1, a.shape[1])


@pytest.mark.parametrize("dc", [300, 40])
def test_norm_constraint_is_enforced(oracle, dc):
    """After every step sqrt(wG c^T K c) <= normConstraint (apsm_train.cu:192-256), with a bound
    at 30 % of the unconstrained final norm so that it binds."""
    d = small_problem(N=4, T=120, seed=4)
    base = dict(windowSize=8, dictionarySize=dc, eB=0.01, gaussianKernelVariance=0.1, gaussianKernelWeight=0.5)
    rows_all = _real_rows(d["rx_train"])
    free = oracle.Apsm(base, 4)
    free.train(d["rx_train"], d["tx_train"][0], incremental=True)
    sf = free.state()
    rows = rows_all[sf["lo"]:sf["used"]]
    nc = 0.3 * np.sqrt(0.5 * _gauss_quad(rows, sf["coeffs"], sf["var"]))
    o = oracle.Apsm(dict(base, normConstraint=nc), 4)
    o.train(d["rx_train"], d["tx_train"][0], incremental=True)
    st = o.state()
    n = np.sqrt(0.5 * _gauss_quad(rows, st["coeffs"], st["var"]))
    assert n <= nc * (1 + 1e-9)
    assert not np.allclose(st["theta"], sf["theta"])


def test_norm_constraint_inactive_is_neutral(oracle):
    d = small_problem(N=4, T=100, seed=6)
    base = dict(windowSize=6, dictionarySize=60, eB=0.01, gaussianKernelVariance=0.1)
    a, b = oracle.Apsm(base, 4), oracle.Apsm(dict(base, normConstraint=1e9), 4)
    a.train(d["rx_train"], d["tx_train"][0], incremental=True)
    b.train(d["rx_train"], d["tx_train"][0], incremental=True)
    np.testing.assert_allclose(a.detect(d["rx_data"]), b.detect(d["rx_data"]), rtol=0, atol=1e-12)


def test_norm_constraint_first_step_closed_form(oracle):
    """From the zero state the first step gives beta_t = soft(y_t, eB) / 2 / (wL|x|^2 + wG) for
    the two rows of sample 0; with the bound binding, c = beta * nc / sqrt(wG beta^T K beta)
    and theta keeps the unscaled beta (only the Gaussian coefficients are scaled)."""
    d = small_problem(N=3, T=1, seed=2)
    eB, v, wG, nc = 0.01, 0.1, 0.4, 0.05
    cfg = dict(windowSize=1, dictionarySize=4, eB=eB, gaussianKernelVariance=v, gaussianKernelWeight=wG, normConstraint=nc)
    o = oracle.Apsm(cfg, 3)
    o.train(d["rx_train"], d["tx_train"][0], incremental=False)
    st = o.state()
    rows = _real_rows(d["rx_train"])
    y = np.array([d["tx_train"][0][0].real, d["tx_train"][0][0].imag], np.float64)
    xn = (np.abs(d["rx_train"][:, 0].astype(np.complex128)) ** 2).sum()
    beta = np.sign(y) * np.maximum(np.abs(y) - eB, 0) / 2 / ((1 - wG) * xn + wG)
    var = v * 9
    n = np.sqrt(wG * _gauss_quad(rows, beta, var))
    assert n > nc
    np.testing.assert_allclose(st["coeffs"], beta * nc / n, rtol=1e-12)
    np.testing.assert_allclose(st["theta"], beta @ rows, rtol=1e-12)


def test_norm_ratio_single_entry_closed_form(oracle):
    """One complex entry x: c^T K c = ca^2 + cb^2 + 2 ca cb exp(-|x|^2 / var) because
    |a - b|^2 = 2|x|^2; the ratio is wL |theta|^2 over wG times that."""
    rng = np.random.default_rng(1)
    N, wG, v = 4, 0.3, 0.2
    o = oracle.Apsm(dict(windowSize=1, dictionarySize=4, gaussianKernelVariance=v, gaussianKernelWeight=wG), N)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    rows = _real_rows(x[:, None])
    c = np.array([0.7, -0.4])
    theta = rng.standard_normal(2 * N)
    var = v * N * N
    o.set_state(rows, c, theta, var)
    q = c[0] ** 2 + c[1] ** 2 + 2 * c[0] * c[1] * np.exp(-np.sum(np.abs(x) ** 2) / var)
    assert o.norm_ratio() == pytest.approx((1 - wG) * theta @ theta / (wG * q), rel=1e-12)
