"""Pins of the CPU oracle (oracle/apsm_oracle.cpp) before it is trusted as the checker.

The reference holds no golden vectors, known-answer tests or expected outputs for this
path (it has no tests at all); what pins the oracle here:
  * two independently organised train algorithms (reference schedule vs incremental
    window-Gram form) that must agree to FP64 rounding;
  * closed forms: the first projection step from the zero state; two steps through the
    window ramp and from a full-window start; the hyperslab projection (excess over eB
    halved, rows inside untouched); the purely linear model (wG = 0) reduces to
    (p - jq)^T x; the LS solution equals numpy's complex lstsq;
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


def _embed(x):
    """complex [N] -> real rows a = [Re; Im], b = [Im; -Re] (float64)"""
    x = np.asarray(x, np.complex128)
    return np.concatenate([x.real, x.imag]), np.concatenate([x.imag, -x.real])


def _soft(r, eB):
    return r - eB if r > eB else (r + eB if r < -eB else 0.0)


def _two_complex_pilots(seed):
    rng = np.random.default_rng(seed)
    N = 3
    x = (rng.standard_normal((N, 3)) + 1j * rng.standard_normal((N, 3))).astype(np.complex64)
    s = np.array([0.7 - 0.2j, -0.4 + 0.9j, 0.3 + 0.5j], np.complex64)
    return N, x, s


def test_window_ramp_two_steps_closed_form(oracle):
    # Two steps of sampleStep 2 with windowSize >= 2: the window ramps W = 2 -> 4 real rows.
    # Step 1 (rows 0, 1; W = 2), zero state: f = 0, beta1_t = soft(y_t)/2/(wL|x_t|^2 + wG).
    # Step 2 (rows 0..3; W = 4), state after step 1 (c_0, c_1 = beta1; c_2 = c_3 = 0):
    #   f_t = wL * (beta1_0 x_0 + beta1_1 x_1).x_t + wG * (beta1_0 g(0,t) + beta1_1 g(1,t)),
    #   beta2_t = soft(y_t - f_t)/4/(wL|x_t|^2 + wG), c_t = beta1_t + beta2_t (t < 2) or beta2_t,
    #   theta = sum_t (beta1_t + beta2_t) x_t.
    N, x, s = _two_complex_pilots(11)
    wG, eB, v = 0.45, 0.02, 0.4
    wL = 1 - wG
    cfg = dict(windowSize=3, dictionarySize=8, eB=eB, gaussianKernelVariance=v, gaussianKernelWeight=wG)
    o = oracle.Apsm(cfg, N)
    o.train(x[:, :2], s[:2])
    st = o.state()
    var = st["var"]
    a0, b0 = _embed(x[:, 0])
    a1, b1 = _embed(x[:, 1])
    X = [a0, b0, a1, b1]
    y = [float(s[0].real), float(s[0].imag), float(s[1].real), float(s[1].imag)]
    g = lambda i, j: np.exp(-0.5 / var * np.sum((X[i] - X[j]) ** 2))  # noqa: E731
    den = [wL * X[t] @ X[t] + wG for t in range(4)]

    def two_steps(W1, W2):
        beta1 = [_soft(y[t], eB) / W1 / den[t] for t in range(2)]
        th1 = beta1[0] * X[0] + beta1[1] * X[1]
        f = [wL * th1 @ X[t] + wG * (beta1[0] * g(0, t) + beta1[1] * g(1, t)) for t in range(4)]
        beta2 = [_soft(y[t] - f[t], eB) / W2 / den[t] for t in range(4)]
        c = [beta1[0] + beta2[0], beta1[1] + beta2[1], beta2[2], beta2[3]]
        return np.array(c), sum(c[t] * X[t] for t in range(4))

    c, th = two_steps(2, 4)
    np.testing.assert_allclose(st["coeffs"], c, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(st["theta"], th, rtol=1e-12, atol=1e-15)
    # the pin tells a wrong ramp apart: a constant W (2 or 4) or a window taken as the
    # whole stream in step 1 gives other coefficients
    for W1, W2 in ((2, 2), (4, 4), (1, 4)):
        cw, _ = two_steps(W1, W2)
        assert np.abs(cw - c).max() > 1e-3 * np.abs(c).max()
    # detection with the final state, written out: f(u) = wL theta.u + wG sum_j c_j g(d_j, u)
    u_a, u_b = _embed(x[:, 2])
    fa = wL * th @ u_a + wG * sum(c[j] * np.exp(-0.5 / var * np.sum((X[j] - u_a) ** 2)) for j in range(4))
    fb = wL * th @ u_b + wG * sum(c[j] * np.exp(-0.5 / var * np.sum((X[j] - u_b) ** 2)) for j in range(4))
    np.testing.assert_allclose(o.detect(x[:, 2:3])[0], [fa, fb], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("incremental", [False, True])
def test_full_window_start_closed_form(oracle, incremental):
    # startWithFullWindow, windowSize 2 (2w = 4 real rows), sampleStep 2, three pilots:
    # step 1 at s = 4 projects rows 0..3 from the zero state with W = 4; step 2 at s = 6
    # projects rows 2..5 (the newest 4) with W = 4 against the state after step 1.
    N, x, s = _two_complex_pilots(12)
    wG, eB, v = 0.5, 0.01, 0.3
    wL = 1 - wG
    cfg = dict(windowSize=2, dictionarySize=8, eB=eB, gaussianKernelVariance=v, gaussianKernelWeight=wG,
               startWithFullWindow=True)
    o = oracle.Apsm(cfg, N)
    o.train(x, s, incremental=incremental)
    st = o.state()
    var = st["var"]
    X = [r for k in range(3) for r in _embed(x[:, k])]
    y = [v_ for k in range(3) for v_ in (float(s[k].real), float(s[k].imag))]
    g = lambda i, j: np.exp(-0.5 / var * np.sum((X[i] - X[j]) ** 2))  # noqa: E731
    den = [wL * X[t] @ X[t] + wG for t in range(6)]

    def two_steps(W1, W2, rows2):
        beta1 = [_soft(y[t], eB) / W1 / den[t] for t in range(4)]
        th1 = sum(beta1[t] * X[t] for t in range(4))
        c = beta1 + [0.0, 0.0]
        f = {t: wL * th1 @ X[t] + wG * sum(beta1[j] * g(j, t) for j in range(4)) for t in rows2}
        for t in rows2:
            c[t] += _soft(y[t] - f[t], eB) / W2 / den[t]
        return np.array(c), sum(c[t] * X[t] for t in range(6))

    c, th = two_steps(4, 4, range(2, 6))
    np.testing.assert_allclose(st["coeffs"], c, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(st["theta"], th, rtol=1e-12, atol=1e-15)
    for W1, W2, rows2 in ((2, 4, range(2, 6)), (4, 6, range(0, 6)), (4, 2, range(4, 6))):
        cw, _ = two_steps(W1, W2, rows2)
        assert np.abs(cw - c).max() > 1e-3 * np.abs(c).max()


MULTI_CFGS = [
    dict(windowSize=10, dictionarySize=200, eB=0.01, gaussianKernelVariance=0.1),
    dict(windowSize=10, dictionarySize=40, eB=0.01, gaussianKernelVariance=0.1),  # FIFO wrap
    dict(windowSize=6, dictionarySize=30, eB=0.01, gaussianKernelVariance=0.3, sampleStep=4),
    dict(windowSize=10, dictionarySize=200, eB=0.01, gaussianKernelVariance=0.1, startWithFullWindow=True),
    dict(windowSize=8, dictionarySize=100, eB=0.001, gaussianKernelVariance=0.0),  # variance heuristic
]


@pytest.mark.parametrize("cfg", MULTI_CFGS)
def test_streamed_multi_user_form_equals_single_user_forms(oracle, cfg):
    # The streamed form (shared stream, window Gram kept between steps) against the
    # incremental and naive single-user forms, user by user, over two train calls.
    d = small_problem(N=4, U=3, T=160, M=60, seed=9)
    m = oracle.ApsmMulti(cfg, 4, 3)
    m.train(d["rx_train"][:, :90], d["tx_train"][:, :90])
    m.train(d["rx_train"][:, 90:], d["tx_train"][:, 90:])
    em, sm = m.detect(d["rx_data"]), m.state()
    for u in range(3):
        for inc in (True, False):
            o = oracle.Apsm(cfg, 4)
            o.train(d["rx_train"][:, :90], d["tx_train"][u, :90], incremental=inc)
            o.train(d["rx_train"][:, 90:], d["tx_train"][u, 90:], incremental=inc)
            so = o.state()
            assert so["used"] == sm["used"] and so["lo"] == sm["lo"] and so["var"] == sm["var"]
            tol = 0.0 if inc else 1e-12
            np.testing.assert_allclose(sm["coeffs"][u], so["coeffs"], rtol=0, atol=tol)
            np.testing.assert_allclose(sm["theta"][u], so["theta"], rtol=0, atol=tol)
            np.testing.assert_allclose(em[u], o.detect(d["rx_data"]), rtol=0, atol=tol)


def test_streamed_form_rejects_what_it_does_not_restate(oracle):
    for extra in (dict(trainPasses=2), dict(normConstraint=0.3), dict(llsInitialization=True), dict(trainingShuffle=True)):
        m = oracle.ApsmMulti(dict(windowSize=4, dictionarySize=50, gaussianKernelVariance=0.1, **extra), 4, 2)
        d = small_problem(N=4, U=2, T=50, M=4, seed=1)
        with pytest.raises(ValueError):
            m.train(d["rx_train"], d["tx_train"][:2])


@pytest.mark.parametrize("s", [np.complex64(0.8 - 0.6j), np.complex64(-0.5 + 0.004j), np.complex64(0.003 - 0.002j)])
def test_projection_contracts_the_excess_over_the_hyperslab(oracle, s):
    # Linear model only (gaussianKernelWeight 0, so wL = 1), one complex pilot, one step from
    # the zero state with W = 2 real rows: a.b = 0 for the two embedded rows, so each row's
    # f after the step is beta_t |x|^2 = soft(y_t, eB) / 2. A row outside its hyperslab keeps
    # its side and halves its excess: |y - f| - eB = (|y| - eB) / 2; a row inside gets beta = 0
    # (coefficient and share of theta exactly zero).
    rng = np.random.default_rng(21)
    N, eB = 4, 0.01
    x = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    cfg = dict(windowSize=1, dictionarySize=4, eB=eB, gaussianKernelVariance=0.3, gaussianKernelWeight=0.0)
    o = oracle.Apsm(cfg, N)
    o.train(x[:, None], np.array([s]))
    st = o.state()
    f = o.detect(x[:, None])[0]
    a, b = _embed(x)
    y = np.array([float(s.real), float(s.imag)])
    for t in range(2):
        if abs(y[t]) > eB:
            r = y[t] - f[t]
            assert np.sign(r) == np.sign(y[t])
            assert abs((abs(r) - eB) - (abs(y[t]) - eB) / 2) < 1e-12
        else:
            assert st["coeffs"][t] == 0.0
            assert abs(f[t]) < 1e-15
    # theta carries exactly the outside rows' coefficients
    beta = [_soft(y[t], eB) / 2 / (a @ a) for t in range(2)]
    np.testing.assert_allclose(st["theta"], beta[0] * a + beta[1] * b, rtol=1e-12, atol=1e-15)
