"""Detect arithmetic modes (detectPrecision) against the FP64 oracle.

* fp32-vector and fp32-matrix (cross terms on the vector ALUs or on the matrix cores, FP32
  inputs and accumulation): the FP32 parity bar of test_gpu_parity (soft outputs within 2e-3,
  hard decisions equal away from decision boundaries), for every antenna bucket and for user
  groups that fill and overflow a 16-user launch.
* bf16 and fp16 (cross-term inputs rounded to 16 bits, FP32 accumulation, FP32 norms,
  exponentials and contraction): the soft-output RMS error relative to the RMS soft output
  and the symbol-error count against the oracle's, with the bounds written below.
"""
import zlib

import numpy as np
import pytest

from helpers import small_problem, to_ri
from test_gpu_parity import compare, orc_run

pytestmark = pytest.mark.gpu

FP32_MODES = ["fp32-vector", "fp32-matrix"]
# relative soft-output RMS error bound and relative SER deviation bound (plus 3 symbols of slack)
HALF_BOUNDS = {"bf16": (0.05, 0.05), "fp16": (0.01, 0.01)}


def run_mode(cfg, N, users, d, mode):
    from libapsm_amd import ApsmNomaDetectorFactory
    det = ApsmNomaDetectorFactory.build(dict(cfg, otype="APSM", detectPrecision=mode), N, users)
    det.train(d["rx_train"], d["tx_train"][:users] if users > 1 else d["tx_train"][0])
    est = det.detect(d["rx_data"])
    return est[None, :] if users == 1 else est


@pytest.mark.parametrize("mode", FP32_MODES)
@pytest.mark.parametrize("N", [3, 4, 8, 13, 16, 32, 64])
def test_fp32_modes_match_oracle(require_gpu, oracle, mode, N):
    print(f"mode {mode} N={N}", flush=True)
    cfg = dict(windowSize=20, dictionarySize=300, eB=0.01, gaussianKernelVariance=0.25 / (N * N))
    d = small_problem(N=N, U=3, T=300, M=777, seed=100 + N)
    est = run_mode(cfg, N, 1, d, mode)
    _, est_o = orc_run(oracle, cfg, N, d["rx_train"], d["tx_train"][0], d["rx_data"])
    compare(est[0], est_o, 2)


@pytest.mark.parametrize("mode", FP32_MODES + ["bf16", "fp16"])
@pytest.mark.parametrize("users", [5, 19])
def test_user_groups_equal_single_user_handles(require_gpu, mode, users):
    """Users are launched in groups of up to 16 with the kernel instantiated for the group's
    padded user count; every user still gets exactly what a one-user handle computes."""
    print(f"mode {mode} users={users}", flush=True)
    cfg = dict(windowSize=16, dictionarySize=200, eB=0.01, gaussianKernelVariance=0.25 / 64)
    d = small_problem(N=8, U=users, T=250, M=300, seed=users)
    est = run_mode(cfg, 8, users, d, mode)
    for u in (0, users // 2, users - 1):
        one = dict(d, tx_train=d["tx_train"][u:u + 1])
        np.testing.assert_array_equal(est[u], run_mode(cfg, 8, 1, one, mode)[0], err_msg=f"user {u}")


@pytest.mark.parametrize("mode", ["bf16", "fp16"])
@pytest.mark.parametrize("N,mod,var", [(4, "QPSK", 0.05), (4, "QPSK", "sigma0.5"), (8, "QAM16", 0.05), (8, "QAM16", "sigma0.5"),
                                       (16, "QAM16", 0.05), (32, "QAM16", "sigma0.5"), (64, "QPSK", 0.05), (64, "QPSK", "sigma0.5")])
def test_half_precision_error_vs_oracle(require_gpu, oracle, mode, N, mod, var):
    """Six superimposed users at 12 dB so that symbol errors occur; the reference's default
    kernel variance (a wide kernel: every dictionary entry contributes) and sigma = 0.5 after
    power normalisation (a narrow kernel)."""
    from libapsm_amd import datagen
    print(f"mode {mode} N={N} {mod} var={var}", flush=True)
    bps = datagen.BITS_PER_SYMBOL[mod]
    v = datagen.kernel_variance_for_sigma(0.5, N) if var == "sigma0.5" else var
    cfg = dict(windowSize=40, dictionarySize=800, eB=0.01, gaussianKernelVariance=v)
    d = small_problem(N=N, U=6, T=1500, M=12000, seed=zlib.crc32(f"{mode}{N}{var}".encode()) % 1000, mod=mod, snr_db=12.0)
    est = run_mode(cfg, N, 1, d, mode)[0]
    ref32 = run_mode(cfg, N, 1, d, "fp32")[0]
    _, est_o = orc_run(oracle, cfg, N, d["rx_train"], d["tx_train"][0], d["rx_data"])
    g = to_ri(est).astype(np.float64)
    rms_o = np.sqrt(np.mean(est_o ** 2))
    rel_rms = np.sqrt(np.mean((g - est_o) ** 2)) / rms_o
    rel_rms32 = np.sqrt(np.mean((to_ri(ref32).astype(np.float64) - est_o) ** 2)) / rms_o
    tx = to_ri(d["tx_data"][0])
    e_p = oracle.symbol_errors(g, tx, bps)
    e_o = oracle.symbol_errors(est_o, tx, bps)
    e_32 = oracle.symbol_errors(to_ri(ref32), tx, bps)
    print(f"{mode} N={N} {mod} var={var}: rel soft RMS err {rel_rms:.3e} (fp32 {rel_rms32:.3e}), "
          f"max {np.abs(g - est_o).max():.3e}; symbol errors {e_p} (fp32 {e_32}, oracle {e_o}) of {tx.shape[0]}")
    rms_bound, ser_bound = HALF_BOUNDS[mode]
    assert rel_rms <= rms_bound
    assert abs(e_p - e_o) <= ser_bound * e_o + 3


def test_switching_precision_keeps_the_trained_state(require_gpu):
    """setDetectPrecision on a trained handle gives what a handle created with that precision
    computes after the same train (the train does not depend on the detect arithmetic)."""
    from libapsm_amd import ApsmNomaDetectorFactory
    cfg = dict(windowSize=16, dictionarySize=300, eB=0.01, gaussianKernelVariance=0.05)
    d = small_problem(N=16, U=4, T=400, M=2000, seed=77)
    det = ApsmNomaDetectorFactory.build(dict(cfg, otype="APSM"), 16)
    det.train(d["rx_train"], d["tx_train"][0])
    for mode in ["bf16", "fp32-vector", "fp16", "fp32-matrix"]:
        det.setDetectPrecision(mode)
        np.testing.assert_array_equal(det.detect(d["rx_data"]), run_mode(cfg, 16, 1, d, mode)[0], err_msg=mode)
    with pytest.raises(ValueError):
        det.setDetectPrecision("int8")


# Bounds of the relative soft-output RMS error against the FP64 oracle where the Gaussian term
# carries at least GSHARE_MIN of f: 16-bit cross-term inputs perturb each squared distance by
# about 2^-8 (bf16) or 2^-11 (fp16) of |x||z|, i.e. each kernel value by that much times
# |x||z| / (2 sigma^2); FP32 stays at rounding level. The 16-bit error must also stand clearly
# above the FP32 error (otherwise the test could not see a wrong rounding).
ACTIVE_BOUNDS = {"fp32": 1e-4, "bf16": 1e-2, "fp16": 2e-3}
GSHARE_MIN = 0.10


@pytest.mark.parametrize("N", [32, 64])
def test_half_precision_where_the_gaussian_term_acts(require_gpu, oracle, N):
    """12 QPSK users at 12 dB (the mixed-precision workload's signal), gaussianKernelWeight
    0.99 and sigma^2 at half the median squared pilot-pair distance, so that the Gaussian term
    is a large share of f and the 16-bit cross terms show in the soft outputs."""
    from libapsm_amd import ApsmNomaDetectorFactory, datagen
    d = datagen.generate(N, 12, 2048, 8192, modulation="QPSK", snr_db=12.0, seed=40 + N)
    sigma2 = datagen.median_pair_variance(d["rx_train"])
    cfg = dict(windowSize=80, dictionarySize=2048, eB=0.01, gaussianKernelVariance=sigma2 / (N * N),
               gaussianKernelWeight=0.99)
    print(f"N={N} sigma2 {sigma2:.2f}", flush=True)
    o = oracle.Apsm(cfg, N)
    o.train(d["rx_train"], d["tx_train"][0], incremental=True)
    est_o = o.detect(d["rx_data"])
    th = o.state()["theta"]
    x = d["rx_data"].astype(np.complex128)
    lin = 0.01 * np.stack([th[:N] @ x.real + th[N:] @ x.imag, th[:N] @ x.imag - th[N:] @ x.real], -1)
    rms_o = np.sqrt(np.mean(est_o ** 2))
    gshare = np.sqrt(np.mean((est_o - lin) ** 2)) / rms_o
    print(f"Gaussian share of f: {gshare:.3f}", flush=True)
    assert gshare >= GSHARE_MIN
    det = ApsmNomaDetectorFactory.build(dict(cfg, otype="APSM"), N)
    det.train(d["rx_train"], d["tx_train"][0])
    tx = to_ri(d["tx_data"][0])
    e_o = oracle.symbol_errors(est_o, tx, 2)
    rel = {}
    for mode in ("fp32", "bf16", "fp16"):
        det.setDetectPrecision(mode)
        g = to_ri(det.detect(d["rx_data"])).astype(np.float64)
        rel[mode] = np.sqrt(np.mean((g - est_o) ** 2)) / rms_o
        e = oracle.symbol_errors(g, tx, 2)
        print(f"N={N} {mode}: rel soft RMS err {rel[mode]:.3e}, max {np.abs(g - est_o).max():.3e}; "
              f"symbol errors {e} (oracle {e_o})", flush=True)
        assert rel[mode] <= ACTIVE_BOUNDS[mode], mode
        assert abs(e - e_o) <= 0.01 * e_o + 3, mode
    assert rel["bf16"] > 5 * rel["fp32"]
    assert rel["bf16"] > rel["fp16"]
