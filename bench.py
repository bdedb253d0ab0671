#!/usr/bin/env python3
"""Benchmark: plain APSM train + detect on one MI355X.

Default workload (16 antennas, 10 superimposed 16QAM users with large-scale gains
uniform in [-6, 0] dB, 15 dB SNR; all 10 users detected; windowSize 160; dictionary
capped at 4096 entries; 200k pilots and 2M detect samples): one step = reset, train on
the pilots and detect all data samples for all users, inputs resident in HBM.
Metric: detected symbols/s = data samples x users / step time.

Prints ONE JSON line. `roofline` covers the dominant kernel (by device time) with its
algorithmic FLOPs; `cpu_baseline` times the FP32 restatement of the reference algorithm
(the reference's own schedule) on host cores over a bounded sample and scales it to the
whole workload.
"""
import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3


# HBM bytes of one train call of a workload (all its dispatches: sequencer, block precompute,
# combine, staging, theta), from two rocprofv3 --pmc passes over `bench.py --steps 1 --warmup 0
# --no-cpu-baseline --no-lls --no-batched --no-mixed --no-serving --no-small`: FETCH_SIZE
# 6 603 568 KiB x 2.000 (a 1 GiB device copy, `bench.py --only copy`, reads 524 315 KiB of
# FETCH_SIZE) + WRITE_SIZE 4 593 643 KiB x 1.000 (the copy reads 1 048 576 KiB).
TRAIN_HBM_BYTES = {"plain-16ant-16qam": (6603568 * 2.000 + 4593643 * 1.000) * 1024}
# HBM bytes of one apsm_train_seq2 launch, from the same calibrated passes restricted to the
# sequencer kernel: FETCH_SIZE 8 717.1 KiB x 2.000 + WRITE_SIZE 53.7 KiB x 1.000 per dispatch.
SEQ_HBM_BYTES_PER_LAUNCH = {"plain-16ant-16qam": (8717.1 * 2.000 + 53.7 * 1.000) * 1024}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="plain-16ant-16qam")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-lls", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-instances measurement")
    ap.add_argument("--batched-instances", type=int, default=512)
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision large-instance measurement")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--no-small", action="store_true", help="skip the small plain workload (GPU, oracle SER, host cores)")
    ap.add_argument("--serving-frames", type=int, default=1200)
    ap.add_argument("--serving-links", type=int, default=2048)
    ap.add_argument("--only", choices=["batched", "mixed", "serving", "copy"], default=None,
                    help="run one further section alone (for profiling); prints {section: result}; "
                         "copy: a device-to-device copy of known size (calibrates HBM byte counters)")
    ap.add_argument("--mixed-samples", type=int, default=None, help="override the mixed workload's detect samples")
    return ap.parse_args()


def train_flops_per_user(W2, H, N, R2, q=2):
    """Algorithmic FLOPs of the incremental train for one user: each entering complex row
    against the dictionary (8N cross + 16 per complex pair) and the window Gram product
    (2 W^2 per step)."""
    total = 0.0
    steps = R2 // q
    s = np.arange(1, steps + 1, dtype=np.float64) * q
    G = np.minimum(s - q, H) / 2.0  # complex dictionary entries seen by the entering row
    W = np.minimum(W2, s)
    total = np.sum((q / 2.0) * G * (8.0 * N + 16.0) + 2.0 * W * W)
    return float(total)


def seq_flops_per_user(W2, R2, q=2):
    """Algorithmic FLOPs of the train sequencer's own work for one user over a whole train:
    per projection step the window Gram product (W^2 FMA over the W real window rows) and
    the entering rows' share against the window (4 FMA per complex pair: 2qW FLOP)."""
    s = np.arange(1, R2 // q + 1, dtype=np.float64) * q
    W = np.minimum(W2, s)
    return float(np.sum(2.0 * W * W + 2.0 * q * W))


def naive_evals(W2, H, R2, q=2):
    """Real kernel evaluations of the reference schedule for one user: every step evaluates
    every window row against every dictionary row (apsm_train.cu:333-356)."""
    s = np.arange(1, R2 // q + 1, dtype=np.float64) * q
    return float(np.sum(np.minimum(W2, s) * np.minimum(s, H)))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(w, data, budget_s, threads=None):
    """The reference algorithm (FP32 restatement, the reference's schedule: every step
    evaluates every window row against the whole dictionary) timed on host cores over a
    bounded sample: one user trained on the first Ts pilots and Ms samples detected, Ts and Ms
    doubled until a part takes a quarter of the budget (or covers the whole workload); the
    per-kernel-evaluation times are scaled to the whole workload (all users)."""
    from libapsm_amd import datagen
    from oracle import pyoracle

    cfg = {k: v for k, v in datagen.workload_config(w).items() if k != "otype"}
    N, T, M = w["N"], w["T"], w["M"]
    W2, H = 2 * w["window"], 2 * w["dict_size"]
    user = w["users"][0]
    prev = pyoracle.threads()
    if threads:
        pyoracle.set_threads(threads)
    try:
        used_threads = pyoracle.threads()
        Ts = min(256, T)
        while True:
            o = pyoracle.Apsm(cfg, N, fp32=True)
            t0 = time.perf_counter()
            o.train(data["rx_train"][:, :Ts], data["tx_train"][user, :Ts])
            t_train = time.perf_counter() - t0
            if t_train > budget_s / 4 or Ts >= min(8192, T):
                break
            Ts = min(2 * Ts, T)
        Ms = min(512, M)
        while True:
            t0 = time.perf_counter()
            o.detect(data["rx_data"][:, :Ms])
            t_det = time.perf_counter() - t0
            if t_det > budget_s / 4 or Ms >= min(65536, M):
                break
            Ms = min(2 * Ms, M)
    finally:
        pyoracle.set_threads(prev)
    per_train_eval = t_train / naive_evals(W2, H, 2 * Ts)
    per_det_eval = t_det / (2.0 * Ms * min(2 * Ts, H))
    U = len(w["users"])
    full_train = naive_evals(W2, H, 2 * T) * per_train_eval
    full_det = 2.0 * M * min(2 * T, H) * per_det_eval
    total = U * (full_train + full_det)
    return {
        "value": M * U / total,
        "unit": "detected symbols/s",
        "cores": used_threads,
        "kind": "port",
        "cpu_model": cpu_model(),
        "measured_s": t_train + t_det,
        "extrapolated_s": total,
        "sample": (f"FP32 restatement, reference schedule, {used_threads} thread(s); user {user} trained on the first {Ts} "
                   f"of {T} pilots ({t_train:.2f} s) and {Ms} of {M} samples detected against that dictionary "
                   f"({t_det:.2f} s); per-kernel-evaluation times scaled to {U} user(s) x ({T} pilots, {M} samples): "
                   f"train {U * full_train:.1f} s + detect {U * full_det:.1f} s"),
    }


def bench_small_plain(steps, budget_s):
    """The small plain workload (8 antennas, 6 superimposed QPSK users at 10 dB, user 1
    detected, windowSize 80, sigma 0.5, 20k pilots, 80k samples) on the GPU, its symbol errors
    beside the FP64 oracle's (tests/golden/scale_c1.npz, all 80k samples), and the reference
    algorithm on the host cores (all cores and one core)."""
    from libapsm_amd import ApsmNomaDetectorFactory, datagen, lib
    from libapsm_amd._lib import DeviceArray, check

    w = datagen.WORKLOADS["plain-cpu-8ant"]
    N, U, T, M = w["N"], len(w["users"]), w["T"], w["M"]
    data = datagen.generate(N, w["U_tx"], T, M, modulation=w["modulation"], snr_db=w["snr_db"], seed=w["seed"],
                            large_scale_db=w["large_scale_db"])

    def to_dev(z):
        return DeviceArray(np.ascontiguousarray(z, dtype=np.complex64).view(np.float32))

    d_rxt, d_txt, d_rxd = to_dev(data["rx_train"]), to_dev(data["tx_train"][w["users"]]), to_dev(data["rx_data"])
    d_est = DeviceArray.empty((U, M, 2), np.float32)
    det = ApsmNomaDetectorFactory.build(datagen.workload_config(w), N, U)

    def step():
        det.resetState()
        det.train_device(d_rxt.ptr, d_txt.ptr, T)
        det.detect_device(d_rxd.ptr, M, d_est.ptr)

    step()
    check(lib().apsm_device_synchronize())
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    check(lib().apsm_device_synchronize())
    ms_step = 1000.0 * (time.perf_counter() - t0) / steps
    est = d_est.numpy()
    errs = sum(datagen.symbol_errors(est[k, :, 0] + 1j * est[k, :, 1], data["tx_data"][u], data["bps"])
               for k, u in enumerate(w["users"]))
    fx = np.load(os.path.join(ROOT, "tests", "golden", "scale_c1.npz"), allow_pickle=False)
    e_orc = int(fx["symbol_errors"].sum())
    return {
        "workload": "plain-cpu-8ant",
        "value": M * U / (ms_step / 1000.0),
        "unit": "detected symbols/s",
        "ms_per_step": ms_step,
        "ser": errs / (M * U),
        "ser_oracle": e_orc / (M * U),
        "symbol_errors": int(errs),
        "symbol_errors_oracle": e_orc,
        "cpu_baseline": cpu_baseline(w, data, budget_s),
        "cpu_baseline_one_core": cpu_baseline(w, data, budget_s / 2, threads=1),
    }


def bench_batched(count, steps):
    """Batched workload: `count` independent instances (antennas in {4, 8, 16, 32}, QPSK or
    16QAM, dictionary 200..4000, windowSize 40/80/160, 10k..100k detect samples, pilots =
    dictionary size) run as one planned, graph-captured batch; the same instances through the
    plain single-handle path one after another give the comparison. Value: detected symbols/s
    of the whole batch (reset + train + detect of every instance per run)."""
    from concurrent.futures import ThreadPoolExecutor

    from libapsm_amd import ApsmBatch, ApsmNomaDetectorFactory, datagen, lib
    from libapsm_amd._lib import DeviceArray, check

    inst = datagen.batched_instances(count)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        data = list(ex.map(datagen.batched_payload, inst))
    t_gen = time.perf_counter() - t0
    cfgs = [datagen.batched_config(x) for x in inst]
    up = lambda z: DeviceArray(np.ascontiguousarray(z, dtype=np.complex64).view(np.float32))  # noqa: E731
    d_rxt = [up(d["rx_train"]) for d in data]
    d_txt = [up(d["tx_train"][x["user"]][None, :]) for d, x in zip(data, inst)]
    d_rxd = [up(d["rx_data"]) for d in data]
    d_est = [DeviceArray.empty((1, x["M"], 2), np.float32) for x in inst]
    T = [x["T"] for x in inst]
    M = [x["M"] for x in inst]
    symbols = float(sum(M))

    b = ApsmBatch(cfgs, [x["N"] for x in inst])
    t0 = time.perf_counter()
    b.plan_device([d.ptr for d in d_rxt], [d.ptr for d in d_txt], T, [d.ptr for d in d_rxd], M, [d.ptr for d in d_est])
    b.synchronize()
    t_plan = time.perf_counter() - t0
    b.run()
    b.synchronize()
    t0 = time.perf_counter()
    dev_ms, split = [], []
    for _ in range(steps):
        b.run()
        dev_ms.append(b.last_ms())
        split.append(b.last_split())
    b.synchronize()
    batch_s = (time.perf_counter() - t0) / steps
    est_b = [d.numpy() for d in d_est]

    # plain path: one handle per instance, one instance after another
    dets = [ApsmNomaDetectorFactory.build(dict(c, otype="APSM"), x["N"]) for c, x in zip(cfgs, inst)]
    check(lib().apsm_device_synchronize())
    t0 = time.perf_counter()
    for i, det in enumerate(dets):
        det.resetState()
        det.train_device(d_rxt[i].ptr, d_txt[i].ptr, T[i])
        det.detect_device(d_rxd[i].ptr, M[i], d_est[i].ptr)
        det.synchronize()
    plain_s = time.perf_counter() - t0
    same = all(np.array_equal(est_b[i], d_est[i].numpy()) for i in range(count))
    errs = 0
    for i, (d, x) in enumerate(zip(data, inst)):
        e = est_b[i][0]
        errs += datagen.symbol_errors(e[:, 0] + 1j * e[:, 1], d["tx_data"][x["user"]], d["bps"])
    return {
        "workload": f"batched-{count}",
        "instances": count,
        "detected_symbols": int(symbols),
        "value": symbols / batch_s,
        "unit": "detected symbols/s",
        "ms_per_run": 1000 * batch_s,
        "device_ms_per_run": float(np.mean(dev_ms)),
        "device_train_ms": float(np.mean([t for t, _ in split])),
        "device_detect_ms": float(np.mean([d for _, d in split])),
        "plain_value": symbols / plain_s,
        "plain_ms": 1000 * plain_s,
        "speedup_vs_plain": plain_s / batch_s,
        "identical_to_plain": bool(same),
        "ser": errs / symbols,
        "plan_s": t_plan,
        "datagen_s": t_gen,
        "tuning": b.tuning(),
    }


def fp64_detect(x, dict_c, coeffs, theta, var, wL, wG):
    """FP64 evaluation of a trained detector's output for samples x [N][S] (complex):
    wL theta.x + wG sum_j (c_a K + ...) with the complex-pair kernel terms
    E0 = exp(-(|x|^2 + |z|^2 - 2p) / 2var), E1 = ...(- 2s), E2 = ...(+ 2s), p = Re(x.conj z),
    s = -Im(x.conj z). Returns [U][S][2]."""
    x = x.astype(np.complex128)
    z = dict_c.astype(np.complex128)
    N = x.shape[0]
    th = theta.astype(np.float64)
    c = coeffs.astype(np.float64)
    xn = np.sum(np.abs(x) ** 2, axis=0)
    zn = np.sum(np.abs(z) ** 2, axis=0)
    xz = x.T @ np.conj(z)  # [S][G]
    p, sp = xz.real, -xz.imag
    nn = xn[:, None] + zn[None, :]
    E0, E1, E2 = (np.exp(-np.maximum(nn - 2 * p, 0) / (2 * var)), np.exp(-np.maximum(nn - 2 * sp, 0) / (2 * var)),
                  np.exp(-np.maximum(nn + 2 * sp, 0) / (2 * var)))
    out = np.zeros((th.shape[0], x.shape[1], 2))
    for u in range(th.shape[0]):
        ca, cb = c[u, 0::2], c[u, 1::2]
        la = th[u, :N] @ x.real + th[u, N:] @ x.imag
        lb = th[u, :N] @ x.imag - th[u, N:] @ x.real
        out[u, :, 0] = wL * la + wG * (E0 @ ca + E1 @ cb)
        out[u, :, 1] = wL * lb + wG * (E2 @ ca + E0 @ cb)
    return out


def bench_mixed(steps, samples=None):
    """Mixed-precision large instance: 64 antennas, 12 QPSK users at 12 dB, dictionary of
    16384 Gaussian centres (16384 pilots), 8M detect samples. The kernel is set where the
    Gaussian term carries a large share of f: gaussianKernelWeight 0.99 and sigma^2 = half the
    median squared pilot-pair distance (datagen.median_pair_variance; the reference's default
    variance or its heuristic leave the Gaussian share near zero at 64 antennas, where the
    cross-term precision could not show). One train, then the detect of all users in FP32
    (matrix cores), BF16 and FP16 cross terms. Per dtype: detect time and symbols/s, symbol
    errors over all samples, and the soft-output error against an FP64 evaluation of the same
    trained detector on the first 4096 samples; the Gaussian share of that evaluation."""
    from libapsm_amd import ApsmNomaDetectorFactory, datagen
    from libapsm_amd._lib import DeviceArray

    w = dict(datagen.WORKLOADS["mixed-64ant-qpsk"])
    N, U, T = w["N"], len(w["users"]), w["T"]
    M = samples or w["M"]
    t0 = time.perf_counter()
    data = datagen.generate(N, w["U_tx"], T, M, modulation=w["modulation"], snr_db=w["snr_db"], seed=w["seed"])
    t_gen = time.perf_counter() - t0
    up = lambda z: DeviceArray(np.ascontiguousarray(z, dtype=np.complex64).view(np.float32))  # noqa: E731
    d_rxt, d_txt, d_rxd = up(data["rx_train"]), up(data["tx_train"][w["users"]]), up(data["rx_data"])
    d_est = DeviceArray.empty((U, M, 2), np.float32)
    sigma2 = datagen.median_pair_variance(data["rx_train"])
    wG = 0.99
    cfg = {"otype": "APSM", "windowSize": w["window"], "eB": w["eB"], "dictionarySize": w["dict_size"],
           "gaussianKernelVariance": sigma2 / (N * N), "gaussianKernelWeight": wG}
    det = ApsmNomaDetectorFactory.build(cfg, N, U)
    det.train_device(d_rxt.ptr, d_txt.ptr, T)
    det.synchronize()
    train_ms = det.kernel_times()[0]
    st = det.getTrainingState()
    dict_c = det.getDictionary()
    G = dict_c.shape[1]
    sub = min(4096, M)
    ref = fp64_detect(data["rx_data"][:, :sub], dict_c, st["coeffs"], st["theta"], st["variance"],
                      det.par.linearKernelWeight, det.par.gaussianKernelWeight)
    rms_ref = float(np.sqrt(np.mean(ref ** 2)))
    lin = fp64_detect(data["rx_data"][:, :sub], dict_c, np.zeros_like(st["coeffs"]), st["theta"], st["variance"],
                      det.par.linearKernelWeight, det.par.gaussianKernelWeight)
    gshare = float(np.sqrt(np.mean((ref - lin) ** 2)) / rms_ref)
    tx = data["tx_data"][w["users"]]
    ref_c = ref[..., 0] + 1j * ref[..., 1]
    e_ref = datagen.symbol_errors(ref_c, tx[:, :sub], data["bps"])
    out = {"workload": "mixed-64ant-qpsk", "antennas": N, "users": U, "dictionary": G, "detect_samples": M,
           "pilots": T, "gaussianKernelWeight": wG, "sigma2": sigma2, "gaussian_share_fp64": gshare,
           "train_ms": train_ms, "datagen_s": t_gen,
           "fp64_subsample": sub, "fp64_symbol_errors_subsample": int(e_ref), "dtypes": {}}
    for mode in ("fp32", "bf16", "fp16"):
        det.setDetectPrecision(mode)
        det.detect_device(d_rxd.ptr, M, d_est.ptr)
        det.synchronize()
        ms = []
        for _ in range(steps):
            det.detect_device(d_rxd.ptr, M, d_est.ptr)
            det.synchronize()
            ms.append(det.kernel_times()[1])
        est = d_est.numpy()
        est_c = est[..., 0] + 1j * est[..., 1]
        errs = datagen.symbol_errors(est_c, tx, data["bps"])
        e_sub = datagen.symbol_errors(est_c[:, :sub], tx[:, :sub], data["bps"])
        det_ms = float(np.mean(ms))
        out["dtypes"][mode] = {
            "detect_ms": det_ms,
            "symbols_per_s": M * U / (det_ms / 1000.0),
            "gaussian_pairs_per_s": M * G / (det_ms / 1000.0),
            "ser": errs / est_c.size,
            "symbol_errors": int(errs),
            "soft_rms_err_vs_fp64": float(np.sqrt(np.mean((est[:, :sub].astype(np.float64) - ref) ** 2))),
            "soft_rel_rms_err_vs_fp64": float(np.sqrt(np.mean((est[:, :sub].astype(np.float64) - ref) ** 2)) / rms_ref),
            "soft_max_err_vs_fp64": float(np.abs(est[:, :sub].astype(np.float64) - ref).max()),
            "symbol_errors_subsample": int(e_sub),
        }
    f32 = out["dtypes"]["fp32"]
    for mode in ("bf16", "fp16"):
        m = out["dtypes"][mode]
        m["ser_rel_diff_vs_fp32"] = abs(m["symbol_errors"] - f32["symbol_errors"]) / max(1, f32["symbol_errors"])
        m["speedup_vs_fp32"] = f32["detect_ms"] / m["detect_ms"]
        m["soft_err_ratio_vs_fp32"] = m["soft_rms_err_vs_fp64"] / max(f32["soft_rms_err_vs_fp64"], 1e-30)
    for mode in ("fp32", "bf16", "fp16"):
        m = out["dtypes"][mode]
        m["symbol_errors_subsample_vs_fp64"] = m["symbol_errors_subsample"] - int(e_ref)
    return out


def bench_copy(n_bytes=1 << 30, reps=4):
    """Device-to-device copies of n_bytes (torch's copy kernel): the reference reading of the HBM
    byte counters (FETCH_SIZE, WRITE_SIZE) for a kernel that reads and writes n_bytes once."""
    import torch
    a = torch.ones(n_bytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    ms = 1000.0 * (time.perf_counter() - t0) / reps
    return {"bytes_read": n_bytes, "bytes_written": n_bytes, "launches": reps + 1, "ms": ms,
            "gb_per_s": 2.0 * n_bytes / (ms / 1000.0) / 1e9}


def h2d_pinned_gbps(n_bytes=64 << 20, reps=5):
    """Host-to-device copy rate from pinned host memory (the serving uploads' ceiling), through
    the library's own pinned and device buffers (synchronous copies of n_bytes). The library's
    HIP runtime is used, not torch's: torch's wheel carries a HIP runtime of its own, which
    finds no device once the library's runtime holds it."""
    from libapsm_amd._lib import DeviceArray, PinnedArray, check, lib
    a = PinnedArray((n_bytes // 4,), np.float32)
    a.array[...] = 1.0
    b = DeviceArray.empty((n_bytes // 4,), np.float32)
    check(lib().apsm_memcpy_h2d(b.ptr, a.ptr, n_bytes))
    t0 = time.perf_counter()
    for _ in range(reps):
        check(lib().apsm_memcpy_h2d(b.ptr, a.ptr, n_bytes))
    return reps * n_bytes / (time.perf_counter() - t0) / 1e9


def bench_serving(n_frames=600, n_links=2048, rates=(1000.0, 2000.0, 4000.0, 8000.0)):
    """Serving loop: up to n_links links (8..32 antennas, QPSK or 16QAM, windowSize 40,
    dictionary 1024) admitted by an HBM budget; frames of 14 OFDM symbols x 12 x U[1, 273]
    subcarriers plus 5 % pilots on uniformly drawn links, Poisson arrivals swept over `rates`
    (frames/s) and a saturation run (all frames at t = 0). Inputs and estimates live in pinned
    host memory, so the latency includes the copies: latency = estimates in host memory -
    arrival. The plain comparison serves the same frames one after another, each through a
    handle of its own link (detect, then train; synchronous)."""
    from concurrent.futures import ThreadPoolExecutor

    from libapsm_amd import ApsmError, ApsmNomaDetectorFactory, ApsmServer, Frame, PinnedArray, datagen, lib
    from libapsm_amd._lib import check

    links = datagen.serving_links(n_links, seed=5)
    max_pilots = int(round(0.05 * 14 * 12 * 273)) + 1
    budget, arena = 24 << 30, 4 << 30
    t0 = time.perf_counter()
    srv = ApsmServer(budget, arena)
    admitted = 0
    for lp in links:
        try:
            srv.add_link(datagen.serving_link_config(lp), lp["N"], 1, max_pilots)
        except ApsmError as e:
            if e.code != -5:
                raise
            break
        admitted += 1
    t_links = time.perf_counter() - t0
    base = datagen.serving_frames(n_frames, seed=5, links=admitted, rate_hz=1.0)

    def payload(f):
        p = datagen.serving_payload(f)
        out = {}
        for k in ("rx_data", "rx_pilots", "tx_pilots"):
            pa = PinnedArray(p[k].shape, np.complex64)
            pa.array[...] = p[k]
            out[k] = pa
        out["est"] = PinnedArray((1, f["samples"]), np.complex64)
        out["tx_data"] = p["tx_data"]
        out["bps"] = p["bps"]
        return out

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        pay = list(ex.map(payload, base))
    t_gen = time.perf_counter() - t0
    symbols = float(sum(f["samples"] for f in base))

    def frames(scale):
        return [Frame(f["link"], p["rx_data"].array, p["rx_pilots"].array, p["tx_pilots"].array,
                      arrival_s=f["arrival_s"] * scale, est=p["est"].array) for f, p in zip(base, pay)]

    # warm-up: one replay of every frame (code objects, first launches, and the first device
    # access to each pinned frame buffer, which costs milliseconds once per buffer; a server
    # reuses a pinned pool, so the timed runs measure the steady state)
    srv.replay(frames(0.0))
    def run(rate):
        fr = frames(0.0 if rate is None else 1.0 / rate)
        res = srv.replay(fr)
        lat = np.array([r["latency_s"] for r in res]) * 1e3
        span = max(r["output_s"] for r in res) - min(r["arrival_s"] for r in res)
        return {"offered_frames_per_s": rate, "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
                "symbols_per_s": symbols / span, "frames_per_s": len(fr) / span}

    runs = [run(rate) for rate in list(rates) + [None]]
    # the largest offered rate whose p99 latency stays below 1 ms: bisection between the swept
    # rates that bracket it (five halvings)
    swept = [r for r in runs if r["offered_frames_per_s"] is not None]
    under = [r for r in swept if r["p99_ms"] < 1.0]
    lo = max(under, key=lambda r: r["offered_frames_per_s"]) if under else None
    over = [r for r in swept if r["p99_ms"] >= 1.0 and (lo is None or r["offered_frames_per_s"] > lo["offered_frames_per_s"])]
    hi = min(over, key=lambda r: r["offered_frames_per_s"]) if over else None
    bisect = []
    if lo is not None and hi is not None:
        for _ in range(5):
            mid = run(0.5 * (lo["offered_frames_per_s"] + hi["offered_frames_per_s"]))
            bisect.append(mid)
            if mid["p99_ms"] < 1.0:
                lo = mid
            else:
                hi = mid
    errs = sum(datagen.symbol_errors(p["est"].array[0], p["tx_data"], p["bps"]) for p in pay)
    st = srv.stats()

    # plain: the same frames one after another, one handle per link, synchronous
    used = sorted({f["link"] for f in base})
    dets = {l: ApsmNomaDetectorFactory.build(dict(datagen.serving_link_config(links[l]), otype="APSM"), links[l]["N"])
            for l in used}
    for f, p in list(zip(base, pay))[:min(64, n_frames)]:
        dets[f["link"]].detect(p["rx_data"].array)
    check(lib().apsm_device_synchronize())
    t0 = time.perf_counter()
    lat_plain = []
    for f, p in zip(base, pay):
        t1 = time.perf_counter()
        d = dets[f["link"]]
        d.detect(p["rx_data"].array)
        lat_plain.append(time.perf_counter() - t1)
        d.train(p["rx_pilots"].array, p["tx_pilots"].array[0])
    plain_s = time.perf_counter() - t0
    sat = runs[-1]
    plain_sps = symbols / plain_s
    return {
        "workload": "serving-2048-links",
        "links_admitted": admitted,
        "link_bytes": st["link_bytes"],
        "hbm_budget": budget,
        "frame_arena": arena,
        "frames": n_frames,
        "detected_symbols": int(symbols),
        "rates": runs,
        "max_rate_p99_under_1ms": lo["offered_frames_per_s"] if lo else None,
        "max_rate_p99_under_1ms_run": lo,
        "max_rate_p99_under_1ms_vs_plain": (lo["symbols_per_s"] / plain_sps) if lo else None,
        "bisection": bisect,
        "h2d_pinned_GBps": h2d_pinned_gbps(),
        "saturation_symbols_per_s": sat["symbols_per_s"],
        "plain_symbols_per_s": symbols / plain_s,
        "plain_detect_p50_ms": float(np.percentile(np.array(lat_plain) * 1e3, 50)),
        "speedup_vs_plain": sat["symbols_per_s"] / (symbols / plain_s),
        "ser": errs / symbols,
        "rounds": st["rounds"],
        "max_frames_per_round": st["max_frames_per_round"],
        "links_s": t_links,
        "datagen_s": t_gen,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.only is not None:
        if args.only == "batched":
            res = bench_batched(args.batched_instances, max(1, args.steps))
        elif args.only == "mixed":
            res = bench_mixed(max(1, args.steps), args.mixed_samples)
        elif args.only == "copy":
            res = bench_copy()
        else:
            res = bench_serving(args.serving_frames, args.serving_links)
        print(json.dumps({args.only: res}))
        return

    # the serving section runs first: its server expects its four streams to get hardware queues
    # of their own (a process has four), which holds while no other handle streams are alive
    serving = None
    if not args.no_serving:
        serving = bench_serving(args.serving_frames, args.serving_links)
        gc.collect()

    from libapsm_amd import ApsmNomaDetectorFactory, LlsNomaDetector, datagen, lib
    from libapsm_amd._lib import DeviceArray, check

    w = datagen.WORKLOADS[args.workload]
    N, U, T, M = w["N"], len(w["users"]), w["T"], w["M"]
    data = datagen.generate(N, w["U_tx"], T, M, modulation=w["modulation"], snr_db=w["snr_db"], seed=w["seed"],
                            large_scale_db=w["large_scale_db"])
    tx_train = np.ascontiguousarray(data["tx_train"][w["users"]])
    def to_dev(z):
        return DeviceArray(np.ascontiguousarray(z, dtype=np.complex64).view(np.float32))

    d_rxt, d_txt, d_rxd = to_dev(data["rx_train"]), to_dev(tx_train), to_dev(data["rx_data"])
    d_est = DeviceArray.empty((U, M, 2), np.float32)
    cfg = datagen.workload_config(w)
    det = ApsmNomaDetectorFactory.build(cfg, N, U)

    def step():
        det.resetState()
        det.train_device(d_rxt.ptr, d_txt.ptr, T)
        det.detect_device(d_rxd.ptr, M, d_est.ptr)

    for _ in range(args.warmup):
        step()
    det.synchronize()
    check(lib().apsm_device_synchronize())
    # the timed steps run as the product runs them, without per-launch profiling events (three
    # event records per train block cost about 1.4 % of a step)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    det.synchronize()
    check(lib().apsm_device_synchronize())
    elapsed = time.perf_counter() - t0
    ms_step = 1000.0 * elapsed / args.steps
    value = M * U / (ms_step / 1000.0)

    # then the same number of steps again, untimed, for the per-kernel device times: HIP events on
    # the handle's stream around every sequencer, block-precompute and detect launch
    t_train_k, t_det_k = [], []
    det.profile_enable(True)
    det.profile_read()
    prof = {k: [0.0, 0] for k in det.PROFILE_KINDS}
    for _ in range(args.steps):
        step()
        a, b = det.kernel_times()
        t_train_k.append(a)
        t_det_k.append(b)
        for k, (ms_k, n_k) in det.profile_read().items():
            prof[k][0] += ms_k
            prof[k][1] += n_k
    det.synchronize()
    det.profile_enable(False)

    est = d_est.numpy()
    est_c = est[..., 0] + 1j * est[..., 1]
    tx_data = data["tx_data"][w["users"]]
    ser = datagen.symbol_errors(est_c, tx_data, data["bps"]) / est_c.size

    tr_ms, dt_ms = float(np.mean(t_train_k)), float(np.mean(t_det_k))
    W2, H = 2 * w["window"], 2 * w["dict_size"]
    G = min(T, w["dict_size"])
    # per launch: measured average duration (events) and algorithmic work of one launch
    seq_ms, seq_n = prof["sequencer"]
    pre_ms, pre_n = prof["block_pre"]
    det_ms_p, det_n = prof["detect"]
    seq_us = 1000.0 * seq_ms / max(1, seq_n)
    seq_flops = U * seq_flops_per_user(W2, 2 * T) / max(1.0, seq_n / args.steps)
    seq_tf = seq_flops / (seq_us * 1e-6) / 1e12
    cus = 256
    roofline = {"bound": f"fp32-valu latency chain, {U}/{cus} CUs (one workgroup per user, one projection step after another)",
                "kernel": "apsm_train_seq2", "achieved": seq_tf, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": seq_tf / PEAK_FP32_TFLOPS,
                "frac_of_occupied_cus": seq_tf / (PEAK_FP32_TFLOPS * U / cus),
                "avg_launch_us": seq_us, "launches_per_step": seq_n / args.steps, "flops_per_launch": seq_flops,
                "projection_steps_per_launch": T / max(1, seq_n / args.steps),
                "us_per_projection_step": seq_us / (T / max(1, seq_n / args.steps)),
                "traffic": SEQ_HBM_BYTES_PER_LAUNCH.get(args.workload),
                "train_call_traffic": TRAIN_HBM_BYTES.get(args.workload),
                "share_of_step": seq_ms / args.steps / ms_step}
    # per-step floor of the chain: the window-Gram product's vector-ALU issue on the busiest SIMD
    # (4 TP FMA per thread, wave64 at 16 lanes per cycle, packed FP32 off) at 2.4 GHz
    Wc = w["window"]
    parts = 1024 // Wc
    tp = -(-Wc // parts)
    waves_per_simd = -(-(-(-parts * Wc // 64)) // 4)
    floor_us = waves_per_simd * 4 * tp * 4 / 2.4e9 * 1e6
    roofline["per_step_floor_us"] = floor_us
    roofline["frac_of_per_step_floor"] = floor_us / roofline["us_per_projection_step"]
    det_flops = M * G * (8.0 * N + 8.0 * U) + M * U * 8.0 * N
    det_us = 1000.0 * det_ms_p / max(1, det_n)
    det_tf = det_flops / (det_us * 1e-6) / 1e12
    entry_flops = (T - G) * G * (8.0 * N + 8.0 * U) + G * G / 2 * (8.0 * N + 8.0 * U)  # entering share over the train
    pre_us = 1000.0 * pre_ms / max(1, pre_n)

    out = {
        "metric": "detected symbols/s (plain APSM train+detect, all users)",
        "value": value,
        "unit": "detected symbols/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded Rayleigh NOMA uplink, see libapsm_amd/datagen.py)",
        "config": {"workload": args.workload, "antennas": N, "users_detected": U, "pilots": T, "detect_samples": M,
                   "modulation": w["modulation"], "snr_db": w["snr_db"], "windowSize": w["window"],
                   "dictionarySize": w["dict_size"], "parallelism": "single-gpu"},
        "roofline": roofline,
        "kernels_ms": {"train": tr_ms, "detect": dt_ms, "sequencer": seq_ms / args.steps,
                       "block_pre": pre_ms / args.steps},
        "secondary_rooflines": [
            {"kernel": "apsm_detect_mfma<16,12,fp32>", "bound": "fp32 mfma cross terms + vector-ALU exp and contraction",
             "achieved": det_tf, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": det_tf / PEAK_FP32_TFLOPS,
             "avg_launch_us": det_us, "flops_per_launch": det_flops, "exp_per_s": 3.0 * M * G / (det_us * 1e-6)},
            {"kernel": "apsm_block_pre + apsm_entry_combine", "avg_launch_us": pre_us,
             "launches_per_step": pre_n / args.steps,
             "achieved_entry_share": entry_flops / (pre_ms / args.steps * 1e-3) / 1e12, "unit": "TFLOP/s"},
        ],
        "ser": ser,
    }
    fx_path = os.path.join(ROOT, "tests", "golden", "scale_c2.npz")
    if args.workload == "plain-16ant-16qam" and os.path.exists(fx_path):
        # the FP64 oracle's symbol errors on the same data (tests/golden/make_scale_fixtures.py)
        e_orc = int(np.load(fx_path, allow_pickle=False)["symbol_errors"].sum())
        e_gpu = int(round(ser * est_c.size))
        out["ser_oracle"] = e_orc / est_c.size
        out["ser_rel_diff"] = abs(e_gpu - e_orc) / max(1, e_orc)
    if not args.no_lls:
        lls = LlsNomaDetector(N)
        d_txu = to_dev(data["tx_train"][w["users"][0]])
        d_el = DeviceArray.empty((M, 2), np.float32)
        t0 = time.perf_counter()
        lls.train_device(d_rxt.ptr, d_txu.ptr, T)
        lls.detect_device(d_rxd.ptr, M, d_el.ptr)
        lls.synchronize()
        t_lls = time.perf_counter() - t0
        el = d_el.numpy()
        u0 = w["users"][0]
        out["lls"] = {"user": u0, "ms": 1000 * t_lls,
                      "ser": datagen.symbol_errors(el[:, 0] + 1j * el[:, 1], data["tx_data"][u0], data["bps"]) / M,
                      "apsm_ser_same_user": datagen.symbol_errors(est_c[0], tx_data[0], data["bps"]) / M}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, data, args.cpu_seconds)
        out["cpu_baseline"]["one_core"] = cpu_baseline(w, data, args.cpu_seconds / 2, threads=1)
    if not args.no_batched or not args.no_mixed or not args.no_small:
        del d_rxt, d_txt, d_rxd, d_est, det
    if not args.no_small:
        out["small_plain"] = bench_small_plain(max(1, args.steps), args.cpu_seconds)
    if not args.no_batched:
        out["batched"] = bench_batched(args.batched_instances, max(1, args.steps))
    if not args.no_mixed:
        out["mixed_precision"] = bench_mixed(max(1, args.steps), args.mixed_samples)
    if serving is not None:
        out["serving"] = serving
    print(json.dumps(out))


if __name__ == "__main__":
    main()
