"""Seeded synthetic uplink NOMA data (the step before the detector).

The reference ships only a pilot-symbol generator (cpp/cli/common/util.cpp:287-325) and
no channel model, so the channel here is this project's: per-frame Rayleigh gains
H[n,u] ~ CN(0,1), optional per-user large-scale gains uniform in dB, AWGN at a given SNR,
x = H diag(g) s + noise, then per-antenna power normalisation over the pilots.

Symbols use the reference's Gray mapping and unit-energy scales
(cpp/cli/common/modulation.cpp:112-151, modulation.hpp:157-222).
"""
import numpy as np

BITS_PER_SYMBOL = {"BPSK": 1, "QPSK": 2, "QAM16": 4, "QAM64": 6, "QAM256": 8}


def mod_scale(bps):
    """Scale that gives unit average energy (modulation.hpp:157-180)."""
    if bps in (1, 2):
        return 1.0 / np.sqrt(bps)
    return 1.0 / np.sqrt(((2 << bps) - 2) // 3)


def _inverse_gray(b):
    g = b.copy()
    shift = b >> 1
    while np.any(shift):
        g ^= shift
        shift >>= 1
    return g


def constellation(bps):
    """Complex point for every bit pattern 0 .. 2^bps - 1 (Gray mapped, unit energy)."""
    bits = np.arange(1 << bps, dtype=np.int64)
    s = mod_scale(bps)
    if bps == 1:
        return (2.0 * _inverse_gray(bits) - 1.0) * s + 0j
    h = bps // 2
    ns = 1 << h
    re = 2.0 * _inverse_gray(bits >> h) - ns + 1
    im = 2.0 * _inverse_gray(bits & (ns - 1)) - ns + 1
    return (re + 1j * im) * s


def hard_decision(est, bps):
    """Nearest constellation point per real dimension (the decision regions of the SER rule)."""
    s = mod_scale(bps)
    est = np.asarray(est)
    if bps == 1:
        return np.where(est.real >= 0, s, -s) + 0j
    ns = 1 << (bps // 2)
    lim = ns - 1

    def q(v):
        k = np.round((v / s + lim) / 2.0)
        k = np.clip(k, 0, lim)
        return (2.0 * k - lim) * s

    return q(est.real) + 1j * q(est.imag)


def symbol_errors(est, tx, bps):
    """Number of symbols whose hard decision differs from the transmitted symbol (decision
    regions of each real dimension; cpp/cli/common/util.cpp:62-110 with the real-part rule
    applied to both dimensions)."""
    s = mod_scale(bps)
    d = np.abs(hard_decision(est, bps) - np.asarray(tx).astype(np.complex128))
    return int(np.count_nonzero(d > 0.5 * s))


def generate(num_antennas, num_users, num_pilots, num_data, modulation="QPSK", snr_db=10.0,
             seed=1, large_scale_db=None, normalize=True):
    """Returns dict(rx_train [N][T], rx_data [N][M], tx_train [U][T], tx_data [U][M]) complex64.

    Pilots come first, then data, from one random stream.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    bps = BITS_PER_SYMBOL[modulation.upper()]
    S = num_pilots + num_data
    pts = constellation(bps)
    sym = pts[rng.integers(0, 1 << bps, size=(num_users, S))]
    Hc = (rng.standard_normal((num_antennas, num_users)) + 1j * rng.standard_normal((num_antennas, num_users))) / np.sqrt(2)
    if large_scale_db is not None:
        lo, hi = large_scale_db
        g = 10.0 ** (rng.uniform(lo, hi, size=num_users) / 20.0)
        Hc = Hc * g[None, :]
    nvar = 10.0 ** (-snr_db / 10.0)
    noise = (rng.standard_normal((num_antennas, S)) + 1j * rng.standard_normal((num_antennas, S))) * np.sqrt(nvar / 2)
    x = Hc @ sym + noise
    if normalize:
        p = np.sqrt(np.mean(np.abs(x[:, :num_pilots]) ** 2, axis=1, keepdims=True))
        x = x / p
    x = x.astype(np.complex64)
    sym = sym.astype(np.complex64)
    return dict(rx_train=np.ascontiguousarray(x[:, :num_pilots]), rx_data=np.ascontiguousarray(x[:, num_pilots:]),
                tx_train=np.ascontiguousarray(sym[:, :num_pilots]), tx_data=np.ascontiguousarray(sym[:, num_pilots:]),
                bps=bps)


def kernel_variance_for_sigma(sigma, num_antennas):
    """Configuration value v such that the kernel is exp(-|u-d|^2 / (2 sigma^2)): the library
    scales v by N^2 (cpp/lib/apsm/apsm_noma_detector.cu:79-80)."""
    return sigma * sigma / float(num_antennas * num_antennas)


def median_pair_variance(rx, pairs=20000, seed=0):
    """sigma^2 = half the median squared distance |x_i - x_j|^2 over random pairs of the
    complex samples rx [N][T] (so that exp(-d^2 / 2 sigma^2) = 1/e at the median distance).
    The reference's own heuristic (cpp/lib/apsm/apsm_noma_detector.cu:194-218: the 70th
    percentile of sum_n |x_i,n - x_j,n| times T^(-1/5), squared, not scaled with N) gives a
    kernel far narrower than the pilot spacing at many antennas, where the Gaussian term then
    vanishes from f."""
    rx = np.asarray(rx)
    T = rx.shape[1]
    idx = np.random.Generator(np.random.PCG64(seed)).integers(0, T, size=(pairs, 2))
    idx = idx[idx[:, 0] != idx[:, 1]]
    d2 = np.sum(np.abs(rx[:, idx[:, 0]].astype(np.complex128) - rx[:, idx[:, 1]]) ** 2, axis=0)
    return float(np.median(d2) / 2.0)


# Workloads (sizes from the benchmark configurations).
WORKLOADS = {
    # 64 antennas, 12 QPSK users at 12 dB; dictionary of 16384 centres; 8M detect samples
    "mixed-64ant-qpsk": dict(N=64, U_tx=12, users=list(range(12)), T=16384, M=8000000, modulation="QPSK",
                             snr_db=12.0, seed=4, large_scale_db=None, window=80, sigma=0.5, eB=0.01,
                             dict_size=16384),
    # 6 NOMA users, QPSK, 8 antennas, 10 dB; detect user 1; w=80, sigma=0.5, eps=0.01
    "plain-cpu-8ant": dict(N=8, U_tx=6, users=[1], T=20000, M=80000, modulation="QPSK", snr_db=10.0, seed=1,
                           large_scale_db=None, window=80, sigma=0.5, eB=0.01, dict_size=20000),
    # 10 users, 16QAM, 16 antennas, 15 dB, gains U[-6,0] dB; all 10 users; w=160, Dc=4096
    "plain-16ant-16qam": dict(N=16, U_tx=10, users=list(range(10)), T=200000, M=2000000, modulation="QAM16",
                              snr_db=15.0, seed=2, large_scale_db=(-6.0, 0.0), window=160, sigma=0.5, eB=0.01,
                              dict_size=4096),
}


def workload_config(w):
    """Detector JSON configuration (reference keys) of a workload."""
    return {"otype": "APSM", "windowSize": w["window"], "eB": w["eB"],
            "gaussianKernelVariance": kernel_variance_for_sigma(w["sigma"], w["N"]),
            "dictionarySize": w["dict_size"]}


def write_reference_files(data, out_dir, name):
    """Write a generated data set in the reference tool's formats: the synchronized received
    signals (keys rxSigTraining, rxSigData; cpp/cli/common/argparse_helper.cpp:160-178) and the
    reference symbols as unscaled integer constellation levels (keys mod_training, mod_data;
    scaled by the modulation scale when loaded, argparse_helper.cpp:101-102). Returns the two
    paths (rx file, tx file)."""
    import os

    from . import binfile
    os.makedirs(out_dir, exist_ok=True)
    s = mod_scale(data["bps"])
    rx_path = os.path.join(out_dir, f"rxData_{name}.bin")
    tx_path = os.path.join(out_dir, f"NOMA_signals_{name}.bin")
    binfile.write(rx_path, {"rxSigTraining": binfile.from_complex(data["rx_train"]),
                            "rxSigData": binfile.from_complex(data["rx_data"])})
    binfile.write(tx_path, {"mod_training": binfile.from_complex(np.round(data["tx_train"] / s)),
                            "mod_data": binfile.from_complex(np.round(data["tx_data"] / s))})
    return rx_path, tx_path


def batched_instances(count=512, seed=3):
    """Parameter draws of the batched workload: per instance antennas in {4, 8, 16, 32},
    QPSK or 16QAM, dictionary size in [200, 4000], windowSize in {40, 80, 160}, 10k..100k
    detect samples; pilots = dictionary size; instance i is generated from seed 3e6 + i."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(count):
        dc = int(rng.integers(200, 4001))
        out.append(dict(N=int(rng.choice([4, 8, 16, 32])), modulation=str(rng.choice(["QPSK", "QAM16"])), dict_size=dc,
                        window=int(rng.choice([40, 80, 160])), M=int(rng.integers(10000, 100001)), T=dc,
                        seed=3000000 + i, U_tx=4, user=0, snr_db=15.0, sigma=0.5, eB=0.01))
    return out


def batched_payload(x):
    """Samples of one batched instance (a draw of batched_instances): pilots first, then data,
    from the instance's seed."""
    return generate(x["N"], x["U_tx"], x["T"], x["M"], modulation=x["modulation"], snr_db=x["snr_db"], seed=x["seed"])


def batched_config(x):
    """Detector configuration of one batched instance."""
    return {"windowSize": x["window"], "dictionarySize": x["dict_size"], "eB": x["eB"],
            "gaussianKernelVariance": kernel_variance_for_sigma(x["sigma"], x["N"])}


def serving_links(links=2048, seed=5):
    """Per-link parameters of the serving workload: 8..32 antennas, QPSK or 16QAM."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return [dict(N=int(rng.integers(8, 33)), modulation=str(rng.choice(["QPSK", "QAM16"]))) for _ in range(links)]


def serving_frames(count, seed=5, links=2048, rate_hz=1000.0, max_rb=273):
    """Requests of the serving workload: Poisson arrivals at rate_hz; each frame carries
    14 OFDM symbols x (12 x U[1, max_rb]) subcarriers of data plus 5 % as many pilots, on a
    link drawn uniformly from `links` (8..32 antennas, QPSK or 16QAM per link)."""
    link_par = serving_links(links, seed)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    t = 0.0
    frames = []
    for i in range(count):
        t += rng.exponential(1.0 / rate_hz)
        n = 14 * 12 * int(rng.integers(1, max_rb + 1))
        pilots = max(1, int(round(0.05 * n)))
        link = int(rng.integers(0, links))
        frames.append(dict(arrival_s=t, link=link, samples=n, pilots=pilots, seed=5000000 + i, **link_par[link]))
    return frames


def serving_payload(frame, U_tx=4, snr_db=15.0, user=0, link_seed_base=7000000):
    """Samples of one serving frame: the link's Rayleigh channel (fixed per link, from
    link_seed_base + link) carries U_tx superimposed users; symbols and noise come from the
    frame's seed. Each antenna is scaled by its expected power (channel gains + noise), the
    same scale for every frame of the link. Returns rx_pilots [N][P], tx_pilots [1][P] (the
    detected user), rx_data [N][S], tx_data [S]."""
    N, bps = frame["N"], BITS_PER_SYMBOL[frame["modulation"].upper()]
    lr = np.random.Generator(np.random.PCG64(link_seed_base + frame["link"]))
    H = (lr.standard_normal((N, U_tx)) + 1j * lr.standard_normal((N, U_tx))) / np.sqrt(2)
    nvar = 10.0 ** (-snr_db / 10.0)
    scale = 1.0 / np.sqrt(np.sum(np.abs(H) ** 2, axis=1) + nvar)
    rng = np.random.Generator(np.random.PCG64(frame["seed"]))
    P, S = frame["pilots"], frame["samples"]
    pts = constellation(bps)
    sym = pts[rng.integers(0, 1 << bps, size=(U_tx, P + S))]
    noise = (rng.standard_normal((N, P + S)) + 1j * rng.standard_normal((N, P + S))) * np.sqrt(nvar / 2)
    x = ((H @ sym + noise) * scale[:, None]).astype(np.complex64)
    sym = sym.astype(np.complex64)
    return dict(rx_pilots=np.ascontiguousarray(x[:, :P]), tx_pilots=np.ascontiguousarray(sym[user:user + 1, :P]),
                rx_data=np.ascontiguousarray(x[:, P:]), tx_data=np.ascontiguousarray(sym[user, P:]), bps=bps)


def serving_link_config(link_par, window=40, dict_size=1024, sigma=0.5, eB=0.01):
    """Detector configuration of a serving link."""
    return {"windowSize": window, "dictionarySize": dict_size, "eB": eB,
            "gaussianKernelVariance": kernel_variance_for_sigma(sigma, link_par["N"])}


def _main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="Write a synthetic NOMA data set in the reference tool's file formats.")
    ap.add_argument("workload", choices=sorted(WORKLOADS))
    ap.add_argument("out_dir")
    ap.add_argument("--pilots", type=int, help="override the number of pilot samples")
    ap.add_argument("--samples", type=int, help="override the number of detect samples")
    a = ap.parse_args(argv)
    w = dict(WORKLOADS[a.workload])
    T = a.pilots or w["T"]
    M = a.samples or w["M"]
    d = generate(w["N"], w["U_tx"], T, M, modulation=w["modulation"], snr_db=w["snr_db"], seed=w["seed"],
                 large_scale_db=w["large_scale_db"])
    rx, tx = write_reference_files(d, a.out_dir, a.workload)
    cfg = {"algorithm": workload_config(w), "modulation": w["modulation"], "user": w["users"][0]}
    import json
    print(json.dumps({"rx": rx, "tx": tx, "config": cfg}))


if __name__ == "__main__":
    _main()
