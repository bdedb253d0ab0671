"""Writes the FP64-oracle fixtures of the full-size parity tests (tests/test_gpu_scale.py).

Each fixture holds what the streamed FP64 oracle (oracle/apsm_oracle.cpp, the multi-user
form with the window Gram kept between steps) computes on a seeded synthetic workload that
the GPU test regenerates with libapsm_amd.datagen (SHA-256 digests of the generated arrays
are stored, so a generator change fails the test instead of comparing different data):

  scale_c1.npz  small plain workload (the CPU-runnable one): 8 antennas, 6 superimposed QPSK
                users at 10 dB, user 1 detected, windowSize 80, 20k pilots (dictionary
                20000: no eviction), 80k data samples. Symbol errors over all samples, soft
                outputs on the first 16384, theta.
  scale_c2.npz  plain workload: 16 antennas, 10 superimposed 16QAM users with large-scale
                gains in [-6, 0] dB at 15 dB SNR, all 10 users detected by one handle,
                windowSize 160, dictionary 4096 (FIFO over 200k pilots: 200k projection
                steps), 2M data samples. Symbol errors of every user over all 2M samples,
                soft outputs of users 0 and 9 on the first 16384 data samples, theta.
  scale_c3.npz  one draw of the 512-instance batched workload (instance 302: 32 antennas,
                QPSK, dictionary 3690, windowSize 160, 95693 data samples). Symbol errors
                over all samples, soft outputs on the first 16384, theta.
  scale_c5.npz  one serving link (link 1 of a 4-link trace of 440 frames at 4000 frames/s:
                8 antennas, 16QAM, windowSize 40, dictionary 1024) after 100+ frames, each
                frame detected with the link's state and then trained on its 5 % pilots.
                Soft outputs (first 8192 samples) and symbol errors of its frames 100, 105
                and its last frame.

Run from the repository root (about 6 minutes on 8 cores):
    python tests/golden/make_scale_fixtures.py [c1] [c2] [c3] [c5]
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from libapsm_amd import datagen  # noqa: E402
from oracle import pyoracle  # noqa: E402

SOFT_SLICE = 16384
C3_INSTANCE = 302
C5_LINKS, C5_FRAMES, C5_RATE, C5_LINK = 4, 440, 4000.0, 1
C5_COMPARE = (100, 105, -1)  # frame indices within the link's own frames
C5_SLICE = 8192


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def to_c(est):
    """float64 [..][2] -> complex128"""
    return est[..., 0] + 1j * est[..., 1]


def c1_data():
    w = datagen.WORKLOADS["plain-cpu-8ant"]
    d = datagen.generate(w["N"], w["U_tx"], w["T"], w["M"], modulation=w["modulation"], snr_db=w["snr_db"], seed=w["seed"],
                         large_scale_db=w["large_scale_db"])
    return w, d


def make_c1():
    t0 = time.time()
    w, d = c1_data()
    cfg = {k: v for k, v in datagen.workload_config(w).items() if k != "otype"}
    users = w["users"]
    o = pyoracle.ApsmMulti(cfg, w["N"], len(users))
    o.train(d["rx_train"], d["tx_train"][users])
    est = o.detect(d["rx_data"])
    errs = np.array([datagen.symbol_errors(to_c(est[k]), d["tx_data"][u], d["bps"]) for k, u in enumerate(users)], np.int64)
    st = o.state()
    np.savez(os.path.join(HERE, "scale_c1.npz"), digest=np.array(digest(d["rx_train"], d["tx_train"], d["rx_data"])),
             symbol_errors=errs, samples=np.int64(w["M"]), soft=est[:, :SOFT_SLICE].astype(np.float32), theta=st["theta"],
             used=np.int64(st["used"]))
    print(f"c1 done {time.time() - t0:.0f} s: symbol errors {errs.tolist()} of {w['M']}", flush=True)


def c2_data():
    w = datagen.WORKLOADS["plain-16ant-16qam"]
    d = datagen.generate(w["N"], w["U_tx"], w["T"], w["M"], modulation=w["modulation"], snr_db=w["snr_db"], seed=w["seed"],
                         large_scale_db=w["large_scale_db"])
    return w, d


def c3_data():
    x = datagen.batched_instances(512)[C3_INSTANCE]
    return x, datagen.batched_payload(x)


def c5_trace():
    links = datagen.serving_links(C5_LINKS, 5)
    frames = datagen.serving_frames(C5_FRAMES, seed=5, links=C5_LINKS, rate_hz=C5_RATE)
    return links, frames


def make_c2():
    t0 = time.time()
    w, d = c2_data()
    cfg = {k: v for k, v in datagen.workload_config(w).items() if k != "otype"}
    U = len(w["users"])
    o = pyoracle.ApsmMulti(cfg, w["N"], U)
    o.train(d["rx_train"], d["tx_train"][w["users"]])
    print(f"c2 train {time.time() - t0:.0f} s", flush=True)
    est = o.detect(d["rx_data"])  # [U][M][2]
    errs = np.array([datagen.symbol_errors(to_c(est[u]), d["tx_data"][w["users"][u]], d["bps"]) for u in range(U)], np.int64)
    st = o.state()
    soft_users = np.array([0, 9], np.int64)
    np.savez(os.path.join(HERE, "scale_c2.npz"), digest=np.array(digest(d["rx_train"], d["tx_train"], d["rx_data"])),
             symbol_errors=errs, samples=np.int64(w["M"]), soft_users=soft_users,
             soft=est[soft_users, :SOFT_SLICE].astype(np.float32), theta=st["theta"], used=np.int64(st["used"]))
    print(f"c2 done {time.time() - t0:.0f} s: symbol errors {errs.tolist()} of {w['M']} per user", flush=True)


def make_c3():
    t0 = time.time()
    x, d = c3_data()
    cfg = datagen.batched_config(x)
    o = pyoracle.ApsmMulti(cfg, x["N"], 1)
    o.train(d["rx_train"], d["tx_train"][x["user"]][None, :])
    est = o.detect(d["rx_data"])[0]
    errs = datagen.symbol_errors(to_c(est), d["tx_data"][x["user"]], d["bps"])
    st = o.state()
    np.savez(os.path.join(HERE, "scale_c3.npz"), digest=np.array(digest(d["rx_train"], d["tx_train"], d["rx_data"])),
             instance=np.int64(C3_INSTANCE), symbol_errors=np.int64(errs), samples=np.int64(x["M"]),
             soft=est[:SOFT_SLICE].astype(np.float32), theta=st["theta"][0])
    print(f"c3 done {time.time() - t0:.0f} s: {errs} symbol errors of {x['M']}", flush=True)


def make_c5():
    t0 = time.time()
    links, frames = c5_trace()
    lp = links[C5_LINK]
    own = [f for f in frames if f["link"] == C5_LINK]
    compare = sorted({k % len(own) for k in C5_COMPARE})
    o = pyoracle.ApsmMulti(datagen.serving_link_config(lp), lp["N"], 1)
    soft, errs, samples, digests = [], [], [], []
    for k, f in enumerate(own):
        p = datagen.serving_payload(f)
        if k in compare:
            est = o.detect(p["rx_data"])[0]
            soft.append(est[:C5_SLICE].astype(np.float32))
            errs.append(datagen.symbol_errors(to_c(est), p["tx_data"], p["bps"]))
            samples.append(f["samples"])
            digests.append(digest(p["rx_data"], p["rx_pilots"], p["tx_pilots"]))
        o.train(p["rx_pilots"], p["tx_pilots"])
    np.savez(os.path.join(HERE, "scale_c5.npz"), link=np.int64(C5_LINK), frames=np.array(compare, np.int64),
             digests=np.array(digests), soft=np.stack(soft), symbol_errors=np.array(errs, np.int64),
             samples=np.array(samples, np.int64), link_frames=np.int64(len(own)))
    print(f"c5 done {time.time() - t0:.0f} s: link {C5_LINK} ({len(own)} frames), frames {compare}: "
          f"symbol errors {errs} of {samples}", flush=True)


if __name__ == "__main__":
    which = sys.argv[1:] or ["c1", "c2", "c3", "c5"]
    print(f"oracle threads: {pyoracle.threads()}", flush=True)
    for name in which:
        {"c1": make_c1, "c2": make_c2, "c3": make_c3, "c5": make_c5}[name]()
