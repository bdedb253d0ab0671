"""Parity at full workload sizes: the HIP path against FP64-oracle fixtures
(tests/golden/make_scale_fixtures.py writes them with the streamed FP64 oracle).

  * small plain workload (8 antennas, 6 QPSK users, user 1 detected, windowSize 80, 20k
    pilots, 80k data samples);
  * plain workload (16 antennas, 10 superimposed 16QAM users, one 10-user handle,
    windowSize 160, dictionary 4096 as a FIFO over 200k pilots = 200k projection steps,
    2M data samples);
  * one draw of the 512-instance batched workload, run inside a batch;
  * one serving link after 100+ frames, served in a mix with other links.

Tolerances (FP32 device arithmetic after up to 2e5 dependent projection steps against FP64):
  * soft outputs on the fixture's slice: max |est_gpu - est_oracle| <= SOFT_ATOL;
  * hard decisions identical wherever the oracle's soft output is farther than DELTA from a
    decision boundary;
  * symbol errors over all samples: |errors_gpu - errors_oracle| <= 1 % of errors_oracle per
    user (or <= SER_ABS_FLOOR symbols for a user with fewer than 100 oracle errors), and
    <= 1 % over all users together.
"""
import hashlib
import os

import numpy as np
import pytest

from helpers import GOLDEN, to_ri
from test_gpu_parity import DELTA, boundary_margin

pytestmark = pytest.mark.gpu

SOFT_ATOL = 2e-3
SER_REL = 0.01
SER_ABS_FLOOR = 2


def fixture(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def check_soft(est_gpu, ref, bps, what):
    from libapsm_amd.datagen import hard_decision
    g = to_ri(est_gpu).astype(np.float64)
    ref = ref.astype(np.float64)
    err = np.abs(g - ref)
    print(f"{what}: max soft err {err.max():.3e}, rms {np.sqrt(np.mean(err ** 2)):.3e}", flush=True)
    assert err.max() <= SOFT_ATOL, what
    far = boundary_margin(ref, bps) > DELTA
    hg = hard_decision(np.asarray(est_gpu), bps)
    ho = hard_decision(ref[:, 0] + 1j * ref[:, 1], bps)
    assert np.array_equal(hg[far], ho[far]), what


def check_errors(gpu, orc, what):
    gpu, orc = np.atleast_1d(np.asarray(gpu, np.int64)), np.atleast_1d(np.asarray(orc, np.int64))
    print(f"{what}: symbol errors gpu {gpu.tolist()} oracle {orc.tolist()}", flush=True)
    for g, o in zip(gpu, orc):
        bound = SER_REL * o if o >= 100 else SER_ABS_FLOOR
        assert abs(int(g) - int(o)) <= bound, what
    assert abs(int(gpu.sum()) - int(orc.sum())) <= max(SER_REL * orc.sum(), SER_ABS_FLOOR), what


def test_small_plain_workload_matches_oracle(require_gpu):
    from libapsm_amd import ApsmNomaDetectorFactory, datagen
    fx = fixture("scale_c1.npz")
    w = datagen.WORKLOADS["plain-cpu-8ant"]
    d = datagen.generate(w["N"], w["U_tx"], w["T"], w["M"], modulation=w["modulation"], snr_db=w["snr_db"], seed=w["seed"],
                         large_scale_db=w["large_scale_db"])
    assert digest(d["rx_train"], d["tx_train"], d["rx_data"]) == str(fx["digest"]), "generator output changed"
    users = w["users"]
    det = ApsmNomaDetectorFactory.build(datagen.workload_config(w), w["N"], len(users))
    det.train(d["rx_train"], d["tx_train"][users])
    est = np.atleast_2d(det.detect(d["rx_data"]))
    st = det.getTrainingState()
    assert st["used_rows"] == int(fx["used"])
    dth = np.abs(st["theta"].astype(np.float64) - fx["theta"]).max()
    print(f"small plain workload: theta max abs diff {dth:.3e}", flush=True)
    assert dth <= 1e-3
    for k, u in enumerate(users):
        check_soft(est[k, :fx["soft"].shape[1]], fx["soft"][k], d["bps"], f"user {u}")
    errs = [datagen.symbol_errors(est[k], d["tx_data"][u], d["bps"]) for k, u in enumerate(users)]
    check_errors(errs, fx["symbol_errors"], "small plain workload")


def test_plain_workload_matches_oracle(require_gpu):
    from libapsm_amd import ApsmNomaDetectorFactory, datagen
    fx = fixture("scale_c2.npz")
    w = datagen.WORKLOADS["plain-16ant-16qam"]
    d = datagen.generate(w["N"], w["U_tx"], w["T"], w["M"], modulation=w["modulation"], snr_db=w["snr_db"], seed=w["seed"],
                         large_scale_db=w["large_scale_db"])
    assert digest(d["rx_train"], d["tx_train"], d["rx_data"]) == str(fx["digest"]), "generator output changed"
    users = w["users"]
    print(f"plain workload: {w['T']} pilots, {w['M']} samples, {len(users)} users", flush=True)
    det = ApsmNomaDetectorFactory.build(datagen.workload_config(w), w["N"], len(users))
    det.train(d["rx_train"], d["tx_train"][users])
    est = det.detect(d["rx_data"])
    st = det.getTrainingState()
    assert st["used_rows"] == int(fx["used"])
    dth = np.abs(st["theta"].astype(np.float64) - fx["theta"]).max()
    print(f"theta max abs diff {dth:.3e}", flush=True)
    assert dth <= 1e-3
    for k, u in enumerate(fx["soft_users"]):
        check_soft(est[u, :fx["soft"].shape[1]], fx["soft"][k], d["bps"], f"user {u}")
    errs = [datagen.symbol_errors(est[u], d["tx_data"][users[u]], d["bps"]) for u in range(len(users))]
    check_errors(errs, fx["symbol_errors"], "plain workload")


def test_batched_instance_matches_oracle(require_gpu):
    from libapsm_amd import ApsmBatch, datagen
    fx = fixture("scale_c3.npz")
    inst = datagen.batched_instances(512)
    i0 = int(fx["instance"])
    chosen = [i0] + list(range(i0 + 1, i0 + 8))  # the fixture's instance first, in a batch of 8 draws
    xs = [inst[i] for i in chosen]
    data = [datagen.batched_payload(x) for x in xs]
    assert digest(data[0]["rx_train"], data[0]["tx_train"], data[0]["rx_data"]) == str(fx["digest"])
    print(f"batched instance {i0}: {xs[0]}", flush=True)
    b = ApsmBatch([datagen.batched_config(x) for x in xs], [x["N"] for x in xs])
    b.plan([d["rx_train"] for d in data], [d["tx_train"][x["user"]] for d, x in zip(data, xs)], [d["rx_data"] for d in data])
    b.synchronize()
    est = b.estimates(0)[0]
    check_soft(est[:fx["soft"].shape[0]], fx["soft"], data[0]["bps"], f"instance {i0}")
    errs = datagen.symbol_errors(est, data[0]["tx_data"][xs[0]["user"]], data[0]["bps"])
    check_errors(errs, fx["symbol_errors"], f"instance {i0}")


def test_served_link_after_100_frames_matches_oracle(require_gpu):
    from libapsm_amd import ApsmServer, Frame, datagen
    fx = fixture("scale_c5.npz")
    link = int(fx["link"])
    links = datagen.serving_links(4, 5)
    frames = datagen.serving_frames(440, seed=5, links=4, rate_hz=4000.0)
    own_idx = [k for k, f in enumerate(frames) if f["link"] == link]
    assert len(own_idx) == int(fx["link_frames"])
    others = [k for k, f in enumerate(frames) if f["link"] != link][:40]  # a mix at the start of the trace
    keep = sorted(own_idx + others)
    srv = ApsmServer(2048 << 20, 512 << 20)
    for lp in links:
        srv.add_link(datagen.serving_link_config(lp), lp["N"], 1, 4096)
    served, pay = [], {}
    t0 = frames[keep[0]]["arrival_s"]
    for k in keep:
        p = datagen.serving_payload(frames[k])
        pay[k] = p
        served.append(Frame(frames[k]["link"], p["rx_data"], p["rx_pilots"], p["tx_pilots"],
                            arrival_s=frames[k]["arrival_s"] - t0))
    print(f"serving link {link}: {len(own_idx)} frames of the link, {len(others)} of other links", flush=True)
    res = srv.replay(served)
    assert all(r["status"] == 0 for r in res)
    by_idx = dict(zip(keep, served))
    for j, fi in enumerate(fx["frames"]):
        k = own_idx[int(fi)]
        p = pay[k]
        assert digest(p["rx_data"], p["rx_pilots"], p["tx_pilots"]) == str(fx["digests"][j])
        est = by_idx[k].est[0]
        check_soft(est[:fx["soft"].shape[1]], fx["soft"][j], p["bps"], f"link {link} frame {int(fi)}")
        errs = datagen.symbol_errors(est, p["tx_data"], p["bps"])
        check_errors(errs, fx["symbol_errors"][j], f"link {link} frame {int(fi)}")
