"""The committed full-size FP64-oracle fixtures (tests/golden/make_scale_fixtures.py) belong to
the data the generator produces today: the GPU parity tests regenerate the inputs and compare
against these outputs, so a generator change must fail here first, on the CPU."""
import hashlib
import os

import numpy as np

from helpers import GOLDEN


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_small_plain_workload_fixture():
    from libapsm_amd import datagen
    fx = load("scale_c1.npz")
    w = datagen.WORKLOADS["plain-cpu-8ant"]
    d = datagen.generate(w["N"], w["U_tx"], w["T"], w["M"], modulation=w["modulation"], snr_db=w["snr_db"], seed=w["seed"],
                         large_scale_db=w["large_scale_db"])
    assert digest(d["rx_train"], d["tx_train"], d["rx_data"]) == str(fx["digest"])
    assert fx["soft"].shape == (len(w["users"]), 16384, 2) and int(fx["used"]) == 2 * w["T"]


def test_plain_workload_fixture():
    from libapsm_amd import datagen
    fx = load("scale_c2.npz")
    w = datagen.WORKLOADS["plain-16ant-16qam"]
    d = datagen.generate(w["N"], w["U_tx"], w["T"], w["M"], modulation=w["modulation"], snr_db=w["snr_db"], seed=w["seed"],
                         large_scale_db=w["large_scale_db"])
    assert digest(d["rx_train"], d["tx_train"], d["rx_data"]) == str(fx["digest"])
    assert fx["symbol_errors"].shape == (len(w["users"]),) and int(fx["samples"]) == w["M"]
    assert fx["soft"].shape == (len(fx["soft_users"]), 16384, 2) and fx["theta"].shape == (len(w["users"]), 2 * w["N"])
    assert int(fx["used"]) == 2 * w["T"]
    # the stored soft outputs are the oracle's for the stored users: their symbol errors on the
    # slice cannot exceed the users' errors over all samples
    for k, u in enumerate(fx["soft_users"]):
        est = fx["soft"][k, :, 0] + 1j * fx["soft"][k, :, 1]
        e = datagen.symbol_errors(est, d["tx_data"][w["users"][u], :16384], d["bps"])
        assert e <= int(fx["symbol_errors"][u])


def test_batched_instance_fixture():
    from libapsm_amd import datagen
    fx = load("scale_c3.npz")
    x = datagen.batched_instances(512)[int(fx["instance"])]
    d = datagen.batched_payload(x)
    assert digest(d["rx_train"], d["tx_train"], d["rx_data"]) == str(fx["digest"])
    assert int(fx["samples"]) == x["M"] and fx["theta"].shape == (2 * x["N"],)


def test_serving_link_fixture():
    from libapsm_amd import datagen
    fx = load("scale_c5.npz")
    link = int(fx["link"])
    frames = datagen.serving_frames(440, seed=5, links=4, rate_hz=4000.0)
    own = [f for f in frames if f["link"] == link]
    assert len(own) == int(fx["link_frames"]) >= 100
    assert all(int(k) >= 100 for k in fx["frames"])
    for j, k in enumerate(fx["frames"]):
        p = datagen.serving_payload(own[int(k)])
        assert digest(p["rx_data"], p["rx_pilots"], p["tx_pilots"]) == str(fx["digests"][j])
        assert int(fx["samples"][j]) == own[int(k)]["samples"]
