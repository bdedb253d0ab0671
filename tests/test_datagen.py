"""Synthetic data generator and the reference file formats (host only)."""
import numpy as np
import pytest

from helpers import GOLDEN


def test_generator_statistics_and_layout():
    from libapsm_amd import datagen
    d = datagen.generate(8, 6, 4000, 1000, modulation="QAM16", snr_db=20.0, seed=9)
    assert d["rx_train"].shape == (8, 4000) and d["rx_data"].shape == (8, 1000)
    assert d["tx_train"].shape == (6, 4000) and d["tx_data"].shape == (6, 1000)
    # per-antenna power normalisation over the pilots
    assert np.allclose(np.mean(np.abs(d["rx_train"]) ** 2, axis=1), 1.0, rtol=1e-4)
    # symbols on the unit-energy constellation, uniformly used
    pts = datagen.constellation(4)
    assert np.abs(np.abs(d["tx_train"][..., None] - pts).min(-1)).max() < 1e-6
    assert np.mean(np.abs(d["tx_train"]) ** 2) == pytest.approx(1.0, rel=0.05)
    # same seed, same data; another seed, other data
    d2 = datagen.generate(8, 6, 4000, 1000, modulation="QAM16", snr_db=20.0, seed=9)
    assert np.array_equal(d["rx_data"], d2["rx_data"])
    d3 = datagen.generate(8, 6, 4000, 1000, modulation="QAM16", snr_db=20.0, seed=10)
    assert not np.array_equal(d["rx_data"], d3["rx_data"])


def test_reference_files_round_trip(tmp_path):
    from libapsm_amd import binfile, datagen
    d = datagen.generate(4, 3, 50, 70, modulation="QAM16", seed=2)
    rx, tx = datagen.write_reference_files(d, str(tmp_path), "t")
    s = datagen.mod_scale(4)
    assert np.array_equal(binfile.load_complex(rx, "rxSigTraining"), d["rx_train"])
    assert np.array_equal(binfile.load_complex(rx, "rxSigData"), d["rx_data"])
    raw = binfile.read(tx)["mod_training"]
    assert set(np.unique(raw)) <= {-3.0, -1.0, 1.0, 3.0}  # unscaled levels, as the bundled files hold
    assert np.abs(binfile.load_complex(tx, "mod_training", s) - d["tx_train"]).max() < 1e-6
    assert np.abs(binfile.load_complex(tx, "mod_data", s) - d["tx_data"]).max() < 1e-6


def test_bundled_files_have_the_reference_layout():
    import os

    from libapsm_amd import binfile
    r = binfile.read(os.path.join(GOLDEN, "offline", "rxData_QPSK_time.bin"))
    assert r["rxSigTraining"].shape == (685, 16, 2) and r["rxSigData"].shape == (3840, 16, 2)
    t = binfile.read(os.path.join(GOLDEN, "offline", "NOMA_signals_qpsk_complex.bin"))
    assert t["mod_training"].shape[1] == 8 and set(np.unique(t["mod_training"])) <= {-1.0, 0.0, 1.0}


def test_batched_and_serving_draws():
    from libapsm_amd import datagen
    inst = datagen.batched_instances(512)
    assert len(inst) == 512
    assert {i["N"] for i in inst} == {4, 8, 16, 32}
    assert {i["window"] for i in inst} == {40, 80, 160}
    assert all(200 <= i["dict_size"] <= 4000 and 10000 <= i["M"] <= 100000 for i in inst)
    fr = datagen.serving_frames(200, links=64)
    assert all(f["samples"] % 168 == 0 and 168 <= f["samples"] <= 168 * 273 for f in fr)
    assert all(8 <= f["N"] <= 32 for f in fr)
    arr = np.array([f["arrival_s"] for f in fr])
    assert np.all(np.diff(arr) > 0)
