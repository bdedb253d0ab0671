"""NOMA_detect on the bundled reference data (cpp/data/offline), on the GPU.

The tool's scripting line (ber%, ser%, evm) is compared with the same detector run through the
Python API and scored with the oracle's metric restatement (SER in the reference tool's own
rule, cli_compat); the detected-data file with the API's estimates (same kernels: 1e-5), and
the training-history CSV with the step count of the train schedule.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from helpers import equidistant, load_offline, to_ri

pytestmark = pytest.mark.gpu

TOOL = os.path.join(ROOT, "libapsm_amd", "bin", "NOMA_detect")
OFF = os.path.join(ROOT, "tests", "golden", "offline")


def run_tool(*args):
    print("NOMA_detect", " ".join(args), flush=True)
    r = subprocess.run([TOOL, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def scripting_values(out):
    return [tuple(float(v) for v in line.split(",")) for line in out.strip().splitlines() if line.count(",") == 2]


@pytest.mark.parametrize("mod,domain,algo", [("QPSK", "time", "APSM"), ("QAM16", "ofdm", "APSM"), ("QPSK", "ofdm", "LLS")])
def test_tool_matches_api(require_gpu, oracle, tmp_path, mod, domain, algo):
    from libapsm_amd import NomaDetectorFactory, binfile
    cfg = {"otype": algo} if algo == "LLS" else {"otype": "APSM", "windowSize": 80}
    ants = equidistant(8)
    user = 1
    import json
    conf = json.dumps({"algorithm": cfg, "modulation": mod, "antennas": {"otype": "equidistant", "number": 8}, "user": user})
    det_file = str(tmp_path / "est.bin")
    out = run_tool("-s", os.path.join(OFF, f"rxData_{mod}_{domain}.bin"), "-r",
                   os.path.join(OFF, f"NOMA_signals_{mod.lower()}_complex.bin"), "-cs", conf, "-o", "scripting", "-d", det_file)
    (ber, ser, evm), = scripting_values(out)

    d = load_offline(mod, domain, antennas=ants)
    M = d["rx_data"].shape[1]
    det = NomaDetectorFactory.build(cfg, len(ants))
    det.train(d["rx_train"], d["tx_train"][user][:d["rx_train"].shape[1]])
    est = det.detect(d["rx_data"])
    ref = to_ri(d["tx_data"][user][:M])
    e = to_ri(est)
    bps = d["bps"]
    assert ber == pytest.approx(100.0 * oracle.bit_errors(e, ref, bps) / (M * bps), abs=1e-3)
    assert ser == pytest.approx(100.0 * oracle.symbol_errors(e, ref, bps, cli_compat=True) / M, abs=1e-3)
    assert evm == pytest.approx(oracle.evm(e, ref), rel=1e-4)
    got = binfile.load_complex(det_file, "estSigData")
    assert got.shape == (1, M)
    assert np.abs(got[0] - est).max() < 1e-5


def test_tool_parallel_users_and_history(require_gpu, tmp_path):
    conf = '{"algorithm": {"otype": "APSM", "windowSize": 40}, "modulation": "QPSK", "antennas": {"otype": "first", "number": 8}}'
    rx = os.path.join(OFF, "rxData_QPSK_time.bin")
    tx = os.path.join(OFF, "NOMA_signals_qpsk_complex.bin")
    out = run_tool("-s", rx, "-r", tx, "-cs", conf, "-o", "scripting", "-p")
    vals = scripting_values(out)
    assert len(vals) == 6  # users 0..5 of the bundled files
    single = scripting_values(run_tool("-s", rx, "-r", tx, "-cs", conf.replace("}}", "}, \"user\": 3}"), "-o", "scripting"))
    # user 3 detected alone and inside the 6-user handle: same detector, same numbers
    assert abs(vals[3][1] - single[0][1]) <= 100.0 * 2 / 3840
    assert abs(vals[3][2] - single[0][2]) < 1e-3
    hist = str(tmp_path / "hist.csv")
    run_tool("-s", rx, "-r", tx, "-cs", conf, "-th", hist)
    lines = open(hist).read().strip().splitlines()
    assert lines[0] == "step,train_evm,train_ser,train_ber,val_evm,val_ser,val_ber"
    assert len(lines) - 1 == 685  # one row per train step (sampleStep 2: one complex pilot per step)
    last = [float(v) for v in lines[-1].split(",")]
    assert last[1] < 1.0 and 0.0 <= last[2] <= 1.0


def test_tool_random_reference(require_gpu):
    out = run_tool("-s", os.path.join(OFF, "rxData_QPSK_time.bin"), "-rs", "1", "-cs",
                   '{"algorithm": {"otype": "LLS"}, "modulation": "QPSK"}', "-o", "scripting")
    (ber, ser, evm), = scripting_values(out)
    assert 0.0 <= ber <= 100.0 and 0.0 <= ser <= 100.0 and evm > 0.0


def test_tool_on_generated_files(require_gpu, oracle, tmp_path):
    """The generator's files feed the tool directly; the tool's SER equals the API's on the
    same data set."""
    import json

    from libapsm_amd import NomaDetectorFactory, datagen
    d = datagen.generate(8, 6, 1500, 3000, modulation="QPSK", snr_db=10.0, seed=1)
    rx, tx = datagen.write_reference_files(d, str(tmp_path), "small")
    cfg = {"otype": "APSM", "windowSize": 40, "eB": 0.01, "gaussianKernelVariance": datagen.kernel_variance_for_sigma(0.5, 8),
           "dictionarySize": 1500}
    out = run_tool("-s", rx, "-r", tx, "-cs", json.dumps({"algorithm": cfg, "modulation": "QPSK", "user": 1}), "-o", "scripting")
    (ber, ser, evm), = scripting_values(out)
    det = NomaDetectorFactory.build(cfg, 8)
    det.train(d["rx_train"], d["tx_train"][1])
    e = to_ri(det.detect(d["rx_data"]))
    ref = to_ri(d["tx_data"][1])
    assert ser == pytest.approx(100.0 * oracle.symbol_errors(e, ref, 2, cli_compat=True) / 3000, abs=1e-3)
    assert evm == pytest.approx(oracle.evm(e, ref), rel=1e-4)
    assert ser < 5.0
