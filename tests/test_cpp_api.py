"""C++ drop-in headers and NOMA_detect helpers, host side (no GPU).

The checks compile tests/cpp/api_check.cpp with g++ against include/ and compare what the
headers compute with the oracle, the committed fixtures, or what the mathematics fixes
(Mersenne-Twister words for the seeded symbol generator, unit energy of the constellations).
"""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

CHECK_SRC = os.path.join(ROOT, "tests", "cpp", "api_check.cpp")
LIBDIR = os.path.join(ROOT, "libapsm_amd")
TOOL = os.path.join(ROOT, "libapsm_amd", "bin", "NOMA_detect")


@pytest.fixture(scope="module")
def api_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cpp") / "api_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"), CHECK_SRC,
                    "-L" + LIBDIR, "-lapsm_amd", "-Wl,-rpath," + LIBDIR, "-o", exe], check=True)
    return lambda *a: json.loads(subprocess.run([exe, *map(str, a)], check=True, capture_output=True, text=True).stdout)


@pytest.fixture(scope="module")
def tool():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "libapsm_amd", "cli")], check=True)
    return TOOL


def test_factory_defaults_and_overrides(api_check):
    from libapsm_amd import ApsmNomaDetectorFactory
    for cfg in ({"otype": "APSM"},
                {"otype": "APSM", "gaussianKernelWeight": 0.25, "windowSize": 80, "sampleStep": 4, "trainPasses": 2,
                 "startWithFullWindow": True, "llsInitialization": True, "trainingShuffle": True, "dictionarySize": 4096,
                 "normConstraint": 1.5, "eB": 0.01, "gaussianKernelVariance": 0.25, "detectVersion": "shmem",
                 "detectPrecision": "bf16"}):
        got = api_check("factory", json.dumps(cfg))
        ref = ApsmNomaDetectorFactory.params_from_config(cfg)
        for k in ("linearKernelWeight", "gaussianKernelWeight", "gaussianKernelVariance", "normConstraint", "eB"):
            assert got[k] == pytest.approx(getattr(ref, k), rel=1e-6), k
        for k in ("windowSize", "sampleStep", "trainPasses", "dictionarySize"):
            assert got[k] == getattr(ref, k), k
        for k in ("startWithFullWindow", "llsInitialization", "trainingShuffle"):
            assert bool(got[k]) == bool(getattr(ref, k)), k
        assert got["version"] == ref.detectVersion
        assert got["detectPrecision"] == ref.detectPrecision
    assert api_check("factory", '{"otype": "APSM"}')["versionString"] == "APSM Detect Balance memory and computation"


def test_factory_errors(api_check):
    assert api_check("factory", '{"otype": "APSM", "window": 3}') == {"invalid_argument": "Invalid key: window"}
    assert api_check("factory", '{"otype": "APSM", "detectVersion": "fast"}') == {"invalid_argument": "Unknown detect version: fast"}
    assert api_check("factory", '{"otype": "APSM", "detectPrecision": "fp8"}') == {"invalid_argument": "Unknown detect precision: fp8"}


@pytest.mark.parametrize("bps", [1, 2, 4, 6])
def test_train_history_metrics_match_oracle(api_check, oracle, tmp_path, bps):
    rng = np.random.default_rng(bps)
    M = 3000
    scale = oracle.mod_scale(bps)
    ns = 2 if bps == 1 else 1 << (bps // 2)
    lv = lambda n: (2 * rng.integers(0, ns, n) - ns + 1) * scale  # noqa: E731
    true = np.stack([lv(M), lv(M) if bps > 1 else np.zeros(M)], -1).astype(np.float32)
    hat = (true + rng.normal(0, 0.6 * scale, true.shape)).astype(np.float32)
    hat[:5] = true[:5] * 3.0  # beyond the outermost level on the same side: no error
    true.tofile(tmp_path / "t.f32")
    hat.tofile(tmp_path / "h.f32")
    got = api_check("history", bps, repr(float(np.float32(scale))), tmp_path / "t.f32", tmp_path / "h.f32")
    assert got["same"] == 1
    assert got["evm"] == pytest.approx(oracle.evm(hat, true), rel=1e-5)
    assert got["ser"] == pytest.approx(oracle.symbol_errors(hat, true, bps) / M, abs=0.5 / M)
    assert got["ber"] == pytest.approx(oracle.bit_errors(hat, true, bps) / (M * bps), abs=0.5 / (M * bps))
    assert 0 < got["ser"] < 1


def test_shuffle_data_matches_fixture(api_check):
    fixture = json.load(open(os.path.join(ROOT, "tests", "golden", "shuffle_perm.json")))
    for T, head in fixture.items():
        got = api_check("shuffle", T)
        assert got["same"] == 1  # every antenna row and the pilots move together
        assert got["perm"][:len(head)] == head
        assert sorted(got["perm"]) == list(range(int(T)))


@pytest.mark.parametrize("bps", [1, 2, 4, 6, 8])
def test_constellations(api_check, oracle, bps):
    got = api_check("modulation", bps)
    assert got["scale"] == pytest.approx(oracle.mod_scale(bps), rel=1e-6)
    assert got["roundtrip"] == 1
    assert got["energy"] == pytest.approx(1.0, rel=1e-5)
    pts = np.array(got["points"])
    ref = np.array([oracle.mod_map(bps, s) for s in range(1 << bps)])
    ref = np.stack([ref.real, ref.imag], -1)
    assert np.abs(pts - ref).max() < 1e-6


@pytest.mark.parametrize("seed,bps", [(0, 2), (1, 4), (12345, 6)])
def test_seeded_reference_symbols(api_check, oracle, seed, bps):
    """Bit i of the -rs generator is std::bernoulli_distribution(0.5) over std::mt19937(seed):
    with libstdc++ it takes two 32-bit words per bit and is 1 exactly when the second word is
    below 2^31. The Mersenne-Twister words come from numpy's MT19937 with the same
    init_genrand seeding, an implementation independent of the C++ library."""
    users, nt, nd = 3, 7, 5
    got = api_check("refdata", seed, bps, users, nt, nd)
    bg = np.random.MT19937()
    bg._legacy_seeding(seed)
    nbits = users * (nt + nd) * bps
    raw = bg.random_raw(2 * nbits)
    bits = (raw[1::2] < 2 ** 31).astype(np.uint32)
    syms = []
    for k in range(users * (nt + nd)):
        v = 0
        for b in bits[k * bps:(k + 1) * bps]:
            v = (v << 1) | int(b)
        syms.append(oracle.mod_map(bps, v))
    syms = np.array(syms).reshape(users, nt + nd)
    tr = np.array(got["train"]).reshape(users, nt, 2)
    da = np.array(got["data"]).reshape(users, nd, 2)
    assert np.abs(tr[..., 0] + 1j * tr[..., 1] - syms[:, :nt]).max() < 1e-6
    assert np.abs(da[..., 0] + 1j * da[..., 1] - syms[:, nt:]).max() < 1e-6


def test_tool_help_and_usage_errors(tool):
    r = subprocess.run([tool, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--synchronized-data" in r.stdout and "--train-history-file" in r.stdout
    r = subprocess.run([tool], capture_output=True, text=True)
    assert r.returncode == 1 and "--synchronized-data: required." in r.stderr
    off = os.path.join(ROOT, "tests", "golden", "offline")
    r = subprocess.run([tool, "-s", os.path.join(off, "rxData_QPSK_time.bin"), "-o", "loud"], capture_output=True, text=True)
    assert r.returncode == 1 and "Invalid print mode: loud" in r.stderr
    r = subprocess.run([tool, "-s", os.path.join(off, "rxData_QPSK_time.bin"), "-cs", '{"modulation": "QPSK"}'],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "Neither --random-seed/-rs nor --reference-data/-r was specified" in r.stderr
    r = subprocess.run([tool, "-s", os.path.join(off, "missing.bin"), "-rs", "1"], capture_output=True, text=True)
    assert r.returncode == 1 and "File not found" in r.stderr
