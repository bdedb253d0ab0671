"""Batched instances (apsm_batch_*) on the GPU: every instance of a ragged batch gets exactly
what a handle of its own computes (same kernels, same order: bitwise), graph replays repeat
the result, and a sample of instances is checked against the FP64 oracle with the parity
tolerance of test_gpu_parity (soft outputs within 2e-3)."""
import numpy as np
import pytest

from helpers import to_ri

pytestmark = pytest.mark.gpu


def make_batch(count, seed, modes=("fp32",)):
    from libapsm_amd import datagen
    rng = np.random.default_rng(seed)
    inst = []
    for i in range(count):
        N = int(rng.choice([4, 8, 16, 32]))
        w = int(rng.choice([40, 80, 160]))
        dc = int(rng.integers(max(w, 60), 400))
        T = int(rng.integers(50, 450))
        M = int(rng.integers(100, 3000))
        U = int(rng.choice([1, 1, 2, 3]))
        mod = str(rng.choice(["QPSK", "QAM16"]))
        d = datagen.generate(N, max(U, 3), T, M, modulation=mod, snr_db=15.0, seed=1000 + i)
        cfg = dict(windowSize=w, dictionarySize=dc, eB=0.01, gaussianKernelVariance=datagen.kernel_variance_for_sigma(0.5, N),
                   sampleStep=int(rng.choice([2, 2, 4])), detectPrecision=str(rng.choice(list(modes))))
        inst.append(dict(cfg=cfg, N=N, U=U, d=d, bps=d["bps"]))
    return inst


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_equals_single_handles(require_gpu, use_graph):
    from libapsm_amd import ApsmBatch, ApsmNomaDetectorFactory
    inst = make_batch(14, seed=5, modes=("fp32", "fp32-vector", "fp32-matrix", "bf16", "fp16"))
    b = ApsmBatch([x["cfg"] for x in inst], [x["N"] for x in inst], [x["U"] for x in inst])
    b.plan([x["d"]["rx_train"] for x in inst], [x["d"]["tx_train"][:x["U"]] for x in inst], [x["d"]["rx_data"] for x in inst],
           use_graph=use_graph)
    b.synchronize()
    first = [b.estimates(i) for i in range(len(inst))]
    for rep in range(2):
        b.run()
        b.synchronize()
        for i in range(len(inst)):
            np.testing.assert_array_equal(b.estimates(i), first[i])
    assert b.last_ms() > 0
    tun = b.tuning()
    # workgroups per CU: a power of two up to 128, or beyond 128 the one-item-per-workgroup grid
    assert tun and all(t["threads"] in (128, 256, 512) and (t["blocks_per_cu"] in (1, 2, 4, 8, 16, 32, 64, 128) or t["blocks_per_cu"] > 128)
                       for t in tun)
    assert all(t["samples_per_workgroup"] in (64, 128, 256, 512) and t["entries_per_tile"] in (32, 64, 128) for t in tun)
    assert {t["kernel"] for t in tun} >= {"vector", "matrix-bf16"}
    for i, x in enumerate(inst):
        det = ApsmNomaDetectorFactory.build(dict(x["cfg"], otype="APSM"), x["N"], x["U"])
        det.train(x["d"]["rx_train"], x["d"]["tx_train"][:x["U"]])
        ref = det.detect(x["d"]["rx_data"])
        ref = ref[None, :] if x["U"] == 1 else ref
        np.testing.assert_array_equal(first[i], ref.astype(np.complex64), err_msg=f"instance {i} {x['cfg']} N={x['N']} U={x['U']}")


def test_batch_matches_oracle(require_gpu, oracle):
    from libapsm_amd import ApsmBatch
    inst = make_batch(6, seed=8, modes=("fp32", "fp32-vector"))
    b = ApsmBatch([x["cfg"] for x in inst], [x["N"] for x in inst], [x["U"] for x in inst])
    b.plan([x["d"]["rx_train"] for x in inst], [x["d"]["tx_train"][:x["U"]] for x in inst], [x["d"]["rx_data"] for x in inst])
    b.synchronize()
    for i, x in enumerate(inst):
        est = b.estimates(i)
        for u in range(x["U"]):
            o = oracle.Apsm(x["cfg"], x["N"])
            o.train(x["d"]["rx_train"], x["d"]["tx_train"][u], incremental=True)
            ref = o.detect(x["d"]["rx_data"])
            err = np.abs(to_ri(est[u]).astype(np.float64) - ref).max()
            print(f"instance {i} user {u}: max soft err {err:.2e}")
            assert err <= 2e-3


def test_batch_rejects_unsupported_instances(require_gpu):
    from libapsm_amd import ApsmBatch, ApsmError
    b = ApsmBatch([dict(windowSize=300, dictionarySize=400, gaussianKernelVariance=0.1)], [4])
    d = np.zeros((4, 10), np.complex64)
    with pytest.raises(ApsmError):
        b.plan([d], [np.zeros(10, np.complex64)], [d])


def test_replan_right_after_a_run(require_gpu):
    """A plan issued while the previous run is still on the device (no synchronize between
    them), with other pilot and sample counts, gives the single-handle results of the new
    inputs."""
    from libapsm_amd import ApsmBatch, ApsmNomaDetectorFactory, datagen
    inst = make_batch(6, seed=21)
    b = ApsmBatch([x["cfg"] for x in inst], [x["N"] for x in inst], [x["U"] for x in inst])
    b.plan([x["d"]["rx_train"] for x in inst], [x["d"]["tx_train"][:x["U"]] for x in inst], [x["d"]["rx_data"] for x in inst])
    b.run()
    b.run()
    rng = np.random.default_rng(22)
    new = []
    for i, x in enumerate(inst):
        T = int(rng.integers(50, 450))
        M = int(rng.integers(100, 3000))
        new.append(datagen.generate(x["N"], max(x["U"], 3), T, M, modulation="QPSK", snr_db=15.0, seed=2000 + i))
    b.plan([d["rx_train"] for d in new], [d["tx_train"][:x["U"]] for d, x in zip(new, inst)], [d["rx_data"] for d in new])
    b.synchronize()
    for i, (x, d) in enumerate(zip(inst, new)):
        det = ApsmNomaDetectorFactory.build(dict(x["cfg"], otype="APSM"), x["N"], x["U"])
        det.train(d["rx_train"], d["tx_train"][:x["U"]])
        ref = det.detect(d["rx_data"])
        ref = ref[None, :] if x["U"] == 1 else ref
        np.testing.assert_array_equal(b.estimates(i), ref.astype(np.complex64), err_msg=f"instance {i}")
