"""lls_train_lowlevel_noalloc (the allocation-free least-squares entry point) against
lls_train_lowlevel on the same device inputs: the same kernels in the same order, so the
weights must be identical; two calls in a row, the second with a smaller problem (the shared
scratch keeps the first call's contents), and a problem beyond the scratch is rejected
before anything is enqueued."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(m, n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, m)).astype(np.float32)  # column-major m x n: column i at A[i]
    w_true = rng.standard_normal(n)
    y = (w_true @ A.astype(np.float64) + 0.05 * rng.standard_normal(m)).astype(np.float32)
    return A, y


def _handle_weights(dA, dy, n, m):
    from libapsm_amd import LlsNomaDetector
    det = LlsNomaDetector(n // 2 if n % 2 == 0 else 1)
    det.trainLowLevel(dA.ptr, dy.ptr, n, 1, m)
    det.synchronize()
    return det.getWeightMat().copy()


def test_noalloc_least_squares_equals_the_handle_entry_point(require_gpu):
    from libapsm_amd import LlsNomaDetector, lib
    from libapsm_amd._lib import DeviceArray, check
    for m, n, seed in ((60000, 32, 1), (3000, 16, 2)):
        print(f"no-alloc LLS: m={m} n={n}", flush=True)
        A, y = _problem(m, n, seed)
        dA, dy = DeviceArray(A), DeviceArray(y)
        ref = _handle_weights(dA, dy, n, m)
        dw = DeviceArray(np.full(n, np.nan, np.float32))
        ds = DeviceArray(np.array([7], np.int32))
        LlsNomaDetector.trainLowLevelNoAlloc(dA.ptr, dy.ptr, dw.ptr, ds.ptr, n, 1, m)
        check(lib().apsm_device_synchronize())
        w, st = dw.numpy(), ds.numpy()
        assert st[0] == 0
        np.testing.assert_array_equal(w, ref)
        lsq = np.linalg.lstsq(A.T.astype(np.float64), y.astype(np.float64), rcond=None)[0]
        assert np.abs(w - lsq).max() <= 1e-4 * max(1.0, np.abs(lsq).max())


def test_noalloc_least_squares_rejects_a_problem_beyond_its_scratch(require_gpu):
    from libapsm_amd import ApsmError, LlsNomaDetector
    from libapsm_amd._lib import DeviceArray
    d = DeviceArray(np.zeros(16, np.float32))
    with pytest.raises(ApsmError):
        # 128 unknowns over 10^6 rows needs about 520 MB of partial sums; nothing is launched
        LlsNomaDetector.trainLowLevelNoAlloc(d.ptr, d.ptr, d.ptr, d.ptr, 128, 1, 1000000)
