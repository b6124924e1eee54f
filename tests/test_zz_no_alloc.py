"""The allocation-free QPGF batch (bq_qpgf_batch_noalloc) against bq_qpgf_batch.

Device input and output buffers are the caller's (hipMalloc through tests/hipmem.py); the
library allocates nothing in the call.  Two calls in a row, the second with fewer points, must each reproduce
bq_qpgf_batch bit for bit (same kernels), including the error word of a plan with a Wood
anomaly, and sizes beyond the documented maximum are rejected without a launch.
"""
import ctypes

import numpy as np
import pytest

D = ctypes.POINTER(ctypes.c_double)


def _ref(L, qd, variant, r):
    nv = 20 if variant == 2 else 8
    out = np.zeros((len(r), nv))
    rc = L.bq_qpgf_batch(qd.ctypes.data_as(D), variant, len(r), np.ascontiguousarray(r).ctypes.data_as(D),
                         out.ctypes.data_as(D))
    return rc, out


def _noalloc(L, qd, variant, r):
    from hipmem import DeviceArray
    nv = 20 if variant == 2 else 8
    dr = DeviceArray(np.ascontiguousarray(r, dtype=np.float64))
    dout = DeviceArray(np.full((len(r), nv), np.nan))
    L.bq_qpgf_batch_noalloc.argtypes = [D, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p]
    rc = L.bq_qpgf_batch_noalloc(qd.ctypes.data_as(D), variant, len(r), dr.ptr, dout.ptr, None)
    out = dout.get()
    dr.free()
    dout.free()
    return rc, out


@pytest.mark.gpu
@pytest.mark.parametrize("variant", [0, 2])
def test_gpu_qpgf_noalloc_matches_batch(gpu_lib, variant):
    L = gpu_lib.lib()
    rng = np.random.default_rng(11)
    k = 2 * np.pi / 0.55
    qd = np.array([k, 0.3 * k, 0.1 * k, 0.9 * k, 0.9, 0.0, 0.45, 0.779])
    for n in (5000, 300):  # second call smaller than the first
        r = rng.uniform(-0.8, 0.8, (n, 3))
        r[::7, 2] = 0.0
        rc0, ref = _ref(L, qd, variant, r)
        rc1, got = _noalloc(L, qd, variant, r)
        print(variant, n, rc0, rc1, np.abs(got - ref).max())
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(got, ref)


@pytest.mark.gpu
def test_gpu_qpgf_noalloc_error_word_and_limit(gpu_lib):
    L = gpu_lib.lib()
    r = np.array([[0.1, 0.2, 0.5], [0.3, -0.2, 0.05], [0.2, 0.1, 0.0]])
    wood = np.array([1.5, 2 * np.pi - 1.5, 0.0, 0.1, 1.0, 0.0, 0.0, 1.0])
    rc0, ref = _ref(L, wood, 0, r)
    rc1, got = _noalloc(L, wood, 0, r)
    assert rc0 in (-12, -23) and rc1 == rc0
    assert np.array_equal(got, ref)
    # the error word is cleared by the next call
    ok = np.array([1.5, 2 * np.pi - 1.4, 0.0, 0.1, 1.0, 0.0, 0.0, 1.0])
    rc2, _ = _noalloc(L, ok, 0, r)
    assert rc2 == 0
    # beyond 64 chunks of 2^20 points: rejected before any launch (null buffers are not read)
    L.bq_qpgf_batch_noalloc.argtypes = [D, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p]
    big = 64 * (1 << 20) + 1
    rc3 = L.bq_qpgf_batch_noalloc(ok.ctypes.data_as(D), 0, big, ctypes.c_void_p(8), ctypes.c_void_p(8), None)
    assert rc3 == -2 and "exceed the maximum" in L.bq_last_error().decode()
