"""HCORE_dgemm_noalloc (the allocation-free HCORE_dgemm, whose workspace is a
static device array) against HCORE_dgemm on the same device inputs.

Bar: bit-identical factors and rank (the same kernels and the same plan run
on the same inputs), on two calls in a row, the second with a smaller
input; an update beyond the static workspace is rejected with
HCORE_HIP_ERR_ARG before any kernel runs, and the array is still usable
after that.
"""
import ctypes as C

import numpy as np
import pytest

from test_oracle import _random_case

pytestmark = pytest.mark.gpu


def _call(hc, fn, M, ra, rb, rc, data, maxrk, acc, rk=0):
    """fn on device copies of data; returns (CU, CV, rank) read back."""
    AU, AV, BU, BV, CU, CV = data
    K = ra + rc
    bufs = [hc.DeviceBuffer.from_numpy(np.asfortranarray(x)) for x in (AU, AV, BU, BV)]
    cu = hc.DeviceBuffer.from_numpy(np.asfortranarray(CU), K)
    cv = hc.DeviceBuffer.from_numpy(np.asfortranarray(CV), K)
    rks = np.array([ra, rb, rc], dtype=np.float64)
    dp = lambda b: C.cast(C.c_void_p(b.ptr), C.POINTER(C.c_double))
    rp = lambda i: rks[i:i + 1].ctypes.data_as(C.POINTER(C.c_double))
    fl = hc.FlopCounter()
    fn(111, 112, M, M, -1.0, dp(bufs[0]), dp(bufs[1]), rp(0), M, dp(bufs[2]), dp(bufs[3]), rp(1),
       M, 1.0, dp(cu), dp(cv), rp(2), M, rk, maxrk, acc, None, C.byref(fl))
    st = hc.last_status()
    r = int(rks[2])
    out = (cu.numpy(r), cv.numpy(r), r, st, fl.update)
    for b in bufs + [cu, cv]:
        b.free()
    return out


def test_noalloc_matches_dgemm(hc):
    L = hc.lib()
    fn = L.HCORE_dgemm_noalloc
    fn.argtypes = L.HCORE_dgemm.argtypes
    fn.restype = None
    L.HCORE_dgemm_noalloc_workspace_bytes.restype = C.c_size_t
    cap = L.HCORE_dgemm_noalloc_workspace_bytes()
    print("static workspace bytes", cap, flush=True)
    assert cap == 64 << 20
    hc.set_error_policy(hc.ON_ERROR_RETURN)
    try:
        for (seed, M, ra, rb, rc) in [(81, 1024, 60, 50, 100), (82, 256, 12, 9, 20)]:
            print("case: noalloc vs dgemm", seed, M, ra, rb, rc, flush=True)
            data = _random_case(seed, M, ra, rb, rc)
            ref = _call(hc, L.HCORE_dgemm, M, ra, rb, rc, data, M, 1e-8)
            got = _call(hc, fn, M, ra, rb, rc, data, M, 1e-8)
            assert ref[3] == 0 and got[3] == 0
            assert got[2] == ref[2]
            assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
            assert got[4] == ref[4]  # same flop accounting
        # an update whose workspace exceeds the static array: rejected, nothing run
        M, ra, rb, rc = 16384, 256, 8, 256
        print("case: noalloc beyond the static workspace", M, ra + rc, flush=True)
        z = lambda r: np.zeros((M, r))
        got = _call(hc, fn, M, ra, rb, rc, (z(ra), z(ra), z(rb), z(rb), z(rc), z(rc)), M, 1e-8)
        assert got[3] == hc.ERR_ARG
        # the array still serves a call after the rejection
        data = _random_case(83, 300, 7, 5, 9)
        ref = _call(hc, L.HCORE_dgemm, 300, 7, 5, 9, data, 300, 1e-8)
        got = _call(hc, fn, 300, 7, 5, 9, data, 300, 1e-8)
        assert got[3] == 0 and got[2] == ref[2]
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    finally:
        hc.set_error_policy(hc.ON_ERROR_EXIT)
