import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def hc():
    import hcore_amd
    L = hcore_amd.lib()
    st = L.HCORE_hip_init(0)
    assert st == 0, "HCORE_hip_init failed: %d" % st
    return hcore_amd


def glibc_rand_stream():
    """The reference tests' rand()/RAND_MAX stream (glibc, default seed 1)."""
    import ctypes
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    RAND_MAX = 2147483647.0

    def nxt():
        return libc.rand() / RAND_MAX
    return nxt


def rel_fro(X, ref):
    return float(np.linalg.norm(X - ref) / max(np.linalg.norm(ref), 1e-300))


def trunc_rank(sig, acc, rk=0):
    """Truncation rule of reference hcore_dgemm.c:444-464 on sorted sigma."""
    K = len(sig)
    if rk != 0:
        return min(rk, K)
    for i in range(1, K):
        if sig[i] < acc:
            return i
    return K
