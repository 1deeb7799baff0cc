"""The matrix-core GEMM engine (every kernel family, both precisions, all
transposes, full and partial tiles) against numpy.

Bar: max |C - C_ref| <= c * eps * k * max|A| max|B| (c = 4; eps of the
precision), i.e. the rounding of a length-k dot product.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16, 4), (37, 70, 45), (64, 64, 64), (100, 33, 300), (130, 97, 1024)]
TALL = [(256, 32, 16), (300, 70, 64), (1024, 96, 128), (700, 20, 7)]


def _check(hc, dt, m, n, k, ta, tb, path, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((k, m) if ta else (m, k)).astype(dt)
    B = rng.standard_normal((n, k) if tb else (k, n)).astype(dt)
    C = rng.standard_normal((m, n)).astype(dt)
    got = hc.gemm(A, B, C, alpha=-0.5, beta=0.25, transA=ta, transB=tb, path=path)
    a64, b64 = A.astype(np.float64), B.astype(np.float64)
    ref = -0.5 * ((a64.T if ta else a64) @ (b64.T if tb else b64)) + 0.25 * C.astype(np.float64)
    eps = np.finfo(dt).eps
    err = np.abs(got.astype(np.float64) - ref)
    bound = 4 * eps * k * np.abs(A).max() * np.abs(B).max() + 4 * eps * np.abs(ref).max()
    if err.max() > bound:
        bad = np.argwhere(err > bound)
        print("path", path, dt.__name__, (m, n, k, ta, tb), "bad entries", len(bad),
              "first", bad[:8].tolist(), "cols mod 16", sorted(set((bad[:, 1] % 16).tolist())),
              flush=True)
    assert err.max() <= bound


@pytest.mark.parametrize("dt", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("path", [1, 2, 3])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_gemm_tiles(hc, dt, path, ta, tb):
    for i, (m, n, k) in enumerate(SHAPES):
        _check(hc, dt, m, n, k, ta, tb, path, 1000 * path + i)


@pytest.mark.parametrize("dt", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("tb", [False, True])
def test_gemm_tall(hc, dt, tb):
    for i, (m, n, k) in enumerate(TALL):
        _check(hc, dt, m, n, k, False, tb, 4, 77 + i)
