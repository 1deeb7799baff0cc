"""GPU parity of tile generation and compression: HCORE_dcompress,
HCORE_dgytlr and the batched HCORE_hip_gytlr_batch / HCORE_hip_compress_batch
against the oracle (LAPACK dgesvd compression, the covariance problem).

Bars: the rank equal to the oracle's within one (a singular value within
rounding of acc may fall either side); the product U V^T within
sqrt(min(m, n)) * acc (absolute Frobenius) of the dense tile and within the
same of the oracle's product; U orthonormal to 1e-10; dense covariance tiles
within 1e-14 of the oracle's (exp and sqrt on the device round differently
by at most an ulp).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _spectrum_matrix(m, n, decay, seed):
    rng = np.random.default_rng(seed)
    k = min(m, n)
    Q1, _ = np.linalg.qr(rng.standard_normal((m, k)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, k)))
    s = 10.0 ** (-decay * np.arange(k))
    return np.asfortranarray((Q1 * s) @ Q2.T)


def _check(U, V, r, A, oU, oV, ro, acc):
    m, n = A.shape
    bound = np.sqrt(min(m, n)) * acc
    assert abs(r - ro) <= 1, (r, ro)
    assert np.linalg.norm(U @ V.T - A) <= bound + 1e-13 * np.linalg.norm(A)
    assert np.linalg.norm(U @ V.T - oU @ oV.T) <= 2 * bound + 1e-13 * np.linalg.norm(A)
    if r:
        assert np.max(np.abs(U.T @ U - np.eye(r))) <= 1e-10


SHAPES = [(16, 16, 0.5), (40, 33, 0.3), (64, 64, 0.25), (100, 150, 0.1), (192, 192, 0.06),
          (256, 256, 0.05), (300, 200, 0.04), (512, 512, 0.03), (512, 100, 0.05)]


@pytest.mark.parametrize("m,n,decay", SHAPES, ids=["%dx%d" % s[:2] for s in SHAPES])
def test_dcompress_vs_oracle(hc, orc, m, n, decay):
    _say("case: dcompress", m, n)
    A = _spectrum_matrix(m, n, decay, m * 7 + n)
    acc = 1e-8
    U, V, r = hc.dcompress(A, acc=acc)
    oU, oV, ro = orc.dcompress(A, acc=acc)
    _check(U, V, r, A, oU, oV, ro, acc)
    # V carries sigma: its column norms are the singular values
    s = np.linalg.svd(A, compute_uv=False)
    assert np.allclose(np.linalg.norm(V, axis=0), s[:r], rtol=1e-8, atol=1e-12)


def test_dcompress_reference_generator_input(hc, orc):
    """The reference test's own tile (dlatms, sigma_i = 10^-i, iseed
    {0,0,0,1}, testing_dgemm.c:119-236) compressed on the GPU."""
    M, smr, acc = 10, 6, 1e-4
    iseed = np.array([0, 0, 0, 1], dtype=np.int32)
    D, GU, GV, gr = orc.generate_compress(M, M, M, smr, M, acc, iseed)
    _say("case: dcompress reference generator tile")
    U, V, r = hc.dcompress(D, maxrank=smr, acc=acc)
    assert r == gr
    assert np.linalg.norm(U @ V.T - GU[:, :gr] @ GV[:M, :gr].T) <= 1e-12


def test_dcompress_maxrank_cap_and_rank_zero(hc, orc):
    A = _spectrum_matrix(80, 80, 0.05, 3)
    _say("case: dcompress maxrank cap")
    U, V, r = hc.dcompress(A, maxrank=12, acc=1e-8)
    oU, oV, ro = orc.dcompress(A, maxrank=12, acc=1e-8)
    assert r == ro == 12
    assert np.linalg.norm(U @ V.T - oU @ oV.T) <= 1e-10 * np.linalg.norm(A)
    _say("case: dcompress rank zero")
    U, V, r = hc.dcompress(1e-12 * np.ones((30, 20)), acc=1e-8)
    assert r == 0


def test_dcompress_nonfinite_is_an_error(hc):
    A = np.ones((32, 32))
    A[3, 4] = np.nan
    _say("case: dcompress NaN")
    hc.set_error_policy(hc.ON_ERROR_RETURN)
    try:
        hc.dcompress(A, acc=1e-8)
        assert hc.last_status() == hc.ERR_NUMERIC
        hc.dcompress(np.ones((600, 10)), acc=1e-8)  # beyond the 512 limit
        assert hc.last_status() == hc.ERR_ARG
    finally:
        hc.set_error_policy(hc.ON_ERROR_EXIT)


def test_dcompress_device_pointers(hc, orc):
    m, n = 200, 120
    A = _spectrum_matrix(m, n, 0.08, 9)
    Ad = hc.DeviceBuffer.from_numpy(A)
    Ud, Vd = hc.DeviceBuffer(m, n), hc.DeviceBuffer(n, n)
    rk = hc.DeviceBuffer(1, 1)
    _say("case: dcompress device pointers")
    import ctypes as C
    L = hc.lib()
    dp = lambda b: C.cast(C.c_void_p(b.ptr), C.POINTER(C.c_double))
    L.HCORE_dcompress(m, n, dp(Ad), m, dp(Ud), m, dp(Vd), n, dp(rk), n, 1e-8)
    r = int(rk.numpy()[0, 0])
    oU, oV, ro = orc.dcompress(A, acc=1e-8)
    _check(Ud.numpy(r), Vd.numpy(r), r, A, oU, oV, ro, 1e-8)


@pytest.mark.parametrize("m0,n0,nb", [(0, 0, 128), (256, 0, 128), (512, 128, 256), (0, 1536, 512)])
def test_dgytlr_vs_oracle(hc, orc, m0, n0, nb):
    bigM, seed = 2048, 42
    P = orc.covariance_points(bigM, seed)
    ref = orc.covariance_tile(bigM, seed, m0, n0, nb, nb, points=P)
    _say("case: dgytlr", m0, n0, nb)
    if m0 == n0:
        _, _, _, D = hc.dgytlr(nb, nb, bigM, m0, n0, seed)
        assert np.max(np.abs(D - ref)) <= 1e-14
        U, V, r, D2 = hc.dgytlr(nb, nb, bigM, m0, n0, seed, compress_diag=True, dense=True)
    else:
        U, V, r, D2 = hc.dgytlr(nb, nb, bigM, m0, n0, seed, tol=1e-8, dense=True)
    assert np.max(np.abs(D2 - ref)) <= 1e-14
    oU, oV, ro = orc.dcompress(ref, acc=1e-8)
    _check(U, V, r, ref, oU, oV, ro, 1e-8)


def test_gytlr_batch_grid(hc, orc):
    """Every lower tile of an 8 x 8 grid of nb = 256 (bigM = 2048) in one
    batch: diagonal tiles dense, the others compressed; each against the
    oracle."""
    bigM, seed, nb = 2048, 7, 256
    nt = bigM // nb
    P = orc.covariance_points(bigM, seed)
    tiles = [(i, j) for j in range(nt) for i in range(j, nt)]
    AU = [hc.DeviceBuffer(nb, nb) for _ in tiles]
    AV = [hc.DeviceBuffer(nb, nb) for _ in tiles]
    AD = [hc.DeviceBuffer(nb, nb) if i == j else None for i, j in tiles]
    rank = hc.DeviceBuffer(len(tiles), 1, np.int32)
    _say("case: gytlr batch", len(tiles), "tiles")
    st = hc.gytlr_batch(nb, nb, [i * nb for i, j in tiles], [j * nb for i, j in tiles], bigM, seed,
                        [b.ptr for b in AU], nb, [b.ptr for b in AV], nb, rank.ptr,
                        AD=[b.ptr if b else None for b in AD], lda=nb, tol=1e-8)
    assert st == 0
    ranks = rank.numpy()[:, 0]
    for t, (i, j) in enumerate(tiles):
        ref = orc.covariance_tile(bigM, seed, i * nb, j * nb, nb, nb, points=P)
        if i == j:
            assert np.max(np.abs(AD[t].numpy() - ref)) <= 1e-14
            continue
        r = int(ranks[t])
        oU, oV, ro = orc.dcompress(ref, acc=1e-8)
        _check(AU[t].numpy(r), AV[t].numpy(r), r, ref, oU, oV, ro, 1e-8)


def test_compress_batch(hc, orc):
    m, n, k = 300, 300, 12
    A = [_spectrum_matrix(m, n, 0.02 + 0.01 * q, 100 + q) for q in range(k)]
    Ad = [hc.DeviceBuffer.from_numpy(a) for a in A]
    AU = [hc.DeviceBuffer(m, n) for _ in A]
    AV = [hc.DeviceBuffer(n, n) for _ in A]
    rank = hc.DeviceBuffer(k, 1, np.int32)
    _say("case: compress batch", k)
    st = hc.compress_batch(m, n, [b.ptr for b in Ad], m, [b.ptr for b in AU], m,
                           [b.ptr for b in AV], n, rank.ptr, tol=1e-8)
    assert st == 0
    ranks = rank.numpy()[:, 0]
    for q in range(k):
        r = int(ranks[q])
        oU, oV, ro = orc.dcompress(A[q], acc=1e-8)
        _check(AU[q].numpy(r), AV[q].numpy(r), r, A[q], oU, oV, ro, 1e-8)
