"""Pins of the oracle's tile compression and covariance problem — no GPU.

orc.dcompress restates the compression step of the reference's generator
(testing_dgemm.c:180-226); orc.covariance_points / orc.covariance_tile define
the covariance problem of HCORE_dgytlr.  Pinned by:
1. published known answers of splitmix64 (seed 0: e220a8397b1dcdaf,
   6e789e6aa1b965f4, 06c45d188009454f);
2. the Z-order of the points, checked with an independent comparison (the
   dimension whose coordinates differ in the highest bit decides), and the
   point multiset;
3. covariance tiles against scipy's cdist distances, symmetry across the
   diagonal, the diagonal value and positive definiteness;
4. compression against numpy's SVD (singular values, truncation rule, the
   truncated product, orthonormal U, sigma-scaled V), on the reference's own
   generated inputs too; planted mistakes (V unscaled, one singular value too
   many) that these checks catch;
5. the library's host-side point generation equal to the oracle's, bit for bit.
"""
import numpy as np
import pytest

from conftest import rel_fro


def test_splitmix64_known_answers(orc):
    assert orc.splitmix64(0, 1) == 0xE220A8397B1DCDAF
    assert orc.splitmix64(0, 2) == 0x6E789E6AA1B965F4
    assert orc.splitmix64(0, 3) == 0x06C45D188009454F


def _zless(a, b):
    """a before b in Z-order (x in the even bits): the dimension with the
    highest differing bit decides, y on a tie of that bit position."""
    qa = [min(int(t * 65536.0), 65535) for t in a]
    qb = [min(int(t * 65536.0), 65535) for t in b]
    best, dim = -1, 0
    for d in (0, 1):
        x = qa[d] ^ qb[d]
        hb = x.bit_length() - 1
        if hb > best or (hb == best and hb >= 0):
            best, dim = hb, d
    if best < 0:
        return None  # same cell
    return qa[dim] < qb[dim]


def test_points_z_order_and_multiset(orc):
    bigM, seed = 3000, 11
    P = orc.covariance_points(bigM, seed)
    assert P.min() >= 0.0 and P.max() < 1.0
    for i in range(bigM - 1):
        z = _zless(P[i], P[i + 1])
        assert z is None or z, i
    raw = np.array([[(orc.splitmix64(seed, 2 * i + 1) >> 11) * 2.0 ** -53,
                     (orc.splitmix64(seed, 2 * i + 2) >> 11) * 2.0 ** -53] for i in range(bigM)])
    key = lambda A: A[np.lexsort((A[:, 1], A[:, 0]))]
    assert np.array_equal(key(P), key(raw))


def test_covariance_tile_properties(orc):
    from scipy.spatial.distance import cdist
    bigM, seed, beta, nug = 512, 3, 0.1, 1e-4
    P = orc.covariance_points(bigM, seed)
    A = orc.covariance_tile(bigM, seed, 0, 0, bigM, bigM, beta, nug, points=P)
    ref = np.exp(-cdist(P, P) / beta) + nug * np.eye(bigM)
    assert np.max(np.abs(A - ref)) <= 1e-14
    assert np.array_equal(A, A.T)
    assert np.all(np.diag(A) == 1.0 + nug)
    np.linalg.cholesky(A)  # positive definite
    # an off-diagonal tile is the transpose of its mirror
    T1 = orc.covariance_tile(bigM, seed, 128, 0, 64, 96, points=P)
    T2 = orc.covariance_tile(bigM, seed, 0, 128, 96, 64, points=P)
    assert np.array_equal(T1, T2.T)
    assert np.array_equal(T1, A[128:192, 0:96])


def _spectrum_matrix(m, n, decay, seed):
    rng = np.random.default_rng(seed)
    k = min(m, n)
    Q1, _ = np.linalg.qr(rng.standard_normal((m, k)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, k)))
    s = 10.0 ** (-decay * np.arange(k))
    return np.asfortranarray((Q1 * s) @ Q2.T), s


@pytest.mark.parametrize("m,n,maxrank", [(40, 40, 40), (60, 25, 25), (25, 60, 10), (128, 96, 96)])
def test_compress_vs_numpy_svd(orc, m, n, maxrank):
    A, _ = _spectrum_matrix(m, n, 0.4, m + n)
    acc = 1e-8
    U, V, r = orc.dcompress(A, maxrank=maxrank, acc=acc)
    s = np.linalg.svd(A, compute_uv=False)
    assert r == min(int(np.sum(s >= acc)), maxrank)
    Un, sn, Vtn = np.linalg.svd(A)
    best = (Un[:, :r] * sn[:r]) @ Vtn[:r]
    assert rel_fro(U @ V.T, best) <= 1e-12
    assert np.max(np.abs(U.T @ U - np.eye(r))) <= 1e-12
    assert np.allclose(np.linalg.norm(V, axis=0), sn[:r], rtol=1e-12)


def test_compress_rank_zero_and_full(orc):
    A = 1e-12 * np.ones((10, 8))
    U, V, r = orc.dcompress(A, acc=1e-8)
    assert r == 0
    B = np.random.default_rng(1).standard_normal((10, 8))
    U, V, r = orc.dcompress(B, acc=0.0)
    assert r == 8 and rel_fro(U @ V.T, B) <= 1e-13


def test_compress_matches_reference_generator(orc):
    """The reference generator's compression (testing_dgemm.c:119-236 on its
    own dlatms input, iseed {0,0,0,1}) and orc.dcompress of its dense
    matrix agree."""
    M, smr, acc = 10, 6, 1e-4
    iseed = np.array([0, 0, 0, 1], dtype=np.int32)
    D, GU, GV, gr = orc.generate_compress(M, M, M, smr, M, acc, iseed)
    U, V, r = orc.dcompress(D, maxrank=smr, acc=acc)
    assert r == gr
    assert rel_fro(U @ V.T, GU[:, :gr] @ GV[:M, :gr].T) <= 1e-13


def test_planted_compress_mistakes_are_caught(orc):
    A, _ = _spectrum_matrix(50, 40, 0.4, 5)
    Un, sn, Vtn = np.linalg.svd(A)
    U, V, r = orc.dcompress(A, acc=1e-8, mutate=1)  # V not scaled by sigma
    assert not np.allclose(np.linalg.norm(V, axis=0), sn[:r], rtol=1e-12)
    U, V, r = orc.dcompress(A, acc=1e-8, mutate=2)  # one singular value too many
    assert r != int(np.sum(sn >= 1e-8))


def test_library_points_equal_oracle(orc):
    """HCORE_hip_covariance_points is host code: no GPU needed."""
    import hcore_amd
    for bigM, seed in ((1, 0), (777, 5), (4096, 123456789)):
        assert np.array_equal(hcore_amd.covariance_points(bigM, seed),
                              orc.covariance_points(bigM, seed))
