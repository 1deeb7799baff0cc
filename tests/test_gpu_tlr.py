"""The TLR Cholesky update stream on the GPU (HCORE_hip_tlr_dpotrf over a
matrix generated and compressed by HCORE_hip_gytlr_batch) against the same
stream run by the oracle on the CPU, and against the dense FP64 Cholesky.

Bars: the generated ranks within one of the oracle's compression ranks; the
GPU factor's residual ||L L^T - A||_F / ||A||_F at most 10x the oracle
stream's (plus 1e-12); every final tile rank within 3 of the oracle
stream's (ranks can drift by one per step through rounding at the
truncation threshold); the factor within 1e-5 (relative Frobenius) of the
dense Cholesky factor of the same matrix.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def oracle_stream(orc, A_tiles, nt, nb, maxrk, acc):
    """The same right-looking TLR Cholesky with the oracle's routines.
    A_tiles: dict (i, k) -> dense diagonal tile or (U, V)."""
    T = dict(A_tiles)
    for k in range(nt):
        L, info, _ = orc.dpotrf(122, T[k, k])
        assert info == 0
        T[k, k] = np.tril(L)
        for i in range(k + 1, nt):
            U, V = T[i, k]
            X, _ = orc.dtrsm(141, 122, 111, 131, 1.0, T[k, k], V)
            T[i, k] = (U, X)
        for i in range(k + 1, nt):
            U, V = T[i, k]
            D, _ = orc.dsyrk(-1.0, U, V, T[i, i])
            T[i, i] = D
        for j in range(k + 1, nt):
            for i in range(j + 1, nt):
                AU, AV = T[i, k]
                BU, BV = T[j, k]
                CU, CV = T[i, j]
                st, nU, nV, r, _, _ = orc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=maxrk,
                                                acc=acc)
                assert st == orc.OK
                T[i, j] = (nU[:, :r], nV[:, :r])
    return T


def _assemble(T, nt, nb):
    L = np.zeros((nt * nb, nt * nb))
    for (i, k), t in T.items():
        L[i * nb:(i + 1) * nb, k * nb:(k + 1) * nb] = t if i == k else t[0] @ t[1].T
    return L


def test_tlr_cholesky_vs_oracle_stream(hc, orc):
    from hcore_amd.tlr import TlrMatrix
    bigM, nb, seed, acc = 2048, 256, 2024, 1e-8
    nt = bigM // nb
    M = TlrMatrix(bigM, nb, seed, maxrank=nb // 2, acc=acc, cap=nb)
    _say("case: TLR generate", bigM, nb)
    M.generate()
    P = orc.covariance_points(bigM, seed)
    A = orc.covariance_tile(bigM, seed, 0, 0, bigM, bigM, points=P)
    tiles = {}
    for k in range(nt):
        tiles[k, k] = np.asfortranarray(A[k * nb:(k + 1) * nb, k * nb:(k + 1) * nb])
        for i in range(k + 1, nt):
            blk = A[i * nb:(i + 1) * nb, k * nb:(k + 1) * nb]
            U, V, r = orc.dcompress(blk, maxrank=nb // 2, acc=acc)
            assert abs(int(M.ranks[i, k]) - r) <= 1, (i, k, M.ranks[i, k], r)
            tiles[i, k] = M.factors(i, k)  # the oracle stream starts from the GPU's tiles
    _say("case: TLR factor (GPU)")
    st, info, smax = M.factor(maxrk=nb, acc=acc)
    assert st == 0 and info == 0
    L = M.dense_lower()
    _say("case: TLR factor (oracle stream)")
    T = oracle_stream(orc, tiles, nt, nb, nb, acc)
    Lo = _assemble(T, nt, nb)
    nA = np.linalg.norm(A)
    res = np.linalg.norm(L @ L.T - A) / nA
    res_o = np.linalg.norm(Lo @ Lo.T - A) / nA
    _say("residual GPU %.3e oracle %.3e, max rank per step %s" % (res, res_o, list(smax)))
    assert res <= 10 * res_o + 1e-12
    for j in range(nt):
        for i in range(j + 1, nt):
            assert abs(int(M.ranks[i, j]) - T[i, j][0].shape[1]) <= 3, (i, j)
    Ld = np.linalg.cholesky(A)
    assert np.linalg.norm(L - Ld) / np.linalg.norm(Ld) <= 1e-5
    assert M.flops.potrf > 0 and M.flops.update > 0
    M.free()


def test_tlr_cholesky_not_spd_reports_info(hc):
    """A diagonal entry made negative in tile 2: the stream stops at that
    tile with LAPACK's info convention (global index of the failing pivot).
    The problem uses nugget 1e-3 here (HCORE_hip_set_covariance)."""
    from hcore_amd.tlr import TlrMatrix
    bigM, nb = 1024, 128
    assert hc.set_covariance(0.1, 1e-3) == 0
    try:
        M = TlrMatrix(bigM, nb, 5, maxrank=nb // 2, acc=1e-8)
        M.generate()
        # break tile 2's diagonal: -1 at its local pivot 10
        D2 = M.D[2].numpy()
        D2[10, 10] = -1.0
        M.D[2].upload(D2)
        _say("case: TLR factor, indefinite")
        st, info, _ = M.factor(maxrk=nb, acc=1e-8)
        assert st == 0 and 2 * nb < info <= 2 * nb + 11
        M.free()
    finally:
        hc.set_covariance(0.1, 1e-4)


def test_tlr_cholesky_split_precision_vs_fp64(hc, orc):
    """The stream with split-precision recompressions (BF16-digit products,
    HCORE_hip_tlr_dpotrf_form) on the same generated matrix as the FP64
    stream: residual ||L L^T - A||_F / ||A||_F within 2x of the FP64 stream's
    (plus 1e-12), final ranks within 3, the factor within 1e-5 of the dense
    Cholesky; an unknown form or slice count is rejected."""
    from hcore_amd.tlr import TlrMatrix
    bigM, nb, seed, acc = 2048, 256, 2024, 1e-8
    P = orc.covariance_points(bigM, seed)
    A = orc.covariance_tile(bigM, seed, 0, 0, bigM, bigM, points=P)
    nA = np.linalg.norm(A)
    out = {}
    for name, prec in (("fp64", hc.FP64), ("split", hc.SPLIT)):
        M = TlrMatrix(bigM, nb, seed, maxrank=nb // 2, acc=acc, cap=nb)
        M.generate()
        _say("case: TLR factor", name)
        st, info, _ = M.factor(maxrk=nb, acc=acc, precision=prec)
        assert st == 0 and info == 0
        L = M.dense_lower()
        out[name] = (np.linalg.norm(L @ L.T - A) / nA, M.ranks.copy(), L)
        M.free()
    _say("residual fp64 %.3e split %.3e" % (out["fp64"][0], out["split"][0]))
    assert out["split"][0] <= 2 * out["fp64"][0] + 1e-12
    low = np.tril_indices(bigM // nb, -1)
    assert np.max(np.abs(out["split"][1][low].astype(int) - out["fp64"][1][low])) <= 3
    Ld = np.linalg.cholesky(A)
    assert np.linalg.norm(out["split"][2] - Ld) / np.linalg.norm(Ld) <= 1e-5
    M = TlrMatrix(512, 256, seed, maxrank=128, acc=acc, cap=256)
    M.generate()
    for prec, sl in ((hc.MIXED, 0), (hc.SPLIT, 7)):
        st, _, _ = M.factor(maxrk=256, acc=acc, precision=prec, slices=sl)
        assert st == hc.ERR_ARG
    M.free()
