"""GPU parity of the dense neighbours of the recompression: HCORE_dpotrf,
HCORE_dtrsm, HCORE_dgemmbdcd and HCORE_duncompress against the oracle
(LAPACK/BLAS) on the same inputs.

Bars: potrf factor equal to the oracle's within 1e-12 relative (the factor
with a positive diagonal is unique), the untouched triangle bit-identical,
the same info for a matrix that is not positive definite; trsm solution
within 1e-12 relative of the oracle's (well-conditioned triangles); the
dense products within 1e-13 relative; flop counters equal to the oracle's.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n))
    return np.asfortranarray(G @ G.T / n + np.eye(n))


@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("n", [1, 64, 200, 512, 1024])
def test_dpotrf_vs_oracle(hc, orc, uplo, n):
    _say("case: dpotrf uplo", uplo, "n", n)
    A = _spd(n, n + uplo)
    fl = hc.FlopCounter()
    F, info = hc.dpotrf(uplo, A, flops=fl)
    R, oinfo, ofl = orc.dpotrf(uplo, A)
    assert info == oinfo == 0
    tri = np.tril if uplo == 122 else np.triu
    other = (lambda X: np.triu(X, 1)) if uplo == 122 else (lambda X: np.tril(X, -1))
    assert np.max(np.abs(tri(F) - tri(R))) <= 1e-12 * np.abs(tri(R)).max()
    assert np.array_equal(other(F), other(A))
    assert fl.potrf == ofl


@pytest.mark.parametrize("uplo", [121, 122])
def test_dpotrf_not_positive_definite(hc, orc, uplo):
    """A negative pivot in the third diagonal block: the same info as
    LAPACK's (the order of the failing leading minor)."""
    n = 300
    A = _spd(n, 5)
    A[150, 150] = -10.0
    _say("case: dpotrf not positive definite, uplo", uplo)
    F, info = hc.dpotrf(uplo, A)
    _, oinfo, _ = orc.dpotrf(uplo, A)
    assert info == oinfo == 151


def test_dpotrf_invalid_uplo(hc):
    """An uplo that is neither triangle: info = -1, nothing runs."""
    F, info = hc.dpotrf(0, np.eye(4))
    assert info == -1


TRSM = [(s, u, t, d) for s in (141, 142) for u in (121, 122) for t in (111, 112)
        for d in (131, 132)]


@pytest.mark.parametrize("case", TRSM, ids=["s%d_u%d_t%d_d%d" % c for c in TRSM])
def test_dtrsm_vs_oracle(hc, orc, case):
    side, uplo, trans, diag = case
    rng = np.random.default_rng(sum(case))
    M, N = 300, 70  # several diagonal blocks on either side
    n = M if side == 141 else N
    A = np.asfortranarray(rng.standard_normal((n, n)) / n + 2 * np.eye(n))
    B = np.asfortranarray(rng.standard_normal((M, N)))
    _say("case: dtrsm", case)
    fl = hc.FlopCounter()
    X = hc.dtrsm(side, uplo, trans, diag, -0.75, A, B, flops=fl)
    R, ofl = orc.dtrsm(side, uplo, trans, diag, -0.75, A, B)
    assert np.max(np.abs(X - R)) <= 1e-12 * np.abs(R).max()
    assert fl.trsm == ofl


def test_dtrsm_tlr_panel(hc, orc):
    """The TLR Cholesky use: the V factor of a panel tile (nb x rank)
    against the lower factor of the diagonal tile, L^{-1} V."""
    nb, r = 1024, 96
    L = np.linalg.cholesky(_spd(nb, 9))
    V = np.asfortranarray(np.random.default_rng(9).standard_normal((nb, r)))
    _say("case: dtrsm panel nb", nb, "rank", r)
    X = hc.dtrsm(141, 122, 111, 131, 1.0, L, V)
    R, _ = orc.dtrsm(141, 122, 111, 131, 1.0, L, V)
    assert np.max(np.abs(X - R)) <= 1e-12 * np.abs(R).max()


@pytest.mark.parametrize("transA", [111, 112])
@pytest.mark.parametrize("M,r", [(64, 5), (500, 32), (1024, 100)])
def test_dgemmbdcd_vs_oracle(hc, orc, transA, M, r):
    rng = np.random.default_rng(M + r + transA)
    AU, AV = rng.standard_normal((M, r)), rng.standard_normal((M, r))
    BD, CD = rng.standard_normal((M, M)), rng.standard_normal((M, M))
    _say("case: dgemmbdcd", transA, M, r)
    fl = hc.FlopCounter()
    out = hc.dgemmbdcd(transA, -1.0, AU, AV, BD, 0.5, CD, flops=fl)
    ref, ofl = orc.dgemmbdcd(transA, -1.0, AU, AV, BD, 0.5, CD)
    assert np.max(np.abs(out - ref)) <= 1e-13 * np.abs(ref).max() * np.sqrt(M)
    assert fl.update == ofl


@pytest.mark.parametrize("M,r", [(64, 5), (1024, 100)])
def test_duncompress_vs_oracle(hc, orc, M, r):
    rng = np.random.default_rng(M + r)
    AU, BV = rng.standard_normal((M, r)), rng.standard_normal((M, r))
    _say("case: duncompress", M, r)
    out = hc.duncompress(AU, BV)
    ref = orc.duncompress(AU, BV)
    assert np.max(np.abs(out - ref)) <= 1e-13 * np.abs(ref).max() * np.sqrt(r)


# ---- the two-phase split API (HCORE_dgemm_qr_svd / _b_dense + HCORE_dgemm_ormqr)
@pytest.mark.parametrize("M,ra,rb,rc", [(64, 8, 8, 8), (500, 32, 32, 32), (1024, 60, 50, 100)])
def test_split_matches_dgemm(hc, orc, M, ra, rb, rc):
    """Phase 1 reveals the rank (the oracle's), phase 2 writes factors equal
    to one HCORE_dgemm call's bit for bit; the caller's buffers may change
    between the phases; the flop counts add up to HCORE_dgemm's."""
    from test_oracle import _random_case
    _say("case: split API", M, ra, rb, rc)
    AU, AV, BU, BV, CU, CV = _random_case(101 + M, M, ra, rb, rc)
    fl1, fl2 = hc.FlopCounter(), hc.FlopCounter()
    ref = hc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=M, acc=1e-8, flops=fl1)
    seen = []
    got = hc.dgemm_split(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=M, acc=1e-8, flops=fl2,
                         between=seen.append)
    st, oU, oV, ro, _, _ = orc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=M, acc=1e-8)
    assert seen == [ref[2]] and got[2] == ref[2] and abs(got[2] - ro) <= 1
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    assert fl2.update == fl1.update


def test_split_b_dense(hc):
    """The dense-B variant: G = B AV with B = BU BV^T given densely.  Pinned
    by the mathematics (no restatement of this variant in the oracle): the
    product within 10 acc of the dense update, the rank by the truncation
    rule on numpy's singular values of the update, CU orthonormal."""
    from test_oracle import _random_case
    from conftest import trunc_rank, rel_fro
    M, ra, rb, rc = 400, 24, 30, 20
    _say("case: split API, dense B")
    AU, AV, BU, BV, CU, CV = _random_case(111, M, ra, rb, rc)
    Bd = BU @ BV.T
    got = hc.dgemm_split(-1.0, AU, AV, None, None, 1.0, CU, CV, maxrk=M, acc=1e-8, B_dense=Bd)
    E = CU @ CV.T - (AU @ AV.T) @ Bd.T
    s_np = np.linalg.svd(E, compute_uv=False)
    assert abs(got[2] - trunc_rank(s_np[:ra + rc], 1e-8)) <= 1
    assert rel_fro(got[0] @ got[1].T, E) <= 10 * 1e-8
    assert np.max(np.abs(got[0].T @ got[0] - np.eye(got[2]))) < 1e-10


def test_split_storage_maxrank(hc):
    """Phase 2 rejects a revealed rank above storage_maxrank
    (hcore_dgemm.c:1648-1651) and the state is released."""
    from test_oracle import _random_case
    _say("case: split API storage maxrank")
    hc.set_error_policy(hc.ON_ERROR_RETURN)
    try:
        AU, AV, BU, BV, CU, CV = _random_case(121, 200, 16, 16, 16)
        got = hc.dgemm_split(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=200, acc=1e-8,
                             storage_maxrank=2)
        assert got[2] == -hc.ERR_R_GT_MAXRK
    finally:
        hc.set_error_policy(hc.ON_ERROR_EXIT)
