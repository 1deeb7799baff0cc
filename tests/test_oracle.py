"""Pins of the CPU oracle (oracle/hcore_oracle.c) — no GPU.

1. The reference's own tests, reproduced on the reference's own inputs:
   testing_dgemm.c:250-392 (dlatms, iseed {0,0,0,1}; ||.||_inf < 10 acc),
   testing_sgemm.c (FP32 twin), testing_dsyrk.c:104-181 and
   testing_dgemmcd.c:105-205 (glibc rand(), seed 1; ||.||_inf < 1000 eps).
2. What the mathematics fixes for the recompression: the core singular values
   are those of the dense update beta*C + alpha*A*B^T (Q_A, Q_B orthonormal),
   the rank follows the truncation rule on them, the truncated product is the
   Eckart-Young optimum, CU is orthonormal and CV carries sigma.
3. Planted mistakes in the oracle that these pins must catch.
"""
import numpy as np
import pytest

from conftest import glibc_rand_stream, rel_fro, trunc_rank


def _dense(U, V):
    return U @ V.T


# ---------------------------------------------------------------- reference tests
def test_reference_testing_dgemm(orc):
    """testing_dgemm.c:250-392 with the oracle in place of HCORE_dgemm."""
    M, smr, acc = 10, 6, 1e-4
    iseed = np.array([0, 0, 0, 1], dtype=np.int32)
    A, AU, AV, ra = orc.generate_compress(M, M, M, smr, M, acc, iseed)
    B, BU, BV, rb = orc.generate_compress(M, M, M, smr, M, acc, iseed)
    Cd, CU, CV, rc = orc.generate_compress(M, M, M, smr, M, acc, iseed)
    ref = Cd + A @ B.T
    st, CUn, CVn, r, sig, fl = orc.dgemm(1.0, AU[:, :ra], AV[:M, :ra], BU[:, :rb], BV[:M, :rb],
                                         1.0, CU[:, :rc], CV[:M, :rc], rk=0, maxrk=M, acc=acc)
    assert st == orc.OK
    rnorm = np.abs(ref - CUn @ CVn.T).sum(axis=1).max()
    assert rnorm < 10 * acc
    assert 1 <= r <= ra + rc


def test_reference_testing_sgemm(orc):
    """testing_sgemm.c: FP32 generator (slatms) and storage_maxrank 10."""
    M, smr, acc = 10, 10, np.float32(1e-4)
    iseed = np.array([0, 0, 0, 1], dtype=np.int32)
    f32 = np.float32
    A, AU, AV, ra = orc.generate_compress(M, M, M, smr, M, acc, iseed, dtype=f32)
    B, BU, BV, rb = orc.generate_compress(M, M, M, smr, M, acc, iseed, dtype=f32)
    Cd, CU, CV, rc = orc.generate_compress(M, M, M, smr, M, acc, iseed, dtype=f32)
    ref = (Cd + A @ B.T).astype(np.float64)
    st, CUn, CVn, r, sig, fl = orc.dgemm(1.0, AU[:, :ra], AV[:M, :ra], BU[:, :rb], BV[:M, :rb],
                                         1.0, CU[:, :rc], CV[:M, :rc], rk=0, maxrk=M, acc=acc,
                                         dtype=f32)
    assert st == orc.OK
    rnorm = np.abs(ref - CUn.astype(np.float64) @ CVn.astype(np.float64).T).sum(axis=1).max()
    assert rnorm < 10 * acc


def reference_dsyrk_inputs(nrows=10, rank=7):
    rnd = glibc_rand_stream()
    C = np.zeros((nrows, nrows), order="F")
    for i in range(nrows):
        for j in range(i + 1):
            C[i, j] = rnd()
    AU = np.zeros((nrows, rank), order="F")
    AV = np.zeros((nrows, rank), order="F")
    for i in range(nrows):
        for j in range(rank):
            AU[i, j] = rnd()
            AV[i, j] = rnd()
    return C, AU, AV


def reference_dgemmcd_inputs(nrows=10, ra=7, rb=6):
    rnd = glibc_rand_stream()
    C = np.zeros((nrows, nrows), order="F")
    for i in range(nrows):
        for j in range(i + 1):
            C[i, j] = rnd()
    AU = np.zeros((nrows, ra), order="F")
    AV = np.zeros((nrows, ra), order="F")
    BU = np.zeros((nrows, rb), order="F")
    BV = np.zeros((nrows, rb), order="F")
    for i in range(nrows):
        for j in range(ra):
            AU[i, j] = rnd()
            AV[i, j] = rnd()
        for j in range(rb):
            BU[i, j] = rnd()
            BV[i, j] = rnd()
    return C, AU, AV, BU, BV


EPS = np.finfo(np.float64).eps


def test_reference_testing_dsyrk(orc):
    """testing_dsyrk.c:104-181: lower triangle vs C + A A^T, < 1000 eps."""
    C, AU, AV = reference_dsyrk_inputs()
    A = AU @ AV.T
    ref = C + np.tril(A @ A.T)
    out, fl = orc.dsyrk(1.0, AU, AV, C)
    assert np.abs(np.tril(ref - out)).sum(axis=1).max() < 1000 * EPS
    assert np.all(np.triu(out, 1) == 0.0)  # hcore_dsyrk.c:174-181
    assert fl == 2 * 7 * 7 * 10 + 2 * 10 * 7 * 7 + 2 * 10 * 10 * 7


def test_reference_testing_dgemmcd(orc):
    """testing_dgemmcd.c:105-205: C - A B^T, < 1000 eps (both associations)."""
    for ra, rb in ((7, 6), (6, 7)):
        C, AU, AV, BU, BV = reference_dgemmcd_inputs(10, ra, rb)
        ref = C - (AU @ AV.T) @ (BU @ BV.T).T
        out = orc.dgemmcd(AU, AV, BU, BV, C)
        assert np.abs(ref - out).sum(axis=1).max() < 1000 * EPS


# ------------------------------------------------------------------ math pins
def _random_case(seed, M, ra, rb, rc, decay=8.0):
    from hcore_amd.tiles import lowrank_tile
    AU, AV = lowrank_tile(M, ra, decay, ("orc", seed, "A"))
    BU, BV = lowrank_tile(M, rb, decay, ("orc", seed, "B"))
    CU, CV = lowrank_tile(M, rc, decay, ("orc", seed, "C"))
    return AU, AV, BU, BV, CU, CV


def check_recompression(AU, AV, BU, BV, CU, CV, alpha, beta, CUn, CVn, r, sig, acc, rk=0,
                        sigma_tol=1e-12):
    """The mathematical pins of one recompression result."""
    E = beta * _dense(CU, CV) + alpha * _dense(AU, AV) @ _dense(BU, BV).T
    s_np = np.linalg.svd(E, compute_uv=False)
    K = AU.shape[1] + CU.shape[1]
    s0 = s_np[0]
    # core singular values = singular values of the dense update (first K)
    assert np.max(np.abs(np.sort(sig)[::-1] - s_np[:K])) <= sigma_tol * s0
    # rank = truncation rule on the (independent) numpy singular values
    assert r == trunc_rank(s_np[:K], acc, rk)
    # Eckart-Young: the error equals the dropped tail
    tail = np.sqrt(np.sum(s_np[r:] ** 2))
    err = np.linalg.norm(CUn @ CVn.T - E)
    assert err <= tail * (1 + 1e-6) + 1e-12 * s0
    # output convention: CU orthonormal, CV columns carry sigma
    assert np.max(np.abs(CUn.T @ CUn - np.eye(r))) < 1e-10
    assert np.max(np.abs(np.linalg.norm(CVn, axis=0) - s_np[:r])) < 1e-10 * s0


CASES = [  # (seed, M, ra, rb, rc, alpha, beta, rk, acc)
    (1, 64, 8, 8, 8, -1.0, 1.0, 0, 1e-8),
    (2, 100, 12, 20, 7, 1.0, 1.0, 0, 1e-6),
    (3, 128, 30, 25, 40, -1.0, 1.0, 0, 1e-8),
    (4, 200, 16, 16, 16, 0.5, 2.0, 0, 1e-8),
    (5, 96, 10, 5, 12, -1.0, 1.0, 9, 1e-8),
    (6, 500, 32, 32, 32, -1.0, 1.0, 0, 1e-8),
]


@pytest.mark.parametrize("case", CASES, ids=[str(c[0]) for c in CASES])
def test_oracle_math_pins(orc, case):
    seed, M, ra, rb, rc, alpha, beta, rk, acc = case
    AU, AV, BU, BV, CU, CV = _random_case(seed, M, ra, rb, rc)
    st, CUn, CVn, r, sig, fl = orc.dgemm(alpha, AU, AV, BU, BV, beta, CU, CV, rk=rk, maxrk=M,
                                         acc=acc)
    assert st == orc.OK
    check_recompression(AU, AV, BU, BV, CU, CV, alpha, beta, CUn, CVn, r, sig, acc, rk)


def test_oracle_brute_force_tiny(orc):
    """nb = 6: the update is compared with a dense truncated SVD directly."""
    rng = np.random.default_rng(7)
    M = 6
    AU, AV, BU, BV = (rng.standard_normal((M, 2)) for _ in range(4))
    CU, CV = rng.standard_normal((M, 3)), rng.standard_normal((M, 3))
    E = CU @ CV.T - (AU @ AV.T) @ (BU @ BV.T).T
    U, s, Vt = np.linalg.svd(E)
    st, CUn, CVn, r, sig, fl = orc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, rk=0, maxrk=M,
                                         acc=1e-9)
    assert st == orc.OK and r == 5
    assert np.allclose(CUn @ CVn.T, (U[:, :r] * s[:r]) @ Vt[:r], atol=1e-12)


def test_oracle_error_paths(orc):
    AU, AV, BU, BV, CU, CV = _random_case(9, 32, 4, 4, 4)
    st = orc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU[:, :0], CV[:, :0], maxrk=32)[0]
    assert st == orc.ERR_ZERO_RANK
    st = orc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=3)[0]
    assert st == orc.ERR_SUM_RANKS
    st = orc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=6)[0]
    assert st == orc.ERR_K_GT_MAXRK


def test_oracle_flop_count_worked_value(orc):
    """The 500x500 rank-32 update is quoted at about 23.5 MFLOP at r = 32."""
    M, ra, rb, rc, r = 500, 32, 32, 32, 32
    K = ra + rc
    f = (2 * orc.flop_counts('q', M, K) + orc.flop_counts('m', ra, rb, M) +
         orc.flop_counts('m', M, ra, rb) + orc.flop_counts('s', K) + r * K +
         orc.flop_counts('o', M, r, K, 0 + 1) + orc.flop_counts('o', r, M, K, 2))
    assert abs(f / 1e6 - 23.5) < 0.1
    assert orc.flop_counts('m', 3, 4, 5) == 120
    assert orc.flop_counts('s', 10) == 22000


def test_flop_counts_hand_values(orc):
    """flop_counts.c:5-51 evaluated by hand, including its float-precision
    thirds, for the oracle and for the library (host code, no GPU):
    q(1024, 1000): 2*1024*1000^2 = 2048000000; 2*1000^3 = 2e9 is exact in
      float, 2e9/3.0f rounds to the float 666666688 (spacing 64 there, not
      the exact 666666666); 2*1024*1000 = 2048000; 17000/3.0f -> 5666;
      2048000000 - 666666688 + 2048000 + 5666 = 1383386978.
    q(1024, 187): 71616512 - 4359468 (13078406/3.0f = 4359468.5) + 382976
      + 1059 = 67641079.   q(500, 64): 4096000 - 174762 + 64000 + 362 = 3985600.
    c(70): 343000/3 = 114333 (integer), - 4900/2.0, + 70/6 = 11 -> 111894.
    o(10, 7, 5, 1): 4*7*10*5 - 2*7*25 + 3*7*5 = 1400 - 350 + 105 = 1155.
    o(10, 7, 5, 2): 1400 - 2*10*25 + 2*10*5 + 7*5 - 25/2 + 5/2
      = 1400 - 500 + 100 + 35 - 12 + 2 = 1025 (integer halves).
    t(3, 4, 1) = 4*3*3 = 36, t(3, 4, 2) = 3*4*4 = 48; m(3, 4, 5) = 120;
    s(10) = 22000."""
    import hcore_amd
    cases = [(("q", 1024, 1000), 1383386978), (("q", 1024, 187), 67641079),
             (("q", 500, 64), 3985600), (("c", 70), 111894), (("o", 10, 7, 5, 1), 1155),
             (("o", 10, 7, 5, 2), 1025), (("t", 3, 4, 1), 36), (("t", 3, 4, 2), 48),
             (("m", 3, 4, 5), 120), (("s", 10), 22000)]
    for args, want in cases:
        assert orc.flop_counts(*args) == want, args
        assert hcore_amd.flop_counts(*args) == want, args


# --------------------------------------------------------------- planted mistakes
def _pins_fail(orc, mutate, case=(3, 128, 24, 24, 40, -1.0, 1.0, 0, 1e-8)):
    seed, M, ra, rb, rc, alpha, beta, rk, acc = case
    AU, AV, BU, BV, CU, CV = _random_case(seed, M, ra, rb, rc)
    st, CUn, CVn, r, sig, fl = orc.dgemm(alpha, AU, AV, BU, BV, beta, CU, CV, rk=rk, maxrk=M,
                                         acc=acc, mutate=mutate)
    try:
        check_recompression(AU, AV, BU, BV, CU, CV, alpha, beta, CUn, CVn, r, sig, acc, rk)
    except AssertionError:
        return True
    return False


@pytest.mark.parametrize("mut", ["MUTATE_ALPHA_SIGN", "MUTATE_NO_G_TRANSPOSE",
                                 "MUTATE_TRUNC_OFF_BY_ONE", "MUTATE_SIGMA_TO_U"])
def test_planted_mistakes_are_caught(orc, mut):
    assert not _pins_fail(orc, orc.MUTATE_NONE)
    assert _pins_fail(orc, getattr(orc, mut))


def test_oracle_dsyrk_dgemmcd_closed_forms(orc):
    rng = np.random.default_rng(3)
    M, K = 70, 9
    AU, AV = rng.standard_normal((M, K)), rng.standard_normal((M, K))
    CD = rng.standard_normal((M, M))
    out, _ = orc.dsyrk(-1.0, AU, AV, CD)
    ref = CD - AU @ (AV.T @ AV) @ AU.T
    assert rel_fro(np.tril(out), np.tril(ref)) < 1e-13
    assert np.all(np.triu(out, 1) == 0)
    Bu, Bv = rng.standard_normal((M, 5)), rng.standard_normal((M, 5))
    out = orc.dgemmcd(AU, AV, Bu, Bv, CD)
    assert rel_fro(out, CD - (AU @ AV.T) @ (Bu @ Bv.T).T) < 1e-13


# ---- dense neighbours: potrf, trsm, dgemmbdcd, duncompress -------------------
def _spd(n, seed):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n))
    return np.asfortranarray(G @ G.T + n * np.eye(n))


@pytest.mark.parametrize("uplo", [121, 122])
def test_oracle_dpotrf_pins(orc, uplo):
    """The factor reproduces A (L L^T or U^T U) to 1e-13 and equals numpy's
    Cholesky (the unique factor with a positive diagonal); the other
    triangle is untouched; a matrix that is not positive definite reports
    the order of the failing minor; flops->potrf gets the 'c' count."""
    A = _spd(70, uplo)
    F, info, fl = orc.dpotrf(uplo, A)
    assert info == 0
    L = np.linalg.cholesky(A)
    if uplo == 122:
        assert np.max(np.abs(np.tril(F) - L)) <= 1e-12 * np.abs(L).max()
        assert np.array_equal(np.triu(F, 1), np.triu(A, 1))
    else:
        assert np.max(np.abs(np.triu(F) - L.T)) <= 1e-12 * np.abs(L).max()
        assert np.array_equal(np.tril(F, -1), np.tril(A, -1))
    assert fl == orc.flop_counts('c', 70)
    B = A.copy(order="F")
    B[40, 40] = -1e6
    _, info, _ = orc.dpotrf(uplo, B)
    assert info == 41


@pytest.mark.parametrize("side", [141, 142])
@pytest.mark.parametrize("uplo", [121, 122])
@pytest.mark.parametrize("trans", [111, 112])
@pytest.mark.parametrize("diag", [131, 132])
def test_oracle_dtrsm_pins(orc, side, uplo, trans, diag):
    """X solves op(A) X = alpha B (left) or X op(A) = alpha B (right) for the
    stored triangle of A (unit diagonal when diag = Unit): residual at the
    rounding level; flops->trsm gets the 't' count of the side."""
    rng = np.random.default_rng(side + uplo + trans + diag)
    M, N = 37, 11
    n = M if side == 141 else N
    A = np.asfortranarray(rng.standard_normal((n, n)) / n + 4 * np.eye(n))
    B = np.asfortranarray(rng.standard_normal((M, N)))
    X, fl = orc.dtrsm(side, uplo, trans, diag, 0.5, A, B)
    T = np.triu(A) if uplo == 121 else np.tril(A)
    if diag == 132:
        T = T - np.diag(np.diag(T)) + np.eye(n)
    opT = T.T if trans == 112 else T
    R = opT @ X if side == 141 else X @ opT
    assert np.max(np.abs(R - 0.5 * B)) <= 1e-13 * np.abs(B).max()
    # numpy's general solve with the explicit triangle (zeros elsewhere)
    ref = np.linalg.solve(opT, 0.5 * B) if side == 141 else np.linalg.solve(opT.T, 0.5 * B.T).T
    assert np.max(np.abs(X - ref)) <= 1e-12 * np.abs(ref).max()
    assert fl == orc.flop_counts('t', M, N, 1 if side == 141 else 2)


@pytest.mark.parametrize("transA", [111, 112])
def test_oracle_dgemmbdcd_uncompress_pins(orc, transA):
    """dgemmbdcd: CD = alpha AU AV^T BD + beta CD (closed form, both of the
    reference's branches at ld = M); duncompress: AU BV^T."""
    rng = np.random.default_rng(transA)
    M, N, r = 40, 40, 7
    AU, AV = rng.standard_normal((M, r)), rng.standard_normal((M, r))
    BD, CD = rng.standard_normal((M, N)), rng.standard_normal((M, N))
    out, fl = orc.dgemmbdcd(transA, -1.0, AU, AV, BD, 0.5, CD)
    ref = -1.0 * AU @ (AV.T @ BD) + 0.5 * CD
    assert np.max(np.abs(out - ref)) <= 1e-13 * np.abs(ref).max()
    assert fl == orc.flop_counts('m', r, N, M) + orc.flop_counts('m', M, N, r)
    BV = rng.standard_normal((N, r))
    assert np.max(np.abs(orc.duncompress(AU, BV) - AU @ BV.T)) <= 1e-13 * np.abs(AU @ BV.T).max()
