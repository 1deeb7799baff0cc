"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Bar (stated per test): the recompressed product within 10*acc relative
Frobenius error of the dense FP64 update, truncated rank within +-1 of the
oracle's (exact in practice), core singular values equal to the oracle's to
1e-12 * sigma_0 absolute; dense outputs (dsyrk, dgemmcd) within the
reference's own 1000*eps bound.
"""
import numpy as np
import pytest

from conftest import rel_fro, trunc_rank
from test_oracle import (check_recompression, reference_dgemmcd_inputs, reference_dsyrk_inputs,
                         _random_case)

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _say(*a):
    print(*a, flush=True)


def test_reference_testing_dgemm_gpu(hc, orc):
    """testing_dgemm.c:250-392 on the GPU: ||ref - CU CV^T||_inf < 10 acc and
    the oracle's rank."""
    _say("case: reference testing_dgemm nb=10")
    M, smr, acc = 10, 6, 1e-4
    iseed = np.array([0, 0, 0, 1], dtype=np.int32)
    A, AU, AV, ra = orc.generate_compress(M, M, M, smr, M, acc, iseed)
    B, BU, BV, rb = orc.generate_compress(M, M, M, smr, M, acc, iseed)
    Cd, CU, CV, rc = orc.generate_compress(M, M, M, smr, M, acc, iseed)
    args = (AU[:, :ra], AV[:M, :ra], BU[:, :rb], BV[:M, :rb], 1.0, CU[:, :rc], CV[:M, :rc])
    CUn, CVn, r = hc.dgemm(1.0, *args, rk=0, maxrk=M, acc=acc)
    assert hc.last_status() == 0
    ref = Cd + A @ B.T
    assert np.abs(ref - CUn @ CVn.T).sum(axis=1).max() < 10 * acc
    st, oU, oV, ro, sig, fl = orc.dgemm(1.0, *args, rk=0, maxrk=M, acc=acc)
    assert r == ro


def test_reference_testing_sgemm_gpu(hc, orc):
    """testing_sgemm.c on the GPU (FP32 throughout): < 10 acc."""
    _say("case: reference testing_sgemm nb=10")
    M, smr = 10, 10
    acc = np.float32(1e-4)
    f32 = np.float32
    iseed = np.array([0, 0, 0, 1], dtype=np.int32)
    A, AU, AV, ra = orc.generate_compress(M, M, M, smr, M, acc, iseed, dtype=f32)
    B, BU, BV, rb = orc.generate_compress(M, M, M, smr, M, acc, iseed, dtype=f32)
    Cd, CU, CV, rc = orc.generate_compress(M, M, M, smr, M, acc, iseed, dtype=f32)
    args = (AU[:, :ra], AV[:M, :ra], BU[:, :rb], BV[:M, :rb], 1.0, CU[:, :rc], CV[:M, :rc])
    CUn, CVn, r = hc.dgemm(1.0, *args, rk=0, maxrk=M, acc=float(acc), dtype=f32)
    ref = (Cd + A @ B.T).astype(np.float64)
    rn = np.abs(ref - CUn.astype(np.float64) @ CVn.astype(np.float64).T).sum(axis=1).max()
    assert rn < 10 * acc
    st, oU, oV, ro, sig, fl = orc.dgemm(1.0, *args, rk=0, maxrk=M, acc=acc, dtype=f32)
    assert abs(r - ro) <= 1


GPU_CASES = [  # (seed, M, ra, rb, rc, alpha, beta, rk, acc)
    (11, 16, 3, 2, 4, -1.0, 1.0, 0, 1e-8),
    (12, 64, 8, 8, 8, -1.0, 1.0, 0, 1e-8),
    (13, 100, 12, 20, 7, 1.0, 1.0, 0, 1e-6),
    (14, 128, 30, 25, 40, -1.0, 1.0, 0, 1e-8),
    (15, 200, 16, 16, 16, 0.5, 2.0, 0, 1e-8),
    (16, 96, 10, 5, 12, -1.0, 1.0, 9, 1e-8),
    (17, 257, 9, 33, 21, -1.0, 1.0, 0, 1e-8),
    (18, 500, 32, 32, 32, -1.0, 1.0, 0, 1e-8),
    (19, 1024, 80, 40, 107, -1.0, 1.0, 0, 1e-8),
    (20, 1024, 70, 60, 90, -1.0, 1.0, 0, 1e-8),   # K = 160: Jacobi outside LDS
    (21, 2048, 24, 24, 40, -1.0, 1.0, 0, 1e-8),
    # register QR shapes with several rows per thread: K = 16 at 4 rows (nb 1024,
    # and nb 700 with the last rows masked), K = 30 and 32 at 2 rows
    (22, 1024, 8, 8, 8, -1.0, 1.0, 0, 1e-8),
    (23, 700, 6, 9, 10, -1.0, 1.0, 0, 1e-8),
    (24, 1024, 12, 20, 18, -1.0, 1.0, 0, 1e-8),
    (25, 600, 20, 10, 12, 1.0, 1.0, 0, 1e-8),
]


@pytest.mark.parametrize("case", GPU_CASES, ids=["s%d_nb%d" % (c[0], c[1]) for c in GPU_CASES])
def test_dgemm_vs_oracle(hc, orc, case):
    seed, M, ra, rb, rc, alpha, beta, rk, acc = case
    _say("case: dgemm", case)
    AU, AV, BU, BV, CU, CV = _random_case(seed, M, ra, rb, rc)
    fl = hc.FlopCounter()
    CUn, CVn, r = hc.dgemm(alpha, AU, AV, BU, BV, beta, CU, CV, rk=rk, maxrk=M, acc=acc,
                           flops=fl)
    assert hc.last_status() == 0
    st, oU, oV, ro, sig, ofl = orc.dgemm(alpha, AU, AV, BU, BV, beta, CU, CV, rk=rk, maxrk=M,
                                         acc=acc)
    E = beta * CU @ CV.T + alpha * (AU @ AV.T) @ (BU @ BV.T).T
    err = rel_fro(CUn @ CVn.T, E)
    _say("   rank gpu %d oracle %d  relerr gpu %.3e oracle %.3e" %
         (r, ro, err, rel_fro(oU @ oV.T, E)))
    assert abs(r - ro) <= 1
    if rk == 0:
        assert err <= 10 * acc
    else:  # fixed rank: the error is the dropped tail, the same as the oracle's
        assert err <= rel_fro(oU @ oV.T, E) * (1 + 1e-6) + 10 * acc
    assert np.max(np.abs(CUn.T @ CUn - np.eye(r))) < 1e-10
    if r == ro:
        assert fl.update == ofl  # same closed-form accounting as the reference
        assert rel_fro(CUn @ CVn.T, oU @ oV.T) <= 10 * acc


def colmajor(hc, x, cap=None):
    """Device copy of x (M x r, column-major) with room for cap columns."""
    return hc.DeviceBuffer.from_numpy(np.asfortranarray(x), cap)


def test_plan_sigma_and_rank_match_oracle(hc, orc):
    """One update through the plan API: core singular values equal to the
    oracle's within 1e-12 sigma_0, same rank, all math pins."""
    for (seed, M, ra, rb, rc) in [(31, 500, 32, 32, 32), (32, 256, 40, 20, 60),
                                  (33, 1024, 60, 50, 100)]:
        _say("case: plan sigma", seed, M, ra, rb, rc)
        AU, AV, BU, BV, CU, CV = _random_case(seed, M, ra, rb, rc)
        K = ra + rc
        dA = [colmajor(hc, x) for x in (AU, AV, BU, BV)]
        cu, cv = colmajor(hc, CU, K), colmajor(hc, CV, K)
        p = hc.Plan(M)
        p.add_dgemm([dA[0].ptr], [dA[1].ptr], [ra], M, [dA[2].ptr], [dA[3].ptr], [rb], M,
                    [cu.ptr], [cv.ptr], [rc], M, -1.0, 1.0, 0, M, 1e-8)
        p.finalize()
        p.execute()
        hc.device_synchronize()
        r = int(p.ranks()[0])
        assert p.status()[0] == 0
        sig = p.sigma(0)
        st, oU, oV, ro, osig, _ = orc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=M, acc=1e-8)
        assert np.max(np.abs(sig - osig)) <= 1e-12 * osig[0]
        assert r == ro
        CUn, CVn = cu.numpy(r), cv.numpy(r)
        check_recompression(AU, AV, BU, BV, CU, CV, -1.0, 1.0, CUn, CVn, r, sig, 1e-8)
        p.destroy()


def _trailing_plan(hc, step, out_of_place):
    """Device buffers + plan for a TrailingStep (gemm updates and dsyrk)."""
    nb = step.nb
    ra, rb, rc = step.ranks()
    K = ra + rc
    panel = {i: (colmajor(hc, u), colmajor(hc, v)) for i, (u, v) in step.panel.items()}
    CUd = [colmajor(hc, c[0], k) for c, k in zip(step.C, K)]
    CVd = [colmajor(hc, c[1], k) for c, k in zip(step.C, K)]
    outU = [hc.DeviceBuffer(nb, k) for k in K] if out_of_place else None
    outV = [hc.DeviceBuffer(nb, k) for k in K] if out_of_place else None
    p = hc.Plan(nb)
    p.add_dgemm([panel[i][0].ptr for i, j in step.pairs], [panel[i][1].ptr for i, j in step.pairs],
                ra, nb, [panel[j][0].ptr for i, j in step.pairs],
                [panel[j][1].ptr for i, j in step.pairs], rb, nb,
                [t.ptr for t in CUd], [t.ptr for t in CVd], rc, nb, -1.0, 1.0, 0, 512, 1e-8,
                CUout=[t.ptr for t in outU] if outU else None,
                CVout=[t.ptr for t in outV] if outV else None, ldco=nb)
    Dd = [colmajor(hc, D) for D in step.D]
    if step.D:
        p.add_dsyrk([panel[i][0].ptr for i in step.diag_idx],
                    [panel[i][1].ptr for i in step.diag_idx],
                    [step.panel[i][0].shape[1] for i in step.diag_idx], nb,
                    [t.ptr for t in Dd], nb, -1.0)
    p.finalize()
    keep = (panel, CUd, CVd, outU, outV, Dd)
    return p, keep


@pytest.mark.parametrize("grid", [8, 13])
def test_trailing_step_batch_vs_oracle(hc, orc, grid):
    """A small TLR Cholesky trailing step (grid x grid tiles, nb=256) in one
    plan: every update within 10*acc of the dense update and the oracle's
    rank; every diagonal tile equal to the oracle's dsyrk to 1e-12 relative.
    grid 8 (21 updates) runs as one lane; grid 13 (66 updates, stacked ranks
    up to 187) splits into the three rank lanes, the largest with the panel
    QR and the register-resident core."""
    from hcore_amd.tiles import TrailingStep
    step = TrailingStep(grid=grid, nb=256, key="gpu-test")
    _say("case: trailing step grid", grid, "nb 256, updates", len(step.pairs))
    p, keep = _trailing_plan(hc, step, out_of_place=False)
    panel, CUd, CVd, _, _, Dd = keep
    p.execute()
    hc.device_synchronize()
    ranks = p.ranks()
    assert np.all(p.status() == 0)
    ra, rb, rc = step.ranks()
    lanes = p.lanes()
    if len(step.pairs) > 64:
        assert len(set(lanes.tolist())) == 3 and np.all(lanes[ra + rc > 128] == 0)
    for u, (i, j) in enumerate(step.pairs):
        AU, AV = step.panel[i]
        BU, BV = step.panel[j]
        CU, CV = step.C[u]
        r = int(ranks[u])
        st, oU, oV, ro, osig, _ = orc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=512,
                                            acc=1e-8)
        E = CU @ CV.T - (AU @ AV.T) @ (BU @ BV.T).T
        got = CUd[u].numpy(r) @ CVd[u].numpy(r).T
        assert abs(r - ro) <= 1, (u, r, ro)
        assert rel_fro(got, E) <= 10 * 1e-8, u
    for s, i in enumerate(step.diag_idx):
        AU, AV = step.panel[i]
        ref, _ = orc.dsyrk(-1.0, AU, AV, step.D[s])
        got = Dd[s].numpy()
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.abs(ref).max()
    p.destroy()


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_reduced_precision_trailing_step(hc, orc, precision):
    """The grid-13 trailing step (66 updates, three rank lanes) on FP32 tiles
    at acc = 1e-5: the mixed-precision plan (FP32 products, QR and
    reassembly; FP64 core SVD and truncation) and the all-FP32 plan (the
    batched HCORE_sgemm).  Every update within 10*acc of the dense update of
    the same FP32 tiles (evaluated in FP64) and within one of the oracle's
    rank on those tiles."""
    from hcore_amd.tiles import TrailingStep
    acc = 1e-5
    f32 = np.float32
    step = TrailingStep(grid=13, nb=256, key="mixed-test", diag=False)
    _say("case: reduced precision", precision, "grid 13 nb 256")
    nb = step.nb
    ra, rb, rc = step.ranks()
    K = ra + rc
    panel = {i: (colmajor(hc, u.astype(f32)), colmajor(hc, v.astype(f32)))
             for i, (u, v) in step.panel.items()}
    CUd = [colmajor(hc, c[0].astype(f32), k) for c, k in zip(step.C, K)]
    CVd = [colmajor(hc, c[1].astype(f32), k) for c, k in zip(step.C, K)]
    p = hc.Plan(nb, hc.MIXED if precision == "mixed" else hc.FP32)
    p.add_dgemm([panel[i][0].ptr for i, j in step.pairs], [panel[i][1].ptr for i, j in step.pairs],
                ra, nb, [panel[j][0].ptr for i, j in step.pairs],
                [panel[j][1].ptr for i, j in step.pairs], rb, nb,
                [t.ptr for t in CUd], [t.ptr for t in CVd], rc, nb, -1.0, 1.0, 0, 512, acc)
    p.finalize()
    p.execute()
    hc.device_synchronize()
    assert np.all(p.status() == 0)
    ranks = p.ranks()
    lanes = p.lanes()
    assert len(set(lanes.tolist())) == 3
    for u, (i, j) in enumerate(step.pairs):
        AU, AV = (x.astype(f32).astype(np.float64) for x in step.panel[i])
        BU, BV = (x.astype(f32).astype(np.float64) for x in step.panel[j])
        CU, CV = (x.astype(f32).astype(np.float64) for x in step.C[u])
        E = CU @ CV.T - (AU @ AV.T) @ (BU @ BV.T).T
        r = int(ranks[u])
        got = CUd[u].numpy(r).astype(np.float64) @ CVd[u].numpy(r).astype(np.float64).T
        assert rel_fro(got, E) <= 10 * acc, (u, rel_fro(got, E))
        st, _, _, ro, _, _ = orc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=512, acc=acc)
        assert abs(r - ro) <= 1, (u, r, ro)
    p.destroy()


@pytest.mark.parametrize("grid", [6, 13])
def test_graph_replays_match_eager(hc, grid):
    """The plan captured in a hipGraph and replayed several times (out of
    place; host copies and synchronisation between replays): every replay
    equals the eager result bit for bit.  grid 13 captures the fork/join of
    the three rank lanes over their streams."""
    from hcore_amd.tiles import TrailingStep
    step = TrailingStep(grid=grid, nb=256, key="graph-test", diag=False)
    _say("case: graph replay grid", grid, "nb 256")
    p, keep = _trailing_plan(hc, step, out_of_place=True)
    _, _, _, outU, outV, _ = keep
    p.execute()
    hc.device_synchronize()
    eager = [t.numpy() for t in outU + outV]
    r0 = p.ranks().copy()
    for rep in range(3):
        for t in outU + outV:
            t.zero()
        hc.device_synchronize()
        p.execute(graph=True)
        hc.device_synchronize()
        got = [t.numpy() for t in outU + outV]
        assert np.array_equal(p.ranks(), r0)
        for a, b in zip(eager, got):
            assert np.array_equal(a, b), rep
    p.destroy()


def test_reference_testing_dsyrk_gpu(hc):
    _say("case: reference testing_dsyrk")
    C, AU, AV = reference_dsyrk_inputs()
    A = AU @ AV.T
    ref = C + np.tril(A @ A.T)
    fl = hc.FlopCounter()
    out = hc.dsyrk(1.0, AU, AV, C, flops=fl)
    assert np.abs(np.tril(ref - out)).sum(axis=1).max() < 1000 * EPS
    assert np.all(np.triu(out, 1) == 0.0)
    assert fl.syrk == 2 * 7 * 7 * 10 + 2 * 10 * 7 * 7 + 2 * 10 * 10 * 7


def test_reference_testing_dgemmcd_gpu(hc):
    for ra, rb in ((7, 6), (6, 7)):
        _say("case: reference testing_dgemmcd", ra, rb)
        C, AU, AV, BU, BV = reference_dgemmcd_inputs(10, ra, rb)
        ref = C - (AU @ AV.T) @ (BU @ BV.T).T
        out = hc.dgemmcd(AU, AV, BU, BV, C)
        assert np.abs(ref - out).sum(axis=1).max() < 1000 * EPS


@pytest.mark.parametrize("M,K", [(64, 5), (300, 37), (1024, 107)])
def test_dsyrk_vs_oracle(hc, orc, M, K):
    _say("case: dsyrk", M, K)
    rng = np.random.default_rng(M + K)
    AU, AV = rng.standard_normal((M, K)), rng.standard_normal((M, K))
    CD = rng.standard_normal((M, M))
    out = hc.dsyrk(-1.0, AU, AV, CD)
    ref, _ = orc.dsyrk(-1.0, AU, AV, CD)
    scale = np.abs(ref).max()
    assert np.max(np.abs(out - ref)) <= 1000 * EPS * scale * K
    assert np.all(np.triu(out, 1) == 0.0)


@pytest.mark.parametrize("M,ra,rb", [(64, 5, 9), (300, 37, 20), (1024, 60, 107)])
def test_dgemmcd_vs_oracle(hc, orc, M, ra, rb):
    _say("case: dgemmcd", M, ra, rb)
    rng = np.random.default_rng(M + ra)
    Au, Av = rng.standard_normal((M, ra)), rng.standard_normal((M, ra))
    Bu, Bv = rng.standard_normal((M, rb)), rng.standard_normal((M, rb))
    CD = rng.standard_normal((M, M))
    out = hc.dgemmcd(Au, Av, Bu, Bv, CD)
    ref = orc.dgemmcd(Au, Av, Bu, Bv, CD)
    assert np.max(np.abs(out - ref)) <= 1000 * EPS * np.abs(ref).max() * max(ra, rb)


def test_nonfinite_input_leaves_factors_untouched(hc):
    """A NaN tile: status NUMERIC (the reference's LAPACK-failure class) and
    the caller's CU/CV are not written (reference clones, hcore_dgemm.c:971)."""
    _say("case: NaN input")
    hc.set_error_policy(hc.ON_ERROR_RETURN)
    try:
        AU, AV, BU, BV, CU, CV = _random_case(41, 64, 4, 4, 4)
        AU = AU.copy()
        AU[3, 1] = np.nan
        CUn, CVn, r = hc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=64)
        assert hc.last_status() == hc.ERR_NUMERIC
        assert r == 4  # Crk untouched
        assert np.array_equal(CUn, CU) and np.array_equal(CVn, CV)
    finally:
        hc.set_error_policy(hc.ON_ERROR_EXIT)
