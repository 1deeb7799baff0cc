"""GPU parity beyond the basic shapes: large stacked ranks (the global-memory
core SVD variants and the 16-column compact-WY solve), zero singular values,
argument handling of the drop-in and plan entry points, and the full nb=1024
trailing step (1953 HCORE_dgemm + 63 HCORE_dsyrk) in FP64 and in the mixed
form, every update against the oracle.

Bars (stated per test): products within 10*acc relative Frobenius error of
the dense FP64 update (or of the oracle's product, for the batch), truncated
rank within +-1 of the oracle's, new CU orthonormal to 1e-10.
"""
import numpy as np
import pytest

from conftest import rel_fro
from test_oracle import _random_case

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _dense(alpha, AU, AV, BU, BV, beta, CU, CV):
    return beta * CU @ CV.T + alpha * (AU @ AV.T) @ (BU @ BV.T).T


def _check_single(hc, orc, case, device=False):
    seed, M, ra, rb, rc, rk, acc = case
    AU, AV, BU, BV, CU, CV = _random_case(seed, M, ra, rb, rc)
    maxrk = max(M // 2, ra + rc)
    CUn, CVn, r = hc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, rk=rk, maxrk=maxrk, acc=acc,
                           device=device)
    assert hc.last_status() == 0
    st, oU, oV, ro, sig, _ = orc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, rk=rk, maxrk=maxrk,
                                       acc=acc)
    E = _dense(-1.0, AU, AV, BU, BV, 1.0, CU, CV)
    err, oerr = rel_fro(CUn @ CVn.T, E), rel_fro(oU @ oV.T, E)
    _say("   rank gpu %d oracle %d  relerr gpu %.3e oracle %.3e" % (r, ro, err, oerr))
    assert abs(r - ro) <= 1
    assert err <= 10 * acc
    assert np.max(np.abs(CUn.T @ CUn - np.eye(r))) < 1e-10


# stacked ranks K = ra + rc beyond 192: the global-memory core variants (K <= 256
# and K <= 512), the compact-WY solve with 64 columns per workgroup up to K = 256
# and with 16 beyond it
LARGE_K = [  # (seed, M, ra, rb, rc, rk, acc)
    (51, 1024, 100, 90, 93, 0, 1e-8),    # K = 193
    (52, 1024, 128, 100, 128, 0, 1e-8),  # K = 256
    (53, 1024, 140, 100, 160, 0, 1e-8),  # K = 300
    (54, 2048, 150, 120, 150, 0, 1e-8),  # K = 300
    (55, 2048, 256, 200, 256, 0, 1e-8),  # K = 512, the largest supported
]


@pytest.mark.parametrize("case", LARGE_K, ids=["K%d_nb%d" % (c[2] + c[4], c[1]) for c in LARGE_K])
def test_large_stacked_rank_single(hc, orc, case):
    _say("case: large K single call", case)
    _check_single(hc, orc, case)


# the register QR shape with two rows per thread for K <= 16 (nb 257..512)
SMALL_K = [(56, 400, 6, 9, 8, 0, 1e-8), (57, 512, 8, 8, 8, 0, 1e-8)]


@pytest.mark.parametrize("case", SMALL_K, ids=["K%d_nb%d" % (c[2] + c[4], c[1]) for c in SMALL_K])
def test_register_qr_two_rows_k16(hc, orc, case):
    _say("case: K <= 16 at nb", case[1])
    _check_single(hc, orc, case)


def test_large_stacked_rank_plan(hc, orc):
    """Four updates with K = 193, 256, 300 and 512 at nb = 2048 in one plan:
    one lane, core variants by K, the 16-column compact-WY solve for the lane."""
    M = 2048
    cases = [(61, 100, 90, 93), (62, 128, 100, 128), (63, 150, 120, 150), (64, 256, 200, 256)]
    _say("case: large K plan nb", M)
    data = [_random_case(s, M, ra, rb, rc) for s, ra, rb, rc in cases]
    dev = [[hc.DeviceBuffer.from_numpy(np.asfortranarray(x)) for x in d[:4]] for d in data]
    K = [ra + rc for _, ra, _, rc in cases]
    cu = [hc.DeviceBuffer.from_numpy(np.asfortranarray(d[4]), k) for d, k in zip(data, K)]
    cv = [hc.DeviceBuffer.from_numpy(np.asfortranarray(d[5]), k) for d, k in zip(data, K)]
    p = hc.Plan(M)
    p.add_dgemm([d[0].ptr for d in dev], [d[1].ptr for d in dev], [c[1] for c in cases], M,
                [d[2].ptr for d in dev], [d[3].ptr for d in dev], [c[2] for c in cases], M,
                [t.ptr for t in cu], [t.ptr for t in cv], [c[3] for c in cases], M,
                -1.0, 1.0, 0, 512, 1e-8)
    p.finalize()
    p.execute()
    hc.device_synchronize()
    assert np.all(p.status() == 0)
    ranks = p.ranks()
    for u, d in enumerate(data):
        AU, AV, BU, BV, CU, CV = d
        r = int(ranks[u])
        st, oU, oV, ro, osig, _ = orc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=512,
                                            acc=1e-8)
        sig = p.sigma(u)
        assert np.max(np.abs(sig - osig)) <= 1e-12 * osig[0]
        E = _dense(-1.0, AU, AV, BU, BV, 1.0, CU, CV)
        CUn, CVn = cu[u].numpy(r), cv[u].numpy(r)
        _say("   K %d rank gpu %d oracle %d relerr %.3e" % (K[u], r, ro, rel_fro(CUn @ CVn.T, E)))
        assert abs(r - ro) <= 1
        assert rel_fro(CUn @ CVn.T, E) <= 10 * 1e-8
        assert np.max(np.abs(CUn.T @ CUn - np.eye(r))) < 1e-10
    p.destroy()


def _cancelling_case(M, k):
    """C = A B^T exactly (identity columns): E = C - A B^T is exactly zero in
    floating point, so T = R_A R_B^T is exactly zero."""
    I = np.eye(M)
    AU, AV = I[:, :k].copy(order="F"), I[:, :k].copy(order="F")
    BU, BV = I[:, k:2 * k].copy(order="F"), I[:, :k].copy(order="F")
    CU, CV = AU.copy(order="F"), BU.copy(order="F")
    return AU, AV, BU, BV, CU, CV


@pytest.mark.parametrize("k,rk", [(4, 0), (4, 6), (24, 0), (24, 30), (70, 0)])
def test_zero_singular_values_orthonormal_u(hc, orc, k, rk):
    """An exactly cancelling update (all singular values zero) in accuracy
    mode (r = 1) and in fixed-rank mode (r = rk): the new CU is orthonormal,
    as dgesvd('A','A') makes it in the reference (hcore_dgemm.c:403-406), and
    the product is zero.  k = 4: LDS core; k = 24, 70: register cores."""
    M = 256
    _say("case: exact cancellation k", k, "rk", rk)
    AU, AV, BU, BV, CU, CV = _cancelling_case(M, k)
    CUn, CVn, r = hc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, rk=rk, maxrk=M, acc=1e-8)
    assert hc.last_status() == 0
    st, oU, oV, ro, sig, _ = orc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, rk=rk, maxrk=M,
                                       acc=1e-8)
    assert r == ro == (rk if rk else 1)
    assert np.max(np.abs(CUn.T @ CUn - np.eye(r))) < 1e-12
    assert np.max(np.abs(CUn @ CVn.T)) == 0.0
    assert np.max(np.abs(oU @ oV.T)) == 0.0


def test_fixed_rank_above_numerical_rank(hc, orc):
    """Fixed rank K on an update of numerical rank K/2 (C shares A's column
    space): U stays orthonormal and the product matches the oracle's."""
    M, k = 300, 12
    _say("case: fixed rank above the numerical rank")
    rng = np.random.default_rng(71)
    AU = np.asfortranarray(np.linalg.qr(rng.standard_normal((M, k)))[0])
    AV = np.asfortranarray(rng.standard_normal((M, k)))
    BU = np.asfortranarray(rng.standard_normal((M, k)))
    BV = np.asfortranarray(rng.standard_normal((M, k)))
    CU = np.asfortranarray(AU @ rng.standard_normal((k, k)))
    CV = np.asfortranarray(rng.standard_normal((M, k)))
    CUn, CVn, r = hc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, rk=2 * k, maxrk=M, acc=1e-8)
    assert hc.last_status() == 0 and r == 2 * k
    st, oU, oV, ro, _, _ = orc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, rk=2 * k, maxrk=M)
    E = _dense(-1.0, AU, AV, BU, BV, 1.0, CU, CV)
    assert ro == r
    assert np.max(np.abs(CUn.T @ CUn - np.eye(r))) < 1e-10
    assert rel_fro(CUn @ CVn.T, E) <= 1e-12


@pytest.mark.parametrize("device", [False, True])
def test_ldc_above_m(hc, orc, device):
    """LDC > M on the drop-in HCORE_dgemm: the result equals the LDC = M call
    bit for bit (host and device pointers)."""
    M, ldc = 300, 317
    _say("case: LDC", ldc, "M", M, "device", device)
    AU, AV, BU, BV, CU, CV = _random_case(72, M, 20, 15, 25)
    ref = hc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=M, acc=1e-8)
    got = hc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=M, acc=1e-8, ldc=ldc, device=device)
    assert got[2] == ref[2]
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


@pytest.mark.parametrize("M,ra,rb,rc", [(256, 20, 24, 30), (500, 32, 32, 32)])
def test_sgemm_drop_in(hc, orc, M, ra, rb, rc):
    """HCORE_sgemm through its C entry point at tile sizes of the configs
    (FP32 throughout, acc 1e-4): within 10*acc of the dense update of the
    same FP32 tiles, rank within one of the oracle's FP32 path."""
    f32 = np.float32
    acc = 1e-4
    _say("case: sgemm", M, ra, rb, rc)
    AU, AV, BU, BV, CU, CV = (x.astype(f32) for x in _random_case(73, M, ra, rb, rc))
    CUn, CVn, r = hc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=M, acc=acc, dtype=f32)
    assert hc.last_status() == 0
    st, oU, oV, ro, _, _ = orc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=M, acc=f32(acc),
                                     dtype=f32)
    d = lambda x: x.astype(np.float64)
    E = _dense(-1.0, d(AU), d(AV), d(BU), d(BV), 1.0, d(CU), d(CV))
    assert abs(r - ro) <= 1
    assert rel_fro(d(CUn) @ d(CVn).T, E) <= 10 * acc


def test_plan_argument_errors(hc):
    """Invalid plan arguments fail in add_dgemm / add_dsyrk with
    HCORE_HIP_ERR_ARG before any kernel runs: only one of CUout/CVout, a
    leading dimension below M, a dsyrk rank of 0 or above LDAU."""
    import ctypes as C
    M = 64
    _say("case: plan argument errors")
    L = hc.lib()
    bufs = [hc.DeviceBuffer(M, 8) for _ in range(8)]
    P = lambda *b: (C.c_void_p * 1)(*[x.ptr for x in b])
    I = lambda v: (C.c_int * 1)(v)
    h = C.c_void_p()
    assert L.HCORE_hip_plan_create(C.byref(h), M, hc.FP64) == 0
    a = [P(bufs[i]) for i in range(6)]

    def add(cuo, cvo, lda=M, ldc=M, ldco=M):
        return L.HCORE_hip_plan_add_dgemm(h, 1, a[0], a[1], I(4), lda, a[2], a[3], I(4), M,
                                          a[4], a[5], I(4), ldc, cuo, cvo, ldco, -1.0, 1.0,
                                          0, M, 1e-8)
    assert add(P(bufs[6]), None) == hc.ERR_ARG
    assert add(None, P(bufs[7])) == hc.ERR_ARG
    assert add(None, None, lda=M - 1) == hc.ERR_ARG
    assert add(None, None, ldc=M - 1) == hc.ERR_ARG
    assert add(P(bufs[6]), P(bufs[7]), ldco=M - 1) == hc.ERR_ARG
    sy = lambda k, ldau=M, ldcd=M: L.HCORE_hip_plan_add_dsyrk(h, 1, a[0], a[1], I(k), ldau,
                                                              P(bufs[6]), ldcd, -1.0)
    assert sy(0) == hc.ERR_ARG
    assert sy(M + 1) == hc.ERR_ARG
    assert sy(4, ldcd=M - 1) == hc.ERR_ARG
    assert sy(4, ldau=M - 1) == hc.ERR_ARG
    assert add(P(bufs[6]), P(bufs[7])) == 0  # a valid call is still accepted
    L.HCORE_hip_plan_destroy(h)
    for b in bufs:
        b.free()


# ---------------------------------------------------------------------------------
# the full trailing step at nb = 1024, every update against the oracle
# ---------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def full_step():
    from hcore_amd.tiles import TrailingStep
    return TrailingStep(grid=64, nb=1024, key="trailing")


def _factored_rel_diff(XU, XV, YU, YV):
    """||XU XV^T - YU YV^T||_F / ||YU YV^T||_F without forming nb x nb."""
    Q, R = np.linalg.qr(np.hstack([XU, -YU]))
    num = np.linalg.norm(R @ np.hstack([XV, YV]).T)
    _, R2 = np.linalg.qr(YU)
    return num / max(np.linalg.norm(R2 @ YV.T), 1e-300)


def _run_batch(hc, step, prec, acc, dt, tune=None):
    nb = step.nb
    ra, rb, rc = step.ranks()
    K = ra + rc
    B = lambda x, cap=None: hc.DeviceBuffer.from_numpy(np.asfortranarray(x.astype(dt)), cap)
    panel = {i: (B(u), B(v)) for i, (u, v) in step.panel.items()}
    CUd = [B(c[0], k) for c, k in zip(step.C, K)]
    CVd = [B(c[1], k) for c, k in zip(step.C, K)]
    OUd = [hc.DeviceBuffer(nb, k, dt) for k in K]
    OVd = [hc.DeviceBuffer(nb, k, dt) for k in K]
    Dd = [B(d) for d in step.D] if prec == hc.FP64 else []
    p = hc.Plan(nb, prec)
    p.add_dgemm([panel[i][0].ptr for i, j in step.pairs], [panel[i][1].ptr for i, j in step.pairs],
                ra, nb, [panel[j][0].ptr for i, j in step.pairs],
                [panel[j][1].ptr for i, j in step.pairs], rb, nb,
                [t.ptr for t in CUd], [t.ptr for t in CVd], rc, nb, -1.0, 1.0, 0, 512, acc,
                CUout=[t.ptr for t in OUd], CVout=[t.ptr for t in OVd])
    if Dd:
        p.add_dsyrk([panel[i][0].ptr for i in step.diag_idx],
                    [panel[i][1].ptr for i in step.diag_idx],
                    [step.panel[i][0].shape[1] for i in step.diag_idx], nb,
                    [t.ptr for t in Dd], nb, -1.0)
    p.finalize()
    if tune is not None:  # after one execution, then the graph of the tuned plan
        p.execute()
        hc.device_synchronize()
        tune(p)
    p.execute(graph=True)
    hc.device_synchronize()
    status, ranks = p.status(), p.ranks()
    outs = [(OUd[u].numpy(int(ranks[u])), OVd[u].numpy(int(ranks[u]))) for u in range(len(K))]
    D = [t.numpy() for t in Dd]
    p.destroy()
    for t in CUd + CVd + OUd + OVd + Dd + [x for pr in panel.values() for x in pr]:
        t.free()
    return status, ranks, outs, D


_ORACLE_BATCH = {}


def _oracle_batch(orc, step, acc, dt):
    key = (id(step), acc, np.dtype(dt).name)
    if key not in _ORACLE_BATCH:
        _ORACLE_BATCH[key] = _oracle_batch_run(orc, step, acc, dt)
    return _ORACLE_BATCH[key]


def _oracle_batch_run(orc, step, acc, dt):
    nb = step.nb
    ra, rb, rc = step.ranks()
    K = ra + rc
    f = lambda x: np.asfortranarray(x.astype(dt).astype(np.float64))
    AU = [f(step.panel[i][0]) for i, j in step.pairs]
    AV = [f(step.panel[i][1]) for i, j in step.pairs]
    BU = [f(step.panel[j][0]) for i, j in step.pairs]
    BV = [f(step.panel[j][1]) for i, j in step.pairs]
    CU, CV = [], []
    for (u, v), k in zip(step.C, K):
        a = np.zeros((nb, k), order="F"); a[:, :u.shape[1]] = f(u); CU.append(a)
        b = np.zeros((nb, k), order="F"); b[:, :v.shape[1]] = f(v); CV.append(b)
    rcc = rc.copy()
    bad, _ = orc.dgemm_batch(-1.0, AU, AV, BU, BV, 1.0, CU, CV, rcc, 0, 512, acc)
    assert bad == 0
    return rcc, [(CU[u][:, :rcc[u]], CV[u][:, :rcc[u]]) for u in range(len(K))]


def _compare_batch(step, ranks, outs, oranks, oouts, acc, label):
    from concurrent.futures import ThreadPoolExecutor
    drift = np.abs(ranks.astype(int) - oranks.astype(int))
    with ThreadPoolExecutor(8) as ex:
        diffs = list(ex.map(lambda u: _factored_rel_diff(outs[u][0].astype(np.float64),
                                                         outs[u][1].astype(np.float64),
                                                         *oouts[u]), range(len(outs))))
    diffs = np.array(diffs)
    _say("   %s: %d updates, max rank drift %d (%d differ), worst rel diff vs oracle %.3e" %
         (label, len(outs), drift.max(), int(np.count_nonzero(drift)), diffs.max()))
    assert drift.max() <= 1
    assert diffs.max() <= 10 * acc


@pytest.mark.timeout(900)
def test_full_trailing_step_fp64_vs_oracle(hc, orc, full_step):
    """The nb=1024 trailing step (1953 gemm + 63 syrk, three rank lanes,
    hipGraph execution): every update's product within 10*acc of the oracle's
    product and its rank within one; every diagonal tile equal to the oracle's
    dsyrk to 1e-12 relative; a sample within 10*acc of the dense update."""
    step = full_step
    _say("case: full trailing step nb 1024 FP64")
    status, ranks, outs, D = _run_batch(hc, step, hc.FP64, 1e-8, np.float64)
    assert np.all(status == 0)
    oranks, oouts = _oracle_batch(orc, step, 1e-8, np.float64)
    _compare_batch(step, ranks, outs, oranks, oouts, 1e-8, "fp64")
    for s, i in enumerate(step.diag_idx):
        ref, _ = orc.dsyrk(-1.0, step.panel[i][0], step.panel[i][1], step.D[s])
        assert np.max(np.abs(D[s] - ref)) <= 1e-12 * np.abs(ref).max()
    for u in range(0, len(step.pairs), 97):
        i, j = step.pairs[u]
        E = _dense(-1.0, *step.panel[i], *step.panel[j], 1.0, *step.C[u])
        assert rel_fro(outs[u][0] @ outs[u][1].T, E) <= 10 * 1e-8


@pytest.mark.timeout(900)
def test_full_trailing_step_mixed_acc1e6_vs_oracle(hc, orc, full_step):
    """The mixed-precision plan (FP32 tiles, products, QR and reassembly;
    FP64 core SVD and truncation) on the 1953-update gemm batch at acc 1e-6:
    every product within 10*acc of the oracle's on the same FP32-rounded
    tiles, ranks within one."""
    step = full_step
    _say("case: full trailing step nb 1024 mixed acc 1e-6")
    status, ranks, outs, _ = _run_batch(hc, step, hc.MIXED, 1e-6, np.float32)
    assert np.all(status == 0)
    oranks, oouts = _oracle_batch(orc, step, 1e-6, np.float32)
    _compare_batch(step, ranks, outs, oranks, oouts, 1e-6, "mixed")


@pytest.mark.timeout(900)
def test_full_trailing_step_split_acc1e8_vs_oracle(hc, orc, full_step):
    """The split-precision plan (FP64 tiles, Householder QR, core SVD and
    truncation; the products F, G, T, T^T Zu and the reassembly as BF16 digit
    slices on the matrix cores, digit counts per lane from the slice
    autotuner) on the 1953-update gemm batch at acc 1e-8: every product within
    10*acc of the oracle's on the same FP64 tiles, ranks within one."""
    step = full_step
    _say("case: full trailing step nb 1024 split acc 1e-8")
    picked = []
    status, ranks, outs, _ = _run_batch(hc, step, hc.SPLIT, 1e-8, np.float64,
                                        tune=lambda p: picked.extend(p.autotune_slices()))
    _say("   slices per lane:", [(int(t[0]), int(t[1]), round(t[2], 4)) for t in picked])
    assert np.all(status == 0)
    assert all(2 <= t[1] <= 6 and t[2] <= 0.1 for t in picked)
    oranks, oouts = _oracle_batch(orc, step, 1e-8, np.float64)
    _compare_batch(step, ranks, outs, oranks, oouts, 1e-8, "split")


def test_concurrent_single_calls_match_sequential(hc):
    """HCORE_dgemm from four host threads at once (each with its own stream
    and workspace inside the library; ctypes releases the GIL during the
    call): every result equals the same call made alone, bit for bit."""
    from concurrent.futures import ThreadPoolExecutor
    cases = [(91 + i, 512, 20 + 3 * i, 16, 30 - 2 * i) for i in range(8)]
    _say("case: concurrent single calls", len(cases))
    data = [_random_case(s, M, ra, rb, rc) for s, M, ra, rb, rc in cases]
    seq = [hc.dgemm(-1.0, *d[:4], 1.0, *d[4:], maxrk=512, acc=1e-8) for d in data]
    with ThreadPoolExecutor(4) as ex:
        par = list(ex.map(lambda d: hc.dgemm(-1.0, *d[:4], 1.0, *d[4:], maxrk=512, acc=1e-8),
                          data + data))
    for u, got in enumerate(par):
        ref = seq[u % len(data)]
        assert got[2] == ref[2]
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


def test_single_calls_hold_no_growing_state(hc):
    """Thousands of small HCORE_dgemm calls (host pointers, as a runtime's
    worker threads make them) leave host RSS and free device memory flat:
    a call creates no events or streams and allocates nothing once the
    thread's buffers have grown to the tile size."""
    import psutil
    from test_oracle import _random_case
    AU, AV, BU, BV, CU, CV = _random_case(7, 16, 3, 3, 3)
    _say("case: 3000 nb 16 calls, host RSS and device memory")
    for _ in range(200):
        hc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=16, acc=1e-8)
    proc = psutil.Process()
    rss0, free0 = proc.memory_info().rss, hc.mem_info()[0]
    for _ in range(3000):
        hc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=16, acc=1e-8)
    rss1, free1 = proc.memory_info().rss, hc.mem_info()[0]
    _say("   RSS %+d KiB, device free %+d KiB" % ((rss1 - rss0) >> 10, (free1 - free0) >> 10))
    assert rss1 - rss0 < 8 << 20
    assert free0 - free1 < 8 << 20
