"""Plans with device-side ranks (one captured graph for successive steps) and
the autotuner.

Bars: every update of both steps within 10*acc of the dense update and its
rank within one of the oracle's; the new ranks published to the device rank
words; an out-of-range rank gives status HCORE_HIP_ERR_ARG and leaves the
update's output and rank untouched; after autotuning, the plan's results
equal the untuned ones to rounding (10*acc) with the same ranks.
"""
import numpy as np
import pytest

from conftest import rel_fro

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _dense(AU, AV, BU, BV, CU, CV):
    return CU @ CV.T - (AU @ AV.T) @ (BU @ BV.T).T


def _trim(x, r):
    return np.asfortranarray(x[:, :r])


def test_device_ranks_two_steps_one_graph(hc, orc):
    """A grid-13 step at nb = 256 (66 updates in three rank lanes, plus the
    dsyrk updates) recorded once with device-side ranks and captured once;
    step 2 rewrites the factors in the same buffers with smaller ranks and
    replays the same graph."""
    from hcore_amd.tiles import TrailingStep
    nb = 256
    s1 = TrailingStep(grid=13, nb=nb, key="dev-step-1")
    s2 = TrailingStep(grid=13, nb=nb, key="dev-step-2")
    pairs, rows, diag = s1.pairs, sorted(s1.panel), s1.diag_idx
    n = len(pairs)
    pan_cap = {i: s1.panel[i][0].shape[1] for i in rows}
    c_cap = [c[0].shape[1] for c in s1.C]
    ra_cap = [pan_cap[i] for i, j in pairs]
    rb_cap = [pan_cap[j] for i, j in pairs]
    K_cap = [a + c for a, c in zip(ra_cap, c_cap)]
    B = hc.DeviceBuffer
    pU = {i: B(nb, pan_cap[i]) for i in rows}
    pV = {i: B(nb, pan_cap[i]) for i in rows}
    CUd = [B(nb, k) for k in K_cap]
    CVd = [B(nb, k) for k in K_cap]
    OU = [B(nb, k) for k in K_cap]
    OV = [B(nb, k) for k in K_cap]
    Dd = [B(nb, nb) for _ in diag]
    # device rank words: one per panel tile, one per C tile
    prank = B(len(rows), 1, np.int32)
    crank = B(n, 1, np.int32)
    pos = {i: k for k, i in enumerate(rows)}
    p_addr = lambda i: prank.ptr + 4 * pos[i]
    plan = hc.Plan(nb)
    plan.add_dgemm_dev([pU[i].ptr for i, j in pairs], [pV[i].ptr for i, j in pairs],
                       [p_addr(i) for i, j in pairs], ra_cap, nb,
                       [pU[j].ptr for i, j in pairs], [pV[j].ptr for i, j in pairs],
                       [p_addr(j) for i, j in pairs], rb_cap, nb,
                       [t.ptr for t in CUd], [t.ptr for t in CVd],
                       [crank.ptr + 4 * u for u in range(n)], c_cap, nb, -1.0, 1.0, 0, nb, 1e-8,
                       CUout=[t.ptr for t in OU], CVout=[t.ptr for t in OV])
    plan.add_dsyrk_dev([pU[i].ptr for i in diag], [pV[i].ptr for i in diag],
                       [p_addr(i) for i in diag], [pan_cap[i] for i in diag], nb,
                       [t.ptr for t in Dd], nb, -1.0)
    plan.finalize()

    def load(step, shrink):
        prk = {i: max(1, pan_cap[i] - shrink) for i in rows}
        crk = [max(1, c - shrink) for c in c_cap]
        pan = {i: (_trim(step.panel[i][0], prk[i]), _trim(step.panel[i][1], prk[i])) for i in rows}
        C = [(_trim(step.C[u][0], crk[u]), _trim(step.C[u][1], crk[u])) for u in range(n)]
        for i in rows:
            pU[i].upload(pan[i][0])
            pV[i].upload(pan[i][1])
        for u in range(n):
            CUd[u].upload(C[u][0])
            CVd[u].upload(C[u][1])
        D = [np.asfortranarray(step.D[s]) for s in range(len(diag))]
        for t, d in zip(Dd, D):
            t.upload(d)
        prank.upload(np.array([prk[i] for i in rows], dtype=np.int32))
        crank.upload(np.array(crk, dtype=np.int32))
        return pan, C, D

    for step, shrink, label in ((s1, 0, "step 1"), (s2, 3, "step 2")):
        _say("case: device ranks,", label)
        pan, C, D = load(step, shrink)
        plan.execute(graph=True)
        hc.device_synchronize()
        assert np.all(plan.status() == 0)
        ranks = plan.ranks()
        pub = crank.numpy()[:, 0]
        assert np.array_equal(pub, ranks)  # published to the device rank words
        for u, (i, j) in enumerate(pairs):
            AU, AV = pan[i]
            BU, BV = pan[j]
            CU, CV = C[u]
            r = int(ranks[u])
            st, oU, oV, ro, _, _ = orc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=nb,
                                             acc=1e-8)
            E = _dense(AU, AV, BU, BV, CU, CV)
            assert abs(r - ro) <= 1, (label, u, r, ro)
            assert rel_fro(OU[u].numpy(r) @ OV[u].numpy(r).T, E) <= 10 * 1e-8, (label, u)
        for s, i in enumerate(diag):
            ref, _ = orc.dsyrk(-1.0, pan[i][0], pan[i][1], D[s])
            got = Dd[s].numpy()
            assert np.max(np.abs(got - ref)) <= 1e-12 * np.abs(ref).max(), (label, s)

    # an out-of-range rank: status ERR_ARG, output and rank word untouched
    _say("case: device ranks, out-of-range rank")
    pan, C, D = load(s1, 0)
    bad = 5
    crk = crank.numpy()[:, 0].copy()
    crk[bad] = c_cap[bad] + 1
    crank.upload(crk)
    OU[bad].zero()
    plan.execute(graph=True)
    hc.device_synchronize()
    st = plan.status()
    assert st[bad] == hc.ERR_ARG and np.all(np.delete(st, bad) == 0)
    assert crank.numpy()[bad, 0] == c_cap[bad] + 1
    assert not OU[bad].numpy().any()
    plan.destroy()


def test_autotune_keeps_results(hc):
    """The autotuner on the grid-13 step: every heavy core launch timed with
    both wavefront counts; afterwards the plan (graph recaptured) gives the
    same ranks and products equal to rounding."""
    from hcore_amd.tiles import TrailingStep
    nb = 256
    step = TrailingStep(grid=13, nb=nb, key="tune-test", diag=False)
    ra, rb, rc = step.ranks()
    K = ra + rc
    B = hc.DeviceBuffer.from_numpy
    panel = {i: (B(u), B(v)) for i, (u, v) in step.panel.items()}
    CUd = [B(c[0], k) for c, k in zip(step.C, K)]
    CVd = [B(c[1], k) for c, k in zip(step.C, K)]
    OU = [hc.DeviceBuffer(nb, k) for k in K]
    OV = [hc.DeviceBuffer(nb, k) for k in K]
    plan = hc.Plan(nb)
    plan.add_dgemm([panel[i][0].ptr for i, j in step.pairs], [panel[i][1].ptr for i, j in step.pairs],
                   ra, nb, [panel[j][0].ptr for i, j in step.pairs],
                   [panel[j][1].ptr for i, j in step.pairs], rb, nb,
                   [t.ptr for t in CUd], [t.ptr for t in CVd], rc, nb, -1.0, 1.0, 0, nb, 1e-8,
                   CUout=[t.ptr for t in OU], CVout=[t.ptr for t in OV])
    plan.finalize()
    plan.execute(graph=True)
    hc.device_synchronize()
    r0 = plan.ranks().copy()
    before = [(OU[u].numpy(int(r0[u])), OV[u].numpy(int(r0[u]))) for u in range(len(K))]
    rep = plan.autotune(reps=2)
    _say("case: autotune", [(int(a), int(b), int(c), round(d, 3), round(e, 3)) for a, b, c, d, e, f in rep])
    assert len(rep) >= 1
    assert np.array_equal(plan.ranks(), r0)  # the re-runs leave the rank words as they were
    plan.execute(graph=True)
    hc.device_synchronize()
    assert np.all(plan.status() == 0)
    assert np.array_equal(plan.ranks(), r0)
    for u in range(len(K)):
        r = int(r0[u])
        a = before[u][0] @ before[u][1].T
        b = OU[u].numpy(r) @ OV[u].numpy(r).T
        assert rel_fro(b, a) <= 10 * 1e-8
    plan.destroy()


@pytest.mark.parametrize("form", ["fp64", "split", "mixed"])
def test_device_rank_queue_routes_by_actual_rank(hc, orc, form):
    """Device-rank plan whose capacities are far above the ranks in use (one
    capacity for every tile, so the lanes by capacity collapse to one): each
    execution's rank queue orders the updates by the stacked rank actually
    read (largest first, every update once) and routes each core SVD by it;
    two steps with different ranks replay one graph, every update within
    10*acc of the dense update (of the FP32-rounded tiles for the mixed form)
    and its rank within one of the oracle's.  FP64 and split precision at
    acc 1e-8, the mixed form (FP32 tiles, FP64 core) at 1e-5."""
    from hcore_amd.tiles import TrailingStep
    prec, dt, acc = {"fp64": (hc.FP64, np.float64, 1e-8), "split": (hc.SPLIT, np.float64, 1e-8),
                     "mixed": (hc.MIXED, np.float32, 1e-5)}[form]
    nb = 256
    step = TrailingStep(grid=13, nb=nb, key="dev-queue")
    pairs, rows = step.pairs, sorted(step.panel)
    n = len(pairs)
    rmax = max(max(step.panel[i][0].shape[1] for i in rows), max(c[0].shape[1] for c in step.C))
    cap = nb // 2  # stacked capacity 2 cap = nb; ranks in use up to rmax
    assert rmax < cap
    B = hc.DeviceBuffer
    pU = {i: B(nb, cap, dt) for i in rows}
    pV = {i: B(nb, cap, dt) for i in rows}
    CUd = [B(nb, 2 * cap, dt) for _ in range(n)]
    CVd = [B(nb, 2 * cap, dt) for _ in range(n)]
    OU = [B(nb, 2 * cap, dt) for _ in range(n)]
    OV = [B(nb, 2 * cap, dt) for _ in range(n)]
    prank = B(len(rows), 1, np.int32)
    crank = B(n, 1, np.int32)
    pos = {i: k for k, i in enumerate(rows)}
    p_addr = lambda i: prank.ptr + 4 * pos[i]
    plan = hc.Plan(nb, prec)
    plan.add_dgemm_dev([pU[i].ptr for i, j in pairs], [pV[i].ptr for i, j in pairs],
                       [p_addr(i) for i, j in pairs], [cap] * n, nb,
                       [pU[j].ptr for i, j in pairs], [pV[j].ptr for i, j in pairs],
                       [p_addr(j) for i, j in pairs], [cap] * n, nb,
                       [t.ptr for t in CUd], [t.ptr for t in CVd],
                       [crank.ptr + 4 * u for u in range(n)], [cap] * n, nb, -1.0, 1.0, 0, nb,
                       acc, CUout=[t.ptr for t in OU], CVout=[t.ptr for t in OV])
    plan.finalize()
    lanes = plan.lanes()
    assert np.all(lanes == 0)
    rnd = lambda x: np.asfortranarray(x.astype(dt).astype(np.float64))
    for shrink, label in ((0, "full ranks"), (6, "ranks - 6")):
        _say("case: device rank queue,", form, label)
        prk = {i: max(1, step.panel[i][0].shape[1] - shrink) for i in rows}
        crk = [max(1, c[0].shape[1] - shrink) for c in step.C]
        pan = {i: (rnd(_trim(step.panel[i][0], prk[i])), rnd(_trim(step.panel[i][1], prk[i])))
               for i in rows}
        C = [(rnd(_trim(step.C[u][0], crk[u])), rnd(_trim(step.C[u][1], crk[u]))) for u in range(n)]
        for i in rows:
            pU[i].upload(pan[i][0])
            pV[i].upload(pan[i][1])
        for u in range(n):
            CUd[u].upload(C[u][0])
            CVd[u].upload(C[u][1])
        prank.upload(np.array([prk[i] for i in rows], dtype=np.int32))
        crank.upload(np.array(crk, dtype=np.int32))
        plan.execute(graph=True)
        hc.device_synchronize()
        assert np.all(plan.status() == 0)
        order = plan.core_queue(0)
        K = np.array([prk[i] + crk[u] for u, (i, j) in enumerate(pairs)])
        assert sorted(order.tolist()) == list(range(n))
        assert np.all(np.diff(K[order]) <= 0), K[order]
        ranks = plan.ranks()
        for u, (i, j) in enumerate(pairs):
            AU, AV = pan[i]
            BU, BV = pan[j]
            CU, CV = C[u]
            r = int(ranks[u])
            st, oU, oV, ro, _, _ = orc.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=nb, acc=acc)
            E = _dense(AU, AV, BU, BV, CU, CV)
            assert abs(r - ro) <= 1, (label, u, r, ro)
            got = OU[u].numpy(r).astype(np.float64) @ OV[u].numpy(r).astype(np.float64).T
            assert rel_fro(got, E) <= 10 * acc, (label, u)
    plan.destroy()
    for t in list(pU.values()) + list(pV.values()) + CUd + CVd + OU + OV + [prank, crank]:
        t.free()
