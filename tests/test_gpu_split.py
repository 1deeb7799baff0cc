"""GPU parity of the split-precision form (precision HCORE_HIP_SPLIT).

The product engine (HCORE_hip_gemm_split: FP64 operands cut into BF16 digits
per k chunk of 32, digit products on the matrix cores with FP32 accumulation,
FP64 sums) is pinned by what its arithmetic fixes:
  * integer operands of at most 8 bits, under arbitrary power-of-two row and
    column scales, are represented exactly by their first digit, so every
    product equals the exact (integer) product bit for bit — for every
    transpose combination, k across chunk and split-k boundaries, and the
    alpha/beta forms;
  * for general operands the error of C(i, j) is at most
    k (S + 3) 2^(1 - 8 S) max|op(A)(i, :)| max|op(B)(:, j)| for S digits
    (the digit remainders and the dropped digit pairs), and S = 0 runs the
    same tiles through the FP64 matrix cores (exact as well on the integer
    operands, with the alpha/beta forms reading C in the GEMM epilogues).
Split plans are compared with the oracle (ranks within one, products within
10*acc); the slice autotuner must pick counts that keep the core singular
values within 0.1*acc of the FP64 ones; graph replays equal eager execution.
"""
import numpy as np
import pytest

from test_gpu_coverage import _compare_batch, _oracle_batch, _run_batch

pytestmark = pytest.mark.gpu

SHAPES = [  # m, n, k, transA, transB
    (1, 1, 1, False, False),
    (16, 16, 32, False, False),
    (33, 70, 31, True, False),
    (70, 33, 33, False, True),
    (64, 64, 100, True, True),
    (35, 35, 1024, True, False),
    (1024, 40, 27, False, True),
    (5, 3, 2048, True, False),
]


def _say(*a):
    print(*a, flush=True)


def _op(X, t):
    return X.T if t else X


@pytest.mark.parametrize("shape", SHAPES)
def test_sliced_product_exact_on_scaled_small_integers(hc, shape):
    m, n, k, ta, tb = shape
    _say("case: sliced product exact", shape)
    rng = np.random.default_rng(m * 7919 + n * 31 + k)
    Ai = rng.integers(-255, 256, size=(m, k))
    Bi = rng.integers(-255, 256, size=(k, n))
    Ai[0, :] = 0  # an all-zero row (scale 1, digits 0)
    ra = rng.integers(-40, 41, size=m).astype(float)  # power-of-two row scales
    cb = rng.integers(-40, 41, size=n).astype(float)
    A = np.ldexp(Ai.astype(float), ra[:, None].astype(int))
    B = np.ldexp(Bi.astype(float), cb[None, :].astype(int))
    exact = np.ldexp((Ai @ Bi).astype(float), (ra[:, None] + cb[None, :]).astype(int))
    C0 = np.ldexp(rng.integers(-1000, 1001, size=(m, n)).astype(float),
                  (ra[:, None] + cb[None, :]).astype(int))
    for S in (0, 2, 6):  # 0: the same tiles through the FP64 matrix-core kernels
        got = hc.gemm_split(_op(A, ta), _op(B, tb), np.zeros((m, n)), S, transA=ta, transB=tb)
        assert np.array_equal(got, exact), "slices %d" % S
        got = hc.gemm_split(_op(A, ta), _op(B, tb), C0, S, alpha=2.0, beta=-1.0, transA=ta,
                            transB=tb)
        assert np.array_equal(got, 2.0 * exact - C0), "slices %d, alpha/beta" % S


@pytest.mark.parametrize("shape", SHAPES[2:])
def test_sliced_product_error_bound(hc, shape):
    m, n, k, ta, tb = shape
    _say("case: sliced product error bound", shape)
    rng = np.random.default_rng(k + 17)
    A = rng.standard_normal((m, k)) * np.logspace(0, -9, m)[:, None]
    B = rng.standard_normal((k, n)) * np.logspace(-3, 3, n)[None, :]
    ref = A @ B
    scale = np.abs(A).max(axis=1)[:, None] * np.abs(B).max(axis=0)[None, :]
    errs = []
    for S in range(2, 7):
        got = hc.gemm_split(_op(A, ta), _op(B, tb), np.zeros((m, n)), S, transA=ta, transB=tb)
        err = np.abs(got - ref)
        # the FP64 reference's own rounding: a few ulps of sum |a||b|
        fp = 4 * k * np.finfo(float).eps * (np.abs(A) @ np.abs(B))
        bound = k * (S + 3) * 2.0 ** (1 - 8 * S) * scale + fp
        assert np.all(err <= bound), "slices %d: worst err/bound %.3g" % (S, (err / bound).max())
        errs.append(float((err / scale).max()))
    _say("   max err / (row max col max) for 2..6 digits:", ["%.2e" % e for e in errs])
    assert errs[0] > errs[2] > errs[4] or errs[2] <= 1e-12
    got0 = hc.gemm_split(_op(A, ta), _op(B, tb), np.zeros((m, n)), 0, transA=ta, transB=tb)
    assert np.all(np.abs(got0 - ref) <= 4 * k * np.finfo(float).eps * (np.abs(A) @ np.abs(B)))


def test_sliced_product_rejects_bad_slice_counts(hc):
    A = np.ones((4, 4))
    for S in (1, 7, -1):
        st = hc.lib().HCORE_hip_gemm_split(S, 111, 111, 4, 4, 4, 1.0, None, 4, None, 4, 0.0, None,
                                           4, None)
        assert st == hc.ERR_ARG
    with pytest.raises(RuntimeError):
        hc.gemm_split(A, A, A, 1)


@pytest.fixture(scope="module")
def lanes_step():
    from hcore_amd.tiles import TrailingStep
    return TrailingStep(grid=13, nb=256, key="split-lanes")


def test_split_plan_three_lanes_vs_oracle(hc, orc, lanes_step):
    """A three-lane nb=256 batch at acc 1e-8 with the default digit counts
    (5 at this acc): statuses OK, ranks within one of the oracle's, products
    within 10*acc of the oracle's."""
    step = lanes_step
    _say("case: split plan, grid 13 nb 256, acc 1e-8, default slices")
    status, ranks, outs, _ = _run_batch(hc, step, hc.SPLIT, 1e-8, np.float64)
    assert np.all(status == 0)
    oranks, oouts = _oracle_batch(orc, step, 1e-8, np.float64)
    _compare_batch(step, ranks, outs, oranks, oouts, 1e-8, "split default")


@pytest.mark.parametrize("acc", [1e-8, 1e-6])
def test_split_plan_autotuned_slices_vs_oracle(hc, orc, lanes_step, acc):
    """The slice autotuner picks, per lane, a count whose core singular values
    stay within 0.1*acc of the FP64 ones (fewer digits at the looser acc);
    the tuned plan still matches the oracle."""
    step = lanes_step
    _say("case: split plan autotuned, grid 13 nb 256, acc", acc)
    picked = []
    status, ranks, outs, _ = _run_batch(hc, step, hc.SPLIT, acc, np.float64,
                                        tune=lambda p: picked.extend(p.autotune_slices()))
    _say("   (lane, digits, dev/acc, dev/acc one fewer):",
         [(int(t[0]), int(t[1]), "%.3g" % t[2], "%.3g" % t[3]) for t in picked])
    assert picked and all(2 <= t[1] <= 6 and 0 <= t[2] <= 0.1 for t in picked)
    # one digit fewer misses the tolerance (unless the smallest count passed)
    assert all(t[1] == 2 or t[3] > 0.1 for t in picked)
    assert np.all(status == 0)
    oranks, oouts = _oracle_batch(orc, step, acc, np.float64)
    _compare_batch(step, ranks, outs, oranks, oouts, acc, "split tuned")


def test_split_plan_graph_replays_match_eager(hc, lanes_step):
    """Three graph replays of a split plan (eager work and synchronisation in
    between) give the eager execution's factors bit for bit."""
    step = lanes_step
    _say("case: split plan graph replays")
    nb = step.nb
    ra, rb, rc = step.ranks()
    K = ra + rc
    Bf = lambda x, cap=None: hc.DeviceBuffer.from_numpy(np.asfortranarray(x), cap)
    panel = {i: (Bf(u), Bf(v)) for i, (u, v) in step.panel.items()}
    CUd = [Bf(c[0], k) for c, k in zip(step.C, K)]
    CVd = [Bf(c[1], k) for c, k in zip(step.C, K)]
    OUd = [hc.DeviceBuffer(nb, k) for k in K]
    OVd = [hc.DeviceBuffer(nb, k) for k in K]
    p = hc.Plan(nb, hc.SPLIT)
    p.add_dgemm([panel[i][0].ptr for i, j in step.pairs], [panel[i][1].ptr for i, j in step.pairs],
                ra, nb, [panel[j][0].ptr for i, j in step.pairs],
                [panel[j][1].ptr for i, j in step.pairs], rb, nb,
                [t.ptr for t in CUd], [t.ptr for t in CVd], rc, nb, -1.0, 1.0, 0, 512, 1e-8,
                CUout=[t.ptr for t in OUd], CVout=[t.ptr for t in OVd])
    p.finalize()
    assert list(p.slices()) == [5] * len(p.slices())
    p.set_slices(4)
    assert list(p.slices()) == [4] * len(p.slices())
    p.execute()
    hc.device_synchronize()
    r0 = p.ranks().copy()
    eager = [(OUd[u].numpy(int(r0[u])), OVd[u].numpy(int(r0[u]))) for u in range(len(K))]
    for rep in range(3):
        for t in OUd + OVd:
            t.zero()
        p.execute(graph=True)
        hc.device_synchronize()
        assert np.array_equal(p.ranks(), r0)
        for u in range(len(K)):
            assert np.array_equal(OUd[u].numpy(int(r0[u])), eager[u][0]), (rep, u)
            assert np.array_equal(OVd[u].numpy(int(r0[u])), eager[u][1]), (rep, u)
    p.destroy()
    for t in CUd + CVd + OUd + OVd + [x for pr in panel.values() for x in pr]:
        t.free()
