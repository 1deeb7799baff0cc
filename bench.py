#!/usr/bin/env python3
"""Benchmark: one TLR Cholesky trailing step of tile updates on one MI355X.

Workload (the nb=1024 uneven-rank trailing-step batch): step 0 of a TLR
Cholesky on a 64 x 64 tile grid, nb = 1024 — 1953 HCORE_dgemm updates
C_ij -= A_i0 A_j0^T (1 <= j < i <= 63) and 63 HCORE_dsyrk diagonal updates,
ranks r(d) = max(8, round(160/(1+|d|/2))), sigma_k = 10^(-10k/r), acc 1e-8,
maxrank 512, FP64.  Tiles are synthetic (seeded), resident in HBM before the
timed region; the recompressed factors are written out of place so every
step performs identical work (the diagonal dsyrk updates run in place).

A "step" is one execution of the whole batch: one plan execution, whose
updates run in up to three lanes by stacked rank, each on its own stream
(the diagonal dsyrk updates follow one of them), captured once into a
hipGraph and replayed.  value = tile updates per second (2016 per step), timed over
graph replays with HIP events on the launching stream.

Also measured in the same run:
  * per-kernel durations: the same step executed eagerly with the lanes one
    after another and a HIP event pair around every launch -> the roofline of
    the dominant kernel (the longest launch);
  * the CPU baseline: the oracle (the reference's LAPACK call sequence,
    scipy-openblas, one update per OpenMP thread) on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12         # B/s
FP64_PEAK = 78.6e12       # FLOP/s (MI355X dense FP64, matrix and vector)
# execution unit of each phase's kernels: the products run on FP64 MFMA, the
# stacking, the Householder QR, the core SVD and the compact-WY solve on the
# vector ALUs
ISSUED_ON = {p: "MFMA" for p in ("F=AV^T BV", "G=BU F^T", "T=RA RB^T+Gram", "Zv=T^T Zu",
                                 "reassembly", "syrk_W", "syrk_X", "syrk_CD")}
PHASES = ["stack", "F=AV^T BV", "G=BU F^T", "geqrf", "T=RA RB^T+Gram", "core_svd",
          "wy_solve", "Zv=T^T Zu", "reassembly", "syrk_W", "syrk_X", "syrk_CD"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def q_flops(m, n):
    # geqrf closed form of the reference (flop_counts.c:5-10), float thirds
    return 2 * m * n * n - int(np.float32(2 * n * n * n) / np.float32(3.0)) + 2 * m * n + \
        int(np.float32(17 * n) / np.float32(3.0))


def graph_ms(plan, stream, steps, warmup):
    """Average step time of a finalized plan over hipGraph replays (HIP events)."""
    import hcore_amd as hc
    plan.execute(stream, graph=True)
    for _ in range(warmup):
        plan.execute(stream, graph=True)
    stream.synchronize()
    e0, e1 = hc.Event(), hc.Event()
    e0.record(stream)
    for _ in range(steps):
        plan.execute(stream, graph=True)
    e1.record(stream)
    stream.synchronize()
    return e0.elapsed_ms(e1) / steps


def device_ranks_leg(step, stream, steps, host_ms, host_ranks, slack=16):
    """The same gemm batch as a plan with device-side ranks whose capacities
    exceed the ranks in use by `slack` per factor (lanes and workspace by
    capacity; each core SVD routed by the rank read at execution through the
    lane's rank queue): graph replays, each after the rank words are reset
    (the plan publishes the new ranks to them), timed one by one with HIP
    events; ranks compared with the host-rank plan's."""
    import hcore_amd as hc
    nb = step.nb
    ra, rb, rc = step.ranks()
    rows = sorted(step.panel)
    n = len(step.pairs)
    B = hc.DeviceBuffer.from_numpy
    panel = {i: (B(u), B(v)) for i, (u, v) in step.panel.items()}
    K = ra + rc
    CUd = [B(c[0], k) for c, k in zip(step.C, K)]
    CVd = [B(c[1], k) for c, k in zip(step.C, K)]
    OUd = [hc.DeviceBuffer(nb, k) for k in K]
    OVd = [hc.DeviceBuffer(nb, k) for k in K]
    pos = {i: k for k, i in enumerate(rows)}
    prk = np.array([step.panel[i][0].shape[1] for i in rows], dtype=np.int32)
    crk = np.asarray(rc, dtype=np.int32)
    prank = B(prk.reshape(-1, 1))
    crank = B(crk.reshape(-1, 1))
    pa = lambda i: prank.ptr + 4 * pos[i]
    plan = hc.Plan(nb, hc.FP64)
    plan.add_dgemm_dev([panel[i][0].ptr for i, j in step.pairs], [panel[i][1].ptr for i, j in step.pairs],
                       [pa(i) for i, j in step.pairs], ra + slack, nb,
                       [panel[j][0].ptr for i, j in step.pairs], [panel[j][1].ptr for i, j in step.pairs],
                       [pa(j) for i, j in step.pairs], rb + slack, nb,
                       [t.ptr for t in CUd], [t.ptr for t in CVd],
                       [crank.ptr + 4 * u for u in range(n)], rc + slack, nb, -1.0, 1.0, 0, 512, 1e-8,
                       CUout=[t.ptr for t in OUd], CVout=[t.ptr for t in OVd], ldco=nb)
    plan.finalize()
    e0, e1 = hc.Event(), hc.Event()
    tot = 0.0
    for rep in range(steps + 2):
        crank.upload(crk.reshape(-1, 1))
        e0.record(stream)
        plan.execute(stream, graph=True)
        e1.record(stream)
        stream.synchronize()
        if rep >= 2:  # capture and one warm-up replay untimed
            tot += e0.elapsed_ms(e1)
    ms = tot / steps
    assert np.all(plan.status() == 0)
    drift = int(np.max(np.abs(plan.ranks().astype(int) - host_ranks.astype(int))))
    lanes = np.bincount(plan.lanes(), minlength=3)
    # per-phase durations with the lanes one after another
    crank.upload(crk.reshape(-1, 1))
    plan.set_timing(2)
    plan.execute(stream)
    stream.synchronize()
    ph_ms, ph_n = plan.phase_times()
    plan.set_timing(False)
    phases = {}
    for l in range(ph_ms.shape[0]):
        row = {PHASES[i]: float(ph_ms[l, i] / ph_n[l, i]) for i in range(len(PHASES)) if ph_n[l, i] > 0}
        if row:
            phases["lane%d" % l] = row
            log("  device ranks lane %d: %s" % (l, " ".join("%s %.3f" % kv for kv in row.items())))
    plan.destroy()
    for t in CUd + CVd + OUd + OVd + [prank, crank] + [x for pr in panel.values() for x in pr]:
        t.free()
    res = {"capacity_slack_per_factor": slack, "ms_per_step": ms, "host_rank_ms_per_step": host_ms,
           "over_host_ranks": ms / host_ms, "rank_drift_vs_host_ranks": drift,
           "updates_per_lane_by_capacity": [int(x) for x in lanes], "phase_ms_serialized": phases}
    log("device ranks (capacities +%d): %.3f ms/step vs host ranks %.3f ms (x%.2f), lanes by "
        "capacity %s, rank drift %d" % (slack, ms, host_ms, ms / host_ms, res["updates_per_lane_by_capacity"],
                                        drift))
    return res


def reduced_precision_leg(step, stream, steps, warmup, acc, sample):
    """FP64, mixed-precision (FP32 tiles) and split-precision (FP64 tiles,
    BF16-digit products, digit counts from the slice autotuner) plans of the
    same gemm batch at one accuracy: step times and the worst relative error
    of sampled updates against the FP64 dense update of the original FP64
    tiles."""
    import hcore_amd as hc
    nb = step.nb
    ra, rb, rc = step.ranks()
    K = ra + rc
    out = {"acc": acc}
    for name, prec, dt in (("fp64", hc.FP64, np.float64), ("mixed", hc.MIXED, np.float32),
                           ("split", hc.SPLIT, np.float64)):
        B = lambda x, cap=None: hc.DeviceBuffer.from_numpy(np.asfortranarray(x.astype(dt)), cap)
        panel = {i: (B(u), B(v)) for i, (u, v) in step.panel.items()}
        CUd = [B(c[0], k) for c, k in zip(step.C, K)]
        CVd = [B(c[1], k) for c, k in zip(step.C, K)]
        OUd = [hc.DeviceBuffer(nb, k, dt) for k in K]
        OVd = [hc.DeviceBuffer(nb, k, dt) for k in K]
        plan = hc.Plan(nb, prec)
        plan.add_dgemm([panel[i][0].ptr for i, j in step.pairs], [panel[i][1].ptr for i, j in step.pairs],
                       ra, nb, [panel[j][0].ptr for i, j in step.pairs],
                       [panel[j][1].ptr for i, j in step.pairs], rb, nb,
                       [t.ptr for t in CUd], [t.ptr for t in CVd], rc, nb, -1.0, 1.0, 0, 512, acc,
                       CUout=[t.ptr for t in OUd], CVout=[t.ptr for t in OVd], ldco=nb)
        plan.finalize()
        plan.execute(stream)
        stream.synchronize()
        tuned = None
        if prec == hc.SPLIT:  # digits per lane: the smallest keeping sigma within 0.1 acc
            tuned = plan.autotune_slices(stream)
            plan.execute(stream)
            stream.synchronize()
        assert np.all(plan.status() == 0)
        ranks = plan.ranks()
        out.setdefault("_ranks", {})[name] = ranks.copy()
        worst = 0.0
        for u in sample:
            i, j = step.pairs[u]
            AU, AV = step.panel[i]
            BU, BV = step.panel[j]
            CU, CV = step.C[u]
            E = CU @ CV.T - (AU @ AV.T) @ (BU @ BV.T).T
            r = int(ranks[u])
            got = OUd[u].numpy(r).astype(np.float64) @ OVd[u].numpy(r).astype(np.float64).T
            worst = max(worst, float(np.linalg.norm(got - E) / np.linalg.norm(E)))
        ms = graph_ms(plan, stream, steps, warmup)
        out[name] = {"ms_per_step": ms, "value": len(step.pairs) / (ms / 1e3),
                     "worst_rel_err": worst, "mean_rank": float(np.mean(ranks))}
        if prec == hc.SPLIT and acc == 1e-8:  # per-phase durations, lanes one after another
            plan.set_timing(2)
            plan.execute(stream)
            stream.synchronize()
            ph_ms, ph_n = plan.phase_times()
            plan.set_timing(False)
            for l in range(ph_ms.shape[0]):
                row = {PHASES[i]: float(ph_ms[l, i] / ph_n[l, i]) for i in range(len(PHASES))
                       if ph_n[l, i] > 0}
                if row:
                    out[name].setdefault("phase_ms_serialized", {})["lane%d" % l] = row
                    log("  split lane %d: %s" % (l, " ".join("%s %.3f" % kv for kv in row.items())))
        if tuned is not None:
            out[name]["slices_per_lane"] = [int(t[1]) for t in tuned]
            out[name]["sigma_dev_over_acc"] = [float(t[2]) for t in tuned]
        plan.destroy()
        for t in CUd + CVd + OUd + OVd + [x for pr in panel.values() for x in pr]:
            t.free()
    rk = out.pop("_ranks")
    for k in ("mixed", "split"):
        out["%s_over_fp64" % k] = out[k]["value"] / out["fp64"]["value"]
        d = np.abs(rk[k].astype(int) - rk["fp64"].astype(int))
        out["rank_drift_%s_vs_fp64" % k] = {"max": int(d.max()),
                                            "updates_differing": int(np.count_nonzero(d))}
    out["meets_acc"] = {k: bool(out[k]["worst_rel_err"] <= 10 * acc) for k in ("fp64", "mixed", "split")}
    return out


def single_tile_leg(reps):
    """The nb=500, rank-32 single update (acc 1e-8, maxrank 250): latency of
    one drop-in HCORE_dgemm call on the GPU (host pointers: staged through
    the library; device pointers: in place) next to the oracle on the same
    tile on one host core."""
    import hcore_amd as hc
    import oracle  # the CPU baseline leg: the oracle timed on one host core
    from hcore_amd.tiles import single_tile_update
    AU, AV, BU, BV, CU, CV = single_tile_update(nb=500, r=32, decay=8.0, key="single")
    args = (-1.0, AU, AV, BU, BV, 1.0, CU, CV)
    E = CU @ CV.T - (AU @ AV.T) @ (BU @ BV.T).T
    out = {}
    # host pointers: the drop-in call as a HiCMA codelet makes it (staging included)
    hc.dgemm(*args, maxrk=250, acc=1e-8)
    t = time.perf_counter()
    for _ in range(reps):
        U, V, r = hc.dgemm(*args, maxrk=250, acc=1e-8)
    out["gpu_host_pointers_us"] = (time.perf_counter() - t) / reps * 1e6
    # the same call from 16 host threads at once (a runtime's workers, one
    # tile each): per-thread streams and buffers let the calls overlap
    from concurrent.futures import ThreadPoolExecutor
    nthr = 16
    with ThreadPoolExecutor(nthr) as ex:
        list(ex.map(lambda _: hc.dgemm(*args, maxrk=250, acc=1e-8), range(nthr)))  # warm-up
        t = time.perf_counter()
        list(ex.map(lambda _: hc.dgemm(*args, maxrk=250, acc=1e-8), range(nthr * reps)))
        out["gpu_host_pointers_16threads_calls_per_s"] = nthr * reps / (time.perf_counter() - t)
    out["rank"] = r
    out["rel_err"] = float(np.linalg.norm(U @ V.T - E) / np.linalg.norm(E))
    # device pointers: factors resident in HBM, C restored from a pristine copy
    # (device-to-device) before each call; only the HCORE_dgemm call is timed
    import ctypes as C
    L = hc.lib()
    M, K = AU.shape[0], CU.shape[1] + AU.shape[1]
    D = {k: hc.DeviceBuffer.from_numpy(v) for k, v in (("AU", AU), ("AV", AV), ("BU", BU),
                                                         ("BV", BV))}
    C0U, C0V = hc.DeviceBuffer.from_numpy(CU, K), hc.DeviceBuffer.from_numpy(CV, K)
    CUd, CVd = hc.DeviceBuffer(M, K), hc.DeviceBuffer(M, K)
    dp = lambda b: C.cast(C.c_void_p(b.ptr), C.POINTER(C.c_double))
    ra, rb = np.array([AU.shape[1]], dtype=np.float64), np.array([BU.shape[1]], dtype=np.float64)
    rcv = np.zeros(1)
    fl = hc.FlopCounter()
    tot = 0.0
    for it in range(reps + 1):
        CUd.copy_from(C0U)
        CVd.copy_from(C0V)
        rcv[0] = CU.shape[1]
        t = time.perf_counter()
        L.HCORE_dgemm(111, 112, M, M, -1.0, dp(D["AU"]), dp(D["AV"]), hc._ptr(ra), M,
                      dp(D["BU"]), dp(D["BV"]), hc._ptr(rb), M, 1.0, dp(CUd), dp(CVd),
                      hc._ptr(rcv), M, 0, 250, 1e-8, None, C.byref(fl))
        if it:
            tot += time.perf_counter() - t
    out["gpu_device_pointers_us"] = tot / reps * 1e6
    rd = int(rcv[0])
    out["rel_err_device"] = float(np.linalg.norm(CUd.numpy(rd) @ CVd.numpy(rd).T - E) /
                                  np.linalg.norm(E))
    for b in list(D.values()) + [C0U, C0V, CUd, CVd]:
        b.free()
    oracle.set_blas_threads(1)
    oracle.dgemm(*args, maxrk=250, acc=1e-8)
    t = time.perf_counter()
    for _ in range(reps):
        st, oU, oV, ro, _, _ = oracle.dgemm(*args, maxrk=250, acc=1e-8)
    out["cpu_oracle_1core_us"] = (time.perf_counter() - t) / reps * 1e6
    out["oracle_rank"] = ro
    out["reps"] = reps
    out["note"] = ("host pointers: staging through pinned memory and PCIe included; device "
                   "pointers: the call alone on HBM-resident factors")
    return out


def _factored_rel_diff(XU, XV, YU, YV):
    """||XU XV^T - YU YV^T||_F / ||YU YV^T||_F without forming the tile."""
    Q, R = np.linalg.qr(np.hstack([XU, -YU]))
    num = np.linalg.norm(R @ np.hstack([XV, YV]).T)
    Q2, R2 = np.linalg.qr(YU)
    return float(num / np.linalg.norm(R2 @ YV.T))


def batch2048_leg(steps, warmup):
    """The 512-update batch on nb = 2048 tiles (ranks 32-256 per tile, stacked
    ranks to 506, acc 1e-8, maxrank 512): graph replays of one FP64 plan, the
    reference's flop accounting with the plan's ranks; then the same batch as
    a split-precision plan (BF16-digit products, digit counts from the slice
    autotuner) with its error per tile on a sample: against the FP64 plan's
    factors, and against the dense FP64 update for the largest stacked ranks."""
    import hcore_amd as hc
    from hcore_amd.tiles import TrailingStep
    t = time.time()
    step = TrailingStep.batch_512()
    t_gen = time.time() - t
    nb = step.nb
    ra, rb, rc = step.ranks()
    K = ra + rc
    B = hc.DeviceBuffer.from_numpy
    panel = {i: (B(u), B(v)) for i, (u, v) in step.panel.items()}
    CUd = [B(c[0], k) for c, k in zip(step.C, K)]
    CVd = [B(c[1], k) for c, k in zip(step.C, K)]
    OUd = [hc.DeviceBuffer(nb, k) for k in K]
    OVd = [hc.DeviceBuffer(nb, k) for k in K]
    plan = hc.Plan(nb, hc.FP64)
    plan.add_dgemm([panel[i][0].ptr for i, j in step.pairs], [panel[i][1].ptr for i, j in step.pairs],
                   ra, nb, [panel[j][0].ptr for i, j in step.pairs],
                   [panel[j][1].ptr for i, j in step.pairs], rb, nb,
                   [t.ptr for t in CUd], [t.ptr for t in CVd], rc, nb, -1.0, 1.0, 0, 512, 1e-8,
                   CUout=[t.ptr for t in OUd], CVout=[t.ptr for t in OVd], ldco=nb)
    plan.finalize()
    stream = hc.Stream()
    ms = graph_ms(plan, stream, steps, warmup)
    assert np.all(plan.status() == 0)
    # per-phase durations with the lanes one after another (HIP events)
    plan.set_timing(2)
    plan.execute(stream)
    stream.synchronize()
    ph_ms, ph_n = plan.phase_times()
    plan.set_timing(False)
    phases = {}
    for l in range(ph_ms.shape[0]):
        row = {PHASES[i]: float(ph_ms[l, i] / ph_n[l, i]) for i in range(len(PHASES))
               if ph_n[l, i] > 0}
        if row:
            phases["lane%d" % l] = row
            log("  batch 2048 lane %d: %s" % (l, " ".join("%s %.2f" % kv for kv in row.items())))
    r = plan.ranks().astype(np.int64)
    Kl = K.astype(np.int64)
    fc = hc.flop_counts
    f = (sum(2 * q_flops(nb, int(k)) for k in Kl) + int(np.sum(4 * ra.astype(np.int64) * rb * nb)) +
         int(np.sum(22 * Kl ** 3)) + int(np.sum(r * Kl)) +
         sum(fc('o', nb, int(x), int(k), 1) + fc('o', int(x), nb, int(k), 2) for x, k in zip(r, Kl)))
    plan.destroy()
    res = {"updates": len(step.pairs), "nb": nb, "K_min": int(K.min()), "K_max": int(K.max()),
           "ms_per_step": ms, "updates_per_s": len(step.pairs) / (ms / 1e3),
           "tflops": f / (ms / 1e3) / 1e12, "flops_per_step": f, "generate_s": t_gen,
           "phase_ms_serialized": phases}
    log("batch 512 x nb 2048 (K %d-%d): %.2f ms/step, %.0f updates/s, %.2f TF/s" %
        (K.min(), K.max(), ms, res["updates_per_s"], res["tflops"]))
    # ---- split precision on the same batch ------------------------------------
    order = np.argsort(-K, kind="stable")
    sample = sorted(set(range(0, len(K), 8)) | set(int(u) for u in order[:4]))
    ref = {u: (OUd[u].numpy(int(r[u])), OVd[u].numpy(int(r[u]))) for u in sample}
    for b in OUd + OVd:
        b.zero()
    sp = hc.Plan(nb, hc.SPLIT)
    sp.add_dgemm([panel[i][0].ptr for i, j in step.pairs], [panel[i][1].ptr for i, j in step.pairs],
                 ra, nb, [panel[j][0].ptr for i, j in step.pairs],
                 [panel[j][1].ptr for i, j in step.pairs], rb, nb,
                 [t.ptr for t in CUd], [t.ptr for t in CVd], rc, nb, -1.0, 1.0, 0, 512, 1e-8,
                 CUout=[t.ptr for t in OUd], CVout=[t.ptr for t in OVd], ldco=nb)
    sp.finalize()
    sp.execute(stream)
    stream.synchronize()
    tuned = sp.autotune_slices(stream)
    ms_s = graph_ms(sp, stream, steps, warmup)
    assert np.all(sp.status() == 0)
    rs = sp.ranks()
    errs = np.array([_factored_rel_diff(OUd[u].numpy(int(rs[u])), OVd[u].numpy(int(rs[u])),
                                        *ref[u]) for u in sample])
    dense = []
    for u in (int(x) for x in order[:4]):
        i, j = step.pairs[u]
        AU, AV = step.panel[i]
        BU, BV = step.panel[j]
        CU, CV = step.C[u]
        E = CU @ CV.T - (AU @ AV.T) @ (BU @ BV.T).T
        got = OUd[u].numpy(int(rs[u])) @ OVd[u].numpy(int(rs[u])).T
        dense.append(float(np.linalg.norm(got - E) / np.linalg.norm(E)))
    sp.destroy()
    for b in CUd + CVd + OUd + OVd + [x for pr in panel.values() for x in pr]:
        b.free()
    drift = np.abs(rs.astype(int) - r.astype(int))
    res["split"] = {"ms_per_step": ms_s, "updates_per_s": len(step.pairs) / (ms_s / 1e3),
                    "split_over_fp64": ms / ms_s,
                    "slices_per_lane": [int(t[1]) for t in tuned],
                    "sigma_dev_over_acc": [float(t[2]) for t in tuned],
                    "sample_tiles": len(sample),
                    "rel_diff_vs_fp64_per_tile": {"max": float(errs.max()),
                                                  "median": float(np.median(errs))},
                    "rel_err_vs_dense_largest_K": dense,
                    "rank_drift_vs_fp64": {"max": int(drift.max()),
                                           "updates_differing": int(np.count_nonzero(drift))}}
    log("batch 512 x nb 2048 split: %.2f ms/step (split/fp64 %.2fx), digits per lane %s, "
        "per-tile rel diff vs fp64 max %.2e (%d tiles), vs dense (4 largest K) max %.2e, "
        "rank drift %d" % (ms_s, ms / ms_s, [int(t[1]) for t in tuned], errs.max(), len(sample),
                           max(dense), drift.max()))
    return res


def tlr_leg(bigM, nb):
    """Generation, compression and TLR Cholesky of the covariance problem
    (bigM Morton-sorted points, exp(-d/0.1) + 1e-4 I, acc 1e-8, compression
    maxrank nb/2, recompression maxrank nb), all on the GPU, with FP64
    recompressions and again (regenerated) with split-precision ones; each
    factor is checked on sampled entries of L L^T against the matrix (numpy,
    from the library's points)."""
    import hcore_amd as hc
    from hcore_amd.tlr import TlrMatrix
    acc = 1e-8
    rng = np.random.default_rng(7)
    rows = sorted(int(g) for g in rng.choice(bigM, 6, replace=False))
    P = hc.covariance_points(bigM, 2024)
    res = None
    for form, prec in (("fp64", hc.FP64), ("split", hc.SPLIT)):
        M = TlrMatrix(bigM, nb, 2024, maxrank=nb // 2, acc=acc, cap=nb)
        t = time.time()
        M.generate()
        t_gen = time.time() - t
        r0 = M.ranks.copy()
        nt = M.nt
        low = [(i, k) for k in range(nt) for i in range(k + 1, nt)]
        t = time.time()
        st, info, smax = M.factor(maxrk=nb, acc=acc, precision=prec)
        t_fac = time.time() - t
        assert st == 0 and info == 0, (st, info)
        fl = M.flops
        Lr = {g: M.row(g) for g in rows}
        worst = 0.0
        for a in rows:
            for b in rows:
                d = np.hypot(*(P[a] - P[b]))
                ref = np.exp(-d / 0.1) + (1e-4 if a == b else 0.0)
                worst = max(worst, abs(Lr[a] @ Lr[b] - ref))
        M.free()
        if form == "fp64":
            res = {"points": bigM, "nb": nb, "tiles": nt,
                   "gemm_updates": nt * (nt - 1) * (nt - 2) // 6,
                   "syrk_updates": nt * (nt - 1) // 2, "generate_s": t_gen, "factor_s": t_fac,
                   "factor_gflops": (fl.update + fl.syrk + fl.trsm + fl.potrf) / t_fac / 1e9,
                   "rank_max_generated": int(max(r0[i, k] for i, k in low)),
                   "rank_mean_generated": float(np.mean([r0[i, k] for i, k in low])),
                   "rank_max_per_step": [int(x) for x in smax[:-1]],
                   "sampled_LLt_max_abs_err": worst}
            log("tlr cholesky %d points nb %d: generate %.2f s, factor %.2f s (%.0f GF/s), ranks "
                "mean %.1f max %d, sampled |LL^T - A| %.2e" %
                (bigM, nb, t_gen, t_fac, res["factor_gflops"], res["rank_mean_generated"],
                 res["rank_max_generated"], worst))
        else:
            res["split"] = {"factor_s": t_fac, "split_over_fp64": res["factor_s"] / t_fac,
                            "sampled_LLt_max_abs_err": worst,
                            "rank_max_per_step": [int(x) for x in smax[:-1]]}
            log("tlr cholesky split recompressions: factor %.2f s (split/fp64 %.2fx), sampled "
                "|LL^T - A| %.2e" % (t_fac, res["factor_s"] / t_fac, worst))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--nb", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the reduced-precision leg")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--single-reps", type=int, default=200)
    ap.add_argument("--check", type=int, default=24, help="updates checked against the oracle")
    ap.add_argument("--tlr", type=int, default=0,
                    help="also run the TLR Cholesky of a covariance problem of this many "
                         "points (tiles of --tlr-nb; 32768 is the 64 x 64-tile stream)")
    ap.add_argument("--tlr-nb", type=int, default=512)
    ap.add_argument("--batch2048", action="store_true",
                    help="also time the 512-update batch on nb = 2048 tiles (stacked ranks to 506)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on exactly one GPU (no sharding, no collectives); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import hcore_amd as hc
    from hcore_amd.tiles import TrailingStep

    L = hc.lib()
    assert L.HCORE_hip_init(0) == 0, "HIP init failed"
    devname, ncu = hc.device_info()
    log("device:", devname, "CUs:", ncu)

    t0 = time.time()
    step = TrailingStep(grid=args.grid, nb=args.nb)
    nb = step.nb
    ra, rb, rc = step.ranks()
    K = ra + rc
    n = len(step.pairs)
    ns = len(step.diag_idx)
    log("generated %d gemm + %d syrk tiles in %.1fs" % (n, ns, time.time() - t0))

    # ---- device residency -------------------------------------------------------
    B = hc.DeviceBuffer.from_numpy
    panel = {i: (B(u), B(v)) for i, (u, v) in step.panel.items()}
    CUd = [B(c[0], k) for c, k in zip(step.C, K)]
    CVd = [B(c[1], k) for c, k in zip(step.C, K)]
    OUd = [hc.DeviceBuffer(nb, k) for k in K]
    OVd = [hc.DeviceBuffer(nb, k) for k in K]
    Dd = [B(d) for d in step.D]
    plan = hc.Plan(nb, hc.FP64)
    plan.add_dgemm([panel[i][0].ptr for i, j in step.pairs], [panel[i][1].ptr for i, j in step.pairs],
                   ra, nb, [panel[j][0].ptr for i, j in step.pairs],
                   [panel[j][1].ptr for i, j in step.pairs], rb, nb,
                   [t.ptr for t in CUd], [t.ptr for t in CVd], rc, nb, -1.0, 1.0, 0, 512, 1e-8,
                   CUout=[t.ptr for t in OUd], CVout=[t.ptr for t in OVd], ldco=nb)
    plan.add_dsyrk([panel[i][0].ptr for i in step.diag_idx], [panel[i][1].ptr for i in step.diag_idx],
                   [step.panel[i][0].shape[1] for i in step.diag_idx], nb,
                   [t.ptr for t in Dd], nb, -1.0)
    t = time.perf_counter()
    plan.finalize()
    finalize_ms = (time.perf_counter() - t) * 1e3
    log("plan workspace %.2f GB, finalize %.1f ms" % (plan.workspace_bytes() / 1e9, finalize_ms))
    stream = hc.Stream()

    # ---- one checked execution ------------------------------------------------
    plan.execute(stream)
    stream.synchronize()
    ranks = plan.ranks()
    status = plan.status()
    assert np.all(status == 0), "device status errors: %s" % np.unique(status)
    sw = plan.sweeps()
    lane_of = plan.lanes()
    # autotuning: every core launch timed with both wavefront counts, the
    # faster kept (results unchanged; later plans of this process reuse it)
    tune = plan.autotune(stream, reps=3)
    for t in tune:
        log("autotune lane %d: core variant %d -> %d (%.3f ms vs %.3f ms, %d tiles)" %
            (int(t[0]), int(t[1]), int(t[2]), t[3], t[4], int(t[5])))
    log("Jacobi sweeps: mean %.1f max %d" % (sw.mean(), sw.max()))
    # numpy check of a sample (largest stacked ranks first): the product within
    # 10*acc of the dense update, the rank by the truncation rule on numpy's SVD
    order = np.argsort(-K, kind="stable")
    sample = sorted(set(list(order[:4]) + list(np.linspace(0, n - 1, max(args.check - 4, 2)).astype(int))))
    worst, rank_drift = 0.0, 0
    for u in sample:
        i, j = step.pairs[u]
        AU, AV = step.panel[i]
        BU, BV = step.panel[j]
        CU, CV = step.C[u]
        r = int(ranks[u])
        E = CU @ CV.T - (AU @ AV.T) @ (BU @ BV.T).T
        s_np = np.linalg.svd(E, compute_uv=False)[:K[u]]
        below = np.nonzero(s_np[1:] < 1e-8)[0]
        r_np = int(below[0]) + 1 if below.size else int(K[u])
        got = OUd[u].numpy(r) @ OVd[u].numpy(r).T
        worst = max(worst, float(np.linalg.norm(got - E) / np.linalg.norm(E)))
        rank_drift = max(rank_drift, abs(r - r_np))
    log("checked %d updates vs the dense update: worst rel err %.3e (bound 1e-7), "
        "max rank drift %d" % (len(sample), worst, rank_drift))
    assert worst <= 1e-7 and rank_drift <= 1

    # ---- flop and byte accounting (the reference's closed forms) ------------------
    fc = hc.flop_counts
    r = ranks.astype(np.int64)
    Kl = K.astype(np.int64)
    ral, rbl = ra.astype(np.int64), rb.astype(np.int64)
    f_qr = sum(2 * q_flops(nb, int(k)) for k in Kl)
    f_F = int(np.sum(2 * ral * rbl * nb))
    f_G = int(np.sum(2 * nb * ral * rbl))
    f_svd = int(np.sum(22 * Kl ** 3))
    f_scale = int(np.sum(r * Kl))
    f_orm = sum(fc('o', nb, int(rr), int(k), 1) + fc('o', int(rr), nb, int(k), 2) for rr, k in zip(r, Kl))
    f_gemm = f_qr + f_F + f_G + f_svd + f_scale + f_orm
    ks = np.array([step.panel[i][0].shape[1] for i in step.diag_idx], dtype=np.int64)
    f_syrk = int(np.sum(2 * ks * ks * nb + 2 * nb * ks * ks + 2 * nb * nb * ks))
    F_total = f_gemm + f_syrk
    # unique compulsory bytes (FP64): C factors in + new factors out, panel
    # factors once, dense diagonal tiles in + out
    panel_bytes = sum(8 * 2 * nb * u.shape[1] for (u, v) in step.panel.values())
    b_unique = int(np.sum(8 * nb * (2 * rc.astype(np.int64) + 2 * r))) + panel_bytes + \
        ns * 8 * 2 * nb * nb

    # ---- timed region: the step replayed through its captured hipGraph --------
    t = time.perf_counter()
    plan.execute(stream, graph=True)     # capture + first replay
    stream.synchronize()
    capture_ms = (time.perf_counter() - t) * 1e3
    for _ in range(args.warmup):
        plan.execute(stream, graph=True)
    stream.synchronize()
    hc.device_synchronize()
    e0, e1 = hc.Event(), hc.Event()
    e0.record(stream)
    for _ in range(args.steps):
        plan.execute(stream, graph=True)
    e1.record(stream)
    stream.synchronize()
    hc.device_synchronize()
    ms_per_step = e0.elapsed_ms(e1) / args.steps
    updates = n + ns
    value = updates / (ms_per_step / 1e3)
    # the same step issued eagerly on the plan's streams (no graph), for comparison
    for _ in range(args.warmup):
        plan.execute(stream)
    stream.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        plan.execute(stream)
    e1.record(stream)
    stream.synchronize()
    eager_conc_ms = e0.elapsed_ms(e1) / args.steps
    log("eager (4 streams, no graph) %.3f ms/step" % eager_conc_ms)

    # ---- per-kernel durations: the same step executed eagerly, lanes one after
    # another on the stream, with a HIP event pair around every launch, so that
    # each duration is the launch's own (not shared with concurrent lanes) ------
    plan.set_timing(2)
    ms0, c0 = plan.phase_times()
    q0, q1 = hc.Event(), hc.Event()
    q0.record(stream)
    for _ in range(args.steps):
        plan.execute(stream)
    q1.record(stream)
    stream.synchronize()
    eager_ms = q0.elapsed_ms(q1) / args.steps
    ms1, c1 = plan.phase_times()
    plan.set_timing(False)
    phase_ms = (ms1 - ms0) / np.maximum(c1 - c0, 1)   # (lane, phase) average per launch
    lane_names = ["lane0", "lane1", "lane2", "syrk"]
    for l in range(3):
        m = lane_of == l
        if m.any():
            lane_names[l] = "K%d-%d" % (K[m].min(), K[m].max())
    for l in range(4):
        row = " ".join("%s %.3f" % (PHASES[i], phase_ms[l, i]) for i in range(12) if c1[l, i] > c0[l, i])
        if row:
            log("  lane %-10s %s" % (lane_names[l], row))
    log("graph %.3f ms/step (timed); lanes serialized, per-launch events: %.3f ms/step" %
        (ms_per_step, eager_ms))

    # ---- dominant kernel roofline ------------------------------------------------
    # the (lane, kernel) with the largest average launch duration; its
    # algorithmic work is that of the updates in the lane
    dl, dp = np.unravel_index(int(np.argmax(phase_ms)), phase_ms.shape)
    dname = PHASES[dp]
    sel = lane_of == dl
    Ks, rs = Kl[sel], r[sel]
    if dname == "syrk_CD":
        traffic_alg = ns * (16 * nb * nb + 8 * 2 * nb * int(ks.mean()))
        ach = traffic_alg / (phase_ms[dl, dp] / 1e3)
        roof = {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": ach / HBM_PEAK, "traffic": None}
    else:
        fl = {"geqrf": sum(2 * q_flops(nb, int(k)) for k in Ks),
              "core_svd": int(np.sum(22 * Ks ** 3)),
              "reassembly": sum(fc('o', nb, int(a), int(k), 1) + fc('o', int(a), nb, int(k), 2)
                                for a, k in zip(rs, Ks)),
              "F=AV^T BV": int(np.sum(2 * ral[sel] * rbl[sel] * nb)),
              "G=BU F^T": int(np.sum(2 * nb * ral[sel] * rbl[sel]))}.get(dname, 0)
        ach = fl / (phase_ms[dl, dp] / 1e3)
        roof = {"bound": "mfma", "achieved": ach / 1e12, "peak": FP64_PEAK / 1e12,
                "unit": "TFLOP/s", "frac": ach / FP64_PEAK, "traffic": None,
                "issued_on": ISSUED_ON.get(dname, "VALU"),
                "peak_note": "MI355X FP64 peak, 78.6 TFLOP/s for both the matrix cores and "
                             "the vector units"}
    roof["kernel"] = dname
    roof["lane"] = lane_names[dl]
    roof["avg_launch_ms"] = float(phase_ms[dl, dp])
    # HBM bytes per launch of that kernel from the committed rocprofv3 counter
    # passes (FETCH_SIZE and WRITE_SIZE, calibrated on the stacking copy kernel)
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tr = json.load(fh)
        if tr.get("phase") == dname and int(K[sel].min()) > 128 and tr.get("lane") == "K>128":
            roof["traffic"] = tr["hbm_bytes_per_launch"]
            roof["traffic_kernel"] = tr["kernel"]
            roof["traffic_source"] = tr["source"]

    # ---- reduced precision: FP64 vs mixed (FP32 tiles) and split (BF16-digit
    # products) plans of the gemm batch at acc 1e-8 and 1e-6 --------------------------
    mixed = None
    if not args.no_mixed:
        mixed = {}
        for acc in (1e-8, 1e-6):
            m = reduced_precision_leg(step, stream, args.steps, args.warmup, acc, sample)
            mixed["acc_%g" % acc] = m
            log("acc %g: split %.3f ms (worst %.2e, digits per lane %s), split/fp64 %.2fx, rank drift %d" %
                (acc, m["split"]["ms_per_step"], m["split"]["worst_rel_err"],
                 m["split"]["slices_per_lane"], m["split_over_fp64"],
                 m["rank_drift_split_vs_fp64"]["max"]))
            log("acc %g: fp64 %.3f ms (worst %.2e), mixed %.3f ms (worst %.2e), mixed/fp64 %.2fx, "
                "rank drift %d" % (acc, m["fp64"]["ms_per_step"], m["fp64"]["worst_rel_err"],
                                   m["mixed"]["ms_per_step"], m["mixed"]["worst_rel_err"],
                                   m["mixed_over_fp64"], m["rank_drift_mixed_vs_fp64"]["max"]))
    # ---- the same gemm batch with device-side ranks (rank queue) ---------------------
    dev_ranks = None
    if mixed:
        dev_ranks = device_ranks_leg(step, stream, args.steps,
                                     mixed["acc_1e-08"]["fp64"]["ms_per_step"], ranks)
    single = None
    if not args.no_cpu:
        single = single_tile_leg(args.single_reps)
        log("single tile nb 500: gpu %.0f us (host ptrs) %.0f us (device ptrs), oracle %.0f us; "
            "16 threads: %.0f calls/s" %
            (single["gpu_host_pointers_us"], single["gpu_device_pointers_us"],
             single["cpu_oracle_1core_us"], single["gpu_host_pointers_16threads_calls_per_s"]))

    # ---- CPU baseline (oracle on the host cores) ------------------------------------
    cpu = None
    if not args.no_cpu:
        import oracle  # the CPU baseline leg: the oracle timed on the host cores
        # every core this process may run on, within the OpenMP thread budget the
        # host sets (16 per GPU on a shared box)
        avail = len(os.sched_getaffinity(0))
        omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        ncores = min(avail, omp) if omp > 0 else avail
        cpu_model = "unknown"
        try:
            with open("/proc/cpuinfo") as fh:
                for line in fh:
                    if line.startswith("model name"):
                        cpu_model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
        AU = [step.panel[i][0] for i, j in step.pairs]
        AV = [step.panel[i][1] for i, j in step.pairs]
        BU = [step.panel[j][0] for i, j in step.pairs]
        BV = [step.panel[j][1] for i, j in step.pairs]
        sAU = [step.panel[i][0] for i in step.diag_idx]
        sAV = [step.panel[i][1] for i in step.diag_idx]
        times = []
        for rep in range(args.cpu_reps + 1):
            CU, CV = [], []
            for (u, v), k in zip(step.C, K):
                a = np.zeros((nb, k), order="F"); a[:, :u.shape[1]] = u; CU.append(a)
                b = np.zeros((nb, k), order="F"); b[:, :v.shape[1]] = v; CV.append(b)
            D = [d.copy(order="F") for d in step.D]
            rcc = rc.copy()
            t = time.time()
            bad, _ = oracle.dgemm_batch(-1.0, AU, AV, BU, BV, 1.0, CU, CV, rcc, 0, 512, 1e-8,
                                        nthreads=ncores)
            oracle.dsyrk_batch(-1.0, sAU, sAV, D, nthreads=ncores)
            dt = time.time() - t
            if rep > 0:
                times.append(dt)
            assert bad == 0
        tmed = float(np.median(times))
        cpu = {"value": updates / tmed, "unit": "tile-updates/s", "cores": ncores,
               "kind": "port",
               "sample": "the full step (%d gemm + %d syrk updates), median of %d runs after "
                         "1 warm-up; %s, 1 BLAS thread per update" %
                         (n, ns, args.cpu_reps, oracle.blas_config().split()[0] + " " +
                          oracle.blas_config().split()[1]),
               "gflops": F_total / tmed / 1e9, "nproc": os.cpu_count(), "affinity_cores": avail,
               "omp_num_threads": omp, "cpu_model": cpu_model,
               "runs_s": [float(x) for x in times]}
        log("cpu baseline: %.3f s/step on %d threads" % (tmed, ncores))

    tlr = tlr_leg(args.tlr, args.tlr_nb) if args.tlr else None
    b2048 = batch2048_leg(max(3, args.steps // 4), 1) if args.batch2048 else None

    out = {
        "metric": "TLR GEMM updates/s at rel.err<=1e-8: mixed-prec vs FP64 GPU, % roofline",
        "value": value,
        "unit": "tile-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded low-rank tiles, generated on the host, resident in HBM)",
        "config": {"workload": "TLR Cholesky trailing step, 64x64 tile grid, nb=1024: "
                               "1953 HCORE_dgemm + 63 HCORE_dsyrk, uneven ranks 8-160, "
                               "acc=1e-8, maxrank=512",
                   "grid": args.grid, "nb": nb, "updates_per_step": updates,
                   "precision": "fp64", "execution": "hipGraph replay of the step",
                   "serialized_lanes_ms_per_step": eager_ms,
                   "eager_streams_ms_per_step": eager_conc_ms,
                   "plan_finalize_ms": finalize_ms,
                   "graph_capture_and_first_replay_ms": capture_ms},
        "roofline": roof,
        "reduced_precision": mixed,
        "device_ranks": dev_ranks,
        "single_tile": single,
        "cpu_baseline": cpu,
        "tlr_cholesky": tlr,
        "batch_nb2048": b2048,
        "gflops": F_total / (ms_per_step / 1e3) / 1e9,
        "step_roofline_frac": max(F_total / (ms_per_step / 1e3) / FP64_PEAK,
                                  b_unique / (ms_per_step / 1e3) / HBM_PEAK),
        "flops_per_step": F_total,
        "unique_bytes_per_step": b_unique,
        "phase_ms": {lane_names[l]: {PHASES[i]: float(phase_ms[l, i]) for i in range(12)
                                     if c1[l, i] > c0[l, i]} for l in range(4)},
        "check": {"updates": len(sample), "worst_rel_err": worst, "max_rank_drift": rank_drift},
        "autotune": [{"lane": int(t[0]), "base_variant": int(t[1]), "chosen_variant": int(t[2]),
                      "ms_base": t[3], "ms_alternative": t[4], "tiles": int(t[5])} for t in tune],
        "device": devname,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
