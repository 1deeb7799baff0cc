"""Synthetic tile workloads (host side, numpy).

Every tile is drawn from a counter-based generator (numpy Philox) keyed by
(workload, i, j, role), so the GPU run and the CPU baseline build identical
FP64 tiles.  A low-rank tile of rank r is stored as U = Q1 diag(sigma),
V = Q2 with Q1, Q2 the orthonormal factors of the QR of seeded N(0,1)
nb x r matrices and sigma_k = 10^(-decay k / r).
"""
import hashlib
import os

import numpy as np


def rng(*key):
    h = hashlib.blake2b(repr(key).encode(), digest_size=16).digest()
    return np.random.Generator(np.random.Philox(key=np.frombuffer(h, dtype=np.uint64)))


def lowrank_tile(nb, r, decay, key, dtype=np.float64):
    g = rng(*key)
    q1, _ = np.linalg.qr(g.standard_normal((nb, r)))
    q2, _ = np.linalg.qr(g.standard_normal((nb, r)))
    sig = 10.0 ** (-decay * np.arange(r) / r)
    return (np.asfortranarray((q1 * sig).astype(dtype)), np.asfortranarray(q2.astype(dtype)))


def single_tile_update(nb=500, r=32, decay=8.0, key="single"):
    """One update C - A B^T with A, B, C of rank r (the nb=500 rank-32 case)."""
    AU, AV = lowrank_tile(nb, r, decay, (key, "A"))
    BU, BV = lowrank_tile(nb, r, decay, (key, "B"))
    CU, CV = lowrank_tile(nb, r, decay, (key, "C"))
    return AU, AV, BU, BV, CU, CV


def cholesky_rank(d, base=160.0, floor=8):
    """Rank of tile (i, j) at distance d = |i - j| of the trailing-step grid."""
    return max(floor, int(np.floor(base / (1.0 + abs(d) / 2.0) + 0.5)))


class TrailingStep:
    """Step 0 of a TLR Cholesky on a g x g tile grid (updates of all C_ij,
    1 <= j < i < g, with the panel column A_i0, plus the g-1 diagonal tiles).

    Defaults: g = 64, nb = 1024, sigma decay 10 (sigma_k = 10^(-10 k / r)).
    Diagonal tiles are dense symmetric N(0,1) + nb*I (only their lower
    triangle matters to the update).
    """

    def __init__(self, grid=64, nb=1024, decay=10.0, key="trailing", diag=True,
                 max_updates=None, rank_of=None):
        self.grid, self.nb, self.decay, self.key = grid, nb, decay, key
        if rank_of is None:  # rank by distance from the diagonal
            rank_of = lambda i, j, role: cholesky_rank(i - j)
        self.pairs = [(i, j) for i in range(1, grid) for j in range(1, i)]
        if max_updates is not None:
            self.pairs = self.pairs[:max_updates]
        rows = sorted({p for pr in self.pairs for p in pr} | set(range(1, grid) if diag else ()))
        self.diag_idx = list(range(1, grid)) if diag else []

        def dense(i):
            a = rng(key, i, i, "D").standard_normal((nb, nb))
            a += a.T
            a[np.diag_indices(nb)] += nb
            return np.asfortranarray(a)

        # tiles are independent: generate them on a thread pool with
        # single-threaded LAPACK inside each task
        from concurrent.futures import ThreadPoolExecutor
        from threadpoolctl import threadpool_limits
        with threadpool_limits(1), ThreadPoolExecutor(max(1, min(16, os.cpu_count() or 1))) as ex:
            pan = list(ex.map(lambda i: lowrank_tile(nb, rank_of(i, 0, "panel"), decay,
                                                     (key, i, 0, "panel")), rows))
            self.C = list(ex.map(lambda ij: lowrank_tile(nb, rank_of(ij[0], ij[1], "C"), decay,
                                                         (key, ij[0], ij[1], "C")), self.pairs))
            self.D = list(ex.map(dense, self.diag_idx))
        self.panel = dict(zip(rows, pan))

    @classmethod
    def batch_512(cls, key="cfg3"):
        """512 updates on nb = 2048 tiles with ranks drawn uniformly from
        [32, 256] per tile (keyed), sigma_k = 10^(-12 k / r): stacked ranks up
        to 512.  Panel tiles of a 34-tile column, the first 512 pairs."""
        def rank(i, j, role):
            return int(rng(key, i, j, role, "rank").integers(32, 257))
        return cls(grid=34, nb=2048, decay=12.0, key=key, diag=False, max_updates=512,
                   rank_of=rank)

    def ranks(self):
        ra = np.array([self.panel[i][0].shape[1] for (i, j) in self.pairs], dtype=np.int32)
        rb = np.array([self.panel[j][0].shape[1] for (i, j) in self.pairs], dtype=np.int32)
        rc = np.array([c[0].shape[1] for c in self.C], dtype=np.int32)
        return ra, rb, rc
