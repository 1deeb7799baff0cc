"""A tile low-rank SPD matrix on the GPU and its Cholesky factorisation.

The layout a HiCMA-style runtime keeps for TLR Cholesky, built on the C ABI:
dense diagonal tiles D[k] (nb x nb); for each tile column k one buffer of
U factors and one of V factors holding the tiles (i, k), i > k, cap columns
apart (nb x cap each), so that a panel's triangular solve is one call; int
ranks per tile.  Tiles come from the covariance problem of HCORE_dgytlr
(generated and compressed on the GPU); the factorisation is
HCORE_hip_tlr_dpotrf (potrf, trsm, dsyrk and batched recompression per step).
"""
import numpy as np

from . import DeviceBuffer, gytlr_batch, tlr_dpotrf, FlopCounter, _check, FP64


class TlrMatrix:
    def __init__(self, bigM, nb, seed, maxrank, acc, cap=None):
        if bigM % nb:
            raise ValueError("bigM must be a multiple of nb")
        self.bigM, self.nb, self.seed = bigM, nb, seed
        self.nt = nt = bigM // nb
        self.cap = cap if cap is not None else nb
        self.maxrank, self.acc = maxrank, acc
        self.D = [DeviceBuffer(nb, nb) for _ in range(nt)]
        self.Ucol = [DeviceBuffer(nb, self.cap * (nt - k - 1)) for k in range(nt - 1)]
        self.Vcol = [DeviceBuffer(nb, self.cap * (nt - k - 1)) for k in range(nt - 1)]
        self.ranks = np.zeros((nt, nt), dtype=np.int32, order="F")
        self.flops = FlopCounter()

    def tile_ptr(self, buf, i, k):
        return int(buf[int(k)].ptr) + 8 * (int(i) - int(k) - 1) * self.cap * self.nb

    def pointer_tables(self):
        nt = self.nt
        U = [0] * (nt * nt)
        V = [0] * (nt * nt)
        for k in range(nt - 1):
            for i in range(k + 1, nt):
                U[i + k * nt] = self.tile_ptr(self.Ucol, i, k)
                V[i + k * nt] = self.tile_ptr(self.Vcol, i, k)
        return [d.ptr for d in self.D], U, V

    def generate(self, stream=None):
        """Every lower tile in one batch: diagonal tiles dense, the others
        compressed at acc with at most maxrank columns."""
        nt, nb = self.nt, self.nb
        tiles = [(i, k) for k in range(nt) for i in range(k, nt)]
        rank = DeviceBuffer(len(tiles), 1, np.int32)
        AU, AV, AD = [], [], []
        for i, k in tiles:
            if i == k:
                AU.append(0); AV.append(0); AD.append(self.D[k].ptr)
            else:
                AU.append(self.tile_ptr(self.Ucol, i, k))
                AV.append(self.tile_ptr(self.Vcol, i, k))
                AD.append(None)
        st = gytlr_batch(nb, nb, [i * nb for i, k in tiles], [k * nb for i, k in tiles], self.bigM,
                         self.seed, AU, nb, AV, nb, rank.ptr, AD=AD, lda=nb,
                         maxrank=self.maxrank, tol=self.acc, stream=stream)
        _check(st, "HCORE_hip_gytlr_batch")
        r = rank.numpy()[:, 0]
        for t, (i, k) in enumerate(tiles):
            if i != k:
                self.ranks[i, k] = r[t]
        rank.free()
        return st

    def factor(self, maxrk=None, acc=None, stream=None, precision=FP64, slices=0):
        """HCORE_hip_tlr_dpotrf in place (the recompressions as split-precision
        plans for precision SPLIT).  Returns (status, info, largest new rank
        per step)."""
        D, U, V = self.pointer_tables()
        return tlr_dpotrf(self.nb, D, U, V, self.ranks, self.cap,
                          self.cap if maxrk is None else maxrk,
                          self.acc if acc is None else acc, stream, self.flops, precision, slices)

    def tile(self, i, k):
        """Host copy of tile (i, k), i >= k: dense diagonal, else U V^T."""
        if i == k:
            return self.D[k].numpy()
        r = int(self.ranks[i, k])
        off = (i - k - 1) * self.cap
        U = self.Ucol[k].numpy()[:, off:off + r]
        V = self.Vcol[k].numpy()[:, off:off + r]
        return U @ V.T

    def factors(self, i, k):
        """Host copies of tile (i, k)'s factors (i > k), its rank columns only."""
        from . import lib
        r = int(self.ranks[i, k])
        out = []
        for buf in (self.Ucol, self.Vcol):
            x = np.zeros((self.nb, r), order="F")
            if r:
                _check(lib().HCORE_hip_memcpy_d2h(x.ctypes.data, self.tile_ptr(buf, i, k),
                                                  x.nbytes), "d2h")
            out.append(x)
        return out[0], out[1]

    def row(self, g):
        """Row g of the lower triangle (after factor(): of L), from the
        factors of the tiles in its tile row."""
        nb = self.nb
        i, li = divmod(int(g), nb)
        out = np.zeros(self.bigM)
        for k in range(i):
            U, V = self.factors(i, k)
            out[k * nb:(k + 1) * nb] = V @ U[li]
        d = self.D[i].numpy()
        out[i * nb:i * nb + li + 1] = d[li, :li + 1]
        return out

    def dense_lower(self, lower_diag=True):
        """The lower triangle as one dense host matrix (diagonal tiles
        through tril when lower_diag)."""
        nb, nt = self.nb, self.nt
        L = np.zeros((self.bigM, self.bigM))
        Ucs = [u.numpy() for u in self.Ucol]
        Vcs = [v.numpy() for v in self.Vcol]
        for k in range(nt):
            d = self.D[k].numpy()
            L[k * nb:(k + 1) * nb, k * nb:(k + 1) * nb] = np.tril(d) if lower_diag else d
            for i in range(k + 1, nt):
                r = int(self.ranks[i, k])
                off = (i - k - 1) * self.cap
                L[i * nb:(i + 1) * nb, k * nb:(k + 1) * nb] = \
                    Ucs[k][:, off:off + r] @ Vcs[k][:, off:off + r].T
        return L

    def free(self):
        for b in self.D + self.Ucol + self.Vcol:
            b.free()
