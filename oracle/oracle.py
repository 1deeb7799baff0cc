"""ctypes front end of the C oracle (oracle/hcore_oracle.c).

TEST INFRASTRUCTURE ONLY (checker and CPU baseline).  Arrays are numpy,
column-major (Fortran order).
"""
import ctypes as C
import os

import numpy as np

from . import build as _build

OK, ERR_ZERO_RANK, ERR_SUM_RANKS, ERR_K_GT_MAXRK, ERR_R_GT_MAXRK, ERR_LAPACK = range(6)
MUTATE_NONE, MUTATE_ALPHA_SIGN, MUTATE_NO_G_TRANSPOSE, MUTATE_TRUNC_OFF_BY_ONE, \
    MUTATE_SIGMA_TO_U = range(5)

_lib = None
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_ulp = C.POINTER(C.c_ulong)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_build.LIB):
            _build.build()
        L = C.CDLL(_build.LIB)
        L.orc_flop_counts.restype = C.c_ulong
        L.orc_flop_counts.argtypes = [C.c_char, C.c_ulong, C.c_ulong, C.c_ulong, C.c_ulong]
        L.orc_dgemm.restype = C.c_int
        L.orc_dgemm.argtypes = [C.c_int, C.c_double, _dp, _dp, C.c_int, C.c_int, _dp, _dp,
                                C.c_int, C.c_int, C.c_double, _dp, _dp, _ip, C.c_int,
                                C.c_int, C.c_int, C.c_double, _ulp, _dp, C.c_int]
        L.orc_sgemm.restype = C.c_int
        L.orc_sgemm.argtypes = [C.c_int, C.c_float, _fp, _fp, C.c_int, C.c_int, _fp, _fp,
                                C.c_int, C.c_int, C.c_float, _fp, _fp, _ip, C.c_int,
                                C.c_int, C.c_int, C.c_float, _ulp, _fp, C.c_int]
        L.orc_dsyrk.restype = C.c_ulong
        L.orc_dsyrk.argtypes = [C.c_int, C.c_int, C.c_double, _dp, C.c_int, _dp, _dp, C.c_int]
        L.orc_dgemmcd.restype = None
        L.orc_dgemmcd.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp, C.c_int, _dp]
        L.orc_generate_compress.restype = C.c_int
        L.orc_generate_compress.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int,
                                            C.c_int, C.c_double, _ip, _ip]
        L.orc_generate_scompress.restype = C.c_int
        L.orc_generate_scompress.argtypes = [C.c_int, C.c_int, C.c_int, _fp, _fp, _fp,
                                             C.c_int, C.c_int, C.c_float, _ip, _ip]
        L.orc_dgemm_batch.restype = C.c_int
        L.orc_dgemm_batch.argtypes = [C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_void_p,
                                      _ip, C.c_int, C.c_void_p, C.c_void_p, _ip, C.c_int,
                                      C.c_double, C.c_void_p, C.c_void_p, _ip, C.c_int,
                                      C.c_int, C.c_int, C.c_double, C.c_int, _ulp]
        L.orc_dsyrk_batch.restype = C.c_int
        L.orc_dsyrk_batch.argtypes = [C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_void_p,
                                      _ip, C.c_int, C.c_void_p, C.c_int, C.c_int, _ulp]
        i, d = C.c_int, C.c_double
        L.orc_dpotrf.restype = i
        L.orc_dpotrf.argtypes = [i, i, _dp, i, _ulp]
        L.orc_dtrsm.restype = i
        L.orc_dtrsm.argtypes = [i, i, i, i, i, i, d, _dp, i, _dp, i, _ulp]
        L.orc_dgemmbdcd.restype = None
        L.orc_dgemmbdcd.argtypes = [i, i, i, d, _dp, _dp, i, i, _dp, i, d, _dp, i, _ulp]
        L.orc_duncompress.restype = i
        L.orc_duncompress.argtypes = [i, i, _dp, i, i, _dp, i, i, _dp, i]
        L.orc_dcompress.restype = i
        L.orc_dcompress.argtypes = [i, i, _dp, i, _dp, i, _dp, i, i, d, i, _ip]
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_splitmix64.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_covariance_points.restype = None
        L.orc_covariance_points.argtypes = [i, C.c_uint64, _dp]
        L.orc_covariance_tile.restype = None
        L.orc_covariance_tile.argtypes = [i, _dp, i, i, i, i, d, d, _dp, i]
        L.orc_blas_config.restype = C.c_char_p
        L.orc_max_threads.restype = C.c_int
        L.orc_set_blas_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _f64(a):
    a = np.asarray(a, dtype=np.float64)
    if not a.flags.f_contiguous:
        a = np.asfortranarray(a)
    return a


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


def blas_config():
    return lib().orc_blas_config().decode()


def flop_counts(op, a, b=0, c=0, d=0):
    return int(lib().orc_flop_counts(op.encode() if isinstance(op, str) else op, a, b, c, d))


def dgemm(alpha, AU, AV, BU, BV, beta, CU, CV, rk=0, maxrk=None, acc=1e-8, mutate=0,
          dtype=np.float64):
    """Oracle HCORE_dgemm / HCORE_sgemm on nb x r factors (ld = nb).

    Returns (status, CU_new, CV_new, rank, sigma, flops_update).  The input
    arrays are not modified.
    """
    M = AU.shape[0]
    ra, rb, rc = AU.shape[1], BU.shape[1], CU.shape[1]
    if maxrk is None:
        maxrk = M
    K = ra + rc
    cap = max(K, rc, 1)
    conv = (lambda a: np.asfortranarray(np.asarray(a, dtype=dtype)))
    AU, AV, BU, BV = conv(AU), conv(AV), conv(BU), conv(BV)
    CUw = np.zeros((M, cap), dtype=dtype, order="F")
    CVw = np.zeros((M, cap), dtype=dtype, order="F")
    CUw[:, :rc] = CU
    CVw[:, :rc] = CV
    rank = C.c_int(rc)
    fl = C.c_ulong(0)
    sig = np.zeros(max(K, 1), dtype=dtype)
    if dtype == np.float64:
        st = lib().orc_dgemm(M, alpha, _ptr(AU), _ptr(AV), ra, M, _ptr(BU), _ptr(BV), rb, M,
                             beta, _ptr(CUw), _ptr(CVw), C.byref(rank), M, rk, maxrk, acc,
                             C.byref(fl), _ptr(sig), mutate)
    else:
        st = lib().orc_sgemm(M, alpha, _ptr(AU, _fp), _ptr(AV, _fp), ra, M, _ptr(BU, _fp),
                             _ptr(BV, _fp), rb, M, beta, _ptr(CUw, _fp), _ptr(CVw, _fp),
                             C.byref(rank), M, rk, maxrk, acc, C.byref(fl), _ptr(sig, _fp),
                             mutate)
    r = rank.value
    return st, CUw[:, :r].copy(order="F"), CVw[:, :r].copy(order="F"), r, sig[:K], fl.value


def dsyrk(alpha, AU, AV, CD):
    """Oracle HCORE_dsyrk: returns (CD_new, flops_syrk)."""
    AU, AV = _f64(AU), _f64(AV)
    CDw = np.array(CD, dtype=np.float64, order="F", copy=True)
    M, K = AU.shape
    fl = lib().orc_dsyrk(M, K, alpha, _ptr(AU), M, _ptr(AV), _ptr(CDw), M)
    return CDw, int(fl)


def dgemmcd(Au, Av, Bu, Bv, CD):
    """Oracle HCORE_dgemmcd: returns CD - (Au Av^T)(Bu Bv^T)^T."""
    Au, Av, Bu, Bv = _f64(Au), _f64(Av), _f64(Bu), _f64(Bv)
    CDw = np.array(CD, dtype=np.float64, order="F", copy=True)
    M = Au.shape[0]
    lib().orc_dgemmcd(M, _ptr(Au), _ptr(Av), Au.shape[1], _ptr(Bu), _ptr(Bv), Bu.shape[1],
                      _ptr(CDw))
    return CDw


def generate_compress(nrows, ncols, ld, storage_maxrank, maxMNK, acc, iseed,
                      dtype=np.float64):
    """The reference tests' generator (xLATMS + xGESVD + truncate).

    iseed: a length-4 int32 numpy array, advanced in place.
    Returns (dense, U, V, rank) with U, V of shape (ld, storage_maxrank).
    """
    dense = np.zeros((ld, ncols), order="F", dtype=dtype)
    U = np.zeros((ld, storage_maxrank), order="F", dtype=dtype)
    V = np.zeros((max(ncols, ld), storage_maxrank), order="F", dtype=dtype)
    rank = C.c_int(0)
    if dtype == np.float64:
        info = lib().orc_generate_compress(nrows, ncols, ld, _ptr(dense), _ptr(U), _ptr(V),
                                           storage_maxrank, maxMNK, acc, _ptr(iseed, _ip),
                                           C.byref(rank))
    else:
        info = lib().orc_generate_scompress(nrows, ncols, ld, _ptr(dense, _fp), _ptr(U, _fp),
                                            _ptr(V, _fp), storage_maxrank, maxMNK, acc,
                                            _ptr(iseed, _ip), C.byref(rank))
    if info:
        raise RuntimeError("dlatms/gesvd failed: info=%d" % info)
    return dense, U, V, rank.value


def _ptr_array(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def dgemm_batch(alpha, AU, AV, BU, BV, beta, CU, CV, rc, rk, maxrk, acc, nthreads=0):
    """CPU baseline: batch of orc_dgemm over OpenMP threads.

    AU..CV are lists of Fortran float64 arrays (nb x capacity, ld = nb); CU/CV
    are updated in place and rc (int32 array) receives the new ranks.
    Returns (n_failed, flops_update).
    """
    n = len(AU)
    M = AU[0].shape[0]
    ra = np.ascontiguousarray([a.shape[1] for a in AU], dtype=np.int32)
    rb = np.ascontiguousarray([b.shape[1] for b in BU], dtype=np.int32)
    fl = C.c_ulong(0)
    keep = [_ptr_array(x) for x in (AU, AV, BU, BV, CU, CV)]
    bad = lib().orc_dgemm_batch(n, M, alpha, keep[0], keep[1], _ptr(ra, _ip), M, keep[2],
                                keep[3], _ptr(rb, _ip), M, beta, keep[4], keep[5],
                                _ptr(rc, _ip), M, rk, maxrk, acc, nthreads, C.byref(fl))
    return bad, fl.value


def dsyrk_batch(alpha, AU, AV, CD, nthreads=0):
    n = len(AU)
    M = AU[0].shape[0]
    K = np.ascontiguousarray([a.shape[1] for a in AU], dtype=np.int32)
    fl = C.c_ulong(0)
    keep = [_ptr_array(x) for x in (AU, AV, CD)]
    lib().orc_dsyrk_batch(n, M, alpha, keep[0], keep[1], _ptr(K, _ip), M, keep[2], M,
                          nthreads, C.byref(fl))
    return fl.value


def set_blas_threads(n):
    """Threads of the oracle's BLAS inside one call."""
    lib().orc_set_blas_threads(int(n))


def max_threads():
    return int(lib().orc_max_threads())


def dpotrf(uplo, A):
    """Oracle HCORE_dpotrf on a copy of A: returns (factor, info, flops_potrf)."""
    Aw = np.array(A, dtype=np.float64, order="F", copy=True)
    fl = C.c_ulong(0)
    info = lib().orc_dpotrf(uplo, Aw.shape[0], _ptr(Aw), Aw.shape[0], C.byref(fl))
    return Aw, int(info), fl.value


def dtrsm(side, uplo, trans, diag, alpha, A, B):
    """Oracle HCORE_dtrsm on a copy of B (M x Brank): returns (X, flops_trsm)."""
    A = _f64(A)
    Bw = np.array(B, dtype=np.float64, order="F", copy=True)
    fl = C.c_ulong(0)
    st = lib().orc_dtrsm(side, uplo, trans, diag, Bw.shape[0], Bw.shape[1], alpha, _ptr(A),
                         A.shape[0], _ptr(Bw), Bw.shape[0], C.byref(fl))
    assert st == 0
    return Bw, fl.value


def dgemmbdcd(transA, alpha, AU, AV, BD, beta, CD):
    """Oracle HCORE_dgemmbdcd (ld = M throughout): returns (CD_new, flops_update)."""
    AU, AV, BD = _f64(AU), _f64(AV), _f64(BD)
    CDw = np.array(CD, dtype=np.float64, order="F", copy=True)
    M, N = CDw.shape
    fl = C.c_ulong(0)
    lib().orc_dgemmbdcd(transA, M, N, alpha, _ptr(AU), _ptr(AV), AU.shape[1], M, _ptr(BD), M,
                        beta, _ptr(CDw), M, C.byref(fl))
    return CDw, fl.value


def duncompress(AU, BV):
    """Oracle HCORE_duncompress: AU BV^T."""
    AU, BV = _f64(AU), _f64(BV)
    M, N = AU.shape[0], BV.shape[0]
    CD = np.zeros((M, N), order="F")
    st = lib().orc_duncompress(M, N, _ptr(AU), AU.shape[1], M, _ptr(BV), BV.shape[1], N,
                               _ptr(CD), M)
    assert st == 0
    return CD


def dcompress(A, maxrank=None, acc=1e-8, mutate=0):
    """Oracle compression (testing_dgemm.c:180-226): returns (U, V, rank)."""
    A = _f64(A)
    m, n = A.shape
    mn = min(m, n)
    U = np.zeros((m, mn), order="F")
    V = np.zeros((n, mn), order="F")
    r = C.c_int(0)
    st = lib().orc_dcompress(m, n, _ptr(A), m, _ptr(U), m, _ptr(V), n,
                             mn if maxrank is None else maxrank, acc, mutate, C.byref(r))
    assert st == 0
    k = r.value
    return np.asfortranarray(U[:, :k]), np.asfortranarray(V[:, :k]), k


def splitmix64(seed, k):
    return int(lib().orc_splitmix64(seed, k))


def covariance_points(bigM, seed):
    """The covariance problem's sorted points, (bigM, 2)."""
    xy = np.zeros(2 * bigM)
    lib().orc_covariance_points(bigM, seed, _ptr(xy))
    return np.stack([xy[:bigM], xy[bigM:]], axis=1)


def covariance_tile(bigM, seed, m0, n0, m, n, beta=0.1, nugget=1e-4, points=None):
    """Tile (m0, n0) (m x n) of the covariance problem."""
    P = covariance_points(bigM, seed) if points is None else points
    xy = np.ascontiguousarray(np.concatenate([P[:, 0], P[:, 1]]))
    out = np.zeros((m, n), order="F")
    lib().orc_covariance_tile(bigM, _ptr(xy), m0, n0, m, n, beta, nugget, _ptr(out), m)
    return out
