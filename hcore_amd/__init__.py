"""hcore_amd — MI355X-native tile low-rank BLAS (HCORE drop-in).

The product is the C-ABI shared library hcore_amd/lib/libhcore_amd.so
(HIP kernels for gfx950 + host orchestration).  This module is a thin ctypes
front end mirroring the reference's C API (HCORE_dgemm, HCORE_sgemm,
HCORE_dsyrk, HCORE_dgemmcd; reference src/include/hcore_d.h,
src/include/hcore_s.h) plus the batched-plan extension of include/hcore_hip.h.

There is no CPU fallback: importing the library without the built .so raises.
"""
import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "lib", "libhcore_amd.so")
INCLUDE_DIR = os.path.join(ROOT, "include")

# status codes (include/hcore_hip.h)
OK, ERR_ZERO_RANK, ERR_SUM_RANKS, ERR_K_GT_MAXRK, ERR_R_GT_MAXRK, ERR_NUMERIC, \
    ERR_RUNTIME, ERR_ARG = range(8)
FP64, FP32, MIXED, SPLIT = 0, 1, 2, 3
ON_ERROR_EXIT, ON_ERROR_RETURN = 0, 1


class FlopCounter(C.Structure):
    """ctypes mirror of flop_counter (include/flop_util_structs.h)."""
    _fields_ = [("__pad1", C.c_ulong), ("__pad2", C.c_ulong), ("copy", C.c_ulong),
                ("flop", C.c_ulong), ("tid", C.c_int), ("potrf", C.c_ulong),
                ("trsm", C.c_ulong), ("syrk", C.c_ulong), ("update", C.c_ulong),
                ("geqrfu", C.c_ulong), ("geqrfv", C.c_ulong), ("geqrfv_gemm", C.c_ulong),
                ("gesvd", C.c_ulong), ("gesvd_gemm", C.c_ulong), ("gesvd_lasetu", C.c_ulong),
                ("gesvd_lasetv", C.c_ulong), ("gesvd_lacpyu", C.c_ulong),
                ("gesvd_lacpyv", C.c_ulong), ("gesvd_dscal", C.c_ulong), ("ormqru", C.c_ulong),
                ("ormqru_laset", C.c_ulong), ("ormqru_lacpy", C.c_ulong), ("ormqrv", C.c_ulong),
                ("ormqrv_laset", C.c_ulong), ("ormqrv_lacpy", C.c_ulong),
                ("omatcopy", C.c_ulong), ("numthreads", C.c_int), ("__pad3", C.c_ulong),
                ("__pad4", C.c_ulong)]


_lib = None
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_vp = C.c_void_p
_pp = C.POINTER(C.c_void_p)
_ip = C.POINTER(C.c_int)


def lib():
    """Load the built library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("hcore_amd: %s is not built; run __graft_entry__.build() "
                               "(there is no CPU fallback)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        i, d, f = C.c_int, C.c_double, C.c_float
        fl = C.POINTER(FlopCounter)
        L.HCORE_dgemm.argtypes = [i, i, i, i, d, _dp, _dp, _dp, i, _dp, _dp, _dp, i, d, _dp, _dp,
                                  _dp, i, i, i, d, _dp, fl]
        L.HCORE_dgemm.restype = None
        L.HCORE_sgemm.argtypes = [i, i, i, i, f, _fp, _fp, _fp, i, _fp, _fp, _fp, i, f, _fp, _fp,
                                  _fp, i, i, i, f, _fp, fl]
        L.HCORE_sgemm.restype = None
        L.HCORE_dsyrk.argtypes = [i, i, i, i, d, _dp, i, _dp, i, d, _dp, i, _dp, fl]
        L.HCORE_dsyrk.restype = None
        L.HCORE_dgemmcd.argtypes = [i, i, i, d, _dp, _dp, i, i, _dp, _dp, i, i, d, _dp, i, i,
                                    _dp, fl]
        L.HCORE_dgemmcd.restype = None
        L.HCORE_dpotrf.argtypes = [i, i, _dp, i, _ip, fl]
        L.HCORE_dpotrf.restype = None
        L.HCORE_dtrsm.argtypes = [i, i, i, i, i, i, d, _dp, i, _dp, i, fl]
        L.HCORE_dtrsm.restype = None
        L.HCORE_dgemmbdcd.argtypes = [i, i, i, i, d, _dp, _dp, _dp, i, _dp, i, d, _dp, i, _dp, fl]
        L.HCORE_dgemmbdcd.restype = None
        L.HCORE_duncompress.argtypes = [i, i, i, i, d, _dp, _dp, i, _dp, _dp, i, d, _dp, i]
        L.HCORE_duncompress.restype = None
        L.HCORE_dcompress.argtypes = [i, i, _dp, i, _dp, i, _dp, i, _dp, i, d]
        L.HCORE_dcompress.restype = None
        L.HCORE_dgytlr.argtypes = [i, i, _dp, _dp, _dp, _dp, i, i, i, i, i, i, C.c_ulonglong, i,
                                   d, i, _dp]
        L.HCORE_dgytlr.restype = None
        L.HCORE_hip_gytlr_batch.argtypes = [i, i, i, _ip, _ip, i, C.c_ulonglong, _pp, i, _pp, i,
                                            _vp, _pp, i, i, d, _vp]
        L.HCORE_hip_compress_batch.argtypes = [i, i, i, _pp, i, _pp, i, _pp, i, _vp, i, d, _vp]
        L.HCORE_hip_set_covariance.argtypes = [d, d]
        L.HCORE_hip_tlr_dpotrf.argtypes = [i, i, _pp, _pp, _pp, _ip, i, i, d, _vp, _ip, fl, _ip]
        L.HCORE_hip_tlr_dpotrf_form.argtypes = [i, i, i, i, _pp, _pp, _pp, _ip, i, i, d, _vp, _ip,
                                                fl, _ip]
        L.HCORE_hip_covariance_points.argtypes = [i, C.c_ulonglong, _dp]
        for fn in ("HCORE_dgemm_qr_svd", "HCORE_dgemm_qr_svd_b_dense", "HCORE_dgemm_ormqr"):
            getattr(L, fn).restype = None
        L.HCORE_hip_gemm.argtypes = [i, i, i, i, i, i, d, _vp, i, _vp, i, d, _vp, i, i, _vp]
        L.HCORE_hip_gemm_split.argtypes = [i, i, i, i, i, i, d, _vp, i, _vp, i, d, _vp, i, _vp]
        L.HCORE_hip_plan_set_slices.argtypes = [_vp, i]
        L.HCORE_hip_plan_get_slices.argtypes = [_vp, _ip, i]
        L.HCORE_hip_plan_autotune_slices.argtypes = [_vp, _vp, d, _dp, i]
        L.flop_counts.argtypes = [C.c_char, C.c_ulong, C.c_ulong, C.c_ulong, C.c_ulong]
        L.flop_counts.restype = C.c_ulong
        L.HCORE_hip_init.argtypes = [i]
        L.HCORE_hip_set_error_policy.argtypes = [i]
        L.HCORE_hip_plan_create.argtypes = [C.POINTER(_vp), i, i]
        L.HCORE_hip_plan_add_dgemm.argtypes = [_vp, i, _pp, _pp, _ip, i, _pp, _pp, _ip, i, _pp,
                                               _pp, _ip, i, _pp, _pp, i, d, d, i, i, d]
        L.HCORE_hip_plan_add_dsyrk.argtypes = [_vp, i, _pp, _pp, _ip, i, _pp, i, d]
        L.HCORE_hip_plan_finalize.argtypes = [_vp]
        L.HCORE_hip_plan_execute.argtypes = [_vp, _vp]
        L.HCORE_hip_plan_execute_graph.argtypes = [_vp, _vp]
        L.HCORE_hip_plan_ranks_device.argtypes = [_vp]
        L.HCORE_hip_plan_ranks_device.restype = _vp
        L.HCORE_hip_plan_get_ranks.argtypes = [_vp, _ip]
        L.HCORE_hip_plan_get_status.argtypes = [_vp, _ip]
        L.HCORE_hip_plan_get_sigma.argtypes = [_vp, i, _dp, _ip]
        L.HCORE_hip_plan_get_sweeps.argtypes = [_vp, _ip]
        L.HCORE_hip_plan_get_core_profile.argtypes = [_vp, C.POINTER(C.c_longlong), _dp]
        L.HCORE_hip_plan_get_lanes.argtypes = [_vp, _ip]
        L.HCORE_hip_plan_get_core_queue.argtypes = [_vp, i, _ip, i]
        L.HCORE_hip_plan_workspace_bytes.argtypes = [_vp]
        L.HCORE_hip_plan_workspace_bytes.restype = C.c_longlong
        L.HCORE_hip_plan_flops.argtypes = [_vp, C.POINTER(C.c_ulong), C.POINTER(C.c_ulong)]
        L.HCORE_hip_plan_autotune.argtypes = [_vp, _vp, i, _dp, i]
        L.HCORE_hip_plan_add_dgemm_dev.argtypes = [_vp, i, _pp, _pp, _pp, _ip, i, _pp, _pp, _pp, _ip,
                                                   i, _pp, _pp, _pp, _ip, i, _pp, _pp, i, d, d, i,
                                                   i, d]
        L.HCORE_hip_plan_add_dsyrk_dev.argtypes = [_vp, i, _pp, _pp, _pp, _ip, i, _pp, i, d]
        L.HCORE_hip_plan_destroy.argtypes = [_vp]
        L.HCORE_hip_plan_destroy.restype = None
        L.HCORE_hip_plan_set_timing.argtypes = [_vp, i]
        L.HCORE_hip_plan_phase_times.argtypes = [_vp, _dp, _ip]
        sz = C.c_size_t
        L.HCORE_hip_malloc.argtypes = [C.POINTER(_vp), sz]
        L.HCORE_hip_free.argtypes = [_vp]
        L.HCORE_hip_memcpy_h2d.argtypes = [_vp, _vp, sz]
        L.HCORE_hip_memcpy_d2h.argtypes = [_vp, _vp, sz]
        L.HCORE_hip_memcpy_d2d.argtypes = [_vp, _vp, sz]
        L.HCORE_hip_memset.argtypes = [_vp, i, sz]
        L.HCORE_hip_stream_create.argtypes = [C.POINTER(_vp)]
        L.HCORE_hip_stream_destroy.argtypes = [_vp]
        L.HCORE_hip_stream_synchronize.argtypes = [_vp]
        L.HCORE_hip_event_create.argtypes = [C.POINTER(_vp)]
        L.HCORE_hip_event_destroy.argtypes = [_vp]
        L.HCORE_hip_event_record.argtypes = [_vp, _vp]
        L.HCORE_hip_event_elapsed_ms.argtypes = [_vp, _vp, C.POINTER(C.c_float)]
        L.HCORE_hip_device_info.argtypes = [C.c_char_p, i, _ip]
        L.HCORE_hip_mem_info.argtypes = [C.POINTER(sz), C.POINTER(sz)]
        _lib = L
    return _lib


def header_functions():
    """Names of every function declared in include/*.h."""
    names = []
    for fn in sorted(os.listdir(INCLUDE_DIR)):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(INCLUDE_DIR, fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M):
            name = m.group(1)
            if name in ("if", "while", "for", "defined") or name.startswith("__"):
                continue
            names.append(name)
    return sorted(set(names))


def set_error_policy(policy):
    return lib().HCORE_hip_set_error_policy(policy)


def last_status():
    return lib().HCORE_hip_last_status()


def flop_counts(op, a, b=0, c=0, d=0):
    return int(lib().flop_counts(op.encode(), a, b, c, d))


def _fortran(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def _ptr(a):
    return a.ctypes.data_as(_fp if a.dtype == np.float32 else _dp)


def dgemm(alpha, AU, AV, BU, BV, beta, CU, CV, rk=0, maxrk=None, acc=1e-8, dtype=np.float64,
          flops=None, ldc=None, device=False):
    """HCORE_dgemm (or HCORE_sgemm for dtype float32) on numpy factors.

    Returns (CU_new, CV_new, rank).  CU/CV are copied into buffers of ldc
    (default M) rows with room for K = rank(C) + rank(A) columns, updated in
    place by the library.  device=True passes device copies of all six
    factors (the library's device-pointer path) instead of host arrays.
    """
    M = AU.shape[0]
    ra, rb, rc = AU.shape[1], BU.shape[1], CU.shape[1]
    if maxrk is None:
        maxrk = M
    if ldc is None:
        ldc = M
    cap = max(ra + rc, 1)
    AU, AV, BU, BV = (_fortran(x, dtype) for x in (AU, AV, BU, BV))
    CUw = np.zeros((ldc, cap), dtype=dtype, order="F")
    CVw = np.zeros((ldc, cap), dtype=dtype, order="F")
    CUw[:M, :rc] = CU
    CVw[:M, :rc] = CV
    rks = np.array([ra, rb, rc], dtype=dtype)
    fc = flops if flops is not None else FlopCounter()
    L = lib()
    fn = L.HCORE_dgemm if dtype == np.float64 else L.HCORE_sgemm
    ptr_t = _fp if np.dtype(dtype) == np.float32 else _dp
    if device:
        bufs = [DeviceBuffer.from_numpy(x) for x in (AU, AV, BU, BV, CUw, CVw)]
        P = [C.cast(C.c_void_p(b.ptr), ptr_t) for b in bufs]
    else:
        P = [_ptr(x) for x in (AU, AV, BU, BV, CUw, CVw)]
    fn(111, 112, M, M, alpha, P[0], P[1], _ptr(rks[0:1]), M, P[2], P[3],
       _ptr(rks[1:2]), M, beta, P[4], P[5], _ptr(rks[2:3]), ldc, rk, maxrk, acc, None,
       C.byref(fc))
    r = int(rks[2])
    if device:
        CUw, CVw = bufs[4].numpy(), bufs[5].numpy()
        for b in bufs:
            b.free()
    return CUw[:M, :r].copy(order="F"), CVw[:M, :r].copy(order="F"), r


def dsyrk(alpha, AU, AV, CD, flops=None):
    """HCORE_dsyrk on host numpy arrays; returns the updated CD."""
    AU, AV = _fortran(AU, np.float64), _fortran(AV, np.float64)
    CDw = np.array(CD, dtype=np.float64, order="F", copy=True)
    M, K = AU.shape
    fc = flops if flops is not None else FlopCounter()
    lib().HCORE_dsyrk(122, 111, M, K, alpha, _ptr(AU), M, _ptr(AV), M, 1.0, _ptr(CDw), M,
                      None, C.byref(fc))
    return CDw


def dgemmcd(Au, Av, Bu, Bv, CD):
    """HCORE_dgemmcd on host numpy arrays; returns CD - (Au Av^T)(Bu Bv^T)^T."""
    Au, Av, Bu, Bv = (_fortran(x, np.float64) for x in (Au, Av, Bu, Bv))
    CDw = np.array(CD, dtype=np.float64, order="F", copy=True)
    M = Au.shape[0]
    fc = FlopCounter()
    lib().HCORE_dgemmcd(111, 112, M, -1.0, _ptr(Au), _ptr(Av), Au.shape[1], M, _ptr(Bu),
                        _ptr(Bv), Bu.shape[1], M, 1.0, _ptr(CDw), M, M, None, C.byref(fc))
    return CDw


def dpotrf(uplo, A, flops=None):
    """HCORE_dpotrf on a host copy of A; returns (factor, info)."""
    Aw = np.array(A, dtype=np.float64, order="F", copy=True)
    info = C.c_int(0)
    fc = flops if flops is not None else FlopCounter()
    lib().HCORE_dpotrf(uplo, Aw.shape[0], _ptr(Aw), Aw.shape[0], C.byref(info), C.byref(fc))
    return Aw, info.value


def dtrsm(side, uplo, trans, diag, alpha, A, B, flops=None):
    """HCORE_dtrsm on host copies; returns X (B is M x Brank)."""
    A = _fortran(A, np.float64)
    Bw = np.array(B, dtype=np.float64, order="F", copy=True)
    fc = flops if flops is not None else FlopCounter()
    lib().HCORE_dtrsm(side, uplo, trans, diag, Bw.shape[0], Bw.shape[1], alpha, _ptr(A),
                      A.shape[0], _ptr(Bw), Bw.shape[0], C.byref(fc))
    return Bw


def dgemmbdcd(transA, alpha, AU, AV, BD, beta, CD, flops=None):
    """HCORE_dgemmbdcd on host copies (ld = M throughout); returns CD_new."""
    AU, AV, BD = (_fortran(x, np.float64) for x in (AU, AV, BD))
    CDw = np.array(CD, dtype=np.float64, order="F", copy=True)
    M, N = CDw.shape
    rk = np.array([AU.shape[1]], dtype=np.float64)
    fc = flops if flops is not None else FlopCounter()
    lib().HCORE_dgemmbdcd(transA, 111, M, N, alpha, _ptr(AU), _ptr(AV), _ptr(rk), M, _ptr(BD),
                          M, beta, _ptr(CDw), M, None, C.byref(fc))
    return CDw


def duncompress(AU, BV):
    """HCORE_duncompress on host arrays; returns AU BV^T."""
    AU, BV = _fortran(AU, np.float64), _fortran(BV, np.float64)
    M, N = AU.shape[0], BV.shape[0]
    CD = np.zeros((M, N), order="F")
    ra = np.array([AU.shape[1]], dtype=np.float64)
    rb = np.array([BV.shape[1]], dtype=np.float64)
    lib().HCORE_duncompress(111, 112, M, N, 1.0, _ptr(AU), _ptr(ra), M, _ptr(BV), _ptr(rb), N,
                            0.0, _ptr(CD), M)
    return CD


def dcompress(A, maxrank=None, acc=1e-8):
    """HCORE_dcompress on a host tile A (m x n); returns (U, V, rank) with
    U (m x r) the left singular vectors and V (n x r) the sigma-scaled right
    ones."""
    A = _fortran(A, np.float64)
    m, n = A.shape
    mn = min(m, n)
    maxrank = mn if maxrank is None else maxrank
    U = np.zeros((m, mn), order="F")
    V = np.zeros((n, mn), order="F")
    rk = np.zeros(1)
    lib().HCORE_dcompress(m, n, _ptr(A), m, _ptr(U), m, _ptr(V), n, _ptr(rk), maxrank, acc)
    r = int(rk[0])
    return np.asfortranarray(U[:, :r]), np.asfortranarray(V[:, :r]), r


def dgytlr(m, n, bigM, m0, n0, seed, maxrank=None, tol=1e-8, compress_diag=False, dense=False):
    """HCORE_dgytlr into host arrays.  Returns (U, V, rank, D): for a diagonal
    tile without compress_diag U, V are None, rank is None and D the dense
    tile; otherwise D is the dense tile when `dense`, else None."""
    mn = min(m, n)
    maxrank = mn if maxrank is None else maxrank
    U = np.zeros((m, mn), order="F")
    V = np.zeros((n, mn), order="F")
    AD = np.zeros((m, n), order="F")
    D = np.zeros((m, n), order="F") if dense else None
    rk = np.array([-1.0])
    lib().HCORE_dgytlr(m, n, _ptr(U), _ptr(V), _ptr(AD), _ptr(rk), m, m, n, bigM, m0, n0, seed,
                       maxrank, tol, int(compress_diag), _ptr(D) if D is not None else None)
    if m0 == n0 and not compress_diag:
        return None, None, None, AD
    r = int(rk[0])
    return np.asfortranarray(U[:, :r]), np.asfortranarray(V[:, :r]), r, D


def gytlr_batch(m, n, m0, n0, bigM, seed, AU, ldu, AV, ldv, rank_ptr, AD=None, lda=0,
                maxrank=None, tol=1e-8, stream=None):
    """HCORE_hip_gytlr_batch on device pointers (lists of addresses; AD
    entries None/0 for compressed tiles).  rank_ptr: device int array."""
    k = len(m0)
    a0, p0 = _iarr(m0)
    a1, p1 = _iarr(n0)
    ad = _parr([x or 0 for x in AD]) if AD is not None else None
    st = lib().HCORE_hip_gytlr_batch(k, m, n, p0, p1, bigM, seed, _parr(AU), ldu, _parr(AV), ldv,
                                     rank_ptr, ad, lda, min(m, n) if maxrank is None else maxrank,
                                     tol, stream)
    return st


def compress_batch(m, n, A, lda, AU, ldu, AV, ldv, rank_ptr, maxrank=None, tol=1e-8, stream=None):
    """HCORE_hip_compress_batch on device pointers."""
    return lib().HCORE_hip_compress_batch(len(A), m, n, _parr(A), lda, _parr(AU), ldu, _parr(AV),
                                          ldv, rank_ptr, min(m, n) if maxrank is None else maxrank,
                                          tol, stream)


def tlr_dpotrf(nb, D, U, V, ranks, cap, maxrk, acc, stream=None, flops=None, precision=FP64,
               slices=0):
    """HCORE_hip_tlr_dpotrf (HCORE_hip_tlr_dpotrf_form for precision SPLIT):
    D (nt device addresses), U, V (nt*nt addresses, tile (i, j) at i + j*nt, 0
    above the diagonal), ranks (nt x nt ints, column-major, updated in
    place).  Returns (status, info, step_max)."""
    nt = len(D)
    rk = np.asfortranarray(ranks, dtype=np.int32)
    flat = np.ascontiguousarray(rk.reshape(-1, order="F"))
    info = C.c_int(0)
    smax = np.zeros(nt, dtype=np.int32)
    fc = flops if flops is not None else FlopCounter()
    if precision == FP64:
        st = lib().HCORE_hip_tlr_dpotrf(nt, nb, _parr(D), _parr(U), _parr(V),
                                        flat.ctypes.data_as(_ip), cap, maxrk, acc, stream,
                                        C.byref(info), C.byref(fc), smax.ctypes.data_as(_ip))
    else:
        st = lib().HCORE_hip_tlr_dpotrf_form(precision, slices, nt, nb, _parr(D), _parr(U),
                                             _parr(V), flat.ctypes.data_as(_ip), cap, maxrk, acc,
                                             stream, C.byref(info), C.byref(fc),
                                             smax.ctypes.data_as(_ip))
    ranks[...] = flat.reshape(nt, nt, order="F")
    return st, info.value, smax


def set_covariance(beta, nugget):
    return lib().HCORE_hip_set_covariance(beta, nugget)


def covariance_points(bigM, seed):
    """The covariance problem's Morton-sorted points, as (bigM, 2)."""
    xy = np.zeros(2 * bigM)
    _check(lib().HCORE_hip_covariance_points(bigM, seed, _ptr(xy)), "covariance_points")
    return np.stack([xy[:bigM], xy[bigM:]], axis=1)


def dgemm_split(alpha, AU, AV, BU, BV, beta, CU, CV, rk=0, maxrk=None, acc=1e-8,
                storage_maxrank=None, B_dense=None, flops=None, between=None):
    """HCORE_dgemm_qr_svd (or _b_dense with a dense B) then HCORE_dgemm_ormqr
    on host numpy factors.  Returns (CU_new, CV_new, rank); `between(r)` is
    called with the revealed rank between the two phases."""
    M = AU.shape[0]
    ra, rc = AU.shape[1], CU.shape[1]
    maxrk = M if maxrk is None else maxrk
    smr = (ra + rc) if storage_maxrank is None else storage_maxrank
    L = lib()
    f64 = lambda x: _fortran(x, np.float64)
    AU, AV = f64(AU), f64(AV)
    CUw = np.zeros((M, ra + rc), order="F")
    CVw = np.zeros((M, ra + rc), order="F")
    CUw[:, :rc], CVw[:, :rc] = CU, CV
    iA, iC = np.array([ra], np.int32), np.array([rc], np.int32)
    ip = lambda a: a.ctypes.data_as(_ip)
    fc = flops if flops is not None else FlopCounter()
    # the out-parameters in the reference's order: work_new, _CU, _CV, CU_ncols,
    # new_UVrk, newU, ld_newU, qrtauA, CV_ncols, newV, ld_newV, qrtauB,
    # use_CUV_clone, CUclone, ld_CUclone, _CU_save, CVclone, ld_CVclone, _CV_save
    keep = {}
    names = ["work_new", "_CU", "_CV", "CU_ncols", "new_UVrk", "newU", "ld_newU", "qrtauA",
             "CV_ncols", "newV", "ld_newV", "qrtauB", "use_CUV_clone", "CUclone", "ld_CUclone",
             "_CU_save", "CVclone", "ld_CVclone", "_CV_save"]
    for nm in names:
        keep[nm] = C.c_int(0) if nm in ("CU_ncols", "new_UVrk", "ld_newU", "CV_ncols",
                                         "ld_newV", "use_CUV_clone", "ld_CUclone",
                                         "ld_CVclone") else C.c_void_p()
    refs = [C.byref(keep[nm]) for nm in names]
    if B_dense is None:
        BU, BV = f64(BU), f64(BV)
        iB = np.array([BU.shape[1]], np.int32)
        L.HCORE_dgemm_qr_svd(111, 112, M, M, C.c_double(alpha), _ptr(AU), _ptr(AV), ip(iA), M,
                             _ptr(BU), _ptr(BV), ip(iB), M, C.c_double(beta), _ptr(CUw), _ptr(CVw),
                             ip(iC), M, rk, smr, maxrk, C.c_double(acc), None, C.byref(fc), *refs)
    else:
        Bd = f64(B_dense)
        L.HCORE_dgemm_qr_svd_b_dense(111, 112, M, M, C.c_double(alpha), _ptr(AU), _ptr(AV),
                                     ip(iA), M, _ptr(Bd), M, C.c_double(beta), _ptr(CUw),
                                     _ptr(CVw), ip(iC), M, rk, smr, maxrk, C.c_double(acc), None,
                                     C.byref(fc), *refs)
    st = last_status()
    if st:
        return None, None, -st
    r = keep["new_UVrk"].value
    if between is not None:
        between(r)
    v = lambda nm: keep[nm].value if isinstance(keep[nm], C.c_int) else C.c_void_p(keep[nm].value)
    L.HCORE_dgemm_ormqr(111, 112, M, M, C.c_double(alpha), _ptr(AU), _ptr(AV), ip(iA), M,
                        C.c_double(beta), _ptr(CUw), _ptr(CVw), ip(iC), M, rk, smr, maxrk,
                        C.c_double(acc), None, C.byref(fc), *[v(nm) for nm in names[1:]])
    st = last_status()
    if st:
        return None, None, -st
    r = int(iC[0])
    return CUw[:, :r].copy(order="F"), CVw[:, :r].copy(order="F"), r


def gemm_split(A, B, C_in, slices, alpha=1.0, beta=0.0, transA=False, transB=False):
    """HCORE_hip_gemm_split on device copies of float64 numpy arrays (slices
    BF16 digits per operand; 0: the same tiles on the FP64 matrix cores);
    returns the new C."""
    a = DeviceBuffer.from_numpy(np.asfortranarray(A, dtype=np.float64))
    b = DeviceBuffer.from_numpy(np.asfortranarray(B, dtype=np.float64))
    c = DeviceBuffer.from_numpy(np.asfortranarray(C_in, dtype=np.float64))
    m, n = C_in.shape
    k = A.shape[0] if transA else A.shape[1]
    st = lib().HCORE_hip_gemm_split(slices, 112 if transA else 111, 112 if transB else 111, m, n,
                                    k, alpha, a.ptr, A.shape[0], b.ptr, B.shape[0], beta, c.ptr,
                                    m, None)
    _check(st, "HCORE_hip_gemm_split")
    out = c.numpy()
    for x in (a, b, c):
        x.free()
    return out


def gemm(A, B, C_in, alpha=1.0, beta=0.0, transA=False, transB=False, path=0):
    """HCORE_hip_gemm on device copies of numpy arrays (float64 or float32);
    returns the new C."""
    dt = C_in.dtype
    a = DeviceBuffer.from_numpy(np.asfortranarray(A, dtype=dt))
    b = DeviceBuffer.from_numpy(np.asfortranarray(B, dtype=dt))
    c = DeviceBuffer.from_numpy(np.asfortranarray(C_in, dtype=dt))
    m, n = C_in.shape
    k = A.shape[0] if transA else A.shape[1]
    st = lib().HCORE_hip_gemm(FP64 if dt == np.float64 else FP32, 112 if transA else 111,
                              112 if transB else 111, m, n, k, alpha, a.ptr, A.shape[0], b.ptr,
                              B.shape[0], beta, c.ptr, m, path, None)
    _check(st, "HCORE_hip_gemm")
    out = c.numpy()
    for x in (a, b, c):
        x.free()
    return out


def _parr(ptrs):
    return (C.c_void_p * max(len(ptrs), 1))(*[int(p) for p in ptrs])


def _iarr(v):
    a = np.ascontiguousarray(v, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def _check(st, what):
    if st:
        raise RuntimeError("%s failed: status %d" % (what, st))


class DeviceBuffer:
    """Device allocation of the library's own HIP runtime.

    Holds a column-major (rows x cols) matrix of `dtype`; `ptr` is the device
    address to pass to plans.  There is deliberately no torch here: the
    library must not share a process with a second HIP runtime."""

    def __init__(self, rows, cols=1, dtype=np.float64):
        self.rows, self.cols, self.dtype = rows, max(cols, 1), np.dtype(dtype)
        self.nbytes = self.rows * self.cols * self.dtype.itemsize
        p = C.c_void_p()
        _check(lib().HCORE_hip_malloc(C.byref(p), max(self.nbytes, 8)), "HCORE_hip_malloc")
        self.ptr = p.value
        self.zero()

    @classmethod
    def from_numpy(cls, x, cap=None):
        x = np.asarray(x)
        rows = x.shape[0]
        cols = x.shape[1] if x.ndim > 1 else 1
        b = cls(rows, cols if cap is None else max(cap, cols), x.dtype)
        b.upload(x)
        return b

    def upload(self, x):
        x = np.asfortranarray(x, dtype=self.dtype)
        assert x.nbytes <= self.nbytes
        _check(lib().HCORE_hip_memcpy_h2d(self.ptr, x.ctypes.data, x.nbytes), "h2d")

    def numpy(self, cols=None):
        cols = self.cols if cols is None else cols
        out = np.zeros((self.rows, cols), dtype=self.dtype, order="F")
        if cols:
            _check(lib().HCORE_hip_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes), "d2h")
        return out

    def copy_from(self, other):
        assert other.nbytes <= self.nbytes
        _check(lib().HCORE_hip_memcpy_d2d(self.ptr, other.ptr, other.nbytes), "d2d")

    def zero(self):
        _check(lib().HCORE_hip_memset(self.ptr, 0, self.nbytes), "memset")

    def free(self):
        if self.ptr:
            lib().HCORE_hip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        h = C.c_void_p()
        _check(lib().HCORE_hip_stream_create(C.byref(h)), "stream_create")
        self.h = h.value

    def synchronize(self):
        _check(lib().HCORE_hip_stream_synchronize(self.h), "stream_synchronize")


class Event:
    def __init__(self):
        h = C.c_void_p()
        _check(lib().HCORE_hip_event_create(C.byref(h)), "event_create")
        self.h = h.value

    def record(self, stream):
        _check(lib().HCORE_hip_event_record(self.h, stream.h if stream else None), "record")

    def elapsed_ms(self, end):
        ms = C.c_float(0)
        _check(lib().HCORE_hip_event_elapsed_ms(self.h, end.h, C.byref(ms)), "elapsed")
        return float(ms.value)


def device_synchronize():
    _check(lib().HCORE_hip_device_synchronize(), "device_synchronize")


def mem_info():
    """(free, total) device memory in bytes."""
    f, t = C.c_size_t(0), C.c_size_t(0)
    _check(lib().HCORE_hip_mem_info(C.byref(f), C.byref(t)), "HCORE_hip_mem_info")
    return f.value, t.value


def device_info():
    name = C.create_string_buffer(256)
    ncu = C.c_int(0)
    _check(lib().HCORE_hip_device_info(name, 256, C.byref(ncu)), "device_info")
    return name.value.decode(), ncu.value


class Plan:
    """Batched plan of one TLR trailing step (include/hcore_hip.h).

    Pointers are integer device addresses (DeviceBuffer.ptr).
    """

    def __init__(self, M, precision=FP64):
        self.L = lib()
        h = C.c_void_p()
        st = self.L.HCORE_hip_plan_create(C.byref(h), M, precision)
        if st:
            raise RuntimeError("HCORE_hip_plan_create failed: %d" % st)
        self.h = h
        self.M = M
        self._keep = []
        self.n = 0

    def add_dgemm(self, AU, AV, ra, lda, BU, BV, rb, ldb, CU, CV, rc, ldc, alpha, beta, rk,
                  maxrk, acc, CUout=None, CVout=None, ldco=None):
        n = len(AU)
        if ldco is None:
            ldco = self.M
        arrs = [_parr(x) for x in (AU, AV, BU, BV, CU, CV)]
        (ra_a, ra_p), (rb_a, rb_p), (rc_a, rc_p) = _iarr(ra), _iarr(rb), _iarr(rc)
        cuo = _parr(CUout) if CUout is not None else None
        cvo = _parr(CVout) if CVout is not None else None
        self._keep += arrs + [ra_a, rb_a, rc_a, cuo, cvo]
        st = self.L.HCORE_hip_plan_add_dgemm(self.h, n, arrs[0], arrs[1], ra_p, lda, arrs[2],
                                             arrs[3], rb_p, ldb, arrs[4], arrs[5], rc_p, ldc,
                                             cuo, cvo, ldco, alpha, beta, rk, maxrk, acc)
        if st:
            raise RuntimeError("HCORE_hip_plan_add_dgemm failed: %d" % st)
        self.n = n

    def add_dsyrk(self, AU, AV, K, ldau, CD, ldcd, alpha):
        n = len(AU)
        arrs = [_parr(x) for x in (AU, AV, CD)]
        k_a, k_p = _iarr(K)
        self._keep += arrs + [k_a]
        st = self.L.HCORE_hip_plan_add_dsyrk(self.h, n, arrs[0], arrs[1], k_p, ldau, arrs[2],
                                             ldcd, alpha)
        if st:
            raise RuntimeError("HCORE_hip_plan_add_dsyrk failed: %d" % st)

    def finalize(self):
        st = self.L.HCORE_hip_plan_finalize(self.h)
        if st:
            raise RuntimeError("HCORE_hip_plan_finalize failed: %d" % st)

    def execute(self, stream=None, graph=False):
        fn = self.L.HCORE_hip_plan_execute_graph if graph else self.L.HCORE_hip_plan_execute
        st = fn(self.h, C.c_void_p(stream.h if stream is not None else None))
        if st:
            raise RuntimeError("plan execute failed: %d" % st)

    def ranks(self):
        out = np.zeros(max(self.n, 1), dtype=np.int32)
        st = self.L.HCORE_hip_plan_get_ranks(self.h, out.ctypes.data_as(_ip))
        if st:
            raise RuntimeError("get_ranks failed: %d" % st)
        return out[:self.n]

    def status(self):
        out = np.zeros(max(self.n, 1), dtype=np.int32)
        st = self.L.HCORE_hip_plan_get_status(self.h, out.ctypes.data_as(_ip))
        if st:
            raise RuntimeError("get_status failed: %d" % st)
        return out[:self.n]

    def lanes(self):
        out = np.zeros(max(self.n, 1), dtype=np.int32)
        _check(self.L.HCORE_hip_plan_get_lanes(self.h, out.ctypes.data_as(_ip)), "get_lanes")
        return out[:self.n]

    def sweeps(self):
        out = np.zeros(max(self.n, 1), dtype=np.int32)
        _check(self.L.HCORE_hip_plan_get_sweeps(self.h, out.ctypes.data_as(_ip)), "get_sweeps")
        return out[:self.n]

    def core_profile(self):
        """(n, 5) array per update: microseconds in the pivoted QR, the Jacobi
        sweeps and the epilogue, the whole core kernel, and the QR's numerical
        rank (zeros for updates whose core kernel records no stamps)."""
        raw = np.zeros((max(self.n, 1), 8), dtype=np.int64)
        tick = C.c_double(0)
        _check(self.L.HCORE_hip_plan_get_core_profile(
            self.h, raw.ctypes.data_as(C.POINTER(C.c_longlong)), C.byref(tick)), "core_profile")
        raw = raw[:self.n]
        us = tick.value / 1e3
        out = np.zeros((self.n, 5))
        out[:, 0] = (raw[:, 1] - raw[:, 0]) * us
        out[:, 1] = (raw[:, 2] - raw[:, 1]) * us
        out[:, 2] = (raw[:, 3] - raw[:, 2]) * us
        out[:, 3] = (raw[:, 3] - raw[:, 0]) * us
        out[:, 4] = raw[:, 4]
        return out

    def sigma(self, u):
        K = C.c_int(0)
        self.L.HCORE_hip_plan_get_sigma(self.h, u, None, C.byref(K))
        out = np.zeros(max(K.value, 1))
        self.L.HCORE_hip_plan_get_sigma(self.h, u, out.ctypes.data_as(_dp), C.byref(K))
        return out[:K.value]

    def ranks_device_ptr(self):
        return self.L.HCORE_hip_plan_ranks_device(self.h)

    def set_timing(self, enable=True):
        _check(self.L.HCORE_hip_plan_set_timing(self.h, int(enable)), "set_timing")

    def phase_times(self):
        """(ms, count) arrays of shape (lanes=4, phases=12), accumulated."""
        ms = np.zeros(4 * 12)
        cnt = np.zeros(4 * 12, dtype=np.int32)
        _check(self.L.HCORE_hip_plan_phase_times(self.h, ms.ctypes.data_as(_dp),
                                                 cnt.ctypes.data_as(_ip)), "phase_times")
        return ms.reshape(4, 12), cnt.reshape(4, 12)

    def autotune(self, stream=None, reps=3):
        """HCORE_hip_plan_autotune: list of (lane, base variant, chosen variant,
        ms base, ms alternative, tiles) per tuned core launch."""
        rep = np.zeros(6 * 32)
        n = self.L.HCORE_hip_plan_autotune(self.h, C.c_void_p(stream.h if stream else None), reps,
                                           rep.ctypes.data_as(_dp), 32)
        if n < 0:
            raise RuntimeError("autotune failed: %d" % -n)
        return [tuple(rep[6 * k:6 * k + 6]) for k in range(min(n, 32))]

    def core_queue(self, lane):
        """HCORE_hip_plan_get_core_queue: the lane's update order of the last
        execution (device-rank plans; None for a host-rank lane)."""
        out = np.zeros(max(self.n, 1), dtype=np.int32)
        k = self.L.HCORE_hip_plan_get_core_queue(self.h, lane, out.ctypes.data_as(_ip), len(out))
        if k < 0:
            raise RuntimeError("HCORE_hip_plan_get_core_queue failed: %d" % -k)
        return out[:k].copy() if k else None

    def set_slices(self, slices):
        """HCORE_hip_plan_set_slices (split-precision plans; 0: from acc)."""
        _check(self.L.HCORE_hip_plan_set_slices(self.h, slices), "HCORE_hip_plan_set_slices")

    def slices(self):
        """HCORE_hip_plan_get_slices: digit count per lane."""
        out = np.zeros(8, dtype=np.int32)
        n = self.L.HCORE_hip_plan_get_slices(self.h, out.ctypes.data_as(_ip), 8)
        if n < 0:
            raise RuntimeError("HCORE_hip_plan_get_slices failed: %d" % -n)
        return out[:n].copy()

    def autotune_slices(self, stream=None, tol=0.1):
        """HCORE_hip_plan_autotune_slices: list of (lane, chosen digits,
        deviation / acc at that count, deviation / acc with one digit fewer)."""
        rep = np.zeros(4 * 8)
        n = self.L.HCORE_hip_plan_autotune_slices(self.h, C.c_void_p(stream.h if stream else None),
                                                  tol, rep.ctypes.data_as(_dp), 8)
        if n < 0:
            raise RuntimeError("autotune_slices failed: %d" % -n)
        return [tuple(rep[4 * k:4 * k + 4]) for k in range(min(n, 8))]

    def add_dgemm_dev(self, AU, AV, ra_ptrs, ra_cap, lda, BU, BV, rb_ptrs, rb_cap, ldb, CU, CV,
                      rc_ptrs, rc_cap, ldc, alpha, beta, rk, maxrk, acc, CUout=None, CVout=None,
                      ldco=None):
        """HCORE_hip_plan_add_dgemm_dev: ranks read on the device from the given
        int addresses, capacities as host ints."""
        n = len(AU)
        ldco = self.M if ldco is None else ldco
        arrs = [_parr(x) for x in (AU, AV, ra_ptrs, BU, BV, rb_ptrs, CU, CV, rc_ptrs)]
        caps = [_iarr(x) for x in (ra_cap, rb_cap, rc_cap)]
        cuo = _parr(CUout) if CUout is not None else None
        cvo = _parr(CVout) if CVout is not None else None
        self._keep += arrs + [c[0] for c in caps] + [cuo, cvo]
        st = self.L.HCORE_hip_plan_add_dgemm_dev(
            self.h, n, arrs[0], arrs[1], arrs[2], caps[0][1], lda, arrs[3], arrs[4], arrs[5],
            caps[1][1], ldb, arrs[6], arrs[7], arrs[8], caps[2][1], ldc, cuo, cvo, ldco, alpha,
            beta, rk, maxrk, acc)
        if st:
            raise RuntimeError("HCORE_hip_plan_add_dgemm_dev failed: %d" % st)
        self.n = n

    def add_dsyrk_dev(self, AU, AV, k_ptrs, k_cap, ldau, CD, ldcd, alpha):
        n = len(AU)
        arrs = [_parr(x) for x in (AU, AV, k_ptrs, CD)]
        kc = _iarr(k_cap)
        self._keep += arrs + [kc[0]]
        st = self.L.HCORE_hip_plan_add_dsyrk_dev(self.h, n, arrs[0], arrs[1], arrs[2], kc[1], ldau,
                                                 arrs[3], ldcd, alpha)
        if st:
            raise RuntimeError("HCORE_hip_plan_add_dsyrk_dev failed: %d" % st)

    def workspace_bytes(self):
        return int(self.L.HCORE_hip_plan_workspace_bytes(self.h))

    def flops(self):
        fu, fs = C.c_ulong(0), C.c_ulong(0)
        st = self.L.HCORE_hip_plan_flops(self.h, C.byref(fu), C.byref(fs))
        if st:
            raise RuntimeError("plan flops failed: %d" % st)
        return fu.value, fs.value

    def destroy(self):
        if self.h:
            self.L.HCORE_hip_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
