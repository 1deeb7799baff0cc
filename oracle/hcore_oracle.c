/*
 * hcore_oracle.c — CPU ORACLE for the HCORE tile low-rank hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the GPU library is
 * compared against and the timed CPU baseline of bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product (hcore_amd/) never links or calls it.
 *
 * It restates the reference's algorithm with the same LAPACK call sequence
 * (geqrf x2, gemm skinny x2, gemm core, gesvd, truncation, ormqr x2), linked
 * to the LAPACK/BLAS that ships inside this image (scipy-openblas 0.3.28,
 * whose Fortran entry points carry a "scipy_" prefix).  The reference itself
 * cannot be compiled here: its sources include <cblas.h> and <lapacke.h>,
 * which the image lacks, so every function below is a restatement.
 *
 * Functions and the reference lines they follow:
 *   orc_flop_counts   src/misc/descutil/flop_counts.c:3-70
 *   orc_dgemm         src/compute/hcore_dgemm.c:703-1012 with __qra 79-137,
 *                     __qrb 149-225, __svd 296-503, __newu 508-617,
 *                     __newv 622-696
 *   orc_sgemm         src/compute/hcore_sgemm.c:703 (FP32 twin)
 *   orc_dsyrk         src/compute/hcore_dsyrk.c:47-189
 *   orc_dgemmcd       src/compute/hcore_dgemmcd.c:36-106
 *   orc_generate_compress / orc_generate_scompress  testing/testing_dgemm.c:119-236,
 *                     testing/testing_sgemm.c:119 (the reference tests' inputs)
 *   orc_dgemm_batch / orc_dsyrk_batch  the CPU baseline: one update per
 *                     OpenMP thread, single-threaded BLAS per call.
 *   orc_dpotrf        src/compute/hcore_dpotrf.c:23-34 (uplo as the intended
 *                     triangle: HCORE 121/'U' upper, 122/'L' lower)
 *   orc_dtrsm         src/compute/hcore_dtrsm.c:24-48
 *   orc_dgemmbdcd     src/compute/hcore_dgemmbdcd.c:35-122
 *   orc_duncompress   src/compute/hcore_duncompress.c:33-84
 *   orc_dcompress     testing/testing_dgemm.c:180-226 (compression step of
 *                     the tile generator)
 *   orc_covariance_points / orc_covariance_tile  the covariance problem of
 *                     HCORE_dgytlr (declared-only HCORE_zgytlr,
 *                     src/include/hcore_z.h:63-75; problem defined here)
 *
 * Parity pins (tests/test_oracle.py): the reference's own assertions
 * (testing_dgemm.c:377-378, testing_sgemm.c:378-379, testing_dsyrk.c:181,
 * testing_dgemmcd.c:205) on the reference's own generated inputs; the dense
 * product and the truncation rank recomputed from numpy's SVD of the dense
 * update; sigma against numpy; planted mistakes (see ORC_MUTATE_*) that the
 * pins must catch.  No function here is "parity unpinned".
 *
 * Error paths: where the reference prints and calls exit(-1), the oracle
 * returns a status code instead (ORC_ERR_*), so tests can check each path.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- LAPACK / BLAS Fortran entry points of scipy-openblas (LP64) ------- */
typedef int lint;
extern void scipy_dgeqrf_(const lint *m, const lint *n, double *a, const lint *lda,
                          double *tau, double *work, const lint *lwork, lint *info);
extern void scipy_sgeqrf_(const lint *m, const lint *n, float *a, const lint *lda,
                          float *tau, float *work, const lint *lwork, lint *info);
extern void scipy_dormqr_(const char *side, const char *trans, const lint *m,
                          const lint *n, const lint *k, const double *a, const lint *lda,
                          const double *tau, double *c, const lint *ldc, double *work,
                          const lint *lwork, lint *info, size_t, size_t);
extern void scipy_sormqr_(const char *side, const char *trans, const lint *m,
                          const lint *n, const lint *k, const float *a, const lint *lda,
                          const float *tau, float *c, const lint *ldc, float *work,
                          const lint *lwork, lint *info, size_t, size_t);
extern void scipy_dgesvd_(const char *jobu, const char *jobvt, const lint *m,
                          const lint *n, double *a, const lint *lda, double *s,
                          double *u, const lint *ldu, double *vt, const lint *ldvt,
                          double *work, const lint *lwork, lint *info, size_t, size_t);
extern void scipy_sgesvd_(const char *jobu, const char *jobvt, const lint *m,
                          const lint *n, float *a, const lint *lda, float *s,
                          float *u, const lint *ldu, float *vt, const lint *ldvt,
                          float *work, const lint *lwork, lint *info, size_t, size_t);
extern void scipy_dgemm_(const char *ta, const char *tb, const lint *m, const lint *n,
                         const lint *k, const double *alpha, const double *a,
                         const lint *lda, const double *b, const lint *ldb,
                         const double *beta, double *c, const lint *ldc, size_t, size_t);
extern void scipy_sgemm_(const char *ta, const char *tb, const lint *m, const lint *n,
                         const lint *k, const float *alpha, const float *a,
                         const lint *lda, const float *b, const lint *ldb,
                         const float *beta, float *c, const lint *ldc, size_t, size_t);
extern void scipy_dlatms_(const lint *m, const lint *n, const char *dist, lint *iseed,
                          const char *sym, double *d, const lint *mode, const double *cond,
                          const double *dmax, const lint *kl, const lint *ku,
                          const char *pack, double *a, const lint *lda, double *work,
                          lint *info, size_t, size_t, size_t);
extern void scipy_dpotrf_(const char *uplo, const lint *n, double *a, const lint *lda,
                          lint *info, size_t);
extern void scipy_dtrsm_(const char *side, const char *uplo, const char *transa,
                         const char *diag, const lint *m, const lint *n, const double *alpha,
                         const double *a, const lint *lda, double *b, const lint *ldb, size_t,
                         size_t, size_t, size_t);
extern void scipy_openblas_set_num_threads(int);
extern char *scipy_openblas_get_config(void);

/* ---- status codes (the reference's exit(-1) sites) --------------------- */
#define ORC_OK 0
#define ORC_ERR_ZERO_RANK 1     /* hcore_dgemm.c:743-746 */
#define ORC_ERR_SUM_RANKS 2     /* hcore_dgemm.c:91-95 (rc+ra > 2*maxrk) */
#define ORC_ERR_K_GT_MAXRK 3    /* hcore_dgemm.c:924-927 */
#define ORC_ERR_R_GT_MAXRK 4    /* hcore_dgemm.c:465-468 */
#define ORC_ERR_LAPACK 5        /* info != 0 anywhere */

/* Planted mistakes for the pin tests: each must make a pin fail. */
#define ORC_MUTATE_NONE 0
#define ORC_MUTATE_ALPHA_SIGN 1   /* drops the sign of alpha on the AU block */
#define ORC_MUTATE_NO_G_TRANSPOSE 2 /* G = BU*F instead of BU*F^T */
#define ORC_MUTATE_TRUNC_OFF_BY_ONE 3 /* keeps one singular value too many */
#define ORC_MUTATE_SIGMA_TO_U 4   /* folds sigma into U instead of V */

const char *orc_blas_config(void) { return scipy_openblas_get_config(); }
void orc_set_blas_threads(int n) { scipy_openblas_set_num_threads(n); }

/* ---- flop counts (flop_counts.c:3-70, same integer/float semantics) ---- */
unsigned long int orc_flop_counts(char op, unsigned long int a, unsigned long int b,
                                  unsigned long int c, unsigned long int d)
{
    unsigned long int r = 0;
    switch (op) {
    case 'q': /* geqrf m x n: 2mn^2 - 2n^3/3 + 2mn + 17n/3, thirds in float */
        r = 2 * a * b * b - (unsigned long int)(2 * b * b * b / 3.0f) + 2 * a * b +
            (unsigned long int)(17 * b / 3.0f);
        break;
    case 'c': r = a * a * a / 3 - a * a / 2.0 + a / 6; break;
    case 't': r = (c == 1) ? b * a * a : (c == 2 ? a * b * b : 0); break;
    case 'm': r = a * b * c * 2; break;
    case 's': r = 22 * a * a * a; break;
    case 'o':
        if (d == 1) r = 4 * b * a * c - 2 * b * c * c + 3 * b * c;
        else if (d == 2) r = 4 * b * a * c - 2 * a * c * c + 2 * a * c + b * c - c * c / 2 + c / 2;
        break;
    case 'r': r = (c == 1) ? b * a * a : (c == 2 ? a * b * b : 0); break;
    default: break;
    }
    return r;
}

/* ======================================================================
 * Generic (double / float) restatement of the recompression.  The macro
 * instantiates it twice so the FP32 twin issues the s-prefixed calls,
 * exactly as hcore_sgemm.c does.
 * ====================================================================== */
#define ORC_DEFINE_GEMM(NAME, REAL, GEQRF, ORMQR, GESVD, GEMM)                          \
static lint NAME##_geqrf(lint m, lint n, REAL *a, lint lda, REAL *tau)                 \
{                                                                                      \
    lint info = 0, lwork = -1;                                                         \
    REAL q;                                                                            \
    GEQRF(&m, &n, a, &lda, tau, &q, &lwork, &info);                                    \
    if (info) return info;                                                             \
    lwork = (lint)q; if (lwork < n) lwork = n;                                         \
    REAL *w = (REAL *)malloc(sizeof(REAL) * (size_t)lwork);                            \
    GEQRF(&m, &n, a, &lda, tau, w, &lwork, &info);                                     \
    free(w);                                                                           \
    return info;                                                                       \
}                                                                                      \
static lint NAME##_ormqr(char side, char trans, lint m, lint n, lint k, const REAL *a, \
                         lint lda, const REAL *tau, REAL *c, lint ldc)                 \
{                                                                                      \
    lint info = 0, lwork = -1;                                                         \
    REAL q;                                                                            \
    ORMQR(&side, &trans, &m, &n, &k, a, &lda, tau, c, &ldc, &q, &lwork, &info, 1, 1);  \
    if (info) return info;                                                             \
    lwork = (lint)q; if (lwork < 1) lwork = 1;                                         \
    REAL *w = (REAL *)malloc(sizeof(REAL) * (size_t)lwork);                            \
    ORMQR(&side, &trans, &m, &n, &k, a, &lda, tau, c, &ldc, w, &lwork, &info, 1, 1);   \
    free(w);                                                                           \
    return info;                                                                       \
}                                                                                      \
static lint NAME##_gesvd(lint n, REAL *a, lint lda, REAL *s, REAL *u, lint ldu,        \
                         REAL *vt, lint ldvt)                                          \
{                                                                                      \
    lint info = 0, lwork = -1;                                                         \
    REAL q;                                                                            \
    GESVD("A", "A", &n, &n, a, &lda, s, u, &ldu, vt, &ldvt, &q, &lwork, &info, 1, 1);  \
    if (info) return info;                                                             \
    lwork = (lint)q; if (lwork < 5 * n) lwork = 5 * n;                                 \
    REAL *w = (REAL *)malloc(sizeof(REAL) * (size_t)lwork);                            \
    GESVD("A", "A", &n, &n, a, &lda, s, u, &ldu, vt, &ldvt, w, &lwork, &info, 1, 1);   \
    free(w);                                                                           \
    return info;                                                                       \
}                                                                                      \
static void NAME##_gemm(char ta, char tb, lint m, lint n, lint k, REAL alpha,          \
                        const REAL *a, lint lda, const REAL *b, lint ldb, REAL beta,   \
                        REAL *c, lint ldc)                                             \
{                                                                                      \
    GEMM(&ta, &tb, &m, &n, &k, &alpha, a, &lda, b, &ldb, &beta, c, &ldc, 1, 1);        \
}                                                                                      \
/* Returns a status; on ORC_OK CU/CV hold the new factors and *rc the rank.   */      \
/* sigma_out (may be NULL) receives the K singular values of the core.        */      \
int NAME(int M, REAL alpha, const REAL *AU, const REAL *AV, int ra, int LDA,          \
         const REAL *BU, const REAL *BV, int rb, int LDB, REAL beta, REAL *CU,        \
         REAL *CV, int *rc_io, int LDC, int rk, int maxrk, REAL acc,                  \
         unsigned long *flops_update, REAL *sigma_out, int mutate)                    \
{                                                                                      \
    const int rc = *rc_io;                                                             \
    if (ra == 0 || rb == 0 || rc == 0) return ORC_ERR_ZERO_RANK;                       \
    if (ra + rc > 2 * maxrk) return ORC_ERR_SUM_RANKS;                                 \
    const int K = rc + ra;                                                             \
    if (maxrk < K) return ORC_ERR_K_GT_MAXRK;                                          \
    if (K > M) return ORC_ERR_K_GT_MAXRK;                                              \
    unsigned long fl = 0;                                                              \
    /* clones of C's factors with room for the stacked columns (760-799) */            \
    REAL *Us = (REAL *)calloc((size_t)M * K, sizeof(REAL));                            \
    REAL *Vs = (REAL *)calloc((size_t)M * K, sizeof(REAL));                            \
    REAL *tauA = (REAL *)calloc((size_t)K, sizeof(REAL));                              \
    REAL *tauB = (REAL *)calloc((size_t)K, sizeof(REAL));                              \
    REAL *F = (REAL *)calloc((size_t)ra * rb, sizeof(REAL));                           \
    REAL *Rb = (REAL *)calloc((size_t)K * K, sizeof(REAL));                            \
    REAL *Ra = (REAL *)calloc((size_t)K * K, sizeof(REAL));                            \
    REAL *T = (REAL *)calloc((size_t)K * K, sizeof(REAL));                             \
    REAL *Ut = (REAL *)calloc((size_t)M * K, sizeof(REAL));                            \
    REAL *Vt = (REAL *)calloc((size_t)K * M, sizeof(REAL));                            \
    REAL *sg = (REAL *)calloc((size_t)K, sizeof(REAL));                                \
    int st = ORC_OK;                                                                   \
    for (int j = 0; j < rc; ++j) {                                                     \
        memcpy(Us + (size_t)j * M, CU + (size_t)j * LDC, sizeof(REAL) * M);           \
        memcpy(Vs + (size_t)j * M, CV + (size_t)j * LDC, sizeof(REAL) * M);           \
    }                                                                                  \
    /* __qra: [beta*CU | alpha*AU] then geqrf (79-137) */                              \
    REAL asc = (mutate == ORC_MUTATE_ALPHA_SIGN) ? (REAL)fabs((double)alpha) : alpha;  \
    for (int j = 0; j < ra; ++j)                                                       \
        for (int i = 0; i < M; ++i)                                                    \
            Us[(size_t)(rc + j) * M + i] = asc * AU[(size_t)j * LDA + i];              \
    for (size_t e = 0; e < (size_t)M * rc; ++e) Us[e] *= beta;                         \
    if (NAME##_geqrf(M, K, Us, M, tauA)) { st = ORC_ERR_LAPACK; goto done; }           \
    fl += orc_flop_counts('q', M, K, 0, 0);                                            \
    /* __qrb: F = AV^T BV; G = BU F^T into V stack; geqrf (149-225) */                 \
    NAME##_gemm('T', 'N', ra, rb, M, 1, AV, LDA, BV, LDB, 0, F, ra);                   \
    fl += orc_flop_counts('m', ra, rb, M, 0);                                          \
    if (mutate == ORC_MUTATE_NO_G_TRANSPOSE && ra == rb)                               \
        NAME##_gemm('N', 'N', M, ra, rb, 1, BU, LDB, F, ra, 0, Vs + (size_t)rc * M, M);\
    else                                                                               \
        NAME##_gemm('N', 'T', M, ra, rb, 1, BU, LDB, F, ra, 0, Vs + (size_t)rc * M, M);\
    fl += orc_flop_counts('m', M, ra, rb, 0);                                          \
    if (NAME##_geqrf(M, K, Vs, M, tauB)) { st = ORC_ERR_LAPACK; goto done; }           \
    fl += orc_flop_counts('q', M, K, 0, 0);                                            \
    /* __svd: upper triangles, T = Ra Rb^T, gesvd, truncate, fold sigma (296-503) */   \
    for (int j = 0; j < K; ++j)                                                        \
        for (int i = 0; i <= j; ++i) {                                                 \
            Ra[(size_t)j * K + i] = Us[(size_t)j * M + i];                             \
            Rb[(size_t)j * K + i] = Vs[(size_t)j * M + i];                             \
        }                                                                              \
    NAME##_gemm('N', 'T', K, K, K, 1, Ra, K, Rb, K, 0, T, K);                          \
    {                                                                                  \
        REAL *Ucore = (REAL *)calloc((size_t)K * K, sizeof(REAL));                     \
        REAL *Vcore = (REAL *)calloc((size_t)K * K, sizeof(REAL));                     \
        lint inf = NAME##_gesvd(K, T, K, sg, Ucore, K, Vcore, K);                      \
        if (inf) { free(Ucore); free(Vcore); st = ORC_ERR_LAPACK; goto done; }         \
        fl += orc_flop_counts('s', K, 0, 0, 0);                                        \
        int r;                                                                         \
        if (rk != 0) r = rk < K ? rk : K;                                              \
        else {                                                                         \
            r = K;                                                                     \
            for (int i = 1; i < K; ++i) if (sg[i] < acc) { r = i; break; }            \
            if (mutate == ORC_MUTATE_TRUNC_OFF_BY_ONE && r < K) r += 1;                \
        }                                                                              \
        if (sigma_out) memcpy(sigma_out, sg, sizeof(REAL) * K);                        \
        if (r > maxrk) { free(Ucore); free(Vcore); st = ORC_ERR_R_GT_MAXRK; goto done; }\
        /* rows of V^T scaled by sigma (486-489); the mutation scales U instead */    \
        for (int k = 0; k < r; ++k)                                                    \
            for (int j = 0; j < K; ++j) {                                              \
                if (mutate == ORC_MUTATE_SIGMA_TO_U) Ucore[(size_t)k * K + j] *= sg[k];\
                else Vcore[(size_t)j * K + k] *= sg[k];                                \
            }                                                                          \
        fl += (unsigned long)r * (unsigned long)K;                                     \
        /* __newu: Ut = Q_A [U_r; 0] via ormqr('L','N') (508-617) */                   \
        for (int j = 0; j < r; ++j)                                                    \
            for (int i = 0; i < K; ++i) Ut[(size_t)j * M + i] = Ucore[(size_t)j * K + i];\
        if (NAME##_ormqr('L', 'N', M, r, K, Us, M, tauA, Ut, M)) {                     \
            free(Ucore); free(Vcore); st = ORC_ERR_LAPACK; goto done; }                \
        fl += orc_flop_counts('o', M, r, K, 1);                                        \
        /* __newv: Vt = [S V^T_r, 0] Q_B^T via ormqr('R','T') then transpose (622-696)*/\
        for (int j = 0; j < K; ++j)                                                    \
            for (int i = 0; i < r; ++i) Vt[(size_t)j * r + i] = Vcore[(size_t)j * K + i];\
        if (NAME##_ormqr('R', 'T', r, M, K, Vs, M, tauB, Vt, r)) {                     \
            free(Ucore); free(Vcore); st = ORC_ERR_LAPACK; goto done; }                \
        fl += orc_flop_counts('o', r, M, K, 2);                                        \
        free(Ucore); free(Vcore);                                                      \
        /* copy back into the caller's factors (971-982) and the rank (1006) */       \
        for (int j = 0; j < r; ++j) {                                                  \
            memcpy(CU + (size_t)j * LDC, Ut + (size_t)j * M, sizeof(REAL) * M);        \
            for (int i = 0; i < M; ++i) CV[(size_t)j * LDC + i] = Vt[(size_t)i * r + j];\
        }                                                                              \
        *rc_io = r;                                                                    \
        if (flops_update) *flops_update += fl;                                         \
    }                                                                                  \
done:                                                                                  \
    free(Us); free(Vs); free(tauA); free(tauB); free(F); free(Ra); free(Rb); free(T);  \
    free(Ut); free(Vt); free(sg);                                                      \
    return st;                                                                         \
}

ORC_DEFINE_GEMM(orc_dgemm, double, scipy_dgeqrf_, scipy_dormqr_, scipy_dgesvd_, scipy_dgemm_)
ORC_DEFINE_GEMM(orc_sgemm, float, scipy_sgeqrf_, scipy_sormqr_, scipy_sgesvd_, scipy_sgemm_)

/* ---- HCORE_dsyrk restatement (hcore_dsyrk.c:47-189) --------------------
 * W = AV^T AV over LDAU rows, X = AU W, CD = alpha X AU^T + CD, then the
 * strict upper triangle of CD is zeroed.  Returns the flop increment. */
unsigned long orc_dsyrk(int M, int K, double alpha, const double *AU, int LDAU,
                        const double *AV, double *CD, int LDCD)
{
    const int N = LDAU;
    double *W = (double *)calloc((size_t)K * K, sizeof(double));
    double *X = (double *)calloc((size_t)M * K, sizeof(double));
    orc_dgemm_gemm('T', 'N', K, K, N, 1.0, AV, LDAU, AV, LDAU, 0.0, W, K);
    orc_dgemm_gemm('N', 'N', M, K, K, 1.0, AU, LDAU, W, K, 0.0, X, M);
    orc_dgemm_gemm('N', 'T', M, M, K, alpha, X, M, AU, LDAU, 1.0, CD, LDCD);
    for (int j = 1; j < M; ++j)
        for (int i = 0; i < j; ++i) CD[(size_t)j * LDCD + i] = 0.0;
    free(W);
    free(X);
    return orc_flop_counts('m', K, K, N, 0) + orc_flop_counts('m', M, K, K, 0) +
           orc_flop_counts('m', M, M, K, 0);
}

/* ---- HCORE_dgemmcd restatement (hcore_dgemmcd.c:36-106) ----------------
 * CD -= (Au Av^T)(Bu Bv^T)^T with the cheaper association; all leading
 * dimensions are M, alpha/beta are not used. */
void orc_dgemmcd(int M, const double *Au, const double *Av, int ra, const double *Bu,
                 const double *Bv, int rb, double *CD)
{
    double *W = (double *)calloc((size_t)ra * rb, sizeof(double));
    double *X = (double *)calloc((size_t)M * (ra > rb ? ra : rb), sizeof(double));
    orc_dgemm_gemm('T', 'N', ra, rb, M, 1.0, Av, M, Bv, M, 0.0, W, ra);
    if (ra > rb) {
        orc_dgemm_gemm('N', 'N', M, rb, ra, 1.0, Au, M, W, ra, 0.0, X, M);
        orc_dgemm_gemm('N', 'T', M, M, rb, -1.0, X, M, Bu, M, 1.0, CD, M);
    } else {
        orc_dgemm_gemm('N', 'T', ra, M, rb, 1.0, W, ra, Bu, M, 0.0, X, ra);
        orc_dgemm_gemm('N', 'N', M, M, ra, -1.0, Au, M, X, ra, 1.0, CD, M);
    }
    free(W);
    free(X);
}

/* ---- the reference tests' input generator (testing_dgemm.c:119-236, and
 * its FP32 twin testing_sgemm.c:119) ----------------------------------------
 * xLATMS with singular values 10^-i, then xGESVD and truncation at acc capped
 * by storage_maxrank: U = first r left vectors, V = (S V^T)^T.  dense
 * (ld x n) receives the generated matrix.  iseed advances in place. */
#define ORC_DEFINE_GENERATOR(NAME, REAL, LATMS, GESVD)                                  \
int NAME(int nrows, int ncols, int ld, REAL *dense, REAL *U, REAL *V,                  \
         int storage_maxrank, int maxMNK, REAL acc, int *iseed, int *prank)            \
{                                                                                      \
    lint m = nrows, n = ncols, lda = ld, info = 0, mode = 0, kl = nrows - 1,           \
         ku = ncols - 1;                                                               \
    REAL cond = -1, dmax = -1;                                                         \
    REAL *d = (REAL *)calloc((size_t)maxMNK, sizeof(REAL));                            \
    for (int i = 0; i < maxMNK; ++i) d[i] = (REAL)pow(10.0, -1.0 * i);                 \
    REAL *work = (REAL *)calloc((size_t)3 * (nrows > ncols ? nrows : ncols) + 16,     \
                                sizeof(REAL));                                         \
    LATMS(&m, &n, "U", iseed, "N", d, &mode, &cond, &dmax, &kl, &ku, "N", dense, &lda, \
          work, &info, 1, 1, 1);                                                       \
    free(work);                                                                        \
    if (info) { free(d); return (int)info; }                                           \
    size_t nel = (size_t)maxMNK * maxMNK;                                              \
    REAL *tU = (REAL *)calloc(nel, sizeof(REAL));                                      \
    REAL *tV = (REAL *)calloc(nel, sizeof(REAL));                                      \
    REAL *cl = (REAL *)calloc((size_t)ld * ncols, sizeof(REAL));                       \
    REAL *S = (REAL *)calloc((size_t)maxMNK, sizeof(REAL));                            \
    memcpy(cl, dense, sizeof(REAL) * (size_t)ld * ncols);                              \
    {                                                                                  \
        lint lwork = -1, ldu = ld, ldvt = ld;                                          \
        REAL q;                                                                        \
        GESVD("A", "A", &m, &n, cl, &lda, S, tU, &ldu, tV, &ldvt, &q, &lwork, &info,   \
              1, 1);                                                                   \
        lwork = (lint)q;                                                               \
        REAL *w = (REAL *)malloc(sizeof(REAL) * (size_t)lwork);                        \
        GESVD("A", "A", &m, &n, cl, &lda, S, tU, &ldu, tV, &ldvt, w, &lwork, &info,    \
              1, 1);                                                                   \
        free(w);                                                                       \
    }                                                                                  \
    int minmn = nrows < ncols ? nrows : ncols;                                         \
    int r = 0;                                                                         \
    if (!info) {                                                                       \
        for (int k = 0; k < ncols; ++k)                                                \
            for (int j = 0; j < nrows; ++j) tV[(size_t)j * ld + k] *= S[k];            \
        while (r < minmn && S[r] >= acc) ++r;                                          \
        if (r > storage_maxrank) r = storage_maxrank;                                  \
        for (int j = 0; j < r; ++j)                                                    \
            for (int i = 0; i < nrows; ++i) U[(size_t)j * ld + i] = tU[(size_t)j * ld + i];\
        for (int j = 0; j < r; ++j)                                                    \
            for (int i = 0; i < ncols; ++i) V[(size_t)j * ld + i] = tV[(size_t)i * ld + j];\
    }                                                                                  \
    *prank = r;                                                                        \
    free(d); free(tU); free(tV); free(cl); free(S);                                    \
    return (int)info;                                                                  \
}

extern void scipy_slatms_(const lint *m, const lint *n, const char *dist, lint *iseed,
                          const char *sym, float *d, const lint *mode, const float *cond,
                          const float *dmax, const lint *kl, const lint *ku,
                          const char *pack, float *a, const lint *lda, float *work,
                          lint *info, size_t, size_t, size_t);
ORC_DEFINE_GENERATOR(orc_generate_compress, double, scipy_dlatms_, scipy_dgesvd_)
ORC_DEFINE_GENERATOR(orc_generate_scompress, float, scipy_slatms_, scipy_sgesvd_)

/* ---- batched CPU baseline ------------------------------------------------
 * One HCORE_dgemm restatement per update, updates spread over OpenMP
 * threads with a dynamic schedule and the BLAS single-threaded per call —
 * how HiCMA runs one HCORE call per runtime worker.  Returns the number of
 * failed updates. */
int orc_dgemm_batch(int n, int M, double alpha, double *const *AU, double *const *AV,
                    const int *ra, int LDA, double *const *BU, double *const *BV,
                    const int *rb, int LDB, double beta, double *const *CU,
                    double *const *CV, int *rc_io, int LDC, int rk, int maxrk, double acc,
                    int nthreads, unsigned long *flops_update)
{
    int bad = 0;
    unsigned long fl = 0;
    scipy_openblas_set_num_threads(1);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad, fl)
#endif
    for (int u = 0; u < n; ++u) {
        unsigned long f = 0;
        int st = orc_dgemm(M, alpha, AU[u], AV[u], ra[u], LDA, BU[u], BV[u], rb[u], LDB, beta,
                           CU[u], CV[u], &rc_io[u], LDC, rk, maxrk, acc, &f, NULL, 0);
        bad += (st != ORC_OK);
        fl += f;
    }
    if (flops_update) *flops_update += fl;
    return bad;
}

int orc_dsyrk_batch(int n, int M, double alpha, double *const *AU, double *const *AV,
                    const int *K, int LDAU, double *const *CD, int LDCD, int nthreads,
                    unsigned long *flops_syrk)
{
    unsigned long fl = 0;
    scipy_openblas_set_num_threads(1);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : fl)
#endif
    for (int u = 0; u < n; ++u) fl += orc_dsyrk(M, K[u], alpha, AU[u], LDAU, AV[u], CD[u], LDCD);
    if (flops_syrk) *flops_syrk += fl;
    return 0;
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- dense neighbours of the recompression: potrf, trsm, the dense-B
 * update and decompression.  HCORE_enum values equal the CBLAS ones
 * (hcore.h:30-44): 111/112/113 trans, 121/122 uplo, 131/132 diag,
 * 141/142 side. */
static char orc_uplo_char(int uplo) {
    /* the reference hands its enum to LAPACKE as the uplo character
     * (hcore_dpotrf.c:29-32), which LAPACK rejects; the restatement follows
     * the intended triangle */
    return (uplo == 121 || uplo == 'U' || uplo == 'u') ? 'U' : 'L';
}

/* hcore_dpotrf.c:23-34: Cholesky of one dense tile; flops->potrf += 'c'. */
int orc_dpotrf(int uplo, int N, double *A, int LDA, unsigned long *flops)
{
    char u = orc_uplo_char(uplo);
    lint n = N, lda = LDA, info = 0;
    scipy_dpotrf_(&u, &n, A, &lda, &info, 1);
    if (flops) *flops += orc_flop_counts('c', N, 0, 0, 0);
    return (int)info;
}

/* hcore_dtrsm.c:24-48: op(A) X = alpha B (left) or X op(A) = alpha B
 * (right), B is M x Brank; flops->trsm += 't' with side 1 (left) or 2. */
int orc_dtrsm(int side, int uplo, int transA, int diag, int M, int Brank, double alpha,
              const double *A, int LDA, double *B, int LDB, unsigned long *flops)
{
    char sd = side == 141 ? 'L' : 'R';
    char up = uplo == 121 ? 'U' : 'L';
    char tr = transA == 111 ? 'N' : 'T';
    char dg = diag == 132 ? 'U' : 'N';
    lint m = M, n = Brank, lda = LDA, ldb = LDB;
    if (side != 141 && side != 142) return -1;
    scipy_dtrsm_(&sd, &up, &tr, &dg, &m, &n, &alpha, A, &lda, B, &ldb, 1, 1, 1, 1);
    if (flops) *flops += orc_flop_counts('t', M, Brank, side == 141 ? 1 : 2, 0);
    return 0;
}

/* hcore_dgemmbdcd.c:35-122: AVBD = AV^T BD (Ark x N), then
 * CD = alpha AU AVBD + beta CD; with transA == HCORE_Trans the inner size
 * of the first product is M and the rows of the second LDA, otherwise LDB
 * and M (the reference's own two branches).  flops->update += both 'm'. */
void orc_dgemmbdcd(int transA, int M, int N, double alpha, const double *AU, const double *AV,
                   int Ark, int LDA, const double *BD, int LDB, double beta, double *CD,
                   int LDC, unsigned long *flops)
{
    double *AVBD = malloc(sizeof(double) * (size_t)(Ark > 0 ? Ark : 1) * (N > 0 ? N : 1));
    const int k1 = transA == 112 ? M : LDB;
    const int m2 = transA == 112 ? LDA : M;
    orc_dgemm_gemm('T', 'N', Ark, N, k1, 1.0, AV, LDA, BD, LDB, 0.0, AVBD, Ark);
    orc_dgemm_gemm('N', 'N', m2, N, Ark, alpha, AU, LDA, AVBD, Ark, beta, CD, LDC);
    if (flops)
        *flops += orc_flop_counts('m', Ark, N, k1, 0) + orc_flop_counts('m', m2, N, Ark, 0);
    free(AVBD);
}

/* hcore_duncompress.c:33-84: CD = AU BV^T (M x N, inner size Ark); alpha
 * and beta are not used and the two ranks must be equal. */
int orc_duncompress(int M, int N, const double *AU, int Ark, int LDA, const double *BV, int Brk,
                    int LDB, double *CD, int LDC)
{
    if (Ark != Brk) return -1;
    orc_dgemm_gemm('N', 'T', M, N, Ark, 1.0, AU, LDA, BV, LDB, 0.0, CD, LDC);
    return 0;
}

/* ---- tile compression and the covariance problem ---------------------------
 * orc_dcompress: the compression step of the reference's generator
 * (testing_dgemm.c:180-226): dgesvd('A','A') of a copy of the m x n tile,
 * r = #{S_k >= acc} (k < min(m, n)) capped at maxrank (:208-212),
 * U = the first r left vectors (:219), V = (S V^T)^T's first r columns
 * (:162-166, :224).  mutate 1: V left unscaled; mutate 2: one singular
 * value too many kept (planted mistakes the pins must catch). */
int orc_dcompress(int m, int n, const double *A, int lda, double *U, int ldu, double *V,
                  int ldv, int maxrank, double acc, int mutate, int *prank)
{
    lint M = m, N = n, ld = m, info = 0, lwork = -1, ldu_ = m, ldvt = n;
    int mn = m < n ? m : n;
    double *cl = (double *)malloc(sizeof(double) * (size_t)m * n);
    double *tU = (double *)malloc(sizeof(double) * (size_t)m * m);
    double *tV = (double *)malloc(sizeof(double) * (size_t)n * n);
    double *S = (double *)malloc(sizeof(double) * (size_t)(mn > 0 ? mn : 1));
    for (int j = 0; j < n; ++j) memcpy(cl + (size_t)j * m, A + (size_t)j * lda, sizeof(double) * m);
    double q;
    scipy_dgesvd_("A", "A", &M, &N, cl, &ld, S, tU, &ldu_, tV, &ldvt, &q, &lwork, &info, 1, 1);
    lwork = (lint)q;
    double *w = (double *)malloc(sizeof(double) * (size_t)lwork);
    scipy_dgesvd_("A", "A", &M, &N, cl, &ld, S, tU, &ldu_, tV, &ldvt, w, &lwork, &info, 1, 1);
    free(w);
    int r = 0;
    if (!info) {
        while (r < mn && S[r] >= acc) ++r;
        if (mutate == 2 && r < mn) ++r;
        if (r > maxrank) r = maxrank;
        for (int k = 0; k < r; ++k) {
            for (int i = 0; i < m; ++i) U[(size_t)k * ldu + i] = tU[(size_t)k * m + i];
            const double s = mutate == 1 ? 1.0 : S[k];
            for (int i = 0; i < n; ++i) V[(size_t)k * ldv + i] = s * tV[(size_t)i * n + k];
        }
    }
    *prank = r;
    free(cl); free(tU); free(tV); free(S);
    return (int)info;
}

/* The covariance problem of HCORE_dgytlr (the real counterpart of the
 * declared-only HCORE_zgytlr, src/include/hcore_z.h:63-75; the reference
 * defines no problem, so this is the build's own definition): bigM points,
 * point i = (u_{2i+1}, u_{2i+2}) with u_k = splitmix64 output k of `seed`
 * mapped to [0, 1) by its top 53 bits, sorted by Morton code on a
 * 2^16 x 2^16 grid (x in the even bits), ties by index. */
static uint64_t orc_splitmix(uint64_t seed, uint64_t k)
{
    uint64_t z = seed + k * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
uint64_t orc_splitmix64(uint64_t seed, uint64_t k) { return orc_splitmix(seed, k); }

static uint32_t orc_morton(double x, double y)
{
    double qx = x * 65536.0, qy = y * 65536.0;
    uint32_t ix = qx >= 65535.0 ? 65535u : (uint32_t)qx, iy = qy >= 65535.0 ? 65535u : (uint32_t)qy;
    uint32_t c = 0;
    for (int b = 0; b < 16; ++b) c |= ((ix >> b) & 1u) << (2 * b) | ((iy >> b) & 1u) << (2 * b + 1);
    return c;
}

static const uint32_t *orc_sort_codes;
static int orc_cmp_code(const void *a, const void *b)
{
    int i = *(const int *)a, j = *(const int *)b;
    if (orc_sort_codes[i] != orc_sort_codes[j]) return orc_sort_codes[i] < orc_sort_codes[j] ? -1 : 1;
    return (i > j) - (i < j);
}

void orc_covariance_points(int bigM, uint64_t seed, double *xy)
{
    double *x = (double *)malloc(sizeof(double) * (size_t)bigM);
    double *y = (double *)malloc(sizeof(double) * (size_t)bigM);
    uint32_t *code = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)bigM);
    int *ord = (int *)malloc(sizeof(int) * (size_t)bigM);
    for (int i = 0; i < bigM; ++i) {
        x[i] = (double)(orc_splitmix(seed, 2 * (uint64_t)i + 1) >> 11) * 0x1.0p-53;
        y[i] = (double)(orc_splitmix(seed, 2 * (uint64_t)i + 2) >> 11) * 0x1.0p-53;
        code[i] = orc_morton(x[i], y[i]);
        ord[i] = i;
    }
    orc_sort_codes = code;
    qsort(ord, (size_t)bigM, sizeof(int), orc_cmp_code);
    for (int i = 0; i < bigM; ++i) {
        xy[i] = x[ord[i]];
        xy[bigM + i] = y[ord[i]];
    }
    free(x); free(y); free(code); free(ord);
}

/* Tile (m0, n0) of size m x n: exp(-|p_i - p_j| / beta) + nugget on the
 * global diagonal; xy as orc_covariance_points wrote it. */
void orc_covariance_tile(int bigM, const double *xy, int m0, int n0, int m, int n, double beta,
                         double nugget, double *out, int ld)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            const int gi = m0 + i, gj = n0 + j;
            const double dx = xy[gi] - xy[gj], dy = xy[bigM + gi] - xy[bigM + gj];
            double v = exp(-sqrt(dx * dx + dy * dy) / beta);
            if (gi == gj) v += nugget;
            out[(size_t)j * ld + i] = v;
        }
}
