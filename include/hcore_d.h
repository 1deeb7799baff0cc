/*
 * hcore_d.h — FP64 tile low-rank kernels of the MI355X-native HCORE.
 *
 * Drop-in for the reference's src/include/hcore_d.h: the three FP64 entry
 * points of the recompression hot path keep their exact prototypes so that a
 * runtime built against the reference (e.g. HiCMA codelets) relinks without
 * source changes.  The arithmetic runs on the GPU (HIP, gfx950); pointers may
 * be host or device memory (host buffers are staged through the device
 * workspace pool).
 *
 * Tile storage: an nb x nb tile T is kept as column-major factors TU (nb x r)
 * and TV (nb x r) with T = TU * TV^T.  Ranks are passed as doubles (the value
 * is cast to int), as in the reference.
 */
#ifndef HCORE_AMD_HCORE_D_H
#define HCORE_AMD_HCORE_D_H
#include <stddef.h>

#include "flop_util_structs.h"
#include "hcore.h"
#ifdef __cplusplus
extern "C" {
#endif

/*
 * C <- beta*C + alpha*A*B^T for low-rank A = AU*AV^T, B = BU*BV^T and
 * C = CU*CV^T, recompressed to the smallest rank whose dropped singular
 * values are below acc (absolute), or to rank rk when rk != 0.
 *
 * Replaces reference src/compute/hcore_dgemm.c:703-1012 (prototype
 * src/include/hcore_d.h:25-46).  transA, transB and N are ignored (the op is
 * always beta*C + alpha*A*B^T).  LDA must equal M (the AU block is read as a
 * contiguous M x rank array, reference hcore_dgemm.c:106).  On return CU and
 * CV (leading dimension LDC) hold the first r columns of the new factors, CU
 * with orthonormal columns and the singular values folded into CV, and
 * Crk[0] = r.  The caller guarantees room for r columns.  Fatal conditions
 * (zero input rank, rc+ra > 2*maxrk, rc+ra > maxrk, r > maxrk) print a
 * message to stderr and call exit(-1), leaving CU/CV untouched.
 * `work` is accepted for ABI compatibility; the device workspace comes from
 * the library's pool.  flops->update is incremented by the reference's own
 * closed-form count.
 */
void HCORE_dgemm(HCORE_enum transA, int transB,
                 int M, int N,
                 double alpha,
                 double *AU, double *AV, double *Ark, int LDA,
                 double *BU, double *BV, double *Brk, int LDB,
                 double beta,
                 double *CU, double *CV, double *Crk, int LDC,
                 int rk, int maxrk, double acc,
                 double *work, flop_counter *flops);

/*
 * Dense-output update CD <- CD - (Au*Av^T) * (Bu*Bv^T)^T.
 * Replaces reference src/compute/hcore_dgemmcd.c:36-106 (prototype
 * src/include/hcore_d.h:68-86).  As in the reference, alpha = -1 and
 * beta = 1 are fixed, transA/transB are ignored, every leading dimension is
 * taken as M, and flops is not updated.
 */
void HCORE_dgemmcd(HCORE_enum transA, HCORE_enum transB,
                   int M,
                   double alpha,
                   double *Au, double *Av, int Arank, int LDA,
                   double *Bu, double *Bv, int Brank, int LDB,
                   double beta,
                   double *CD, int LDC,
                   int maxrank,
                   double *work, flop_counter *flops);

/*
 * Diagonal update CD <- alpha * AU*(AV^T*AV)*AU^T + CD with the strict upper
 * triangle of CD then set to zero.
 * Replaces reference src/compute/hcore_dsyrk.c:47-189 (prototype
 * src/include/hcore_d.h:107-116).  As in the reference, the row count of AV
 * is taken as LDAU, uplo/trans/beta/LDAV are ignored and flops->syrk is
 * incremented.
 */
void HCORE_dsyrk(HCORE_enum uplo, HCORE_enum trans,
                 int M, int K,
                 double alpha,
                 const double *AU, int LDAU,
                 const double *AV, int LDAV,
                 double beta,
                 double *CD, int LDCD,
                 double *work, flop_counter *flops);

/*
 * CD <- alpha * AU*(AV^T*BD) + beta*CD for a low-rank A = AU*AV^T (rank
 * Ark[0]) and a dense BD.  Replaces reference src/compute/hcore_dgemmbdcd.c:35-122
 * (prototype src/include/hcore_d.h:87-101), with its two branches: for
 * transA == HCORE_Trans the inner size of AV^T*BD is M and CD has LDA rows,
 * otherwise the inner size is LDB and CD has M rows.  transB and work are
 * ignored; flops->update gets both products' counts.
 */
void HCORE_dgemmbdcd(HCORE_enum transA, HCORE_enum transB,
                     int M, int N,
                     double alpha,
                     double *AU, double *AV, double *Ark, int LDA,
                     double *BD, int LDB,
                     double beta,
                     double *CD, int LDC,
                     double *work, flop_counter *flops);

/*
 * Cholesky factorisation of one dense N x N tile in place: A = L*L^T (lower)
 * or U^T*U (upper); the other triangle is not touched.  Replaces reference
 * src/compute/hcore_dpotrf.c:23-34 (prototype src/include/hcore_d.h:102-106).
 * The reference hands its HCORE_enum to LAPACKE as the uplo character, which
 * LAPACK rejects; here uplo selects the intended triangle (HCORE_Upper = 121
 * or 'U', HCORE_Lower = 122 or 'L'; anything else gives *info = -1).
 * *info = k > 0 when the leading minor of order k is not positive definite.
 * flops->potrf += the reference's 'c' count.
 */
void HCORE_dpotrf(HCORE_enum uplo,
                  int N,
                  double *A, int LDA,
                  int *info,
                  flop_counter *flops);

/*
 * Triangular solve op(A)*X = alpha*BV (side HCORE_Left, A M x M) or
 * X*op(A) = alpha*BV (HCORE_Right, A Brank x Brank), BV (M x Brank) is
 * overwritten by X.  Replaces reference src/compute/hcore_dtrsm.c:24-48
 * (prototype src/include/hcore_d.h:117-125), which calls cblas_dtrsm with
 * the same enum values.  flops->trsm += the reference's 't' count.
 */
void HCORE_dtrsm(HCORE_enum side, HCORE_enum uplo,
                 HCORE_enum transA, HCORE_enum diag,
                 int M, int Brank,
                 double alpha,
                 double *A, int LDA,
                 double *BV, int LDBV,
                 flop_counter *flops);

/*
 * Decompression CD = AU*BV^T (M x N, inner size Ark[0], which must equal
 * Brk[0]).  Replaces reference src/compute/hcore_duncompress.c:33-84
 * (prototype src/include/hcore_d.h:126-139): transA, transB, alpha and beta
 * are ignored, as there.
 */
void HCORE_duncompress(HCORE_enum transA, HCORE_enum transB,
                       int M, int N,
                       double alpha,
                       double *AU, double *Ark, int LDA,
                       double *BV, double *Brk, int LDB,
                       double beta,
                       double *CD, int LDC);

/*
 * Compression of one dense m x n tile A (max(m, n) <= 512): U (m x r, ldu)
 * receives the first r left singular vectors, V (n x r, ldv) the right
 * singular vectors scaled by the singular values (A ~ U V^T), Ark[0] = r,
 * the number of singular values >= acc, capped by maxrank.  No library
 * symbol of the reference: it restates the compression step of the
 * reference's tile generator generate_compress
 * (testing/testing_dgemm.c:119-236, truncation at :208-212).  Host or
 * device pointers; U and V must hold min(m, n, maxrank) columns.
 */
void HCORE_dcompress(int m, int n, const double *A, int lda,
                     double *U, int ldu, double *V, int ldv,
                     double *Ark, int maxrank, double acc);

/*
 * Generate tile (m0, n0) (m x n) of the bigM x bigM covariance problem and
 * compress it: the real-precision counterpart of HCORE_zgytlr, which the
 * reference declares without defining (src/include/hcore_z.h:63-75), with
 * the same arguments.  The problem: bigM points in [0,1)^2 drawn from the
 * splitmix64 sequence of `seed`, Morton-sorted; entry (i, j) =
 * exp(-|p_i - p_j| / beta) + nugget [i == j] (beta 0.1, nugget 1e-4 unless
 * HCORE_hip_set_covariance says otherwise).  A diagonal tile (m0 == n0) with
 * compress_diag == 0 is written dense into AD (lda) and Ark is not touched;
 * otherwise AU/AV/Ark as HCORE_dcompress with acc = tol.  Dense, when not
 * NULL, also receives the dense tile (lda).
 */
void HCORE_dgytlr(int m, int n,
                  double *AU, double *AV, double *AD, double *Ark,
                  int lda, int ldu, int ldv,
                  int bigM, int m0, int n0, unsigned long long int seed,
                  int maxrank, double tol,
                  int compress_diag,
                  double *Dense);

/*
 * The recompression split at the rank reveal, for runtimes that size C's
 * storage after the rank is known.  Replaces reference
 * src/compute/hcore_dgemm.c:1022-1303 (HCORE_dgemm_qr_svd_b_dense),
 * :1312-1597 (HCORE_dgemm_qr_svd) and :1602-1699 (HCORE_dgemm_ormqr);
 * prototypes src/include/hcore_d.h:146-268.  Ranks are int here, as there.
 *
 * Phase 1 (qr_svd / qr_svd_b_dense) runs stacking, both QRs, the core SVD
 * and the truncation on the GPU and returns the new rank in *_p_new_UVrk; its
 * state stays on the device, reached again through the address it stores in
 * *_p_newU (the other _p_ outputs are set for the reference's bookkeeping and
 * are not dereferenced by the caller).  Phase 2 (ormqr), given those values
 * back, writes the new factors (first new_UVrk columns) into CU/CV, which the
 * caller may have reallocated in between (ld LDC >= M), sets Crk, and
 * releases the state.  A new rank above storage_maxrank is rejected in phase
 * 2 (hcore_dgemm.c:1648-1651) and the state released.  The _b_dense form
 * takes B densely (M x M, ld LDB): G = B AV replaces G = BU (AV^T BV)^T.
 */
void HCORE_dgemm_qr_svd_b_dense(int transA, int transB, int M, int N, double alpha,
                                double *AU, double *AV, int *Ark, int LDA,
                                double *B, int LDB, double beta,
                                double *CU, double *CV, int *Crk, int LDC,
                                int rk, int storage_maxrank, int maxrk, double acc,
                                double *work, flop_counter *flops,
                                double **_p_work_new, double **_p__CU, double **_p__CV,
                                int *_p_CU_ncols, int *_p_new_UVrk, double **_p_newU,
                                int *_p_ld_newU, double **_p_qrtauA, int *_p_CV_ncols,
                                double **_p_newV, int *_p_ld_newV, double **_p_qrtauB,
                                int *_p_use_CUV_clone, double **_p_CUclone, int *_p_ld_CUclone,
                                double **_p__CU_save, double **_p_CVclone, int *_p_ld_CVclone,
                                double **_p__CV_save);
void HCORE_dgemm_qr_svd(int transA, int transB, int M, int N, double alpha,
                        double *AU, double *AV, int *Ark, int LDA,
                        double *BU, double *BV, int *Brk, int LDB, double beta,
                        double *CU, double *CV, int *Crk, int LDC,
                        int rk, int storage_maxrank, int maxrk, double acc,
                        double *work, flop_counter *flops,
                        double **_p_work_new, double **_p__CU, double **_p__CV,
                        int *_p_CU_ncols, int *_p_new_UVrk, double **_p_newU,
                        int *_p_ld_newU, double **_p_qrtauA, int *_p_CV_ncols,
                        double **_p_newV, int *_p_ld_newV, double **_p_qrtauB,
                        int *_p_use_CUV_clone, double **_p_CUclone, int *_p_ld_CUclone,
                        double **_p__CU_save, double **_p_CVclone, int *_p_ld_CVclone,
                        double **_p__CV_save);
void HCORE_dgemm_ormqr(int transA, int transB, int M, int N, double alpha,
                       double *AU, double *AV, int *Ark, int LDA, double beta,
                       double *CU, double *CV, int *Crk, int LDC,
                       int rk, int storage_maxrank, int maxrk, double acc,
                       double *work, flop_counter *flops,
                       double *_CU, double *_CV, int CU_ncols, int new_UVrk,
                       double *newU, int ld_newU, double *qrtauA, int CV_ncols,
                       double *newV, int ld_newV, double *qrtauB, int use_CUV_clone,
                       double *CUclone, int ld_CUclone, double *_CU_save,
                       double *CVclone, int ld_CVclone, double *_CV_save);

/*
 * HCORE_dgemm without any device allocation: the same arguments and results,
 * factors as DEVICE pointers only (ranks may be host or device doubles),
 * run on the null stream.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The one-update
 * workspace is a static device array of HCORE_dgemm_noalloc_workspace_bytes()
 * (64 MiB) per GPU, held until the process ends; an update that needs more
 * (e.g. M = 16384 with K = 512) is rejected with HCORE_HIP_ERR_ARG before any
 * kernel runs.  Calls that overlap on one GPU share the array and race.
 * Call HCORE_hip_init first.
 */
void HCORE_dgemm_noalloc(HCORE_enum transA, int transB,
                         int M, int N,
                         double alpha,
                         double *AU, double *AV, double *Ark, int LDA,
                         double *BU, double *BV, double *Brk, int LDB,
                         double beta,
                         double *CU, double *CV, double *Crk, int LDC,
                         int rk, int maxrk, double acc,
                         double *work, flop_counter *flops);
size_t HCORE_dgemm_noalloc_workspace_bytes(void);

/* Data symbols the reference exports or imports (reference
 * src/compute/hcore_dgemm.c:37-43, hcore_dsyrk.c:21-22,
 * hcore_dgemmcd.c:24-26).  `use_scratch` is defined weakly (0) by the
 * library so applications that define it still link.  None of them changes
 * results on the GPU path. */
extern int use_scratch;
extern int use_trmm;
extern int gemm_print_index;
extern int gemm_print_mat;
extern int hc_nelm_limit;
extern int syrk_print_index;
extern int gemmcd_print_index;
extern int gemmcd_print_mat;
extern int gemmcd_use_scratch;
extern int gemmbdcd_print_index;   /* hcore_dgemmbdcd.c:23-24 */
extern int gemmbdcd_print_mat;
extern int gemmfrk_print_index;    /* read by hcore_duncompress.c:23-24 */
extern int gemmfrk_print_mat;

#ifdef __cplusplus
}
#endif
#endif /* HCORE_AMD_HCORE_D_H */
