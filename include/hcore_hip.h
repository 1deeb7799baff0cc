/*
 * hcore_hip.h — GPU-side extensions of the MI355X-native HCORE (no
 * counterpart in the reference, whose routines are single-tile host calls).
 *
 * 1. Library context: device selection, error policy, last status.
 * 2. Batched plans: the hundreds to thousands of independent tile updates of
 *    one TLR Cholesky trailing step (HCORE_dgemm on every C_ij, HCORE_dsyrk on
 *    every diagonal tile) recorded once, with device-resident factors, and
 *    executed as a fixed sequence of batched kernels on one HIP stream — so
 *    the whole step can be captured into a hipGraph and replayed.
 *
 * All pointers inside a plan are DEVICE pointers; the pointer arrays and the
 * rank arrays passed to the plan builders are HOST arrays (the runtime keeps
 * tile ranks on the host, as HiCMA does).  Streams are hipStream_t passed as
 * void*.  Every function returns 0 on success or a HCORE_HIP_* status.
 */
#ifndef HCORE_AMD_HCORE_HIP_H
#define HCORE_AMD_HCORE_HIP_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference's exit(-1) sites, reference hcore_dgemm.c) */
#define HCORE_HIP_OK 0
#define HCORE_HIP_ERR_ZERO_RANK 1    /* an input rank is 0 (hcore_dgemm.c:743-746) */
#define HCORE_HIP_ERR_SUM_RANKS 2    /* rc+ra > 2*maxrk (hcore_dgemm.c:91-95) */
#define HCORE_HIP_ERR_K_GT_MAXRK 3   /* rc+ra > maxrk or > M (hcore_dgemm.c:924-927) */
#define HCORE_HIP_ERR_R_GT_MAXRK 4   /* truncated rank > maxrk (hcore_dgemm.c:465-468) */
#define HCORE_HIP_ERR_NUMERIC 5      /* NaN/Inf in the core SVD (LAPACK info != 0) */
#define HCORE_HIP_ERR_RUNTIME 6      /* HIP runtime failure */
#define HCORE_HIP_ERR_ARG 7          /* invalid argument */

/* precision forms of the recompression */
#define HCORE_HIP_FP64 0   /* everything in FP64 (the plain HCORE_dgemm) */
#define HCORE_HIP_FP32 1   /* everything in FP32 (HCORE_sgemm) */
#define HCORE_HIP_MIXED 2  /* FP32 tiles, products, QR and reassembly; the core
                              SVD and the truncation test against acc in FP64 */
#define HCORE_HIP_SPLIT 3  /* FP64 tiles, QR, core SVD and truncation; the
                              products F, G, T, T^T Zu and the reassembly as
                              BF16 digit slices on the matrix cores with FP32
                              accumulation (Ozaki-style split precision) */

/* error policy of the drop-in HCORE_* entry points */
#define HCORE_HIP_ON_ERROR_EXIT 0    /* reference behaviour: message + exit(-1) */
#define HCORE_HIP_ON_ERROR_RETURN 1  /* message, record status, return */

int HCORE_hip_init(int device);
int HCORE_hip_finalize(void);
int HCORE_hip_set_error_policy(int policy);
/* status of the last HCORE_* call made by this thread */
int HCORE_hip_last_status(void);

typedef struct hcore_hip_plan hcore_hip_plan;

/* Create an empty plan for tiles of M rows in the given precision form. */
int HCORE_hip_plan_create(hcore_hip_plan **plan, int M, int precision);

/*
 * Record n independent HCORE_dgemm updates C_u <- beta*C_u + alpha*A_u*B_u^T.
 * AU..CV: host arrays of n device pointers (FP64, or FP32 for HCORE_HIP_FP32
 * and HCORE_HIP_MIXED)
 * to column-major factors with leading dimensions LDA, LDB, LDC (each >= M).
 * CUout/CVout: both NULL (the new factors overwrite CU/CV, leading dimension
 * LDC) or both host arrays of n device pointers that receive the new factors
 * with leading dimension LDCout (>= M).  Ark/Brk/Crk: host int ranks.  rk,
 * maxrk, acc as in HCORE_dgemm.  HCORE_HIP_ERR_ARG for a NULL pointer, a
 * leading dimension below M or only one of CUout/CVout.
 * May be called once per plan.
 */
int HCORE_hip_plan_add_dgemm(hcore_hip_plan *plan, int n,
                             void *const *AU, void *const *AV, const int *Ark, int LDA,
                             void *const *BU, void *const *BV, const int *Brk, int LDB,
                             void *const *CU, void *const *CV, const int *Crk, int LDC,
                             void *const *CUout, void *const *CVout, int LDCout,
                             double alpha, double beta, int rk, int maxrk, double acc);

/*
 * Record n HCORE_dsyrk diagonal updates CD_u <- alpha*AU_u (AV_u^T AV_u) AU_u^T
 * + CD_u, strict upper triangle of CD_u set to 0.  Host arrays of device
 * pointers; K: host int ranks, 1 <= K_u <= LDAU; LDAU and LDCD >= M
 * (HCORE_HIP_ERR_ARG otherwise).  May be called once per plan.
 */
int HCORE_hip_plan_add_dsyrk(hcore_hip_plan *plan, int n,
                             void *const *AU, void *const *AV, const int *K, int LDAU,
                             void *const *CD, int LDCD, double alpha);

/*
 * The same, with the input ranks in DEVICE memory, read when the plan
 * executes: Ark/Brk/Crk are host arrays of n device addresses of int ranks,
 * and Ark_cap/Brk_cap/Crk_cap host int upper bounds that size the workspace
 * and choose the lanes.  After every execution the new rank of each update
 * that succeeded replaces *Crk[u] on the device, so a runtime that keeps its
 * tile ranks on the GPU (and their factors at fixed addresses) replays one
 * captured graph for every step.  A rank outside 1..capacity gives the update
 * status HCORE_HIP_ERR_ARG and leaves its factors and rank untouched.
 */
int HCORE_hip_plan_add_dgemm_dev(hcore_hip_plan *plan, int n,
                                 void *const *AU, void *const *AV, const int *const *Ark,
                                 const int *Ark_cap, int LDA,
                                 void *const *BU, void *const *BV, const int *const *Brk,
                                 const int *Brk_cap, int LDB,
                                 void *const *CU, void *const *CV, int *const *Crk,
                                 const int *Crk_cap, int LDC,
                                 void *const *CUout, void *const *CVout, int LDCout,
                                 double alpha, double beta, int rk, int maxrk, double acc);
/* dsyrk updates with device-side ranks K[u] (device addresses) bounded by K_cap. */
int HCORE_hip_plan_add_dsyrk_dev(hcore_hip_plan *plan, int n,
                                 void *const *AU, void *const *AV, const int *const *K,
                                 const int *K_cap, int LDAU, void *const *CD, int LDCD,
                                 double alpha);

/* Upload descriptors and carve the device workspace.  Called once. */
int HCORE_hip_plan_finalize(hcore_hip_plan *plan);

/* Enqueue the whole batch on `stream` (hipStream_t; NULL = default stream). */
int HCORE_hip_plan_execute(hcore_hip_plan *plan, void *stream);

/* Same, through a hipGraph captured on the first call and replayed after.
 * In the graph a lane's core SVD launches of different stacked-rank buckets
 * (host-rank plans) run side by side; a side stream for them is made at the
 * first capture. */
int HCORE_hip_plan_execute_graph(hcore_hip_plan *plan, void *stream);

/*
 * Autotuning (no counterpart in the reference).  After at least one
 * execution, time each core-SVD launch of the plan with its register-resident
 * variant and with the same algorithm at twice (or half) the wavefronts per
 * tile, reps times each, and keep the faster for later executions (a captured
 * graph is captured again).  Results are not changed.  The choice is cached
 * per (M, precision, variant) for plans finalized later in the process, and
 * in the file named by HCORE_TUNE_CACHE across processes.  report (may be
 * NULL) receives up to nreport entries of 6 doubles: lane, base variant,
 * chosen variant, best ms of the base, best ms of the alternative, tiles.
 * Returns the number of launches tuned, or -HCORE_HIP_* on error.
 */
int HCORE_hip_plan_autotune(hcore_hip_plan *plan, void *stream, int reps, double *report,
                            int nreport);

/*
 * Split-precision plans (HCORE_HIP_SPLIT; no counterpart in the reference).
 * Digits per operand of the sliced products, 2..6 (each adds 8 bits; the
 * error of a product term relative to its chunk's row and column maxima is
 * about slices * 2^(-8 slices)), for every lane; 0 restores the default,
 * chosen from acc at finalize (acc 1e-8: 5, 1e-6: 4).  Before or after
 * finalize (a captured graph is captured again).  HCORE_HIP_ERR_ARG for
 * another precision form or a count outside 0, 2..6.
 */
int HCORE_hip_plan_set_slices(hcore_hip_plan *plan, int slices);
/* The digit count of each lane after finalize (up to nlanes entries);
 * returns the number of lanes, or -HCORE_HIP_* on error. */
int HCORE_hip_plan_get_slices(hcore_hip_plan *plan, int *slices_host, int nlanes);
/*
 * Slice autotuner (no counterpart in the reference).  After at least one
 * execution: runs the plan once with its sliced products through the FP64
 * matrix cores and once per digit count 2..6, and keeps for each lane (a
 * stacked-rank bucket) the smallest count whose singular values of every
 * update's core differ from the FP64 ones by at most tol * acc (tol <= 0:
 * 0.1).  The plan's outputs are those of its last execution afterwards
 * (execute again to refresh them).  report (may be NULL) receives up to
 * nreport entries of 4 doubles per lane: lane, chosen count, its largest
 * singular-value deviation / acc, the same for one digit fewer (-1 if none).
 * Returns the number of lanes tuned, or -HCORE_HIP_* on error.
 */
int HCORE_hip_plan_autotune_slices(hcore_hip_plan *plan, void *stream, double tol,
                                   double *report, int nreport);

/* Device array of the n new ranks (int), valid after execution. */
const int *HCORE_hip_plan_ranks_device(const hcore_hip_plan *plan);
/* Copy new ranks / per-update status (HCORE_HIP_*) to host; synchronizes. */
int HCORE_hip_plan_get_ranks(hcore_hip_plan *plan, int *ranks_host);
int HCORE_hip_plan_get_status(hcore_hip_plan *plan, int *status_host);
/*
 * Plans with device-side ranks route each lane's core SVDs by the stacked
 * rank K = ra + rc read at execution (not by the capacities): a one-workgroup
 * kernel sorts the lane's updates by K (largest first) into a device queue
 * and each core shape's launch serves its slice of it.  This copies the
 * order of the last execution (update indices, up to n of them; needs a
 * prior execution) and returns the lane's update count, 0 for a lane of a
 * host-rank plan (sorted at finalize), or -HCORE_HIP_* on error.
 */
int HCORE_hip_plan_get_core_queue(hcore_hip_plan *plan, int lane, int *updates_host, int n);
/* Copy the lane each update runs in (index into the timing lanes). */
int HCORE_hip_plan_get_lanes(hcore_hip_plan *plan, int *lanes_host);
/* Copy the number of Jacobi sweeps each update's core SVD used (diagnostic). */
int HCORE_hip_plan_get_sweeps(hcore_hip_plan *plan, int *sweeps_host);
/* Core SVD profile (diagnostic): 8 values per update, filled by the
 * register-resident core kernels (33 <= K <= 192): wall-clock stamps [0] start,
 * [1] after the pivoted QR, [2] after the Jacobi sweeps, [3] end, and [4] the
 * numerical rank of the pivoted QR; *tick_ns receives the clock period. */
int HCORE_hip_plan_get_core_profile(hcore_hip_plan *plan, long long *out, double *tick_ns);
/* Copy the K singular values of update u (descending, FP64) to host. */
int HCORE_hip_plan_get_sigma(hcore_hip_plan *plan, int u, double *sigma_host, int *K_out);
/* Bytes of device workspace held by the plan. */
long long HCORE_hip_plan_workspace_bytes(const hcore_hip_plan *plan);
/* Flop count of the recorded work by the reference's closed forms, using
 * the ranks of the last execution (requires HCORE_hip_plan_get_ranks). */
int HCORE_hip_plan_flops(hcore_hip_plan *plan, unsigned long *update, unsigned long *syrk);

void HCORE_hip_plan_destroy(hcore_hip_plan *plan);

/* Per-phase timing.  A plan of more than 64 updates splits them into up to
 * three lanes by stacked rank K (> 128, 65..128, <= 64), the first on the
 * caller's stream and the others on streams of their own (four hardware
 * queues with the null stream; those streams come from a per-thread pool and
 * return to it when the plan is destroyed); the dsyrk updates follow the
 * next-to-last lane on its stream.  When timing is enabled,
 * HCORE_hip_plan_execute records a HIP event pair around every kernel launch
 * on the stream it runs on.  Phases: 0 stack, 1 F=AV^T BV, 2 G=BU F^T,
 * 3 geqrf, 4 T=R_A R_B^T, 5 core SVD (all K buckets of the lane, run
 * concurrently), 6 reserved, 7 Zv=T^T Zu, 8 reassembly, 9 syrk W,
 * 10 syrk X, 11 syrk CD.  HCORE_hip_plan_phase_times fills
 * HCORE_HIP_NLANES * HCORE_HIP_NPHASES entries (index lane*NPHASES+phase;
 * lane 3 holds the syrk phases); times accumulate over executions. */
#define HCORE_HIP_NPHASES 12
#define HCORE_HIP_NLANES 4
/* enable: 0 off, 1 on (lanes concurrent, as in untimed execution), 2 on with
 * the lanes run one after another on the caller's stream, so that every
 * launch's duration is its own rather than shared with the other lanes */
int HCORE_hip_plan_set_timing(hcore_hip_plan *plan, int enable);
int HCORE_hip_plan_phase_times(hcore_hip_plan *plan, double *ms, int *count);

/* The library's batched matrix-core GEMM on one problem (no counterpart in
 * the reference): C = alpha*op(A)*op(B) + beta*C on device pointers, FP64 or
 * FP32 (precision HCORE_HIP_FP64 / HCORE_HIP_FP32), column-major, trans
 * values as HCORE_enum (111 no transpose, 112/113 transpose).  path selects
 * the kernel family: 0 the one the plans would choose, 1 64 x 64 tiles,
 * 2 32 x 32 tiles, 3 32 x 32 tiles with the k range split over four
 * wavefronts, 4 the tall-skinny kernel (op(A) = A, n <= 128, k <= 128).
 * stream NULL: the calling thread's library stream.  Synchronizes. */
int HCORE_hip_gemm(int precision, int transA, int transB, int m, int n, int k, double alpha,
                   const void *A, int lda, const void *B, int ldb, double beta, void *C, int ldc,
                   int path, void *stream);
/*
 * The split-precision product of HCORE_HIP_SPLIT plans on its own (no
 * counterpart in the reference): C = alpha op(A) op(B) + beta C for FP64
 * operands, each k chunk of 32 cut into `slices` (2..6) BF16 digits per row of
 * op(A) and column of op(B) under a power-of-two scale, digit products on the
 * matrix cores with FP32 accumulation, FP64 sums; 0 runs the same output
 * tiles on the FP64 matrix cores.  Synchronous.
 */
int HCORE_hip_gemm_split(int slices, int transA, int transB, int m, int n, int k, double alpha,
                         const double *A, int lda, const double *B, int ldb, double beta,
                         double *C, int ldc, void *stream);

/* Tile generation and compression in batches (all pointers device memory,
 * tiles of one size m x nn, max(m, nn) <= 512).  Tile i of the covariance
 * problem (HCORE_dgytlr's) starts at global row m0[i], column n0[i] (host
 * arrays); when AD and AD[i] are not NULL it is written dense to AD[i] (lda),
 * else compressed to AU[i] (ldu), AV[i] (ldv) with its rank in the device
 * int rank[i].  HCORE_hip_compress_batch compresses the caller's dense tiles
 * A[i] (lda).  stream NULL: the calling thread's library stream.  Both
 * synchronize before returning; HCORE_HIP_ERR_NUMERIC when a tile held
 * NaN/Inf. */
int HCORE_hip_gytlr_batch(int n, int m, int nn, const int *m0, const int *n0, int bigM,
                          unsigned long long seed, double *const *AU, int ldu,
                          double *const *AV, int ldv, int *rank, double *const *AD, int lda,
                          int maxrank, double tol, void *stream);
int HCORE_hip_compress_batch(int n, int m, int nn, const double *const *A, int lda,
                             double *const *AU, int ldu, double *const *AV, int ldv, int *rank,
                             int maxrank, double tol, void *stream);
/* Kernel parameters of the covariance problem (defaults beta 0.1, nugget 1e-4). */
int HCORE_hip_set_covariance(double beta, double nugget);
/* The problem's sorted points on the host: xy[0..bigM) = x, xy[bigM..2 bigM) = y. */
int HCORE_hip_covariance_points(int bigM, unsigned long long seed, double *xy);

/* The TLR Cholesky update stream (the loop HiCMA runs on top of HCORE):
 * for k = 0 .. nt-1, HCORE_dpotrf of the dense diagonal tile D[k] (lower),
 * V_ik <- L_kk^{-1} V_ik for i > k, HCORE_dsyrk D_i -= A_ik A_ik^T and the
 * recompression A_ij -= A_ik A_jk^T (i > j > k) as one batched plan per
 * step.  D: nt device pointers (nb x nb, ld nb); U, V: nt*nt device pointers
 * (tile (i, j) at i + j*nt, used for i > j; nb x cap, ld nb, cap >= maxrk);
 * rank: host ints, updated in place; maxrk, acc: as HCORE_dgemm.  When the V
 * factors of a panel lie cap columns apart in one buffer, the panel's trsm
 * is one call.  info: 0, or the global index of the first non-positive
 * pivot.  step_ranks_max (optional, nt ints): largest new rank per step.
 * Synchronous; follows the library's error policy. */
int HCORE_hip_tlr_dpotrf(int nt, int nb, double *const *D, double *const *U, double *const *V,
                         int *rank, int cap, int maxrk, double acc, void *stream, int *info,
                         flop_counter *flops, int *step_ranks_max);
/*
 * The same stream with the recompressions in a chosen precision form:
 * HCORE_HIP_FP64 (exactly HCORE_hip_tlr_dpotrf) or HCORE_HIP_SPLIT (the
 * products as BF16 digit slices; slices 0: the count from acc, else 2..6);
 * potrf, trsm and the dsyrk updates stay FP64.  HCORE_HIP_ERR_ARG for another
 * form or slice count.  No counterpart in the reference.
 */
int HCORE_hip_tlr_dpotrf_form(int precision, int slices, int nt, int nb, double *const *D,
                              double *const *U, double *const *V, int *rank, int cap, int maxrk,
                              double acc, void *stream, int *info, flop_counter *flops,
                              int *step_ranks_max);

/* Device memory and stream helpers (the library's own HIP runtime; callers
 * that share device buffers with the library allocate them here). */
int HCORE_hip_malloc(void **ptr, size_t bytes);
int HCORE_hip_free(void *ptr);
int HCORE_hip_memcpy_h2d(void *dst, const void *src, size_t bytes);
int HCORE_hip_memcpy_d2h(void *dst, const void *src, size_t bytes);
int HCORE_hip_memcpy_d2d(void *dst, const void *src, size_t bytes);
int HCORE_hip_memset(void *ptr, int value, size_t bytes);
int HCORE_hip_device_synchronize(void);
int HCORE_hip_stream_create(void **stream);
int HCORE_hip_stream_destroy(void *stream);
int HCORE_hip_stream_synchronize(void *stream);
int HCORE_hip_event_create(void **event);
int HCORE_hip_event_destroy(void *event);
int HCORE_hip_event_record(void *event, void *stream);
int HCORE_hip_event_elapsed_ms(void *start, void *stop, float *ms);
/* Device name and number of compute units of the current device. */
int HCORE_hip_device_info(char *name, int len, int *num_cu);
/* Free and total device memory of the library's GPU (bytes). */
int HCORE_hip_mem_info(size_t *free_bytes, size_t *total_bytes);

#ifdef __cplusplus
}
#endif
#endif /* HCORE_AMD_HCORE_HIP_H */
