/*
 * hcore_s.h — FP32 tile low-rank kernel of the MI355X-native HCORE.
 * Drop-in for the reference's src/include/hcore_s.h:25-45.
 */
#ifndef HCORE_AMD_HCORE_S_H
#define HCORE_AMD_HCORE_S_H
#include "flop_util_structs.h"
#include "hcore.h"
#ifdef __cplusplus
extern "C" {
#endif

/*
 * FP32 twin of HCORE_dgemm: same algorithm, single precision throughout,
 * float-typed ranks and acc.  Replaces reference
 * src/compute/hcore_sgemm.c:703 (prototype src/include/hcore_s.h:25-45).
 */
void HCORE_sgemm(HCORE_enum transA, HCORE_enum transB,
                 int M, int N,
                 float alpha,
                 float *AU, float *AV, float *Ark, int LDA,
                 float *BU, float *BV, float *Brk, int LDB,
                 float beta,
                 float *CU, float *CV, float *Crk, int LDC,
                 int rk, int maxrk, float acc,
                 float *work, flop_counter *flops);

extern int sgemm_use_trmm;
extern int sgemm_print_index;
extern int sgemm_print_mat;

#ifdef __cplusplus
}
#endif
#endif /* HCORE_AMD_HCORE_S_H */
