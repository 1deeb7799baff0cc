/*
 * flop_counts.h — closed-form operation counts used for the flop_counter.
 * Replaces the reference's src/misc/include/flop_counts.h:1-5; the formulas
 * (src/misc/descutil/flop_counts.c:3-70) are reproduced exactly so that a
 * caller summing flop_counter.update gets the same figures as before.
 *
 * op: 'q' geqrf(m=a, n=b); 'c' potrf(n=a); 't' trsm(m=a, n=b, side=c);
 *     'm' gemm(m=a, n=b, k=c); 's' svd(n=a); 'o' ormqr(m=a, n=b, k=c, side=d);
 *     'r' trmm(m=a, n=b, side=c).  side: 1 = left, 2 = right.
 */
#ifndef HCORE_AMD_FLOP_COUNTS_H
#define HCORE_AMD_FLOP_COUNTS_H
#ifdef __cplusplus
extern "C" {
#endif

unsigned long int flop_counts(char op, unsigned long int a, unsigned long int b,
                              unsigned long int c, unsigned long int d);

#ifdef __cplusplus
}
#endif
#endif /* HCORE_AMD_FLOP_COUNTS_H */
