/*
 * hcore.h — common definitions of the MI355X-native HCORE tile low-rank BLAS.
 *
 * Same enum values and type as the reference's public header
 * (src/include/hcore.h:30-44) so existing callers compile unchanged.  Unlike
 * the reference this header pulls in no CBLAS/LAPACKE header: the library
 * computes on the GPU and needs no host BLAS.
 */
#ifndef HCORE_AMD_HCORE_H
#define HCORE_AMD_HCORE_H

#ifndef hcore_min
#define hcore_min(a, b) ((a) < (b) ? (a) : (b))
#endif
#ifndef hcore_max
#define hcore_max(a, b) ((a) < (b) ? (b) : (a))
#endif

/* Values equal the CBLAS enumerators (reference src/include/hcore.h:30-42). */
#define HCORE_NoTrans    111
#define HCORE_Trans      112
#define HCORE_ConjTrans  113

#define HCORE_Upper      121
#define HCORE_Lower      122
#define HCORE_UpperLower 123

#define HCORE_NonUnit    131
#define HCORE_Unit       132

#define HCORE_Left       141
#define HCORE_Right      142

typedef int HCORE_enum; /* reference src/include/hcore.h:44 */

#endif /* HCORE_AMD_HCORE_H */
