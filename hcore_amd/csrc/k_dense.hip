// k_dense.hip — dense triangular kernels of the Cholesky panel: the diagonal
// blocks of a blocked triangular solve (HCORE_dtrsm, reference
// src/compute/hcore_dtrsm.c:24-48) and of a blocked Cholesky (HCORE_dpotrf,
// reference src/compute/hcore_dpotrf.c:23-34).  The off-diagonal work of
// both runs as batched matrix-core GEMMs (k_gemm.hip); these kernels only
// cover the kTrsmNB x kTrsmNB blocks on the diagonal.
#include <hip/hip_runtime.h>

#include <cmath>

#include "hcore_internal.hpp"
#include "hcore_wave.hpp"

namespace hcore {

// One diagonal block of a triangular solve.  A wavefront per 64 right-hand
// sides, one lane per right-hand side: the block S (kTrsmNB x kTrsmNB) is
// staged in LDS (read as a broadcast), the lane's right-hand side lives in
// registers (the substitution is fully unrolled, so every index is a
// constant), and each column of S is applied to the remaining entries as
// soon as its unknown is known (column-oriented substitution).
template <typename T>
__global__ __launch_bounds__(64) void k_trsm_diag(TrsmDiagProb<T> P) {
  constexpr int NB = kTrsmNB;
  __shared__ T S[NB][NB + 1];
  if (P.skip && *P.skip) return;
  const int lane = threadIdx.x;
  const int nbk = P.nbk;
  // S(k, l), identity beyond nbk
  for (int e = lane; e < NB * NB; e += 64) {
    const int k = e % NB, l = e / NB;
    T v = k == l ? T(1) : T(0);
    if (k < nbk && l < nbk) {
      const int p = P.b0 + (P.trans ? l : k), q = P.b0 + (P.trans ? k : l);
      v = P.A[p + (size_t)q * P.lda];
      if (P.unit && k == l) v = T(1);
    }
    S[k][l] = v;
  }
  __syncthreads();
  const int c = blockIdx.x * 64 + lane;
  const bool act = c < P.nrhs;
  T z[NB];
  T* base = P.B + (size_t)c * P.cs + (size_t)P.b0 * P.es;
#pragma unroll
  for (int k = 0; k < NB; ++k) z[k] = (act && k < nbk) ? P.alpha * base[(size_t)k * P.es] : T(0);
  if (P.lower) {
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      z[k] = z[k] / S[k][k];
#pragma unroll
      for (int l = k + 1; l < NB; ++l) z[l] = fma(-S[l][k], z[k], z[l]);
    }
  } else {
#pragma unroll
    for (int k = NB - 1; k >= 0; --k) {
      z[k] = z[k] / S[k][k];
#pragma unroll
      for (int l = 0; l < k; ++l) z[l] = fma(-S[l][k], z[k], z[l]);
    }
  }
  if (act) {
#pragma unroll
    for (int k = 0; k < NB; ++k)
      if (k < nbk) base[(size_t)k * P.es] = z[k];
  }
}

template <typename T>
void launch_trsm_diag(const TrsmDiagProb<T>& p, hipStream_t s) {
  if (p.nrhs <= 0 || p.nbk <= 0) return;
  hipLaunchKernelGGL(k_trsm_diag<T>, dim3((p.nrhs + 63) / 64), dim3(64), 0, s, p);
}

// Unblocked right-looking Cholesky of one diagonal block, held in LDS as the
// lower factor (the upper case works on the transposed view, U = L^T).
template <typename T>
__global__ __launch_bounds__(256) void k_potf2(T* A, int lda, int b0, int nbk, int upper,
                                               int* info) {
  constexpr int NB = kTrsmNB;
  __shared__ T S[NB][NB + 1];
  __shared__ int bad;
  if (*info) return;
  const int t = threadIdx.x;
  auto at = [&](int k, int l) -> T& {  // stored entry behind S(k, l), k >= l
    const int p = b0 + (upper ? l : k), q = b0 + (upper ? k : l);
    return A[p + (size_t)q * lda];
  };
  for (int e = t; e < nbk * nbk; e += 256) {
    const int k = e % nbk, l = e / nbk;
    if (k >= l) S[k][l] = at(k, l);
  }
  if (t == 0) bad = 0;
  __syncthreads();
  int kdone = nbk;
  for (int k = 0; k < nbk; ++k) {
    if (t == 0) {
      const T d = S[k][k];
      if (!(d > T(0))) bad = 1;  // also NaN
      else S[k][k] = sqrt(d);
    }
    __syncthreads();
    if (bad) {
      kdone = k;
      break;
    }
    const T dk = S[k][k];
    for (int l = k + 1 + t; l < nbk; l += 256) S[l][k] = S[l][k] / dk;
    __syncthreads();
    const int rem = nbk - k - 1;
    for (int e = t; e < rem * rem; e += 256) {
      const int l = k + 1 + e % rem, m = k + 1 + e / rem;
      if (l >= m) S[l][m] = fma(-S[l][k], S[m][k], S[l][m]);
    }
    __syncthreads();
  }
  // the factored columns (all of them, or those before the failing pivot)
  for (int e = t; e < nbk * nbk; e += 256) {
    const int k = e % nbk, l = e / nbk;
    if (k >= l && l < kdone) at(k, l) = S[k][l];
  }
  if (t == 0 && kdone < nbk) *info = b0 + kdone + 1;
}

template <typename T>
void launch_potf2(T* A, int lda, int b0, int nbk, int upper, int* info, hipStream_t s) {
  if (nbk <= 0) return;
  hipLaunchKernelGGL(k_potf2<T>, dim3(1), dim3(256), 0, s, A, lda, b0, nbk, upper, info);
}

template <typename T>
__global__ __launch_bounds__(256) void k_scale(T* B, int ldb, int m, int n, T alpha) {
  const int j = blockIdx.y;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256) {
    T* b = B + i + (size_t)j * ldb;
    *b = alpha == T(0) ? T(0) : alpha * *b;
  }
}

template <typename T>
void launch_scale(T* B, int ldb, int m, int n, T alpha, hipStream_t s) {
  if (m <= 0 || n <= 0) return;
  hipLaunchKernelGGL(k_scale<T>, dim3((m + 255) / 256, n), dim3(256), 0, s, B, ldb, m, n, alpha);
}

template void launch_trsm_diag<double>(const TrsmDiagProb<double>&, hipStream_t);
template void launch_potf2<double>(double*, int, int, int, int, int*, hipStream_t);
template void launch_scale<double>(double*, int, int, int, double, hipStream_t);

}  // namespace hcore
