// core_complete.hpp — completion of the left singular vectors of the core SVD
// to an orthonormal set.
//
// LAPACK's dgesvd('A','A') (reference hcore_dgemm.c:403-406) always returns an
// orthonormal U, also for zero singular values; U_r is then applied to Q_A, so
// the new CU is orthonormal whatever the rank.  The Jacobi cores here derive
// U from W / sigma, which leaves the columns of zero singular values (an
// exactly cancelling update, or a fixed rank above the rank of T) zero.  This
// pass replaces every zero column i < r of Z (K x r, leading dimension K) by a
// unit vector orthogonal to all other columns: e_m for the row m with the
// smallest squared row norm over the columns (the direction least covered),
// orthogonalised twice by classical Gram-Schmidt and normalised.  It runs
// only when the core saw such a column, so the common path pays nothing.
#pragma once
#include <hip/hip_runtime.h>

#include "hcore_wave.hpp"

namespace hcore {

// Block-wide sum; part: NT/64 elements of LDS.  Every thread gets the total.
template <typename T, int NT>
__device__ __forceinline__ T complete_block_sum(T v, T* part) {
  constexpr int NW = NT / 64;
  v = wave_sum(v);
  if constexpr (NW == 1) {
    return v;
  } else {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    T s = T(0);
#pragma unroll
    for (int i = 0; i < NW; ++i) s += part[i];
    __syncthreads();
    return s;
  }
}

// Block-wide argmin of (val, idx), ties to the smaller index.  part/ipart:
// NT/64 elements of LDS each.  Every thread gets the winning index.
template <typename T, int NT>
__device__ __forceinline__ int complete_block_argmin(T val, int idx, T* part, int* ipart) {
  constexpr int NW = NT / 64;
#pragma unroll
  for (int off = 32; off >= 1; off /= 2) {
    const T ov = __shfl_xor(val, off);
    const int oi = __shfl_xor(idx, off);
    if (ov < val || (ov == val && oi < idx)) {
      val = ov;
      idx = oi;
    }
  }
  if constexpr (NW == 1) {
    return idx;
  } else {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
      part[w] = val;
      ipart[w] = idx;
    }
    __syncthreads();
    T bv = part[0];
    int bi = ipart[0];
    for (int i = 1; i < NW; ++i)
      if (part[i] < bv || (part[i] == bv && ipart[i] < bi)) {
        bv = part[i];
        bi = ipart[i];
      }
    __syncthreads();
    return bi;
  }
}

// Z: K x r (ld K) in global memory, written by this block; v: K elements and
// d: r elements of scratch (LDS or global, private to the block); part/ipart:
// NT/64 elements of LDS each.  Must be called by every thread of the block.
template <typename T, int NT>
__device__ void complete_basis(T* Z, int K, int r, T* v, T* d, T* part, int* ipart) {
  const int tid = threadIdx.x;
  __threadfence_block();
  block_sync<NT>();
  for (int i = 0; i < r; ++i) {
    T s = T(0);
    for (int row = tid; row < K; row += NT) s = fma(Z[row + (size_t)i * K], Z[row + (size_t)i * K], s);
    s = complete_block_sum<T, NT>(s, part);
    if (s > T(0)) continue;  // uniform: the column is in use
    // the row least covered by the other columns
    T best = T(3.0e38);
    int bm = K;
    for (int row = tid; row < K; row += NT) {
      T q = T(0);
      for (int c = 0; c < r; ++c) q = fma(Z[row + (size_t)c * K], Z[row + (size_t)c * K], q);
      if (q < best) {
        best = q;
        bm = row;
      }
    }
    const int m = complete_block_argmin<T, NT>(best, bm, part, ipart);
    // first Gram-Schmidt pass against e_m: v = e_m - sum_c Z(m, c) Z(:, c)
    for (int row = tid; row < K; row += NT) {
      T acc = row == m ? T(1) : T(0);
      for (int c = 0; c < r; ++c)
        if (c != i) acc = fma(-Z[m + (size_t)c * K], Z[row + (size_t)c * K], acc);
      v[row] = acc;
    }
    __threadfence_block();
    block_sync<NT>();
    // second pass: d = Z^T v, v -= Z d
    for (int c = tid; c < r; c += NT) {
      T acc = T(0);
      if (c != i)
        for (int row = 0; row < K; ++row) acc = fma(Z[row + (size_t)c * K], v[row], acc);
      d[c] = acc;
    }
    __threadfence_block();
    block_sync<NT>();
    T nv = T(0);
    for (int row = tid; row < K; row += NT) {
      T acc = v[row];
      for (int c = 0; c < r; ++c)
        if (c != i) acc = fma(-d[c], Z[row + (size_t)c * K], acc);
      v[row] = acc;
      nv = fma(acc, acc, nv);
    }
    nv = complete_block_sum<T, NT>(nv, part);
    const T inv = T(1) / sqrt(nv);
    for (int row = tid; row < K; row += NT) Z[row + (size_t)i * K] = v[row] * inv;
    __threadfence_block();
    block_sync<NT>();
  }
}

}  // namespace hcore
