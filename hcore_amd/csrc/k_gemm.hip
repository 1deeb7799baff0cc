// k_gemm.hip — batched GEMM kernels of the recompression path.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <limits>
#include <type_traits>

#include "hcore_internal.hpp"
#include "hcore_wave.hpp"

namespace hcore {

// ------------------------------------------------------------------------------
// batched GEMM on 64x64 tiles, 4x4 outputs per thread, k staged through LDS
// ------------------------------------------------------------------------------
constexpr int GK = 32;  // k chunk

template <typename T>
__device__ __forceinline__ T gemm_masked(const T* X, int ld, int p, int q, bool upper, bool unit_lower) {
  if (upper && p > q) return T(0);
  if (unit_lower && p <= q) return p == q ? T(1) : T(0);
  return X[p + (size_t)q * ld];
}

// Load the k chunk starting at k0 of op(A) (rows row0..) and op(B) (columns
// col0..) into registers: LD elements of each per thread, consecutive lanes
// on consecutive addresses for every transpose combination.
template <typename T, int TILE, int NT, int LD, int GKX = GK>
__device__ __forceinline__ void gemm_fetch(const GemmProb<T>& P, int n, int row0, int col0, int k0,
                                           bool ta, bool tb, bool ua, bool ub, T (&ra)[LD],
                                           T (&rb)[LD], int t) {
  const bool la = (P.flags & GF_A_UNIT_LOWER) != 0, lb = (P.flags & GF_B_UNIT_LOWER) != 0;
#pragma unroll
  for (int q = 0; q < LD; ++q) {
    const int idx = t + NT * q;
    int i, kk;
    if (!ta) { i = idx % TILE; kk = idx / TILE; }
    else     { kk = idx % GKX; i = idx / GKX; }
    const int gi = row0 + i, gk = k0 + kk;
    T v = T(0);
    if (gi < P.m && gk < P.k) {
      const int p = ta ? gk : gi, qq = ta ? gi : gk;  // stored (row, col)
      v = gemm_masked(P.A, P.lda, p, qq, ua, la);
    }
    ra[q] = v;
  }
#pragma unroll
  for (int q = 0; q < LD; ++q) {
    const int idx = t + NT * q;
    int j, kk;
    if (tb) { j = idx % TILE; kk = idx / TILE; }
    else    { kk = idx % GKX; j = idx / GKX; }
    const int gj = col0 + j, gk = k0 + kk;
    T v = T(0);
    if (gj < n && gk < P.k) {
      const int p = tb ? gj : gk, qq = tb ? gk : gj;
      v = gemm_masked(P.B, P.ldb, p, qq, ub, lb);
    }
    rb[q] = v;
  }
}

// C = alpha op(A) op(B) + beta C (or beta C0) for the 4 x 4 outputs of one
// thread of a TD x TD thread tile.
template <typename T, int TD>
__device__ __forceinline__ void gemm_store(const GemmProb<T>& P, const T (&acc)[4][4], int row0,
                                           int col0, int tx, int ty, int n) {
  const bool lzu = (P.flags & GF_LOWER_ZERO_UPPER) != 0;
  const bool bz = (P.flags & GF_BETA_ZERO) != 0;
  const bool c0 = (P.flags & GF_C0) != 0;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int i = row0 + tx + TD * a, j = col0 + ty + TD * b;
      if (i < P.m && j < n) {
        T* c = P.C + i + (size_t)j * P.ldc;
        if (lzu && i < j) { *c = T(0); continue; }
        T v = P.alpha * acc[a][b];
        if (c0) {
          if (i < P.m0) v = fma(P.beta, P.C0[i + (size_t)j * P.ldc0], v);
        } else if (!bz) {
          v = fma(P.beta, *c, v);
        }
        *c = v;
      }
    }
}

// Batched GEMM on TILE x TILE output tiles (64: 256 threads; 32: one
// wavefront), 4 x 4 outputs per thread, k staged through LDS in chunks of GK
// with the next chunk's global loads in flight during the current chunk.
template <typename T, int TILE>
__global__ __launch_bounds__((TILE / 4) * (TILE / 4)) void k_gemm(
    const GemmProb<T>* __restrict__ probs, const GemmTask* __restrict__ tasks) {
  constexpr int TD = TILE / 4;       // threads per tile edge
  constexpr int NT = TD * TD;        // threads
  constexpr int LD = TILE * GK / NT; // elements of each operand loaded per thread per chunk
  // +1 column of padding: the transposed fills write consecutive k at one i
  __shared__ T As[GK][TILE + 1];
  __shared__ T Bs[GK][TILE + 1];
  const GemmTask tk = tasks[blockIdx.x];
  const GemmProb<T> P = probs[tk.prob];
  if (P.skip && *P.skip) return;
  int n = P.n;
  if (P.n_dev) n = min(n, *P.n_dev);
  const int row0 = tk.ti * TILE, col0 = tk.tj * TILE;
  if (row0 >= P.m || col0 >= n) return;
  const int t = threadIdx.x;
  const int tx = t % TD, ty = t / TD;
  const bool lzu = (P.flags & GF_LOWER_ZERO_UPPER) != 0;
  if (lzu && row0 + TILE - 1 < col0) {  // tile strictly above the diagonal: zero fill
    for (int e = t; e < TILE * TILE; e += NT) {
      const int i = row0 + (e % TILE), j = col0 + (e / TILE);
      if (i < P.m && j < n) P.C[i + (size_t)j * P.ldc] = T(0);
    }
    return;
  }
  const bool ta = P.flags & GF_A_TRANS, tb = P.flags & GF_B_TRANS;
  const bool ua = P.flags & GF_A_UPPER, ub = P.flags & GF_B_UPPER;
  T acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = T(0);

  T ra[LD], rb[LD];
  gemm_fetch<T, TILE, NT, LD>(P, n, row0, col0, 0, ta, tb, ua, ub, ra, rb, t);
  for (int k0 = 0; k0 < P.k; k0 += GK) {
#pragma unroll
    for (int q = 0; q < LD; ++q) {
      const int idx = t + NT * q;
      if (!ta) As[idx / TILE][idx % TILE] = ra[q];
      else     As[idx % GK][idx / GK] = ra[q];
      if (tb) Bs[idx / TILE][idx % TILE] = rb[q];
      else    Bs[idx % GK][idx / GK] = rb[q];
    }
    __syncthreads();
    if (k0 + GK < P.k) gemm_fetch<T, TILE, NT, LD>(P, n, row0, col0, k0 + GK, ta, tb, ua, ub, ra, rb, t);
#pragma unroll 8
    for (int kk = 0; kk < GK; ++kk) {
      T av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = As[kk][tx + TD * a];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = Bs[kk][ty + TD * b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
    }
    __syncthreads();
  }
  gemm_store<T, TD>(P, acc, row0, col0, tx, ty, n);
}

// Split-k variant for long reductions (k = nb) into small outputs: KS thread
// groups of TD x TD threads share one output tile, group g taking the k
// chunks g, g + KS, ... (GKS wide, staged through the group's own LDS
// buffers); the groups' partial tiles are then summed in LDS in a fixed
// pairwise order, so the result does not depend on timing.  Every group runs
// the same number of chunk steps (chunks past k load zeros), so the block
// barriers line up.
template <typename T, int TILE, int KS>
__global__ __launch_bounds__((TILE / 4) * (TILE / 4) * KS) void k_gemm_sk(
    const GemmProb<T>* __restrict__ probs, const GemmTask* __restrict__ tasks) {
  constexpr int TD = TILE / 4;
  constexpr int NG = TD * TD;          // threads per group
  constexpr int NT = NG * KS;
  constexpr int GKS = 16;              // k chunk of one group
  constexpr int LD = TILE * GKS / NG;  // elements of each operand per thread per chunk
  constexpr int STG = 2 * GKS * (TILE + 1);
  constexpr int RED = (KS / 2) * TILE * TILE;
  constexpr int SM = KS * STG > RED ? KS * STG : RED;
  static_assert((KS & (KS - 1)) == 0 && KS >= 2 && LD >= 1, "split-k shape");
  __shared__ T sm[SM];
  const GemmTask tk = tasks[blockIdx.x];
  const GemmProb<T> P = probs[tk.prob];
  if (P.skip && *P.skip) return;
  int n = P.n;
  if (P.n_dev) n = min(n, *P.n_dev);
  const int row0 = tk.ti * TILE, col0 = tk.tj * TILE;
  if (row0 >= P.m || col0 >= n) return;
  const int grp = threadIdx.x / NG, t = threadIdx.x % NG;
  const int tx = t % TD, ty = t / TD;
  const bool lzu = (P.flags & GF_LOWER_ZERO_UPPER) != 0;
  if (lzu && row0 + TILE - 1 < col0) {
    for (int e = threadIdx.x; e < TILE * TILE; e += NT) {
      const int i = row0 + (e % TILE), j = col0 + (e / TILE);
      if (i < P.m && j < n) P.C[i + (size_t)j * P.ldc] = T(0);
    }
    return;
  }
  const bool ta = P.flags & GF_A_TRANS, tb = P.flags & GF_B_TRANS;
  const bool ua = P.flags & GF_A_UPPER, ub = P.flags & GF_B_UPPER;
  T (*As)[TILE + 1] = reinterpret_cast<T (*)[TILE + 1]>(sm + grp * STG);
  T (*Bs)[TILE + 1] = reinterpret_cast<T (*)[TILE + 1]>(sm + grp * STG + GKS * (TILE + 1));
  T acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = T(0);
  const int nsteps = (P.k + GKS * KS - 1) / (GKS * KS);
  T ra[LD], rb[LD];
  gemm_fetch<T, TILE, NG, LD, GKS>(P, n, row0, col0, grp * GKS, ta, tb, ua, ub, ra, rb, t);
  for (int st = 0; st < nsteps; ++st) {
#pragma unroll
    for (int q = 0; q < LD; ++q) {
      const int idx = t + NG * q;
      if (!ta) As[idx / TILE][idx % TILE] = ra[q];
      else     As[idx % GKS][idx / GKS] = ra[q];
      if (tb) Bs[idx / TILE][idx % TILE] = rb[q];
      else    Bs[idx % GKS][idx / GKS] = rb[q];
    }
    __syncthreads();
    if (st + 1 < nsteps)
      gemm_fetch<T, TILE, NG, LD, GKS>(P, n, row0, col0, ((st + 1) * KS + grp) * GKS, ta, tb, ua,
                                       ub, ra, rb, t);
#pragma unroll 8
    for (int kk = 0; kk < GKS; ++kk) {
      T av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = As[kk][tx + TD * a];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = Bs[kk][ty + TD * b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
    }
    __syncthreads();
  }
  // pairwise sum of the groups' tiles: at level h, group g (g % 2h == h) hands
  // its tile to group g - h
#pragma unroll
  for (int h = KS / 2; h >= 1; h /= 2) {
    if (grp >= h && grp < 2 * h) {
      T* buf = sm + (size_t)(grp - h) * TILE * TILE;
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) buf[(tx + TD * a) + (ty + TD * b) * TILE] = acc[a][b];
    }
    __syncthreads();
    if (grp < h) {
      const T* buf = sm + (size_t)grp * TILE * TILE;
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] += buf[(tx + TD * a) + (ty + TD * b) * TILE];
    }
    __syncthreads();
  }
  if (grp == 0) gemm_store<T, TD>(P, acc, row0, col0, tx, ty, n);
}

// Tall-skinny batched GEMM (m = the nb rows of a tile, n and k at most a few
// dozen): C = alpha A op(B) + beta C (or beta C0), A not transposed.  One
// workgroup covers kTallRows rows (two per thread) and kTallCols columns; the
// k x kTallCols block of op(B) is staged in LDS in chunks of kTallK rows and
// read as a broadcast, A is read straight from global memory, coalesced along
// the rows.  The many tiny output tiles of such products as 2-D tiles would
// be bound by workgroup dispatch, not by their bytes.
constexpr int kTallQ = 8;  // k steps whose A loads are issued together
static_assert(kTallK % kTallQ == 0, "k chunk of the tall GEMM");
template <typename T>
__global__ __launch_bounds__(256) void k_gemm_tall(const GemmProb<T>* __restrict__ probs,
                                                   const GemmTask* __restrict__ tasks) {
  __shared__ T Bs[kTallK][kTallCols];
  const GemmTask tk = tasks[blockIdx.x];
  const GemmProb<T> P = probs[tk.prob];
  if (P.skip && *P.skip) return;
  int n = P.n;
  if (P.n_dev) n = min(n, *P.n_dev);
  const int row0 = tk.ti * kTallRows, col0 = tk.tj * kTallCols;
  if (row0 >= P.m || col0 >= n) return;
  const int nc = min(kTallCols, n - col0);
  const int t = threadIdx.x;
  const int r0 = row0 + t, r1 = row0 + t + 256;
  const bool tb = (P.flags & GF_B_TRANS) != 0;
  const bool ua = (P.flags & GF_A_UPPER) != 0, ub = (P.flags & GF_B_UPPER) != 0;
  const bool la = (P.flags & GF_A_UNIT_LOWER) != 0, lb = (P.flags & GF_B_UNIT_LOWER) != 0;
  T acc0[kTallCols], acc1[kTallCols];
#pragma unroll
  for (int j = 0; j < kTallCols; ++j) acc0[j] = acc1[j] = T(0);
  for (int k0 = 0; k0 < P.k; k0 += kTallK) {
    const int kc = min(kTallK, P.k - k0);
    for (int e = t; e < kTallK * kTallCols; e += 256) {
      int kk, j;
      if (tb) { j = e % kTallCols; kk = e / kTallCols; }
      else    { kk = e % kTallK; j = e / kTallK; }
      T v = T(0);
      if (kk < kc && j < nc) {
        const int gk = k0 + kk, gj = col0 + j;
        v = tb ? gemm_masked(P.B, P.ldb, gj, gk, ub, lb) : gemm_masked(P.B, P.ldb, gk, gj, ub, lb);
      }
      Bs[kk][j] = v;
    }
    __syncthreads();
    for (int kq = 0; kq < kc; kq += kTallQ) {
      // kTallQ columns of A in flight at once
      T a0[kTallQ], a1[kTallQ];
#pragma unroll
      for (int q = 0; q < kTallQ; ++q) {
        const int gk = k0 + kq + q;
        const bool in = kq + q < kc;
        a0[q] = in && r0 < P.m ? gemm_masked(P.A, P.lda, r0, gk, ua, la) : T(0);
        a1[q] = in && r1 < P.m ? gemm_masked(P.A, P.lda, r1, gk, ua, la) : T(0);
      }
#pragma unroll
      for (int q = 0; q < kTallQ; ++q) {
#pragma unroll
        for (int j = 0; j < kTallCols; ++j) {
          const T b = Bs[kq + q][j];  // zero past kc
          acc0[j] = fma(a0[q], b, acc0[j]);
          acc1[j] = fma(a1[q], b, acc1[j]);
        }
      }
    }
    __syncthreads();
  }
  const bool bz = (P.flags & GF_BETA_ZERO) != 0;
  const bool c0 = (P.flags & GF_C0) != 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int i = h ? r1 : r0;
    if (i >= P.m) continue;
#pragma unroll
    for (int j = 0; j < kTallCols; ++j) {
      if (j < nc) {
        const int gj = col0 + j;
        T* c = P.C + i + (size_t)gj * P.ldc;
        T v = P.alpha * (h ? acc1[j] : acc0[j]);
        if (c0) {
          if (i < P.m0) v = fma(P.beta, P.C0[i + (size_t)gj * P.ldc0], v);
        } else if (!bz) {
          v = fma(P.beta, *c, v);
        }
        *c = v;
      }
    }
  }
}

template <typename T>
void launch_gemm_tall(const GemmProb<T>* probs, const GemmTask* tasks, int ntasks, hipStream_t s) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(k_gemm_tall<T>, dim3(ntasks), dim3(256), 0, s, probs, tasks);
}

template <typename T>
void launch_gemm(const GemmProb<T>* probs, const GemmTask* tasks, int ntasks, int tile,
                 hipStream_t s, bool split) {
  if (ntasks <= 0) return;
  // Only the 32-tiles are split (four groups of one wavefront): a split 64-tile
  // workgroup needs a whole CU's wave slots and LDS at once and delays the
  // other lanes' kernels more than it gains.
  if (split && tile == 32) {
    hipLaunchKernelGGL((k_gemm_sk<T, 32, 4>), dim3(ntasks), dim3(256), 0, s, probs, tasks);
    return;
  }
  if (tile == 32)
    hipLaunchKernelGGL((k_gemm<T, 32>), dim3(ntasks), dim3(64), 0, s, probs, tasks);
  else
    hipLaunchKernelGGL((k_gemm<T, 64>), dim3(ntasks), dim3(256), 0, s, probs, tasks);
}

// ---- explicit instantiations ---------------------------------------------------
#define HCORE_INSTANTIATE(T) \
  template void launch_gemm<T>(const GemmProb<T>*, const GemmTask*, int, int, hipStream_t, bool); \
  template void launch_gemm_tall<T>(const GemmProb<T>*, const GemmTask*, int, hipStream_t);
HCORE_INSTANTIATE(double)
HCORE_INSTANTIATE(float)
#undef HCORE_INSTANTIATE

}  // namespace hcore
