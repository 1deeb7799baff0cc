// hcore_kernels.hip — CDNA4 (gfx950) kernels of the TLR recompression path.
//
// One workgroup (256 threads = 4 wavefronts of 64) per work item.  The hot
// path of HCORE_dgemm (reference src/compute/hcore_dgemm.c:703-1012) becomes
//   k_stack   clone/scale of the factors into stacked panels   (hcore_dgemm.c:760-799, 106-114)
//   k_gemm    skinny products F = AV^T BV, G = BU F^T, the core
//             T = R_A R_B^T and the dense updates of dsyrk/dgemmcd
//   k_geqrf   blocked Householder QR of the nb x K stacks      (hcore_dgemm.c:123, 216)
//   k_core    one-sided Jacobi SVD of T + rank truncation       (hcore_dgemm.c:403-489)
//   k_apply   Q_A [U_r;0] and Q_B [V_r S_r;0] by compact-WY     (hcore_dgemm.c:543, 655)
//
// Reductions over the nb rows of a tile use wavefront shuffles; panels of
// Householder vectors are staged in LDS; the core matrix of the Jacobi SVD
// lives in LDS when it fits (gfx950: 160 KiB per workgroup).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <limits>
#include <type_traits>

#include "hcore_internal.hpp"

namespace hcore {

// ------------------------------------------------------------------------------
// wavefront helpers
// ------------------------------------------------------------------------------
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane exchanges without the LDS crossbar.  All of them require a full wave
// (every lane active), which is how every caller below uses them.
//   offsets 1, 2: DPP quad_perm; 4: DPP row_half_mirror (lane i <-> 7 - i in
//   each group of 8, which flips bit 2 and keeps bits 3..5); 8: DPP row_ror:8;
//   16, 32: v_permlane16_swap / v_permlane32_swap (gfx950).
template <int OFF>
__device__ __forceinline__ unsigned dpp_partner(unsigned v) {
  static_assert(OFF == 1 || OFF == 2 || OFF == 4 || OFF == 8, "DPP partner offset");
  if constexpr (OFF == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xf, 0xf, false);
  else if constexpr (OFF == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xf, 0xf, false);
  else if constexpr (OFF == 4) return __builtin_amdgcn_mov_dpp(v, 0x141, 0xf, 0xf, false);
  else return __builtin_amdgcn_mov_dpp(v, 0x128, 0xf, 0xf, false);
}

// (x, y) -> in lanes whose bit OFF is clear: x + x(partner); where it is set:
// y + y(partner), for OFF = 16 or 32: one permlane swap per 32-bit word.
template <int OFF>
__device__ __forceinline__ float swap_sum(float x, float y) {
  static_assert(OFF == 16 || OFF == 32, "swap offset");
  const unsigned xb = __float_as_uint(x), yb = __float_as_uint(y);
  if constexpr (OFF == 32) {
    const auto r = __builtin_amdgcn_permlane32_swap(xb, yb, false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
  } else {
    const auto r = __builtin_amdgcn_permlane16_swap(xb, yb, false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
  }
}
template <int OFF>
__device__ __forceinline__ double swap_sum(double x, double y) {
  static_assert(OFF == 16 || OFF == 32, "swap offset");
  const unsigned long long xb = __double_as_longlong(x), yb = __double_as_longlong(y);
  const unsigned xl = (unsigned)xb, xh = (unsigned)(xb >> 32);
  const unsigned yl = (unsigned)yb, yh = (unsigned)(yb >> 32);
  unsigned a0, a1, b0, b1;
  if constexpr (OFF == 32) {
    const auto rl = __builtin_amdgcn_permlane32_swap(xl, yl, false, false);
    const auto rh = __builtin_amdgcn_permlane32_swap(xh, yh, false, false);
    a0 = rl[0]; b0 = rl[1]; a1 = rh[0]; b1 = rh[1];
  } else {
    const auto rl = __builtin_amdgcn_permlane16_swap(xl, yl, false, false);
    const auto rh = __builtin_amdgcn_permlane16_swap(xh, yh, false, false);
    a0 = rl[0]; b0 = rl[1]; a1 = rh[0]; b1 = rh[1];
  }
  const double xn = __longlong_as_double(((unsigned long long)a1 << 32) | a0);
  const double yn = __longlong_as_double(((unsigned long long)b1 << 32) | b0);
  return xn + yn;
}

// Value of the partner lane at offset OFF <= 8 (see dpp_partner).
template <int OFF>
__device__ __forceinline__ float lane_partner(float v) {
  return __uint_as_float(dpp_partner<OFF>(__float_as_uint(v)));
}
template <int OFF>
__device__ __forceinline__ double lane_partner(double v) {
  const unsigned long long b = __double_as_longlong(v);
  const unsigned lo = dpp_partner<OFF>((unsigned)b), hi = dpp_partner<OFF>((unsigned)(b >> 32));
  return __longlong_as_double(((unsigned long long)hi << 32) | lo);
}

// v + v(partner at OFF), for any OFF = 1..32.
template <int OFF, typename T>
__device__ __forceinline__ T add_partner(T v) {
  if constexpr (OFF >= 16) return swap_sum<OFF>(v, v);
  else return v + lane_partner<OFF>(v);
}

template <int OFF, typename T>
__device__ __forceinline__ T butterfly_sum(T v) {
  if constexpr (OFF >= 1) return butterfly_sum<OFF / 2>(add_partner<OFF>(v));
  else return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  return butterfly_sum<32>(v);
}

// Reduce N values (N a power of two <= 64) across the 64 lanes: each exchange
// step hands half of the remaining values to the partner lane, so the cost is
// about N + 6 exchanges instead of 6N.  On return v[0] holds the wave total of
// value index reduce_scatter_index<N>(lane).  The halving is written as a
// compile-time recursion so every array index is a constant (no scratch).
template <int HALF, int OFF, int N, typename T>
__device__ __forceinline__ void rs_level(T (&v)[N], int lane) {
  if constexpr (HALF >= 1) {
    if constexpr (OFF >= 16) {
      // lanes with bit OFF set keep the upper half: the permlane swap of
      // (v[i], v[i + HALF]) delivers both partners' halves at once
#pragma unroll
      for (int i = 0; i < HALF; ++i) v[i] = swap_sum<OFF>(v[i], v[i + HALF]);
    } else {
      const bool upper = (lane & OFF) != 0;
#pragma unroll
      for (int i = 0; i < HALF; ++i) {
        const T send = upper ? v[i] : v[i + HALF];
        const T keep = upper ? v[i + HALF] : v[i];
        v[i] = keep + lane_partner<OFF>(send);
      }
    }
    rs_level<HALF / 2, OFF / 2>(v, lane);
  } else {
    v[0] = butterfly_sum<OFF>(v[0]);
  }
}
template <int N, typename T>
__device__ __forceinline__ void reduce_scatter(T (&v)[N], int lane) {
  static_assert((N & (N - 1)) == 0 && N <= 64, "N must be a power of two <= 64");
  rs_level<N / 2, 32>(v, lane);
}
template <int N>
__device__ __forceinline__ int reduce_scatter_index(int lane) {
  constexpr int LOG = __builtin_ctz(N);
  return (lane >> (6 - LOG)) & (N - 1);
}

template <typename T>
__device__ __forceinline__ T eps_of();
template <>
__device__ __forceinline__ double eps_of<double>() { return DBL_EPSILON * 0.5; }
template <>
__device__ __forceinline__ float eps_of<float>() { return FLT_EPSILON * 0.5f; }

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

// ------------------------------------------------------------------------------
// stacked factors (clone + scale): Us = [beta*CU | alpha*AU], Vs(:, :rc) = CV
// ------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_stack(const StackProb<T>* __restrict__ probs, int nb,
                                                    T alpha, T beta) {
  // blockIdx.y selects kStackCols consecutive columns of the 2rc+ra output columns
  const StackProb<T> P = probs[blockIdx.x];
  const int ncol = 2 * P.rc + P.ra;
  const int c0 = blockIdx.y * kStackCols;
  if (c0 >= ncol) return;
  const int cn = min(kStackCols, ncol - c0);
  for (int cc = 0; cc < cn; ++cc) {  // uniform: one source column at a time
    const int col = c0 + cc;
    const T* src;
    T* dst;
    T sc;
    if (col < P.rc) {
      src = P.CU + (size_t)col * P.ldc; dst = P.Us + (size_t)col * nb; sc = beta;
    } else if (col < P.rc + P.ra) {
      src = P.AU + (size_t)(col - P.rc) * P.lda; dst = P.Us + (size_t)col * nb; sc = alpha;
    } else {
      const int c = col - P.rc - P.ra;
      src = P.CV + (size_t)c * P.ldc; dst = P.Vs + (size_t)c * nb; sc = T(1);
    }
#pragma unroll 4
    for (int row = threadIdx.x; row < nb; row += kThreads) dst[row] = sc * src[row];
  }
}

template <typename T>
void launch_stack(const StackProb<T>* probs, int n, int nb, int maxcols, T alpha, T beta,
                  hipStream_t s) {
  if (n <= 0 || maxcols <= 0) return;
  const int ny = (maxcols + kStackCols - 1) / kStackCols;
  hipLaunchKernelGGL(k_stack<T>, dim3(n, ny), dim3(kThreads), 0, s, probs, nb, alpha, beta);
}

// ------------------------------------------------------------------------------
// Householder QR (LAPACK convention: H = I - tau v v^T, v(j) = 1, R(j,j) = beta)
// ------------------------------------------------------------------------------
// Panel factorization by one wavefront on the LDS panel Yp (PB columns of
// NBP = 64*NBW rows).  Lane l owns rows l + 64 i.  Per column: one reduction
// for the norm, one batched reduce-scatter for all remaining panel columns.
template <typename T, int NBW, int PB>
__device__ void panel_factor(T* Yp, T* tau_s, T* beta_s, T* wsc, int j0, int pb, int lane) {
  constexpr int NBP = 64 * NBW;
  const T* Yl = Yp + lane;
  for (int c = 0; c < pb; ++c) {
    const int j = j0 + c;
    T* col = Yp + c * NBP;
    T ss = T(0);
#pragma unroll
    for (int i = 0; i < NBW; ++i) {
      const int row = lane + 64 * i;
      const T x = col[row];
      if (row > j) ss = fma(x, x, ss);
    }
    ss = wave_sum(ss);
    const T alpha = col[j];
    T beta = alpha, tau = T(0);
    if (ss != T(0)) {
      const T nrm = sqrt(fma(alpha, alpha, ss));
      beta = alpha >= T(0) ? -nrm : nrm;
      tau = (beta - alpha) / beta;
      const T scal = T(1) / (alpha - beta);
      T v[NBW];
#pragma unroll
      for (int i = 0; i < NBW; ++i) {
        const int row = lane + 64 * i;
        const T x = col[row];
        v[i] = row > j ? x * scal : (row == j ? T(1) : T(0));
        if (row > j) col[row] = v[i];
      }
      // w_k = v^T Yp(:, k) for every remaining panel column, in one reduction
      T w[PB];
#pragma unroll
      for (int k = 0; k < PB; ++k) {
        T acc = T(0);
        if (k > c && k < pb) {
#pragma unroll
          for (int i = 0; i < NBW; ++i) acc = fma(v[i], Yl[k * NBP + 64 * i], acc);
        }
        w[k] = acc;
      }
      reduce_scatter<PB>(w, lane);
      if ((lane & ((64 / PB) - 1)) == 0) wsc[reduce_scatter_index<PB>(lane)] = w[0];
      wave_sync();
#pragma unroll
      for (int k = 0; k < PB; ++k) {
        if (k > c && k < pb) {
          const T wk = wsc[k] * tau;
#pragma unroll
          for (int i = 0; i < NBW; ++i) {
            T* p = Yp + k * NBP + lane + 64 * i;
            *p = fma(-v[i], wk, *p);
          }
        }
      }
    }
    if (lane == 0) {
      tau_s[c] = tau;
      beta_s[c] = beta;
    }
    wave_sync();
  }
}

template <typename T, int NBW, int PB, int CB>
__global__ __launch_bounds__(kThreads) void k_geqrf(const QrProb<T>* __restrict__ probs, int nb) {
  constexpr int NBP = 64 * NBW;
  constexpr int NW = PB * CB;  // W block per chunk
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* Yp = reinterpret_cast<T*>(smem_raw);  // PB x NBP
  T* Tt = Yp + PB * NBP;                   // PB x PB, column-major
  T* Sg = Tt + PB * PB;                    // PB x PB
  T* tau_s = Sg + PB * PB;                 // PB
  T* beta_s = tau_s + PB;                  // PB
  T* wsc = beta_s + PB;                    // PB (panel reduction exchange)
  T* scr = wsc + PB;                       // 4 waves x 2 x NW

  const QrProb<T> P = probs[blockIdx.x];
  const int K = P.K;
  T* __restrict__ A = P.A;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  T* wscr0 = scr + wid * 2 * NW;
  T* wscr1 = wscr0 + NW;
  const T* Yl = Yp + lane;  // this lane's rows: Yl[c * NBP + 64 * i]

  for (int j0 = 0; j0 < K; j0 += PB) {
    const int pb = min(PB, K - j0);
    #pragma unroll 1
    for (int e = tid; e < PB * NBP; e += kThreads) {
      const int c = e / NBP, row = e - c * NBP;
      T v = T(0);
      if (c < pb && row < nb) v = A[row + (size_t)(j0 + c) * nb];
      Yp[e] = v;
    }
    if (tid < PB) {
      tau_s[tid] = T(0);
      beta_s[tid] = T(0);
    }
    __syncthreads();
    if (wid == 0) panel_factor<T, NBW, PB>(Yp, tau_s, beta_s, wsc, j0, pb, lane);
    __syncthreads();
    // write the factored panel back: R above/on the diagonal, reflectors below
    #pragma unroll 1
    for (int e = tid; e < pb * NBP; e += kThreads) {
      const int c = e / NBP, row = e - c * NBP;
      if (row >= j0 && row < nb) {
        const T v = (row == j0 + c) ? beta_s[c] : Yp[e];
        A[row + (size_t)(j0 + c) * nb] = v;
      }
    }
    if (tid < pb) P.tau[j0 + tid] = tau_s[tid];
    __syncthreads();
    // turn the LDS panel into the unit lower-trapezoidal reflector block
    #pragma unroll 1
    for (int e = tid; e < PB * NBP; e += kThreads) {
      const int c = e / NBP, row = e - c * NBP;
      if (c >= pb || row < j0 + c) Yp[e] = T(0);
      else if (row == j0 + c) Yp[e] = T(1);
    }
    __syncthreads();
    // Gram S(a,b) = v_a^T v_b for the PB(PB-1)/2 pairs a < b: each wavefront
    // takes a quarter of the pairs and reduces them together
    {
      constexpr int NPAIR = PB * (PB - 1) / 2;
      constexpr int PPW = (NPAIR + 3) / 4;  // pairs per wavefront
      constexpr int NR = PPW <= 8 ? 8 : 16;
      T acc[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        acc[q] = T(0);
        const int pidx = wid * PPW + q;
        if (q < PPW && pidx < NPAIR) {
          int b = 1, a = pidx;
          while (a >= b) { a -= b; ++b; }
#pragma unroll
          for (int i = 0; i < NBW; ++i) acc[q] = fma(Yl[a * NBP + 64 * i], Yl[b * NBP + 64 * i], acc[q]);
        }
      }
      reduce_scatter<NR>(acc, lane);
      const int q = reduce_scatter_index<NR>(lane);
      const int pidx = wid * PPW + q;
      if ((lane & ((64 / NR) - 1)) == 0 && q < PPW && pidx < NPAIR) {
        int b = 1, a = pidx;
        while (a >= b) { a -= b; ++b; }
        Sg[a + b * PB] = acc[0];
      }
    }
    __syncthreads();
    // compact-WY factor (forward, column-wise): T(:c,c) = -tau_c T(:c,:c) S(:c,c)
    if (wid == 0) {
#pragma unroll 1
      for (int c = 0; c < PB; ++c) {
        if (lane < PB) {
          T val = T(0);
          if (lane < c) {
            T s = T(0);
            for (int c2 = lane; c2 < c; ++c2) s = fma(Tt[lane + c2 * PB], Sg[c2 + c * PB], s);
            val = -tau_s[c] * s;
          } else if (lane == c) {
            val = tau_s[c];
          }
          Tt[lane + c * PB] = val;
        }
        wave_sync();
      }
    }
    __syncthreads();
    if (tid < PB * PB) P.Tp[(size_t)(j0 / PB) * PB * PB + tid] = Tt[tid];

    // trailing update A(:, j0+PB:) <- (I - Y T Y^T)^T A, CB columns per wavefront
    const int ncols = K - (j0 + PB);
    if (ncols > 0) {
      const int nch = (ncols + CB - 1) / CB;
      for (int q = wid; q < nch; q += 4) {
        const int k0 = j0 + PB + q * CB;
        const int cb = min(CB, K - k0);
        T X[NBW][CB];
#pragma unroll
        for (int i = 0; i < NBW; ++i) {
          const int row = lane + 64 * i;
#pragma unroll
          for (int kk = 0; kk < CB; ++kk)
            X[i][kk] = (row >= j0 && row < nb && kk < cb) ? A[row + (size_t)(k0 + kk) * nb] : T(0);
        }
        T W[NW];
#pragma unroll
        for (int c = 0; c < PB; ++c) {
          T y[NBW];
#pragma unroll
          for (int i = 0; i < NBW; ++i) y[i] = Yl[c * NBP + 64 * i];
#pragma unroll
          for (int kk = 0; kk < CB; ++kk) {
            T acc = T(0);
#pragma unroll
            for (int i = 0; i < NBW; ++i) acc = fma(y[i], X[i][kk], acc);
            W[c * CB + kk] = acc;
          }
          __builtin_amdgcn_sched_barrier(0);
        }
        reduce_scatter<NW>(W, lane);
        wscr0[reduce_scatter_index<NW>(lane)] = W[0];
        wave_sync();
        if (lane < NW) {
          const int c = lane / CB, kk = lane - c * CB;
          T s = T(0);
          for (int c2 = 0; c2 <= c; ++c2) s = fma(Tt[c2 + c * PB], wscr0[c2 * CB + kk], s);
          wscr1[lane] = s;
        }
        wave_sync();
#pragma unroll
        for (int c = 0; c < PB; ++c) {
          T y[NBW];
#pragma unroll
          for (int i = 0; i < NBW; ++i) y[i] = Yl[c * NBP + 64 * i];
#pragma unroll
          for (int kk = 0; kk < CB; ++kk) {
            const T w2 = wscr1[c * CB + kk];
#pragma unroll
            for (int i = 0; i < NBW; ++i) X[i][kk] = fma(-y[i], w2, X[i][kk]);
          }
          __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int i = 0; i < NBW; ++i) {
          const int row = lane + 64 * i;
          if (row >= j0 && row < nb) {
#pragma unroll
            for (int kk = 0; kk < CB; ++kk)
              if (kk < cb) A[row + (size_t)(k0 + kk) * nb] = X[i][kk];
          }
        }
        wave_sync();
      }
    }
    __syncthreads();
  }
}

template <int NBW>
constexpr int qr_cb() { return NBW >= 32 ? 1 : (NBW >= 16 ? 2 : 4); }

template <typename T, int NBW>
static size_t qr_smem() {
  constexpr int CB = qr_cb<NBW>();
  return sizeof(T) * ((size_t)kPB * 64 * NBW + 2 * kPB * kPB + 3 * kPB + 4 * 2 * kPB * CB);
}

template <typename T, int NBW>
static void launch_geqrf_nbw(const QrProb<T>* probs, int n, int nb, hipStream_t s) {
  constexpr int CB = qr_cb<NBW>();
  const size_t sm = qr_smem<T, NBW>();
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_geqrf<T, NBW, kPB, CB>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    attr = true;
  }
  if (n <= 0) return;
  hipLaunchKernelGGL((k_geqrf<T, NBW, kPB, CB>), dim3(n), dim3(kThreads), sm, s, probs, nb);
}

template <typename T>
void launch_geqrf(const QrProb<T>* probs, int n, int nb, hipStream_t s) {
  if (nb <= 128) launch_geqrf_nbw<T, 2>(probs, n, nb, s);
  else if (nb <= 256) launch_geqrf_nbw<T, 4>(probs, n, nb, s);
  else if (nb <= 512) launch_geqrf_nbw<T, 8>(probs, n, nb, s);
  else if (nb <= 1024) launch_geqrf_nbw<T, 16>(probs, n, nb, s);
  else launch_geqrf_nbw<T, 32>(probs, n, nb, s);
}

// ------------------------------------------------------------------------------
// register-resident Householder QR of the panel of columns [j0, j0 + KMAX) of
// an nb x K stack (nb <= NT * RPT): thread i holds rows i, i + NT, ...,
// i + (RPT - 1) NT of the panel.  Per column g one fused reduction gives
//   s_t = sum_{i>g} x_ig x_it  for every panel column t,
// from which, with row g broadcast through LDS: the norm (t = g), the
// reflector products w_t = tau (x_gt + s_t / (alpha - beta)) (t > g) and the
// Gram entries y_t^T v_g (t < g).  A thread first sums its RPT rows, so the
// cross-lane reduction (the larger part of a column's instructions at one row
// per thread) is paid once per RPT rows.  Every wavefront derives the
// reflector from the reduced partials itself (lane t: entry t), so a column
// costs one block barrier: the partials and row g are double-buffered by
// column parity (a wavefront can only rewrite a buffer after the next
// column's barrier, which every reader of the buffer has passed).  Outputs:
// the panel in LAPACK layout (R on and above the diagonal, reflectors below),
// tau, the panel's Gram lower triangle Sl and the kPB x kPB compact-WY blocks
// on the diagonal of the panel's factor (Sl and taus are written by wavefront
// 0; a block barrier must follow before other wavefronts read them).
// ------------------------------------------------------------------------------
template <int J, typename T, int NT, int RPT, int KMAX>
__device__ __forceinline__ void qr_reg_col(T (&x)[RPT][KMAX], int K, int j0, int tid, int lane,
                                           int wid, T (*part)[NT / 64][KMAX], T (*rowj)[KMAX],
                                           T (*wvw)[KMAX], T (*Sl)[KMAX + 1], T* taus) {
  constexpr int NW = NT / 64;
  constexpr int CH = 16;       // values per batched wave reduction
  constexpr int BUF = J & 1;   // partials / row g buffer of this column
  const int g = j0 + J;        // global column
  if (g >= K) return;          // uniform
  T xj[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) xj[r] = tid + r * NT > g ? x[r][J] : T(0);
#pragma unroll
  for (int c0 = 0; c0 < KMAX; c0 += CH) {
    T sv[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      T acc = xj[0] * x[0][c0 + q];
#pragma unroll
      for (int r = 1; r < RPT; ++r) acc = fma(xj[r], x[r][c0 + q], acc);
      sv[q] = acc;
    }
    reduce_scatter<CH>(sv, lane);
    if ((lane & ((64 / CH) - 1)) == 0)
      part[BUF][wid][c0 + reduce_scatter_index<CH>(lane)] = sv[0];
  }
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    if (tid + r * NT == g) {
#pragma unroll
      for (int k = 0; k < KMAX; ++k) rowj[BUF][k] = x[r][k];
    }
  }
  __syncthreads();
  const int t = lane < KMAX ? lane : KMAX - 1;
  T tot = T(0), ss = T(0);
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    tot += part[BUF][w][t];
    ss += part[BUF][w][J];
  }
  const T alpha = rowj[BUF][J];
  T beta = alpha, tau = T(0), scal = T(0);
  if (ss > T(0)) {  // uniform: every lane reads the same partials
    const T nrm = sqrt(fma(alpha, alpha, ss));
    beta = alpha >= T(0) ? -nrm : nrm;
    tau = (beta - alpha) / beta;
    scal = T(1) / (alpha - beta);
  }
  const T dotv = fma(scal, tot, rowj[BUF][t]);  // y_t^T v_g (t < J), v_g^T x_t (t > J)
  if (lane < KMAX) wvw[wid][lane] = (lane > J && j0 + lane < K) ? tau * dotv : T(0);
  if (wid == 0) {
    if (lane < J) Sl[J][lane] = dotv;
    if (lane == 0) taus[J] = tau;
  }
  wave_sync();
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int row = tid + r * NT;
    if (row > g) {
      const T v = x[r][J] * scal;
      x[r][J] = v;
#pragma unroll
      for (int k = J + 1; k < KMAX; ++k) x[r][k] = fma(-v, wvw[wid][k], x[r][k]);
    } else if (row == g) {
      x[r][J] = beta;
#pragma unroll
      for (int k = J + 1; k < KMAX; ++k) x[r][k] -= wvw[wid][k];
    }
  }
  if constexpr (J + 1 < KMAX)
    qr_reg_col<J + 1, T, NT, RPT, KMAX>(x, K, j0, tid, lane, wid, part, rowj, wvw, Sl, taus);
}

// The panel's compact-WY factor from its Gram entries, column by column:
// T(:c,c) = -tau_c T(:c,:c) S(:c,c), T(c,c) = tau_c (one wavefront), and its
// diagonal kPB x kPB blocks to the blocked format of the stack.
template <typename T, int PB>
__device__ __forceinline__ void qr_panel_factor(const QrProb<T>& P, int K, int j0, int pw,
                                                int lane, T (*Sl)[PB + 1], const T* taus,
                                                T (*Tf)[PB + 1]) {
  for (int c = 0; c < PB; ++c) {
    if (lane < PB) {
      T val = T(0);
      if (c < pw) {
        if (lane < c) {
          T acc = T(0);
          for (int b = lane; b < c; ++b) acc = fma(Tf[lane][b], Sl[c][b], acc);
          val = -taus[c] * acc;
        } else if (lane == c) {
          val = taus[c];
        }
      }
      Tf[lane][c] = val;
    }
    wave_sync();
  }
  for (int e = lane; e < PB * kPB; e += 64) {
    const int blk = e / (kPB * kPB), w2 = e % (kPB * kPB);
    const int a = w2 % kPB, c = w2 / kPB;
    if (j0 + blk * kPB < K)
      P.Tp[(size_t)(j0 / kPB + blk) * kPB * kPB + w2] = Tf[blk * kPB + a][blk * kPB + c];
  }
}

template <typename T, int NT, int RPT, int KMAX>
__global__ __launch_bounds__(NT) void k_qr_reg(const QrProb<T>* __restrict__ probs, int nb, int j0) {
  constexpr int NW = NT / 64;
  static_assert(KMAX % 16 == 0 && KMAX <= 64 && NT >= KMAX && KMAX % kPB == 0, "register QR shape");
  __shared__ T part[2][NW][KMAX];
  __shared__ T rowj[2][KMAX];
  __shared__ T wvw[NW][KMAX];       // each wavefront's copy of the reflector products
  __shared__ T Sl[KMAX][KMAX + 1];  // Sl[j][t] = y_t^T y_j, t < j (panel columns)
  __shared__ T taus[KMAX];
  __shared__ T Tf[KMAX][KMAX + 1];  // the panel's compact-WY factor
  const QrProb<T> P = probs[blockIdx.x];
  const int K = P.K;
  if (j0 >= K) return;
  const int pw = min(KMAX, K - j0);  // panel width
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  T x[RPT][KMAX];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int row = tid + r * NT;
#pragma unroll
    for (int k = 0; k < KMAX; ++k)
      x[r][k] = (row < nb && k < pw) ? P.A[row + (size_t)(j0 + k) * nb] : T(0);
  }
  qr_reg_col<0, T, NT, RPT, KMAX>(x, K, j0, tid, lane, wid, part, rowj, wvw, Sl, taus);
  __syncthreads();
  if (tid < pw) P.tau[j0 + tid] = taus[tid];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int row = tid + r * NT;
    if (row < nb && row >= j0) {
#pragma unroll
      for (int k = 0; k < KMAX; ++k)
        if (k < pw) P.A[row + (size_t)(j0 + k) * nb] = x[r][k];
    }
  }
  if (P.S) {  // whole stack in one panel: its Gram lower triangle
    for (int e = tid; e < K * K; e += NT) {
      const int jj = e % K, t = e / K;
      if (t < jj) P.S[jj + (size_t)t * K] = Sl[jj][t];
    }
  }
  if (wid == 0) qr_panel_factor<T, KMAX>(P, K, j0, pw, lane, Sl, taus, Tf);
}

// Whole-stack QR in one workgroup (nb <= NT * RPT, any K): the panels of PB
// columns are factored in registers by the column recursion above; after each
// panel its compact-WY factor (LDS) updates the trailing columns, four at a
// time (three block barriers per step):
// X <- (I - Y T Y^T)^T X = X - Y T^T (Y^T X), with the reflectors kept in
// the registers that held the panel.  Each thread only ever touches its own
// rows of the stack, so the panels and the trailing columns need no
// inter-thread ordering through global memory.
template <typename T, int NT, int RPT, int PB>
__global__ __launch_bounds__(NT) void k_qr_fused(const QrProb<T>* __restrict__ probs, int nb) {
  constexpr int NW = NT / 64;
  constexpr int CB = 4;
  constexpr int CH = 16;
  static_assert(PB % CH == 0 && PB <= 64 && NT >= PB * CB && PB % kPB == 0, "fused QR shape");
  __shared__ T part[2][NW][PB];
  __shared__ T rowj[2][PB];
  __shared__ T wvw[NW][PB];
  __shared__ T Sl[PB][PB + 1];
  __shared__ T taus[PB];
  __shared__ T Tf[PB][PB + 1];
  __shared__ T tpart[NW][CB][PB];
  __shared__ T Wm[PB][CB];
  __shared__ T W2[PB][CB];
  const QrProb<T> P = probs[blockIdx.x];
  const int K = P.K;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  for (int j0 = 0; j0 < K; j0 += PB) {
    const int pw = min(PB, K - j0);
    T x[RPT][PB];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int row = tid + r * NT;
#pragma unroll
      for (int k = 0; k < PB; ++k)
        x[r][k] = (row < nb && k < pw) ? P.A[row + (size_t)(j0 + k) * nb] : T(0);
    }
    qr_reg_col<0, T, NT, RPT, PB>(x, K, j0, tid, lane, wid, part, rowj, wvw, Sl, taus);
    __syncthreads();
    if (tid < pw) P.tau[j0 + tid] = taus[tid];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int row = tid + r * NT;
      if (row < nb && row >= j0) {
#pragma unroll
        for (int k = 0; k < PB; ++k)
          if (k < pw) P.A[row + (size_t)(j0 + k) * nb] = x[r][k];
      }
    }
    if (wid == 0) qr_panel_factor<T, PB>(P, K, j0, pw, lane, Sl, taus, Tf);
    __syncthreads();
    if (j0 + PB >= K) break;
    // the panel's reflectors (unit lower trapezoidal) in place of its columns
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int row = tid + r * NT;
#pragma unroll
      for (int a = 0; a < PB; ++a) {
        const int g = j0 + a;
        x[r][a] = (a < pw && row < nb) ? (row > g ? x[r][a] : (row == g ? T(1) : T(0))) : T(0);
      }
    }
    for (int c0 = j0 + PB; c0 < K; c0 += CB) {
      const int cb = min(CB, K - c0);
      T xc[RPT][CB];
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int row = tid + r * NT;
#pragma unroll
        for (int c = 0; c < CB; ++c)
          xc[r][c] = (row < nb && row >= j0 && c < cb) ? P.A[row + (size_t)(c0 + c) * nb] : T(0);
      }
#pragma unroll
      for (int c = 0; c < CB; ++c) {
#pragma unroll
        for (int a0 = 0; a0 < PB; a0 += CH) {
          T sv[CH];
#pragma unroll
          for (int q = 0; q < CH; ++q) {
            T acc = x[0][a0 + q] * xc[0][c];
#pragma unroll
            for (int r = 1; r < RPT; ++r) acc = fma(x[r][a0 + q], xc[r][c], acc);
            sv[q] = acc;
          }
          reduce_scatter<CH>(sv, lane);
          if ((lane & ((64 / CH) - 1)) == 0)
            tpart[wid][c][a0 + reduce_scatter_index<CH>(lane)] = sv[0];
          __builtin_amdgcn_sched_barrier(0);  // one reduction's registers live at a time
        }
      }
      __syncthreads();
      if (tid < PB * CB) {
        const int a = tid % PB, c = tid / PB;
        T acc = T(0);
#pragma unroll
        for (int w = 0; w < NW; ++w) acc += tpart[w][c][a];
        Wm[a][c] = acc;
      }
      __syncthreads();
      if (tid < PB * CB) {  // W2 = T^T W (T upper triangular)
        const int a = tid % PB, c = tid / PB;
        T acc = T(0);
        for (int b = 0; b <= a; ++b) acc = fma(Tf[b][a], Wm[b][c], acc);
        W2[a][c] = acc;
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int row = tid + r * NT;
        if (row < nb && row >= j0) {
#pragma unroll
          for (int c = 0; c < CB; ++c) {
            T v = xc[r][c];
#pragma unroll
            for (int a = 0; a < PB; ++a) v = fma(-x[r][a], W2[a][c], v);
            if (c < cb) P.A[row + (size_t)(c0 + c) * nb] = v;
          }
        }
      }
    }
    __syncthreads();
  }
}

// Shapes (threads x rows per thread) of the register QR: the stack's rows are
// spread over NT threads, RPT rows each.  More rows per thread pay the
// cross-lane reductions once per RPT rows; fewer keep more waves per stack.
constexpr int kQr16Threads = 256;  // K <= 16: 4 rows per thread at nb = 1024 (two stacks per CU)
constexpr int kQr32Threads = 512;  // K <= 32: 2 rows per thread at nb = 1024
constexpr int kQrfThreads = 512;   // panels: 2 rows per thread at nb = 1024
constexpr int kQrfPanel = 16;      // panel width of the fused QR (VGPR budget at 2 rows)

template <typename T, int NT, int RPT, int KMAX>
static void launch_qr_reg_shape(const QrProb<T>* probs, int n, int nb, hipStream_t s) {
  hipLaunchKernelGGL((k_qr_reg<T, NT, RPT, KMAX>), dim3(n), dim3(NT), 0, s, probs, nb, 0);
}

// Only the shapes below are instantiated (nb <= 1024): K <= 16 at 256
// threads x 1, 2 or 4 rows; K <= 32 at 256 x 1, 512 x 1 or 512 x 2 rows.
template <typename T>
void launch_qr_reg(const QrProb<T>* probs, int n, int nb, int kmax, hipStream_t s) {
  if (n <= 0) return;
  static_assert(kQr16Threads == 256 && kQr32Threads == 512, "register QR shapes");
  if (kmax <= 16) {
    if (nb <= 256) launch_qr_reg_shape<T, 256, 1, 16>(probs, n, nb, s);
    else if (nb <= 512) launch_qr_reg_shape<T, 256, 2, 16>(probs, n, nb, s);
    else launch_qr_reg_shape<T, 256, 4, 16>(probs, n, nb, s);
  } else {
    if (nb <= 256) launch_qr_reg_shape<T, 256, 1, 32>(probs, n, nb, s);
    else if (nb <= 512) launch_qr_reg_shape<T, 512, 1, 32>(probs, n, nb, s);
    else launch_qr_reg_shape<T, 512, 2, 32>(probs, n, nb, s);
  }
}

template <typename T>
void launch_qr_fused(const QrProb<T>* probs, int n, int nb, hipStream_t s) {
  if (n <= 0) return;
  constexpr int NT = kQrfThreads, PB = kQrfPanel;
  const int rpt = (nb + NT - 1) / NT;
  // nb <= 2 NT (the plans route stacks of at most 1024 rows here)
  if (nb <= 256) hipLaunchKernelGGL((k_qr_fused<T, 256, 1, 32>), dim3(n), dim3(256), 0, s, probs, nb);
  else if (rpt <= 1) hipLaunchKernelGGL((k_qr_fused<T, NT, 1, PB>), dim3(n), dim3(NT), 0, s, probs, nb);
  else hipLaunchKernelGGL((k_qr_fused<T, NT, 2, PB>), dim3(n), dim3(NT), 0, s, probs, nb);
}

// ------------------------------------------------------------------------------
// core SVD: one-sided (Hestenes) Jacobi on the columns of T, round-robin
// ordering, then sort, truncation (hcore_dgemm.c:444-464) and U_r = G_r / sigma
// ------------------------------------------------------------------------------
constexpr int kMaxSweeps = 40;

template <int NT>
__device__ __forceinline__ void block_sync() {
  if constexpr (NT == 64) wave_sync();
  else __syncthreads();
}

// Sum over a group of G consecutive lanes (G a power of two <= 64).
template <int G, typename T>
__device__ __forceinline__ T gsum(T v) {
  return butterfly_sum<G / 2>(v);
}

// Squared column norms of the ncol columns of A (K rows, leading dim ld):
// a group of G lanes per column, rows in registers (K <= G*RPL).
template <typename T, int NT, int G, int RPL>
__device__ void column_norms(const T* A, int ld, int ncol, int K, T* nrm) {
  constexpr int NG = NT / G;
  const int grp = threadIdx.x / G, gl = threadIdx.x % G;
  for (int j = grp; j < ncol; j += NG) {
    const T* c = A + (size_t)j * ld;
    T s = T(0);
#pragma unroll
    for (int i = 0; i < RPL; ++i) {
      const int row = gl + G * i;
      const T x = row < K ? c[row] : T(0);
      s = fma(x, x, s);
    }
    s = gsum<G>(s);
    if (gl == 0) nrm[j] = s;
  }
}

// One cyclic sweep of one-sided Jacobi over the ncol columns of A (round-robin
// ordering; n = ncol rounded up to even; n/2 disjoint pairs per round, each
// handled by a group of G lanes holding RPL rows of both columns in
// registers).  nrm holds the squared column norms, updated per rotation
// (a'a' = aa - t ab, b'b' = bb + t ab).  Pairs whose smaller column is below
// thr2 (squared) are skipped: such columns are truncated anyway and leave an
// error far below acc.  Sets *rotated if any pair was rotated.
template <typename T, int NT, int G, int RPL>
__device__ void jacobi_sweep(T* A, int ld, int ncol, int K, T tol, T thr2, T* nrm, int* rotated) {
  const int tid = threadIdx.x;
  const int n = ncol + (ncol & 1);
  const int npairs = n / 2;
  constexpr int NG = NT / G;
  const int grp = tid / G, gl = tid % G;
  for (int round = 0; round < n - 1; ++round) {
    for (int pid = grp; pid < npairs; pid += NG) {
      int a = 0, b;
      if (pid > 0) {
        a = pid - 1 + round;
        if (a >= n - 1) a -= n - 1;
        a += 1;
      }
      b = n - 2 - pid + round;
      if (b >= n - 1) b -= n - 1;
      b += 1;
      if (a >= ncol || b >= ncol) continue;  // virtual column (uniform in the group)
      T* ca = A + (size_t)a * ld;
      T* cb = A + (size_t)b * ld;
      T x[RPL], y[RPL];
      T apq = T(0);
#pragma unroll
      for (int i = 0; i < RPL; ++i) {
        const int row = gl + G * i;
        const bool ok = row < K;
        x[i] = ok ? ca[row] : T(0);
        y[i] = ok ? cb[row] : T(0);
        apq = fma(x[i], y[i], apq);
      }
      apq = gsum<G>(apq);
      const T app = nrm[a], aqq = nrm[b];
      if (fmin(app, aqq) > thr2 && fabs(apq) > tol * sqrt(app * aqq)) {
        const T zeta = (aqq - app) / (T(2) * apq);
        T t;
        if (fabs(zeta) > T(1e30)) t = T(0.5) / zeta;
        else t = (zeta >= T(0) ? T(1) : T(-1)) / (fabs(zeta) + sqrt(fma(zeta, zeta, T(1))));
        const T cs = T(1) / sqrt(fma(t, t, T(1)));
        const T sn = cs * t;
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
          const int row = gl + G * i;
          if (row < K) {
            ca[row] = cs * x[i] - sn * y[i];
            cb[row] = sn * x[i] + cs * y[i];
          }
        }
        if (gl == 0) {
          nrm[a] = fma(-t, apq, app);
          nrm[b] = fma(t, apq, aqq);
          *rotated = 1;
        }
      }
    }
    block_sync<NT>();
  }
}

// Householder QR with column pivoting of the K x K matrix A (column-major,
// leading dimension ld), in place, reflectors discarded.  Stops at the first
// step whose largest remaining column norm^2 is <= thr2; returns that step k
// (the numerical rank).  On return rdiag[c] = R(c,c) and the strict upper
// triangle of A's first k rows holds R; piv[j] is the original column now in
// position j.  Groups of G lanes update one column each, with the column's
// rows (gl + G i, i < RPL, K <= G*RPL) in registers so that all its loads are
// in flight together; two block barriers per step.
template <typename T, int NT, int G, int RPL>
__device__ int qrcp(T* A, int K, int ld, T thr2, T* cn, T* rdiag, int* piv, T* gv, int* gi, T* red) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  constexpr int NG = NT / G;
  constexpr int NW = NT / 64;
  const int grp = tid / G, gl = tid % G;
  // exact column norms and per-group maxima
  {
    T best = T(-1);
    int bi = -1;
    for (int j = grp; j < K; j += NG) {
      const T* c = A + (size_t)j * ld;
      T sq = T(0);
#pragma unroll
      for (int i = 0; i < RPL; ++i) {
        const int row = gl + G * i;
        const T x = row < K ? c[row] : T(0);
        sq = fma(x, x, sq);
      }
      sq = gsum<G>(sq);
      if (gl == 0) {
        cn[j] = sq;
        piv[j] = j;
      }
      if (sq > best) { best = sq; bi = j; }
    }
    if (gl == 0) { gv[grp] = best; gi[grp] = bi; }
  }
  block_sync<NT>();
  int k = K;
  for (int s = 0; s < K; ++s) {
    // pivot: largest remaining column (ties: smallest index)
    T best = T(-1);
    int p = -1;
    for (int g = 0; g < NG; ++g) {
      const T v = gv[g];
      const int i = gi[g];
      if (i >= s && (v > best || (v == best && i < p))) { best = v; p = i; }
    }
    if (p < 0 || !(best > thr2)) { k = s; break; }
    // swap columns s and p; partial norm^2 of the new column s below row s
    T part = T(0);
    T* cs_ = A + (size_t)s * ld;
    T* cp_ = A + (size_t)p * ld;
    for (int i = tid; i < K; i += NT) {
      const T as = cs_[i], ap = cp_[i];
      if (p != s) { cs_[i] = ap; cp_[i] = as; }
      if (i > s) part = fma(ap, ap, part);
    }
    if (tid == 0 && p != s) {
      const T t = cn[s]; cn[s] = cn[p]; cn[p] = t;
      const int q = piv[s]; piv[s] = piv[p]; piv[p] = q;
    }
    part = wave_sum(part);
    if (NW > 1) {
      if (lane == 0) red[wid] = part;
    }
    block_sync<NT>();
    T xn2 = part;
    if (NW > 1) {
      xn2 = T(0);
      for (int w = 0; w < NW; ++w) xn2 += red[w];
    }
    const T alpha = cs_[s];
    T beta = alpha, tau = T(0), scal = T(0);
    if (xn2 > T(0)) {
      const T nrm = sqrt(fma(alpha, alpha, xn2));
      beta = alpha >= T(0) ? -nrm : nrm;
      tau = (beta - alpha) / beta;
      scal = T(1) / (alpha - beta);
    }
    // this lane's rows of the reflector v (v_s = 1, v_i = A(i,s)*scal, 0 above s)
    T v[RPL];
#pragma unroll
    for (int i = 0; i < RPL; ++i) {
      const int row = gl + G * i;
      v[i] = row > s && row < K ? cs_[row] * scal : (row == s ? T(1) : T(0));
    }
    // apply H = I - tau v v^T to columns j > s, recomputing each column's
    // remaining norm exactly
    T gbest = T(-1);
    int gidx = -1;
    for (int j = s + 1 + grp; j < K; j += NG) {
      T* cj = A + (size_t)j * ld;
      T x[RPL];
      T dot = T(0);
#pragma unroll
      for (int i = 0; i < RPL; ++i) {
        const int row = gl + G * i;
        x[i] = row >= s && row < K ? cj[row] : T(0);
        dot = fma(v[i], x[i], dot);
      }
      const T w = gsum<G>(dot) * tau;
      T sq = T(0);
#pragma unroll
      for (int i = 0; i < RPL; ++i) {
        const int row = gl + G * i;
        if (row >= s && row < K) {
          const T a = fma(-v[i], w, x[i]);
          cj[row] = a;
          if (row > s) sq = fma(a, a, sq);
        }
      }
      sq = gsum<G>(sq);
      if (gl == 0) cn[j] = sq;
      if (sq > gbest) { gbest = sq; gidx = j; }
    }
    if (gl == 0) { gv[grp] = gbest; gi[grp] = gidx; }
    if (tid == 0) rdiag[s] = beta;
    block_sync<NT>();
  }
  return k;
}

// Overwrite A (holding R's first k rows in its upper triangle, rdiag on the
// diagonal) with R^T restricted to k columns: column c < k of the result is
// row c of R (zeros above the diagonal).
template <typename T, int NT>
__device__ void r_transpose(T* A, int K, int ld, int k, const T* rdiag) {
  const size_t npairs = (size_t)K * (K - 1) / 2;
  for (size_t e = threadIdx.x; e < npairs; e += NT) {
    // pair e -> (lo, hi), lo < hi, enumerated column by column of the upper triangle
    int hi = (int)((1.0 + sqrt(1.0 + 8.0 * (double)e)) * 0.5);
    while ((size_t)hi * (hi - 1) / 2 > e) --hi;
    while ((size_t)(hi + 1) * hi / 2 <= e) ++hi;
    const int lo = (int)(e - (size_t)hi * (hi - 1) / 2);
    if (lo >= k) continue;
    const T v = A[lo + (size_t)hi * ld];  // R(lo, hi)
    A[hi + (size_t)lo * ld] = v;          // R^T(hi, lo)
    A[lo + (size_t)hi * ld] = T(0);
  }
  for (int c = threadIdx.x; c < k; c += NT) A[c + (size_t)c * ld] = rdiag[c];
}

// Shared epilogue: singular values (norms of the k orthogonalised columns of
// W = R^T V), descending order, truncation (hcore_dgemm.c:444-464) and the
// left singular vectors U = P W / sigma (rows permuted back by piv).
template <typename T, int NT>
__device__ void core_epilogue(const CoreProb<T>& P, const T* W, int ld, int k, const int* piv, T* sig,
                              int* perm, int* flags, int rk, int maxrk, T acc) {
  const int K = P.K;
  const int tid = threadIdx.x;
  if (tid == 0) flags[1] = K;  // flags[2] (non-finite input) was set at load time
  block_sync<NT>();
  for (int j = tid; j < K; j += NT) {
    T s = T(0);
    if (j < k) {
      const T* cj = W + (size_t)j * ld;
      for (int row = 0; row < K; ++row) s = fma(cj[row], cj[row], s);
      s = sqrt(s);
    }
    sig[j] = s;
    if (!(s == s) || s > T(1e300)) flags[2] = 1;
  }
  block_sync<NT>();
  if (flags[2]) {  // NaN/Inf: the reference's gesvd would fail (info != 0)
    if (tid == 0) {
      *P.status = ST_NONFINITE;
      *P.rank = 0;
    }
    return;
  }
  for (int j = tid; j < K; j += NT) {  // descending order, ties by index
    const T sj = sig[j];
    int pos = 0;
    for (int i = 0; i < K; ++i) {
      const T si = sig[i];
      pos += (si > sj) || (si == sj && i < j);
    }
    perm[pos] = j;
  }
  block_sync<NT>();
  int r;
  if (rk != 0) {
    r = rk < K ? rk : K;
  } else {  // first i >= 1 with sigma_i < acc (r >= 1)
    for (int i = tid + 1; i < K; i += NT)
      if (sig[perm[i]] < acc) atomicMin(&flags[1], i);
    block_sync<NT>();
    r = flags[1];
  }
  for (int i = tid; i < K; i += NT) P.sigma[i] = sig[perm[i]];
  if (tid == 0) {
    *P.rank = r;
    *P.status = r > maxrk ? ST_R_GT_MAXRK : ST_OK;
  }
  for (int e = tid; e < r * K; e += NT) {
    const int i = e / K, row = e - i * K;
    const int j = perm[i];
    const T sj = sig[j];
    P.Zu[piv[row] + (size_t)i * K] = (sj > T(0) && j < k) ? W[row + (size_t)j * ld] / sj : T(0);
  }
}

__host__ __device__ constexpr size_t core_head_bytes(size_t elem, int kmax) {
  // sig, nrm, rdiag (elem) + perm, piv (int) per column, 256 group maxima
  // (elem + int), 16 wave partials (elem), 4 flags; 16-byte aligned
  return (((3 * elem + 2 * sizeof(int)) * (size_t)kmax + 256 * (elem + sizeof(int)) +
           16 * elem + 16 + 15) / 16) * 16;
}

// SVD of the K x K core T: QR with column pivoting of T^T (T^T P = Q R, so
// T = P R^T Q^T), truncation of R at the negligible level, then one-sided
// Jacobi on the columns of R^T (few sweeps: R^T is strongly graded).
// IN_LDS: the work matrix lives in LDS; otherwise in the global scratch P.Tj.
template <typename T, int NT, int G, int RPL, bool IN_LDS>
__global__ __launch_bounds__(NT) void k_core(const CoreProb<T>* __restrict__ probs, int kmax,
                                             int rk, int maxrk, T acc) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sig = reinterpret_cast<T*>(smem_raw);
  T* nrm = sig + kmax;
  T* rdiag = nrm + kmax;
  T* gv = rdiag + kmax;  // per-group pivot candidates (NT / G <= 256 groups)
  T* red = gv + 256;
  int* perm = reinterpret_cast<int*>(red + 16);
  int* piv = perm + kmax;
  int* gi = piv + kmax;
  int* flags = gi + 256;  // [0] rotated, [1] r, [2] nonfinite
  static_assert(NT / G <= 256 && NT / 64 <= 16, "LDS scratch sized for 256 groups, 16 waves");
  const CoreProb<T> P = probs[blockIdx.x];
  const int K = P.K;
  const int tid = threadIdx.x;
  T* A = IN_LDS ? reinterpret_cast<T*>(smem_raw + core_head_bytes(sizeof(T), kmax)) : P.Tj;
  // A = T^T; flag NaN/Inf (pivoting would silently skip such columns)
  if (tid == 0) flags[2] = 0;
  block_sync<NT>();
  bool bad = false;
  const int ld = core_ld(K, G);
  for (int e = tid; e < K * K; e += NT) {
    const int c = e / K, row = e - c * K;
    const T v = P.Tc[c + (size_t)row * K];
    bad |= !(fabs(v) <= T(1e300));
    A[row + (size_t)c * ld] = v;
  }
  if (bad) flags[2] = 1;
  block_sync<NT>();
  T thr2 = T(0);
  if (rk == 0 && acc > T(0)) {
    const T thr = T(1e-3) * acc / sqrt(T(K));
    thr2 = thr * thr;
  }
  const int k = qrcp<T, NT, G, RPL>(A, K, ld, thr2, nrm, rdiag, piv, gv, gi, red);
  block_sync<NT>();
  r_transpose<T, NT>(A, K, ld, k, rdiag);
  block_sync<NT>();
  const T tol = T(K) * eps_of<T>();
  int sweep = 0;
  if (k > 1) {
    for (; sweep < kMaxSweeps; ++sweep) {
      column_norms<T, NT, G, RPL>(A, ld, k, K, nrm);
      if (tid == 0) flags[0] = 0;
      block_sync<NT>();
      jacobi_sweep<T, NT, G, RPL>(A, ld, k, K, tol, thr2, nrm, &flags[0]);
      const int rotated = flags[0];
      block_sync<NT>();
      if (!rotated) break;
    }
  }
  if (tid == 0 && P.sweeps) *P.sweeps = sweep + 1;
  core_epilogue<T, NT>(P, A, ld, k, piv, sig, perm, flags, rk, maxrk, acc);
}

int core_lds_max_k(size_t elem) {
  const size_t lds = 160 * 1024 - 256;
  int k = 1;
  while (core_head_bytes(elem, k + 1) + elem * (size_t)(k + 1) * core_ld(k + 1, kCoreLdsG) <= lds) ++k;
  return k;
}

// variant -> (threads, group, rows per lane, storage)
#define HCORE_CORE_VARIANTS(X)                          \
  X(CORE_K16, 64, 8, 2, true)                           \
  X(CORE_K32, 64, 4, 8, true)                           \
  X(CORE_K64, 64, 2, 32, true)                          \
  X(CORE_K128, 256, 4, 32, true)                        \
  X(CORE_KLDS, 512, kCoreLdsG, 36, true)                \
  X(CORE_K192, 768, 8, 24, false)                       \
  X(CORE_K256, 1024, 16, 16, false)                     \
  X(CORE_K512, 1024, 32, 16, false)

// ------------------------------------------------------------------------------
// register-resident core SVD (K <= 64*RPL): the pivoted QR of T^T and the
// one-sided Jacobi sweeps keep the whole matrix in VGPRs.
//
// QRCP: wave w owns the columns j = c*W + w (slot c < CPW) of T^T, lane l the
// rows l + 64 i.  Per step the owner of the pivot column builds the reflector
// and hands it to the other waves through LDS; every wave updates its own
// columns with one batched wave reduction and downdates their norms (exact
// recomputation once a norm has lost half of its digits).  Two block
// barriers per step.
//
// Jacobi: R^T (K x k) is cut into 2W blocks of B columns; wave w holds two
// blocks (a, b).  An outer step rotates every pair between its two blocks
// (B rounds of B disjoint pairs, one batched reduction each, the rotation
// parameters computed by one lane group per pair), then the blocks move one
// place along the ring top_1 .. top_{W-1}, bot_{W-1} .. bot_0 (top_0 fixed;
// the circle method), so that 2W-1 outer steps meet every pair of blocks
// once.  The pairs inside each block are rotated in the first outer step of a
// sweep.  Only the block exchange touches LDS.
// ------------------------------------------------------------------------------
__host__ __device__ constexpr int pow2_ceil(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

template <typename T>
__device__ __forceinline__ bool is_finite_val(T v) {
  return fabs(v) <= std::numeric_limits<T>::max();
}

// Sum each of d[0..NV) over the 64 lanes; every lane receives all NV totals.
// xs: the wave's own LDS exchange area (N values).
template <int N, int NV, typename T>
__device__ __forceinline__ void wave_sums(T (&d)[NV], T* xs, int lane) {
  static_assert(NV <= N, "reduction width");
  T v[N];
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] = q < NV ? d[q] : T(0);
  reduce_scatter<N>(v, lane);
  if ((lane & ((64 / N) - 1)) == 0) xs[reduce_scatter_index<N>(lane)] = v[0];
  wave_sync();
#pragma unroll
  for (int q = 0; q < NV; ++q) d[q] = xs[q];
  wave_sync();
}

// Hestenes rotation angle for squared norms app, aqq and inner product apq:
// the smaller of the two angles (|t| <= 1).
template <typename T>
__device__ __forceinline__ bool jacobi_angle(T app, T aqq, T apq, T& cs, T& sn, T& t) {
  const T zeta = (aqq - app) / (T(2) * apq);
  if (fabs(zeta) > T(1e30)) t = T(0.5) / zeta;
  else t = (zeta >= T(0) ? T(1) : T(-1)) / (fabs(zeta) + sqrt(fma(zeta, zeta, T(1))));
  cs = T(1) / sqrt(fma(t, t, T(1)));
  sn = cs * t;
  return true;
}

// Rotation that orthogonalises two columns with squared norms app, aqq and
// inner product apq (Hestenes): returns false when the pair is left alone
// (converged, or a column below the truncation floor thr2).
template <typename T>
__device__ __forceinline__ bool jacobi_params(T app, T aqq, T apq, T tol, T thr2, T& cs, T& sn,
                                              T& t) {
  cs = T(1);
  sn = T(0);
  t = T(0);
  if (!(fmin(app, aqq) > thr2 && fabs(apq) > tol * sqrt(app * aqq))) return false;
  return jacobi_angle(app, aqq, apq, cs, sn, t);
}

// Same, also folding the pair's relative off-diagonal |apq| / sqrt(app aqq)
// (pairs above the floor) into the running maximum rho.
template <typename T>
__device__ __forceinline__ bool jacobi_params(T app, T aqq, T apq, T tol, T thr2, T& cs, T& sn,
                                              T& t, float& rho) {
  cs = T(1);
  sn = T(0);
  t = T(0);
  if (!(fmin(app, aqq) > thr2)) return false;
  const T nn = sqrt(app * aqq);
  if (nn > T(0)) rho = fmaxf(rho, (float)fabs(apq) / (float)nn);
  if (!(fabs(apq) > tol * nn)) return false;
  return jacobi_angle(app, aqq, apq, cs, sn, t);
}

// Round-robin (circle method) pairing of n players, round t, pair q < n/2.
__host__ __device__ constexpr int rr_u(int n, int t, int q) {
  return q == 0 ? 0 : ((q - 1 + t) % (n - 1)) + 1;
}
__host__ __device__ constexpr int rr_v(int n, int t, int q) { return ((n - 2 - q + t) % (n - 1)) + 1; }

// Pair q of round t over the wave's 2B columns (0..B-1: block a, B..2B-1:
// block b).  CROSS: (a_q, b_{(q+t) mod B}); otherwise the round-robin inside
// each block (pairs q < B/2 in a, the others in b).
template <bool CROSS, int B>
__device__ __forceinline__ void jacobi_pair(int t, int q, int& u, int& v) {
  if (CROSS) {
    u = q;
    v = B + (q + t) % B;
  } else {
    constexpr int H = B / 2;
    const int base = q < H ? 0 : B, qq = q < H ? q : q - H;
    u = base + rr_u(B, t, qq);
    v = base + rr_v(B, t, qq);
  }
}

template <typename T, int B, int RPL>
__device__ __forceinline__ T& jreg(T (&a)[B][RPL], T (&b)[B][RPL], int idx, int i) {
  return idx < B ? a[idx][i] : b[idx - B][i];
}

// One round of B disjoint pairs.  nrm: the wave's 2B squared column norms
// (LDS), rot: 2B values of LDS scratch.  The lanes that hold a pair's reduced
// inner product compute its rotation and keep the largest relative
// off-diagonal |a_pq| / sqrt(a_pp a_qq) they have seen in rho.
template <bool CROSS, typename T, int B, int RPL>
__device__ __forceinline__ void jacobi_round(T (&a)[B][RPL], T (&b)[B][RPL], int t, T* nrm, T* rot,
                                             int lane, T tol, T thr2, int* rotflag, float& rho) {
  constexpr int NB = pow2_ceil(B);
  T d[NB];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    d[q] = T(0);
    if (q < B) {
      int u, v;
      jacobi_pair<CROSS, B>(t, q, u, v);
#pragma unroll
      for (int i = 0; i < RPL; ++i) d[q] = fma(jreg(a, b, u, i), jreg(a, b, v, i), d[q]);
    }
  }
  reduce_scatter<NB>(d, lane);
  const int idx = reduce_scatter_index<NB>(lane);
  if ((lane & ((64 / NB) - 1)) == 0 && idx < B) {
    int u, v;
    jacobi_pair<CROSS, B>(t, idx, u, v);
    const T app = nrm[u], aqq = nrm[v], apq = d[0];
    T cs, sn, tt;
    if (jacobi_params(app, aqq, apq, tol, thr2, cs, sn, tt, rho)) {
      nrm[u] = fma(-tt, apq, app);
      nrm[v] = fma(tt, apq, aqq);
      *rotflag = 1;
    }
    rot[2 * idx] = cs;
    rot[2 * idx + 1] = sn;
  }
  wave_sync();
#pragma unroll
  for (int q = 0; q < B; ++q) {
    const T cs = rot[2 * q], sn = rot[2 * q + 1];
    if (sn != T(0)) {
      int u, v;
      jacobi_pair<CROSS, B>(t, q, u, v);
#pragma unroll
      for (int i = 0; i < RPL; ++i) {
        const T xu = jreg(a, b, u, i), xv = jreg(a, b, v, i);
        jreg(a, b, u, i) = cs * xu - sn * xv;
        jreg(a, b, v, i) = sn * xu + cs * xv;
      }
    }
  }
  wave_sync();
}

// Move the blocks one place along the ring (top_0 stays): through LDS, one
// 64-row slice at a time.  Squared norms travel with their blocks.
template <typename T, int W, int B, int RPL>
__device__ __forceinline__ void block_exchange(T (&a)[B][RPL], T (&b)[B][RPL], T* big, T* nst,
                                               T* nrm, int w, int lane) {
  const int posA = w, posB = 2 * W - 1 - w;
  const int dA = w == 0 ? 0 : posA + 1;
  const int dB = posB == 2 * W - 1 ? 1 : posB + 1;
  if (lane < B) {
    nst[dA * B + lane] = nrm[lane];
    nst[dB * B + lane] = nrm[B + lane];
  }
#pragma unroll
  for (int i = 0; i < RPL; ++i) {
#pragma unroll
    for (int q = 0; q < B; ++q) {
      big[(dA * B + q) * 64 + lane] = a[q][i];
      big[(dB * B + q) * 64 + lane] = b[q][i];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < B; ++q) {
      a[q][i] = big[(posA * B + q) * 64 + lane];
      b[q][i] = big[(posB * B + q) * 64 + lane];
    }
    if (i == 0 && lane < B) {
      nrm[lane] = nst[posA * B + lane];
      nrm[B + lane] = nst[posB * B + lane];
    }
    __syncthreads();
  }
}

// Pivot candidate of one wave: the largest remaining squared norm among its
// unprocessed slots (slot c's norm in lane c), ties to the smallest column.
template <typename T, int W, int CPW>
__device__ __forceinline__ void qrcp_candidate(T nl, unsigned long long done_mask, int w, int lane,
                                               T* cand_v, int* cand_i) {
  T bv = lane < CPW && !((done_mask >> lane) & 1) ? nl : T(-1);
  int bi = bv >= T(0) ? lane * W + w : -1;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const T ov = __shfl_xor(bv, off, 64);
    const int oi = __shfl_xor(bi, off, 64);
    if (oi >= 0 && (ov > bv || (ov == bv && (bi < 0 || oi < bi)))) {
      bv = ov;
      bi = oi;
    }
  }
  if (lane == 0) {
    cand_v[w] = bv;
    cand_i[w] = bi;
  }
}

template <typename T, int W, int RPL, int CPW, int B>
struct CoreReg {
  static constexpr int NT = 64 * W;
  static constexpr int KM = 64 * RPL;
  static constexpr int NQ = pow2_ceil(CPW);
  static constexpr int N2 = pow2_ceil(2 * B);
  static constexpr int NBLK = 2 * W;
  static constexpr int XS = NQ > N2 ? NQ : N2;
  static constexpr int BIG = NBLK * B * 64;
  static_assert(CPW * W >= KM && NBLK * B >= KM && B % 2 == 0 && N2 <= 64 && NQ <= 64,
                "register core shape");
  static constexpr size_t n_elem = (size_t)BIG + 2 * KM + W * XS + 2 * W * B + NBLK * B +
                                   2 * W * B + 2 * W + 2;
  static constexpr size_t n_int = 2 * W + 4 * KM + 8;
  static constexpr size_t smem = n_elem * sizeof(T) + n_int * sizeof(int);
};

template <typename T, int W, int RPL, int CPW, int B>
__global__ __launch_bounds__(64 * W) void k_core_reg(const CoreProb<T>* __restrict__ probs, int rk,
                                                      int maxrk, T acc) {
  using C = CoreReg<T, W, RPL, CPW, B>;
  constexpr int NT = C::NT, KM = C::KM;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* big = reinterpret_cast<T*>(smem_raw);  // transposition / block exchange
  T* vbuf = big + C::BIG;                   // KM: reflector
  T* sig = vbuf + KM;                       // KM
  T* xs_all = sig + KM;                     // W x XS
  T* nrm_all = xs_all + W * C::XS;          // W x 2B block norms
  T* nst = nrm_all + W * 2 * B;             // NBLK x B norms in transit
  T* rot_all = nst + C::NBLK * B;           // W x 2B rotations
  T* cand_v = rot_all + W * 2 * B;          // 2 x W pivot candidates
  T* misc = cand_v + 2 * W;                 // [0] tau
  int* cand_i = reinterpret_cast<int*>(misc + 2);  // 2 x W
  int* piv = cand_i + 2 * W;                       // KM: column at each position
  int* posv = piv + KM;                            // KM: position of each column
  int* perm = posv + KM;                           // KM: sorted order
  int* rnk = perm + KM;                            // KM: sorted position of each column
  int* flags = rnk + KM;  // [0] r, [1] nonfinite, [2..4] rotated, [5..7] max rho (sweep mod 3)

  const CoreProb<T> P = probs[blockIdx.x];
  const int K = P.K;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  T* xs = xs_all + w * C::XS;
  if (tid == 0) {
    flags[1] = 0;
    flags[2] = 0;
    flags[5] = 0;
    if (P.clk) P.clk[0] = (long long)wall_clock64();
  }
  for (int j = tid; j < KM; j += NT) posv[j] = -1;
  __syncthreads();

  // ---- T^T into registers, exact column norms ----
  T x[CPW][RPL];
  bool bad = false;
#pragma unroll
  for (int c = 0; c < CPW; ++c) {
    const int j = c * W + w;
#pragma unroll
    for (int i = 0; i < RPL; ++i) {
      const int row = lane + 64 * i;
      T v = T(0);
      if (j < K && row < K) v = P.Tc[j + (size_t)row * K];
      bad |= !is_finite_val(v);
      x[c][i] = v;
    }
  }
  if (bad) flags[1] = 1;
  T thr2 = T(0);
  if (rk == 0 && acc > T(0)) {
    const T thr = T(1e-3) * acc / sqrt(T(K));
    thr2 = thr * thr;
  }
  const T tol3z = sqrt(eps_of<T>() * T(2));
  // Remaining squared norm of slot c lives in lane c (nl; n0: at its last exact
  // computation); done_mask: processed (or absent) slots, uniform in the wave.
  unsigned long long done_mask = 0;
#pragma unroll
  for (int c = 0; c < CPW; ++c)
    if (c * W + w >= K) done_mask |= 1ull << c;
  T nl, n0;
  {
    T d[CPW];
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
      d[c] = T(0);
#pragma unroll
      for (int i = 0; i < RPL; ++i) d[c] = fma(x[c][i], x[c][i], d[c]);
    }
    wave_sums<C::NQ>(d, xs, lane);
    nl = lane < CPW ? xs[lane] : T(0);
    n0 = nl;
    wave_sync();
    qrcp_candidate<T, W, CPW>(nl, done_mask, w, lane, cand_v, cand_i);
  }
  __syncthreads();

  // ---- QR with column pivoting of T^T ----
  int k = K;
  for (int s = 0; s < K; ++s) {
    const int par = s & 1;
    T best = T(-1);
    int p = -1;
    for (int w2 = 0; w2 < W; ++w2) {
      const T v = cand_v[par * W + w2];
      const int ii = cand_i[par * W + w2];
      if (ii >= 0 && (v > best || (v == best && ii < p))) { best = v; p = ii; }
    }
    if (p < 0 || !(best > thr2)) { k = s; break; }
    const int ls = s & 63, is = s >> 6;
    if (w == p % W) {
      const int cp = p / W;
      T col[RPL];
#pragma unroll
      for (int i = 0; i < RPL; ++i) col[i] = T(0);
#pragma unroll
      for (int c = 0; c < CPW; ++c)
        if (c == cp) {
#pragma unroll
          for (int i = 0; i < RPL; ++i) col[i] = x[c][i];
        }
      T part = T(0), sel = T(0);
#pragma unroll
      for (int i = 0; i < RPL; ++i) {
        const int row = lane + 64 * i;
        if (row > s && row < K) part = fma(col[i], col[i], part);
        if (i == is) sel = col[i];
      }
      const T xn2 = wave_sum(part);
      const T alpha = __shfl(sel, ls, 64);
      T beta = alpha, tau = T(0), scal = T(0);
      if (xn2 > T(0)) {
        const T nrm = sqrt(fma(alpha, alpha, xn2));
        beta = alpha >= T(0) ? -nrm : nrm;
        tau = (beta - alpha) / beta;
        scal = T(1) / (alpha - beta);
      }
#pragma unroll
      for (int i = 0; i < RPL; ++i) {
        const int row = lane + 64 * i;
        vbuf[row] = row > s && row < K ? col[i] * scal : (row == s ? T(1) : T(0));
      }
      if (lane == 0) {
        misc[0] = tau;
        piv[s] = p;
      }
#pragma unroll
      for (int c = 0; c < CPW; ++c)
        if (c == cp) {
#pragma unroll
          for (int i = 0; i < RPL; ++i) {
            const int row = lane + 64 * i;
            if (row == s) x[c][i] = beta;
            else if (row > s) x[c][i] = T(0);
          }
        }
      done_mask |= 1ull << cp;
    }
    __syncthreads();
    const T tau = misc[0];
    T v[RPL];
#pragma unroll
    for (int i = 0; i < RPL; ++i) v[i] = vbuf[lane + 64 * i];
    T d[CPW];
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
      d[c] = T(0);
      if (!((done_mask >> c) & 1)) {
#pragma unroll
        for (int i = 0; i < RPL; ++i) d[c] = fma(v[i], x[c][i], d[c]);
      }
    }
    wave_sums<C::NQ>(d, xs, lane);
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
      if (!((done_mask >> c) & 1)) {
        const T wc = tau * d[c];
        T sel = T(0);
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
          x[c][i] = fma(-v[i], wc, x[c][i]);
          if (i == is) sel = x[c][i];
        }
        if (lane == ls) xs[c] = sel;  // row s of the slot: R(s, j)
      }
    }
    wave_sync();
    const bool live = lane < CPW && !((done_mask >> lane) & 1);
    if (live) {
      const T rs = xs[lane];
      nl = fma(-rs, rs, nl);
    }
    wave_sync();
    if (__any(live && n0 > thr2 && nl <= tol3z * n0)) {  // norms lost half their digits
#pragma unroll
      for (int c = 0; c < CPW; ++c) {
        d[c] = T(0);
        if (!((done_mask >> c) & 1)) {
#pragma unroll
          for (int i = 0; i < RPL; ++i) {
            const int row = lane + 64 * i;
            if (row > s && row < K) d[c] = fma(x[c][i], x[c][i], d[c]);
          }
        }
      }
      wave_sums<C::NQ>(d, xs, lane);
      if (live) {
        nl = xs[lane];
        n0 = nl;
      }
      wave_sync();
    }
    qrcp_candidate<T, W, CPW>(nl, done_mask, w, lane, cand_v + (par ^ 1) * W, cand_i + (par ^ 1) * W);
    __syncthreads();
  }

  // positions: the k pivots, then the remaining columns in index order
  if (tid == 0) {
    if (P.clk) {
      P.clk[1] = (long long)wall_clock64();
      P.clk[4] = k;
    }
    for (int s = 0; s < k; ++s) posv[piv[s]] = s;
    int cnt = k;
    for (int j = 0; j < K; ++j)
      if (posv[j] < 0) {
        piv[cnt] = j;
        posv[j] = cnt++;
      }
  }
  __syncthreads();

  // ---- R^T (K x k) into the Jacobi blocks, through LDS in chunks ----
  T a[B][RPL], b[B][RPL];
#pragma unroll
  for (int q = 0; q < B; ++q)
#pragma unroll
    for (int i = 0; i < RPL; ++i) {
      a[q][i] = T(0);
      b[q][i] = T(0);
    }
  const int cA = 2 * w * B, cB = (2 * w + 1) * B;  // first column of the home blocks
  {
    // R^T(jp, c) = R(c, jp) for c < k, through the update's global scratch
    // (ld K): written from the QR layout, read back in the block layout, so
    // the two register sets are never live together
    T* Rt = P.Tj;
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
      const int j = c * W + w;
      if (j < K) {
        const int jp = posv[j];
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
          const int row = lane + 64 * i;
          if (row < k) Rt[jp + (size_t)row * K] = x[c][i];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < B; ++q) {
#pragma unroll
      for (int i = 0; i < RPL; ++i) {
        const int row = lane + 64 * i;
        if (row < K) {
          if (cA + q < k) a[q][i] = Rt[row + (size_t)(cA + q) * K];
          if (cB + q < k) b[q][i] = Rt[row + (size_t)(cB + q) * K];
        }
      }
    }
  }

  // ---- block one-sided Jacobi on R^T ----
  T* nrm = nrm_all + w * 2 * B;
  T* rot = rot_all + w * 2 * B;
  {
    T d[2 * B];
#pragma unroll
    for (int q = 0; q < B; ++q) {
      d[q] = T(0);
      d[B + q] = T(0);
#pragma unroll
      for (int i = 0; i < RPL; ++i) {
        d[q] = fma(a[q][i], a[q][i], d[q]);
        d[B + q] = fma(b[q][i], b[q][i], d[B + q]);
      }
    }
    wave_sums<C::N2>(d, xs, lane);
    if (lane < 2 * B) nrm[lane] = xs[lane];
    wave_sync();
  }
  // Sweeps until no pair rotates, or until a sweep started with every
  // relative off-diagonal below sqrt(tol): one-sided Jacobi converges
  // quadratically, so that sweep's rotations leave them near tol.
  const T tol = T(K) * eps_of<T>();
  const float rho_stop = (float)sqrt(tol);
  int sweep = 0;
  if (k > 1) {
    for (; sweep < kMaxSweeps; ++sweep) {
      int* rf = &flags[2 + sweep % 3];
      unsigned* rm = reinterpret_cast<unsigned*>(&flags[5 + sweep % 3]);
      float rho = 0.f;  // this lane's largest relative off-diagonal of the sweep
      if (tid == 0) {
        flags[2 + (sweep + 1) % 3] = 0;
        flags[5 + (sweep + 1) % 3] = 0;
      }
#pragma unroll 1
      for (int o = 0; o < 2 * W - 1; ++o) {
        if (o == 0) {
#pragma unroll
          for (int t = 0; t < B - 1; ++t)
            jacobi_round<false>(a, b, t, nrm, rot, lane, tol, thr2, rf, rho);
        }
#pragma unroll
        for (int t = 0; t < B; ++t)
          jacobi_round<true>(a, b, t, nrm, rot, lane, tol, thr2, rf, rho);
        if (o == 2 * W - 2) {  // publish the sweep's maximum before the last exchange
#pragma unroll
          for (int off = 32; off >= 1; off >>= 1) rho = fmaxf(rho, __shfl_xor(rho, off, 64));
          if (lane == 0) atomicMax(rm, __float_as_uint(rho));
        }
        block_exchange<T, W, B, RPL>(a, b, big, nst, nrm, w, lane);
      }
      if (!*rf || __uint_as_float(*rm) < rho_stop) {
        ++sweep;
        break;
      }
    }
  }
  if (tid == 0) {
    if (P.sweeps) *P.sweeps = sweep;
    if (P.clk) P.clk[2] = (long long)wall_clock64();
  }

  // ---- singular values, order, truncation, left singular vectors ----
  {
    T d[2 * B];
#pragma unroll
    for (int q = 0; q < B; ++q) {
      d[q] = T(0);
      d[B + q] = T(0);
#pragma unroll
      for (int i = 0; i < RPL; ++i) {
        d[q] = fma(a[q][i], a[q][i], d[q]);
        d[B + q] = fma(b[q][i], b[q][i], d[B + q]);
      }
    }
    wave_sums<C::N2>(d, xs, lane);
    if (lane < 2 * B) {
      const int c = lane < B ? cA + lane : cB + lane - B;
      if (c < K) {
        const T s = c < k ? sqrt(xs[lane]) : T(0);
        sig[c] = s;
        if (!is_finite_val(s)) flags[1] = 1;
      }
    }
  }
  if (tid == 0) flags[0] = K;
  __syncthreads();
  if (flags[1]) {  // NaN/Inf: the reference's gesvd would fail (info != 0)
    if (tid == 0) {
      *P.status = ST_NONFINITE;
      *P.rank = 0;
    }
    return;
  }
  for (int j = tid; j < K; j += NT) {  // descending order, ties by index
    const T sj = sig[j];
    int pos = 0;
    for (int i = 0; i < K; ++i) {
      const T si = sig[i];
      pos += (si > sj) || (si == sj && i < j);
    }
    perm[pos] = j;
    rnk[j] = pos;
  }
  __syncthreads();
  int r;
  if (rk != 0) {
    r = rk < K ? rk : K;
  } else {  // first i >= 1 with sigma_i < acc (r >= 1)  (hcore_dgemm.c:444-464)
    for (int i = tid + 1; i < K; i += NT)
      if (sig[perm[i]] < acc) atomicMin(&flags[0], i);
    __syncthreads();
    r = flags[0];
  }
  for (int i = tid; i < K; i += NT) P.sigma[i] = sig[perm[i]];
  if (tid == 0) {
    *P.rank = r;
    *P.status = r > maxrk ? ST_R_GT_MAXRK : ST_OK;
  }
#pragma unroll
  for (int q = 0; q < 2 * B; ++q) {
    const int c = q < B ? cA + q : cB + q - B;
    if (c < K) {
      const int i = rnk[c];
      if (i < r) {
        const T s = sig[c];
        const T inv = (s > T(0) && c < k) ? T(1) / s : T(0);
#pragma unroll
        for (int ii = 0; ii < RPL; ++ii) {
          const int row = lane + 64 * ii;
          if (row < K) {
            const T val = q < B ? a[q][ii] : b[q - B][ii];
            P.Zu[piv[row] + (size_t)i * K] = val * inv;
          }
        }
      }
    }
  }
  if (tid == 0 && P.clk) P.clk[3] = (long long)wall_clock64();
}

// register core variants: (waves, rows per lane, QRCP columns per wave, Jacobi block width)
#define HCORE_CORE_REG_VARIANTS(X) \
  X(CORE_R64, 4, 1, 16, 8)         \
  X(CORE_R128, 8, 2, 16, 8)        \
  X(CORE_R192, 8, 3, 24, 12)

template <typename T>
static void launch_core_reg(const CoreProb<T>* probs, int n, int variant, int rk, int maxrk, T acc,
                            hipStream_t s) {
  static bool attr = false;
  if (!attr) {
#define HCORE_SET_ATTR_R(V, W, RPL, CPW, B)                                                   \
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_core_reg<T, W, RPL, CPW, B>),    \
                            hipFuncAttributeMaxDynamicSharedMemorySize,                       \
                            (int)CoreReg<T, W, RPL, CPW, B>::smem);
    HCORE_CORE_REG_VARIANTS(HCORE_SET_ATTR_R)
#undef HCORE_SET_ATTR_R
    attr = true;
  }
  if (n <= 0) return;
  switch (variant) {
#define HCORE_LAUNCH_R(V, W, RPL, CPW, B)                                                      \
  case V:                                                                                      \
    hipLaunchKernelGGL((k_core_reg<T, W, RPL, CPW, B>), dim3(n), dim3(64 * W),                \
                       (CoreReg<T, W, RPL, CPW, B>::smem), s, probs, rk, maxrk, acc);          \
    break;
    HCORE_CORE_REG_VARIANTS(HCORE_LAUNCH_R)
#undef HCORE_LAUNCH_R
    default: break;
  }
}

template <typename T>
void launch_core(const CoreProb<T>* probs, int n, int variant, int kmax, int rk, int maxrk, T acc,
                 hipStream_t s) {
  static bool attr = false;
  if (!attr) {
#define HCORE_SET_ATTR(V, NT, G, RPL, LDS)                                                   \
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_core<T, NT, G, RPL, LDS>),      \
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    HCORE_CORE_VARIANTS(HCORE_SET_ATTR)
#undef HCORE_SET_ATTR
    attr = true;
  }
  if (variant >= CORE_R64) {
    launch_core_reg<T>(probs, n, variant, rk, maxrk, acc, s);
    return;
  }
  if (n <= 0) return;
  const size_t head = core_head_bytes(sizeof(T), kmax);
  switch (variant) {
#define HCORE_LAUNCH(V, NT, G, RPL, LDS)                                                     \
  case V: {                                                                                  \
    const size_t sm = head + (LDS ? sizeof(T) * (size_t)kmax * core_ld(kmax, G) : 0);        \
    hipLaunchKernelGGL((k_core<T, NT, G, RPL, LDS>), dim3(n), dim3(NT), sm, s, probs, kmax,  \
                       rk, maxrk, acc);                                                      \
    break;                                                                                   \
  }
    HCORE_CORE_VARIANTS(HCORE_LAUNCH)
#undef HCORE_LAUNCH
    default: break;
  }
}

int core_variant_for(int K, int kl, bool reg) {
  if (K <= 16) return CORE_K16;
  if (K <= 32) return CORE_K32;
  if (reg) {
    if (K <= 64) return CORE_R64;
    if (K <= 128) return CORE_R128;
    if (K <= 192) return CORE_R192;
  }
  if (K <= 64) return CORE_K64;
  if (K <= 128) return CORE_K128;
  if (K <= kl && K <= 144) return CORE_KLDS;
  if (K <= 192) return CORE_K192;
  if (K <= 256) return CORE_K256;
  return CORE_K512;
}

// ------------------------------------------------------------------------------
// reassembly X = Q [Z; 0]: panels applied last to first, X <- (I - Y T Y^T) X
// ------------------------------------------------------------------------------
template <typename T, int NBW, int PB, int CB>
__global__ __launch_bounds__(kThreads) void k_apply(const ApplyProb<T>* __restrict__ probs, int nb) {
  constexpr int NBP = 64 * NBW;
  constexpr int NW = PB * CB;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* Yp = reinterpret_cast<T*>(smem_raw);  // PB x NBP
  T* Tt = Yp + PB * NBP;                   // PB x PB
  T* scr = Tt + PB * PB;                   // 4 x 2 x NW

  const ApplyProb<T> P = probs[blockIdx.x];
  if (*P.status != ST_OK) return;
  const int r = *P.rank;
  const int K = P.K;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  T* __restrict__ X = P.X;
  const size_t ldx = P.ldx;
  T* wscr0 = scr + wid * 2 * NW;
  T* wscr1 = wscr0 + NW;
  const T* Yl = Yp + lane;  // this lane's rows: Yl[c * NBP + 64 * i]

  #pragma unroll 1

  for (int e = tid; e < r * nb; e += kThreads) {
    const int col = e / nb, row = e - col * nb;
    X[row + col * ldx] = row < K ? P.Z[row + (size_t)col * K] : T(0);
  }
  __syncthreads();

  const int npan = (K + PB - 1) / PB;
  const int nch = (r + CB - 1) / CB;
  for (int p = npan - 1; p >= 0; --p) {
    const int j0 = p * PB;
    const int pb = min(PB, K - j0);
    #pragma unroll 1
    for (int e = tid; e < PB * NBP; e += kThreads) {
      const int c = e / NBP, row = e - c * NBP;
      T v = T(0);
      if (c < pb && row < nb) {
        if (row == j0 + c) v = T(1);
        else if (row > j0 + c) v = P.A[row + (size_t)(j0 + c) * nb];
      }
      Yp[e] = v;
    }
    if (tid < PB * PB) Tt[tid] = P.Tp[(size_t)p * PB * PB + tid];
    __syncthreads();
    for (int q = wid; q < nch; q += 4) {
      const int k0 = q * CB;
      const int cb = min(CB, r - k0);
      T Xr[NBW][CB];
#pragma unroll
      for (int i = 0; i < NBW; ++i) {
        const int row = lane + 64 * i;
#pragma unroll
        for (int kk = 0; kk < CB; ++kk)
          Xr[i][kk] = (row >= j0 && row < nb && kk < cb) ? X[row + (k0 + kk) * ldx] : T(0);
      }
      T W[NW];
#pragma unroll
      for (int c = 0; c < PB; ++c) {
        T y[NBW];
#pragma unroll
        for (int i = 0; i < NBW; ++i) y[i] = Yl[c * NBP + 64 * i];
#pragma unroll
        for (int kk = 0; kk < CB; ++kk) {
          T acc = T(0);
#pragma unroll
          for (int i = 0; i < NBW; ++i) acc = fma(y[i], Xr[i][kk], acc);
          W[c * CB + kk] = acc;
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      reduce_scatter<NW>(W, lane);
      wscr0[reduce_scatter_index<NW>(lane)] = W[0];
      wave_sync();
      if (lane < NW) {  // W2 = T W (T upper triangular)
        const int c = lane / CB, kk = lane - c * CB;
        T s = T(0);
        for (int c2 = c; c2 < PB; ++c2) s = fma(Tt[c + c2 * PB], wscr0[c2 * CB + kk], s);
        wscr1[lane] = s;
      }
      wave_sync();
#pragma unroll
      for (int c = 0; c < PB; ++c) {
        T y[NBW];
#pragma unroll
        for (int i = 0; i < NBW; ++i) y[i] = Yl[c * NBP + 64 * i];
#pragma unroll
        for (int kk = 0; kk < CB; ++kk) {
          const T w2 = wscr1[c * CB + kk];
#pragma unroll
          for (int i = 0; i < NBW; ++i) Xr[i][kk] = fma(-y[i], w2, Xr[i][kk]);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int i = 0; i < NBW; ++i) {
        const int row = lane + 64 * i;
        if (row >= j0 && row < nb) {
#pragma unroll
          for (int kk = 0; kk < CB; ++kk)
            if (kk < cb) X[row + (k0 + kk) * ldx] = Xr[i][kk];
        }
      }
      wave_sync();
    }
    __syncthreads();
  }
}

template <typename T, int NBW>
static void launch_apply_nbw(const ApplyProb<T>* probs, int n, int nb, hipStream_t s) {
  constexpr int CB = qr_cb<NBW>();
  const size_t sm = sizeof(T) * ((size_t)kPB * 64 * NBW + kPB * kPB + 4 * 2 * kPB * CB);
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_apply<T, NBW, kPB, CB>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    attr = true;
  }
  if (n <= 0) return;
  hipLaunchKernelGGL((k_apply<T, NBW, kPB, CB>), dim3(n), dim3(kThreads), sm, s, probs, nb);
}

template <typename T>
void launch_apply(const ApplyProb<T>* probs, int n, int nb, hipStream_t s) {
  if (nb <= 128) launch_apply_nbw<T, 2>(probs, n, nb, s);
  else if (nb <= 256) launch_apply_nbw<T, 4>(probs, n, nb, s);
  else if (nb <= 512) launch_apply_nbw<T, 8>(probs, n, nb, s);
  else if (nb <= 1024) launch_apply_nbw<T, 16>(probs, n, nb, s);
  else launch_apply_nbw<T, 32>(probs, n, nb, s);
}

// Set the large-LDS attributes of every kernel variant up front, so that no
// attribute call happens while a stream is being captured into a graph.
// ------------------------------------------------------------------------------
// compact-WY solve M = T Y1^T Z by blocked back substitution (K <= kWyMaxK):
// for the panels b from last to first,
//   M_b = T_bb (B_b - S(b, >b) M_{>b}),  S = Y^T Y,  B = Y1^T Z (in M on entry),
// which is M = T B for the forward compact-WY T (T^{-1} = diag(1/tau) +
// striu(S)).  A workgroup holds 64 columns of M in LDS; wave w computes rows
// 2w, 2w+1 of each panel, lane j column j.  The panel's rows of S (contiguous
// in the stored lower triangle) and its kPB x kPB factor are staged in LDS by
// coalesced loads; the next panel's are fetched into registers while the
// current one is solved.
// ------------------------------------------------------------------------------
constexpr int kWyLd = 65;  // row stride of M in LDS: column-wise fills spread over the banks
template <typename T>
__global__ __launch_bounds__(kThreads) void k_wy_solve(const WySolveProb<T>* __restrict__ probs,
                                                       int ncb) {
  static_assert(kPB == 8 && kThreads == 256 && kWyMaxK <= kThreads,
                "two rows of a panel per wavefront; one S element per thread and row");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const WySolveProb<T> P = probs[blockIdx.x / ncb];
  if (*P.status != ST_OK) return;
  const int r = *P.rank;
  const int j0 = (blockIdx.x % ncb) * 64;
  if (j0 >= r) return;
  const int K = P.K;
  const int nc = min(64, r - j0);
  T* Mb = reinterpret_cast<T*>(smem_raw);  // K x 64, row c at Mb[c * kWyLd]
  T* RH = Mb + (size_t)K * kWyLd;           // kPB x 64 right-hand sides
  T* Sb = RH + kPB * 64;                    // kPB x kWyMaxK: S(hi + i, base + a) at Sb[a * kWyMaxK + i]
  T* Tl = Sb + kPB * kWyMaxK;               // kPB x kPB panel factor
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  // columns j0 .. j0 + 63 of M, read along its columns (consecutive rows per thread)
  for (int e = tid; e < K * 64; e += kThreads) {
    const int c = e % K, j = e / K;
    Mb[c * kWyLd + j] = j < nc ? P.M[c + (size_t)(j0 + j) * K] : T(0);
  }
  const int nblk = (K + kPB - 1) / kPB;
  T pre[kPB], pret = T(0);
  auto fetch = [&](int b) {
    const int base = b * kPB, hi = min(K, base + kPB), len = K - hi;
#pragma unroll
    for (int a = 0; a < kPB; ++a)
      pre[a] = (tid < len && base + a < hi) ? P.S[(size_t)(base + a) * K + hi + tid] : T(0);
    if (tid < kPB * kPB) pret = P.Tp[(size_t)b * kPB * kPB + tid];
  };
  auto stash = [&]() {
#pragma unroll
    for (int a = 0; a < kPB; ++a) Sb[a * kWyMaxK + tid] = pre[a];
    if (tid < kPB * kPB) Tl[tid] = pret;
  };
  fetch(nblk - 1);
  stash();
  __syncthreads();
  for (int b = nblk - 1; b >= 0; --b) {
    const int base = b * kPB, hi = min(K, base + kPB), len = K - hi;
    if (b > 0) fetch(b - 1);  // in flight while this panel is solved
    {
      const int a0 = 2 * w, a1 = 2 * w + 1;
      const T* S0 = Sb + a0 * kWyMaxK;
      const T* S1 = Sb + a1 * kWyMaxK;
      T p0 = T(0), p1 = T(0), q0 = T(0), q1 = T(0);
      int i = 0;
      for (; i + 1 < len; i += 2) {
        const T m0 = Mb[(hi + i) * kWyLd + lane], m1 = Mb[(hi + i + 1) * kWyLd + lane];
        p0 = fma(S0[i], m0, p0);
        q0 = fma(S1[i], m0, q0);
        p1 = fma(S0[i + 1], m1, p1);
        q1 = fma(S1[i + 1], m1, q1);
      }
      if (i < len) {
        const T m0 = Mb[(hi + i) * kWyLd + lane];
        p0 = fma(S0[i], m0, p0);
        q0 = fma(S1[i], m0, q0);
      }
      if (base + a0 < hi) RH[a0 * 64 + lane] = Mb[(base + a0) * kWyLd + lane] - (p0 + p1);
      if (base + a1 < hi) RH[a1 * 64 + lane] = Mb[(base + a1) * kWyLd + lane] - (q0 + q1);
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int a = 2 * w + e, c = base + a;
      if (c < hi) {
        T v = T(0);
        for (int cc = a; cc < hi - base; ++cc) v = fma(Tl[a + cc * kPB], RH[cc * 64 + lane], v);
        Mb[c * kWyLd + lane] = v;
      }
    }
    __syncthreads();
    if (b > 0) {
      stash();
      __syncthreads();
    }
  }
  for (int e = tid; e < K * nc; e += kThreads) {
    const int c = e % K, j = e / K;
    P.M[c + (size_t)(j0 + j) * K] = Mb[c * kWyLd + j];
  }
}

template <typename T>
static size_t wy_smem(int kmax) {
  return sizeof(T) * ((size_t)kmax * kWyLd + kPB * 64 + kPB * kWyMaxK + kPB * kPB);
}

template <typename T>
void launch_wy_solve(const WySolveProb<T>* probs, int n, int kmax, int rmax, hipStream_t s) {
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_wy_solve<T>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)wy_smem<T>(kWyMaxK));
    attr = true;
  }
  if (n <= 0 || rmax <= 0) return;
  const int ncb = (rmax + 63) / 64;
  hipLaunchKernelGGL(k_wy_solve<T>, dim3(n * ncb), dim3(kThreads), wy_smem<T>(kmax), s, probs, ncb);
}

// ------------------------------------------------------------------------------
// element-wise precision conversion, one workgroup per (problem, 4096 elements)
// ------------------------------------------------------------------------------
template <typename S, typename D>
__global__ __launch_bounds__(kThreads) void k_convert(const ConvProb* __restrict__ probs, int nchunk) {
  const ConvProb P = probs[blockIdx.x / nchunk];
  const int e0 = (blockIdx.x % nchunk) * 4096;
  const S* src = reinterpret_cast<const S*>(P.src);
  D* dst = reinterpret_cast<D*>(P.dst);
  const int e1 = min(P.n, e0 + 4096);
  for (int e = e0 + threadIdx.x; e < e1; e += kThreads) dst[e] = (D)src[e];
}

void launch_convert(const ConvProb* probs, int n, int nmax, bool to_double, hipStream_t s) {
  if (n <= 0 || nmax <= 0) return;
  const int nchunk = (nmax + 4095) / 4096;
  if (to_double)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(n * nchunk), dim3(kThreads), 0, s, probs, nchunk);
  else
    hipLaunchKernelGGL((k_convert<double, float>), dim3(n * nchunk), dim3(kThreads), 0, s, probs, nchunk);
}

template <typename T>
void prepare_kernels() {
  launch_wy_solve<T>(nullptr, 0, 1, 1, nullptr);
  launch_geqrf<T>(nullptr, 0, 128, nullptr);
  launch_geqrf<T>(nullptr, 0, 256, nullptr);
  launch_geqrf<T>(nullptr, 0, 512, nullptr);
  launch_geqrf<T>(nullptr, 0, 1024, nullptr);
  launch_geqrf<T>(nullptr, 0, 2048, nullptr);
  launch_apply<T>(nullptr, 0, 128, nullptr);
  launch_apply<T>(nullptr, 0, 256, nullptr);
  launch_apply<T>(nullptr, 0, 512, nullptr);
  launch_apply<T>(nullptr, 0, 1024, nullptr);
  launch_apply<T>(nullptr, 0, 2048, nullptr);
  launch_core<T>(nullptr, 0, CORE_K16, 1, 0, 1, T(0), nullptr);
  launch_core_reg<T>(nullptr, 0, CORE_R64, 0, 1, T(0), nullptr);
}

// ---- explicit instantiations ---------------------------------------------------
#define HCORE_INSTANTIATE(T)                                                                  \
  template void launch_gemm<T>(const GemmProb<T>*, const GemmTask*, int, int, hipStream_t,   \
                                bool);                                                        \
  template void launch_gemm_tall<T>(const GemmProb<T>*, const GemmTask*, int, hipStream_t);  \
  template void launch_stack<T>(const StackProb<T>*, int, int, int, T, T, hipStream_t);       \
  template void launch_geqrf<T>(const QrProb<T>*, int, int, hipStream_t);                     \
  template void launch_qr_reg<T>(const QrProb<T>*, int, int, int, hipStream_t);               \
  template void launch_qr_fused<T>(const QrProb<T>*, int, int, hipStream_t);                  \
  template void launch_core<T>(const CoreProb<T>*, int, int, int, int, int, T, hipStream_t);  \
  template void launch_apply<T>(const ApplyProb<T>*, int, int, hipStream_t);                 \
  template void launch_wy_solve<T>(const WySolveProb<T>*, int, int, int, hipStream_t);        \
  template void prepare_kernels<T>();
HCORE_INSTANTIATE(double)
HCORE_INSTANTIATE(float)

}  // namespace hcore
