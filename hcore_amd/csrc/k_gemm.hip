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
// batched GEMM on the matrix cores.  Every product of the path (the k = nb
// reductions F = AV^T BV and Y^T Y, the small cores T = R_A R_B^T and T^T Zu,
// the tall products G = BU F^T and X = [Z; 0] - Y M, and dsyrk's W, X and CD)
// runs through v_mfma_f64_16x16x4_f64 (FP32 plans: v_mfma_f32_16x16x4_f32;
// the products of split-precision plans: BF16 digit slices, k_gemm_oz below):
// one wavefront owns a 32 x 32 (or 16-row-block x 32) piece of the output as
// 16 x 16 accumulator blocks; operands are staged through LDS (2-D tiles) or
// read straight from HBM with coalesced 16-row segments (tall products).
// ------------------------------------------------------------------------------
constexpr int GK = 32;  // k chunk staged in LDS

// One 16 x 16 x 4 matrix-core step, D = A B + C, per precision.  Operand
// layout (lane l): A(m = l % 16, k = l / 16), B(k = l / 16, n = l % 16).
// Element v of lane l's accumulator holds D(m, n = l % 16) with
//   FP32 (v_mfma_f32_16x16x4_f32): m = 4 (l / 16) + v,
//   FP64 (v_mfma_f64_16x16x4_f64): m = l / 16 + 4 v
// (the double-precision result rows are interleaved by four).  The kernels
// below feed op(B) as A and op(A) as B, so that lane l holds
// C(i = l % 16, j = m(l, v)): 16 consecutive rows of 4 columns, a coalesced
// column-major store.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  using V = double __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V mac(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int m_of(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct Mfma<float> {
  using V = float __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V mac(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int m_of(int lane, int v) { return 4 * (lane >> 4) + v; }
};

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

// The k range [lo, hi) of the output tile (TILE rows from row0, TILE columns
// from col0) outside of which op(A)(i, k) or op(B)(k, j) vanish by the
// triangular flags (upper: stored (p, q) zero for p > q; unit lower: zero for
// p < q), lo rounded down to a multiple of the chunk GKX: T = R_A R_B^T needs
// k >= max(i, j), the compact-WY products Y1^T Z k >= i.
template <int TILE, int GKX>
__device__ __forceinline__ void gemm_k_range(int flags, int row0, int col0, int k, int& lo,
                                             int& hi) {
  const bool ta = (flags & GF_A_TRANS) != 0, tb = (flags & GF_B_TRANS) != 0;
  lo = 0;
  hi = k;
  if (flags & GF_A_UPPER) {
    if (ta) hi = min(hi, row0 + TILE); else lo = max(lo, row0);
  }
  if (flags & GF_A_UNIT_LOWER) {
    if (ta) lo = max(lo, row0); else hi = min(hi, row0 + TILE);
  }
  if (flags & GF_B_UPPER) {
    if (tb) lo = max(lo, col0); else hi = min(hi, col0 + TILE);
  }
  if (flags & GF_B_UNIT_LOWER) {
    if (tb) hi = min(hi, col0 + TILE); else lo = max(lo, col0);
  }
  lo = (lo / GKX) * GKX;
  if (hi < lo) hi = lo;
}

// Registers -> LDS: As[kk][i] = op(A)(i, k0 + kk), Bs[kk][j] = op(B)(k0 + kk, j).
template <typename T, int TILE, int NT, int LD, int GKX, int LDT>
__device__ __forceinline__ void gemm_stage(T (*As)[LDT], T (*Bs)[LDT], const T (&ra)[LD],
                                           const T (&rb)[LD], bool ta, bool tb, int t) {
#pragma unroll
  for (int q = 0; q < LD; ++q) {
    const int idx = t + NT * q;
    if (!ta) As[idx / TILE][idx % TILE] = ra[q];
    else     As[idx % GKX][idx / GKX] = ra[q];
    if (tb) Bs[idx / TILE][idx % TILE] = rb[q];
    else    Bs[idx % GKX][idx / GKX] = rb[q];
  }
}

// C(i, j) = alpha * acc + beta * C (or beta * C0 rows < m0, 0 below), with the
// strict upper triangle zeroed under GF_LOWER_ZERO_UPPER: one value.
template <typename T>
__device__ __forceinline__ void gemm_store1(const GemmProb<T>& P, int i, int j, int n, T acc) {
  if (i >= P.m || j >= n) return;
  if (((P.flags & GF_LOWER_ONLY) && i < j) || ((P.flags & GF_UPPER_ONLY) && i > j)) return;
  T* c = P.C + i + (size_t)j * P.ldc;
  if ((P.flags & GF_LOWER_ZERO_UPPER) && i < j) {
    *c = T(0);
    return;
  }
  T v = P.alpha * acc;
  if (P.flags & GF_C0) {
    if (i < P.m0) v = fma(P.beta, P.C0[i + (size_t)j * P.ldc0], v);
  } else if (!(P.flags & GF_BETA_ZERO)) {
    v = fma(P.beta, *c, v);
  }
  *c = v;
}

// The same for the NV values (i[q], j[q], acc[q]) of one lane.  Every
// beta-term load is issued before the first store: as far as the compiler
// knows a store through C may alias any later load, so a load after a store
// would wait out one HBM round trip per value.
template <typename T, int NV>
__device__ __forceinline__ void gemm_store_n(const GemmProb<T>& P, const int (&i)[NV],
                                             const int (&j)[NV], int n, const T (&acc)[NV]) {
  const bool c0 = (P.flags & GF_C0) != 0, rd = !c0 && !(P.flags & GF_BETA_ZERO);
  if (!c0 && !rd) {  // uniform: nothing to read, each value stored as computed
#pragma unroll
    for (int q = 0; q < NV; ++q) gemm_store1(P, i[q], j[q], n, acc[q]);
    return;
  }
  const bool lo = (P.flags & GF_LOWER_ONLY) != 0, up = (P.flags & GF_UPPER_ONLY) != 0;
  const bool lzu = (P.flags & GF_LOWER_ZERO_UPPER) != 0;
  bool wr[NV], nd[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    wr[q] = i[q] < P.m && j[q] < n && !((lo && i[q] < j[q]) || (up && i[q] > j[q]));
    nd[q] = wr[q] && !(lzu && i[q] < j[q]) && (c0 ? i[q] < P.m0 : rd);
  }
  // the loads are unconditional (a value not needed reads the source's first
  // element), so they all issue back to back
  const T* src = c0 ? P.C0 : P.C;
  const size_t ld = c0 ? (size_t)P.ldc0 : (size_t)P.ldc;
  T b[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) b[q] = src[nd[q] ? i[q] + (size_t)j[q] * ld : 0];
  // every value before the first store: a use of a loaded value after a store
  // would wait for the store as well (the empty asm statements keep the
  // compiler from sinking the arithmetic into the conditional stores)
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const T v = P.alpha * acc[q];
    b[q] = (lzu && i[q] < j[q]) ? T(0) : (nd[q] ? fma(P.beta, b[q], v) : v);
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) asm volatile("" : "+v"(b[q]));
#pragma unroll
  for (int q = 0; q < NV; ++q)
    if (wr[q]) P.C[i[q] + (size_t)j[q] * P.ldc] = b[q];
}

// The 2 x 2 accumulator blocks of one wavefront's 32 x 32 piece at (r0, c0).
template <typename T>
__device__ __forceinline__ void gemm_store_wave(const GemmProb<T>& P, const typename Mfma<T>::V (&acc)[2][2],
                                                int r0, int c0, int n, int lane) {
  int ii[16], jj[16];
  T vv[16];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int q = 8 * rb + 4 * cb + v;
        ii[q] = r0 + 16 * rb + (lane & 15);
        jj[q] = c0 + 16 * cb + Mfma<T>::m_of(lane, v);
        vv[q] = acc[rb][cb][v];
      }
  gemm_store_n<T, 16>(P, ii, jj, n, vv);
}

// GKX/4 matrix-core steps over the staged chunk for the wavefront's 32 x 32
// piece (rows r0w.., columns c0w.. of the tile).
template <typename T, int GKX, int LDT>
__device__ __forceinline__ void gemm_mfma_chunk(const T (*As)[LDT], const T (*Bs)[LDT],
                                                typename Mfma<T>::V (&acc)[2][2], int r0w, int c0w,
                                                int lane) {
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int kk = 0; kk < GKX; kk += 4) {
    const T a0 = As[kk + lk][r0w + li], a1 = As[kk + lk][r0w + 16 + li];
    const T b0 = Bs[kk + lk][c0w + li], b1 = Bs[kk + lk][c0w + 16 + li];
    acc[0][0] = Mfma<T>::mac(b0, a0, acc[0][0]);
    acc[0][1] = Mfma<T>::mac(b1, a0, acc[0][1]);
    acc[1][0] = Mfma<T>::mac(b0, a1, acc[1][0]);
    acc[1][1] = Mfma<T>::mac(b1, a1, acc[1][1]);
  }
}

// Batched GEMM on TILE x TILE output tiles (64: four wavefronts of 32 x 32;
// 32: one wavefront), k staged through LDS in chunks (GK, or GK/2 for the
// one-wavefront tiles) with the next chunk's global loads in flight during the
// current chunk's matrix-core steps.
// Three wavefronts per SIMD (the launch bound caps the registers at 168; left
// alone the compiler takes about 200 and two fit).
template <typename T, int TILE>
__global__ __launch_bounds__((TILE / 4) * (TILE / 4), 3) void k_gemm(
    const GemmProb<T>* __restrict__ probs, const GemmTask* __restrict__ tasks) {
  using V = typename Mfma<T>::V;
  constexpr int NT = (TILE / 4) * (TILE / 4);  // 256 or 64 threads
  constexpr int GKT = TILE == 64 ? GK : GK / 2;  // k chunk: 8 elements per thread per operand
  constexpr int LD = TILE * GKT / NT;            // elements of each operand per thread per chunk
  // +1 column of padding: the transposed fills write consecutive k at one i
  __shared__ T As[GKT][TILE + 1];
  __shared__ T Bs[GKT][TILE + 1];
  const GemmTask tk = tasks[blockIdx.x];
  const GemmProb<T> P = probs[tk.prob];
  if (P.skip && *P.skip) return;
  int n = P.n;
  if (P.n_dev) n = min(n, *P.n_dev);
  const int row0 = tk.ti * TILE, col0 = tk.tj * TILE;
  if (row0 >= P.m || col0 >= n) return;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (((P.flags & GF_LOWER_ONLY) && row0 + TILE - 1 < col0) ||
      ((P.flags & GF_UPPER_ONLY) && col0 + TILE - 1 < row0))
    return;  // the tile lies in the triangle that is not written
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
  const int r0w = TILE == 64 ? 32 * (w & 1) : 0, c0w = TILE == 64 ? 32 * (w >> 1) : 0;
  V acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = V{T(0), T(0), T(0), T(0)};

  int klo, khi;
  gemm_k_range<TILE, GKT>(P.flags, row0, col0, P.k, klo, khi);
  T ra[LD], rb[LD];
  gemm_fetch<T, TILE, NT, LD, GKT>(P, n, row0, col0, klo, ta, tb, ua, ub, ra, rb, t);
  for (int k0 = klo; k0 < khi; k0 += GKT) {
    gemm_stage<T, TILE, NT, LD, GKT, TILE + 1>(As, Bs, ra, rb, ta, tb, t);
    __syncthreads();
    if (k0 + GKT < khi)
      gemm_fetch<T, TILE, NT, LD, GKT>(P, n, row0, col0, k0 + GKT, ta, tb, ua, ub, ra, rb, t);
    gemm_mfma_chunk<T, GKT, TILE + 1>(As, Bs, acc, r0w, c0w, lane);
    __syncthreads();
  }
  gemm_store_wave<T>(P, acc, row0 + r0w, col0 + c0w, n, lane);
}

// Split-k variant for long reductions (k = nb) into small outputs: KS
// wavefronts share one 32 x 32 output tile, wavefront g taking the k chunks
// g, g + KS, ... (GKS wide, staged through its own LDS buffers); the partial
// tiles are then summed in LDS in a fixed pairwise order, so the result does
// not depend on timing.  Every wavefront runs the same number of chunk steps
// (chunks past k load zeros), so the block barriers line up.  Registers capped
// for three wavefronts per SIMD, as k_gemm.
template <typename T, int TILE, int KS>
__global__ __launch_bounds__(64 * KS, 3) void k_gemm_sk(const GemmProb<T>* __restrict__ probs,
                                                     const GemmTask* __restrict__ tasks) {
  using V = typename Mfma<T>::V;
  static_assert(TILE == 32, "split-k tiles are one wavefront's 32 x 32 piece");
  constexpr int NG = 64;               // threads per group (one wavefront)
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
  const int grp = threadIdx.x / NG, t = threadIdx.x % NG, lane = t;
  if (((P.flags & GF_LOWER_ONLY) && row0 + TILE - 1 < col0) ||
      ((P.flags & GF_UPPER_ONLY) && col0 + TILE - 1 < row0))
    return;  // the tile lies in the triangle that is not written
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
  V acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = V{T(0), T(0), T(0), T(0)};
  int klo, khi;
  gemm_k_range<TILE, GKS>(P.flags, row0, col0, P.k, klo, khi);
  const int nsteps = (khi - klo + GKS * KS - 1) / (GKS * KS);
  T ra[LD], rb[LD];
  gemm_fetch<T, TILE, NG, LD, GKS>(P, n, row0, col0, klo + grp * GKS, ta, tb, ua, ub, ra, rb, t);
  for (int st = 0; st < nsteps; ++st) {
    gemm_stage<T, TILE, NG, LD, GKS, TILE + 1>(As, Bs, ra, rb, ta, tb, t);
    __syncthreads();
    if (st + 1 < nsteps)
      gemm_fetch<T, TILE, NG, LD, GKS>(P, n, row0, col0, klo + ((st + 1) * KS + grp) * GKS, ta, tb, ua,
                                       ub, ra, rb, t);
    gemm_mfma_chunk<T, GKS, TILE + 1>(As, Bs, acc, 0, 0, lane);
    __syncthreads();
  }
  // pairwise sum of the groups' tiles: at level h, group g (h <= g < 2h) hands
  // its tile to group g - h
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int h = KS / 2; h >= 1; h /= 2) {
    if (grp >= h && grp < 2 * h) {
      T* buf = sm + (size_t)(grp - h) * TILE * TILE;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            buf[(16 * a + li) + (16 * b + Mfma<T>::m_of(lane, v)) * TILE] = acc[a][b][v];
    }
    __syncthreads();
    if (grp < h) {
      const T* buf = sm + (size_t)grp * TILE * TILE;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            acc[a][b][v] += buf[(16 * a + li) + (16 * b + Mfma<T>::m_of(lane, v)) * TILE];
    }
    __syncthreads();
  }
  if (grp == 0) gemm_store_wave<T>(P, acc, row0, col0, n, lane);
}

// Tall-skinny batched GEMM (m = the nb rows of a tile, n and k at most a few
// dozen): C = alpha A op(B) + beta C (or beta C0), A not transposed.  One
// workgroup covers kTallRows rows and kTallCols columns: wavefront w a quarter
// of the rows as 16-row blocks by two 16-column blocks of accumulators.  The
// k x kTallCols block of op(B) is staged in LDS in chunks of kTallK rows; A
// goes from HBM straight into the matrix-core operand (lane l: row l % 16 of
// the block, column k + l / 16: four coalesced 128-byte segments per load),
// the next KQ k steps' loads in flight during the current steps' matrix-core
// work.  The many tiny output tiles of such
// products as 2-D tiles would be bound by workgroup dispatch, not by bytes.
template <typename T>
__global__ __launch_bounds__(256) void k_gemm_tall(const GemmProb<T>* __restrict__ probs,
                                                   const GemmTask* __restrict__ tasks) {
  using V = typename Mfma<T>::V;
  constexpr int RB = kTallRows / 4 / 16;  // 16-row blocks per wavefront
  constexpr int CBK = kTallCols / 16;     // 16-column blocks
  constexpr int KQ = 2;                   // matrix-core k steps per batch of A loads
  static_assert(RB * 64 == kTallRows && CBK * 16 == kTallCols && kTallK % (4 * KQ) == 0,
                "tall shape");
  __shared__ T Bs[kTallK][kTallCols + 1];
  const GemmTask tk = tasks[blockIdx.x];
  const GemmProb<T> P = probs[tk.prob];
  if (P.skip && *P.skip) return;
  int n = P.n;
  if (P.n_dev) n = min(n, *P.n_dev);
  const int row0 = tk.ti * kTallRows, col0 = tk.tj * kTallCols;
  if (row0 >= P.m || col0 >= n) return;
  const int nc = min(kTallCols, n - col0);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int rw = row0 + w * (kTallRows / 4);  // first row of this wavefront
  const bool tb = (P.flags & GF_B_TRANS) != 0;
  const bool ua = (P.flags & GF_A_UPPER) != 0, ub = (P.flags & GF_B_UPPER) != 0;
  const bool la = (P.flags & GF_A_UNIT_LOWER) != 0, lb = (P.flags & GF_B_UNIT_LOWER) != 0;
  V acc[RB][CBK];
#pragma unroll
  for (int a = 0; a < RB; ++a)
#pragma unroll
    for (int b = 0; b < CBK; ++b) acc[a][b] = V{T(0), T(0), T(0), T(0)};
  auto load_a = [&](int gk, T (&av)[RB]) {
#pragma unroll
    for (int a = 0; a < RB; ++a) {
      const int row = rw + 16 * a + li;
      av[a] = (gk < P.k && row < P.m) ? gemm_masked(P.A, P.lda, row, gk, ua, la) : T(0);
    }
  };
  for (int k0 = 0; k0 < P.k; k0 += kTallK) {
    const int kc = min(kTallK, P.k - k0);
    // A in batches of KQ matrix-core steps, the next batch in flight during the
    // current one's steps; the first batch is requested before op(B)'s staging
    T av[KQ][RB], an[KQ][RB];
#pragma unroll
    for (int s = 0; s < KQ; ++s) load_a(k0 + 4 * s + lk, av[s]);
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
    for (int kq = 0; kq < kc; kq += 4 * KQ) {
      if (kq + 4 * KQ < kc) {
#pragma unroll
        for (int s = 0; s < KQ; ++s) load_a(k0 + kq + 4 * (KQ + s) + lk, an[s]);
      }
#pragma unroll
      for (int s = 0; s < KQ; ++s) {
        T bv[CBK];
#pragma unroll
        for (int b = 0; b < CBK; ++b) bv[b] = Bs[kq + 4 * s + lk][16 * b + li];  // zero past kc
#pragma unroll
        for (int a = 0; a < RB; ++a)
#pragma unroll
          for (int b = 0; b < CBK; ++b) acc[a][b] = Mfma<T>::mac(bv[b], av[s][a], acc[a][b]);
      }
#pragma unroll
      for (int s = 0; s < KQ; ++s)
#pragma unroll
        for (int a = 0; a < RB; ++a) av[s][a] = an[s][a];
    }
    __syncthreads();
  }
  // one value at a time: only the first rows' workgroups read (C0 rows < m0),
  // and gathering the loads costs this kernel registers (occupancy)
#pragma unroll
  for (int a = 0; a < RB; ++a)
#pragma unroll
    for (int b = 0; b < CBK; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        gemm_store1(P, rw + 16 * a + li, col0 + 16 * b + Mfma<T>::m_of(lane, v), n, acc[a][b][v]);
}

// ------------------------------------------------------------------------------
// Split-precision products (plans of precision HCORE_HIP_SPLIT): FP64 operands
// multiplied on the BF16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32
// accumulation), Ozaki style.  For every k chunk of 32 (one matrix-core step)
// each row i of op(A) and column j of op(B) gets a power of two 2^E just above
// its largest magnitude in the chunk; the scaled values x in (-1, 1) are cut
// into S integer digits of at most 8 bits, d_1 = rint(256 x), then the
// rounded remainders, each exact in BF16:
//   x = sum_{s=1..S} d_s 2^(-8 s) + rho,  |rho| <= 2^(-8 S - 1).
// One chunk's digit products are integers whose sums stay below 2^21, so the
// FP32 accumulation is exact; the pairs (s, t) with s + t <= S + 1 are
// summed per level s + t and added to FP64 accumulators with the scale
// 2^(E_i + F_j - 8 (s + t)).  The error of a product term relative to the
// chunk's row and column maxima is about S 2^(-8 S).
// One workgroup per 32 x 32 output tile; KS wavefronts share its k range
// (chunks g, g + KS, ...), their FP64 tiles summed in LDS in a fixed order.
// ------------------------------------------------------------------------------
using OzAcc = float __attribute__((ext_vector_type(4)));
using OzBf8 = __bf16 __attribute__((ext_vector_type(8)));
using OzU4 = unsigned __attribute__((ext_vector_type(4)));

// Eight integer-valued doubles with |d| <= 256 as BF16: the float of such an
// integer has at most 8 significant bits, so its upper half is the BF16.
__device__ __forceinline__ OzBf8 oz_pack(const double (&d)[8]) {
  OzU4 w;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    w[q] = (__float_as_uint((float)d[2 * q]) >> 16) |
           (__float_as_uint((float)d[2 * q + 1]) & 0xffff0000u);
  return __builtin_bit_cast(OzBf8, w);
}

// The S digits of 8 values scaled by 2^-e.
template <int S>
__device__ __forceinline__ void oz_digits(const double (&v)[8], int e, OzBf8 (&out)[S]) {
  double x[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) x[q] = ldexp(v[q], -e);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    double d[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double t = x[q] * 256.0;  // exact
      d[q] = rint(t);
      x[q] = t - d[q];                // exact: the rounded-off part of t
    }
    out[s] = oz_pack(d);
  }
}

// Exponent E with max|v| < 2^E over a row (or column) whose 32 chunk values
// are spread over the four lanes li, li + 16, li + 32, li + 48 (0 when all
// are zero).
__device__ __forceinline__ int oz_exponent(const double (&v)[8]) {
  double m = 0.0;
#pragma unroll
  for (int q = 0; q < 8; ++q) m = fmax(m, fabs(v[q]));
  int e = 0;
  (void)frexp(m, &e);
  if (m == 0.0) e = -1100;  // below every other exponent: the max picks the rest
  e = max(e, __shfl_xor(e, 16));
  e = max(e, __shfl_xor(e, 32));
  return e == -1100 ? 0 : e;
}

// One staged k chunk (32 wide) into the wavefront's FP64 32 x 32 accumulators:
// acc[rb][cb][v] = C(i = 16 rb + lane % 16, j = 16 cb + 4 (lane / 16) + v).
template <int S, int LDT>
__device__ __forceinline__ void oz_chunk(const double (*As)[LDT], const double (*Bs)[LDT],
                                         double (&acc)[2][2][4], int lane) {
  const int li = lane & 15, lg = lane >> 4;
  // op(B)'s two 16-column blocks: digits kept for both row blocks of op(A)
  OzBf8 sb[2][S];
  int eb[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    double vb[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) vb[q] = Bs[8 * lg + q][16 * h + li];  // op(B)(k = 8 lg + q, j)
    eb[h] = oz_exponent(vb);
    oz_digits<S>(vb, eb[h], sb[h]);
  }
  // column 16 cb + 4 lg + v's exponent is held by lane 4 lg + v
  int ej[2][4];
#pragma unroll
  for (int cb = 0; cb < 2; ++cb)
#pragma unroll
    for (int v = 0; v < 4; ++v) ej[cb][v] = __shfl(eb[cb], 4 * lg + v);
#pragma unroll 1
  for (int rb = 0; rb < 2; ++rb) {
    double va[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) va[q] = As[8 * lg + q][16 * rb + li];  // op(A)(i, k = 8 lg + q)
    const int ea = oz_exponent(va);
    OzBf8 sa[S];
    oz_digits<S>(va, ea, sa);
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      OzAcc lv[S];
#pragma unroll
      for (int L = 0; L < S; ++L) {  // level L: the digit pairs (s, L - s), 0-based
        lv[L] = OzAcc{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s <= L; ++s)
          lv[L] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sb[cb][L - s], sa[s], lv[L], 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // sum_L lv[L] 2^(-8 (L + 2)), Horner in FP64
        double sum = (double)lv[S - 1][v];
#pragma unroll
        for (int L = S - 2; L >= 0; --L) sum = fma(sum, 0x1p-8, (double)lv[L][v]);
        const double c = ldexp(sum, ea + ej[cb][v] - 16);
        if (rb == 0) acc[0][cb][v] += c;
        else acc[1][cb][v] += c;
      }
    }
  }
}

template <int S, int KS>
__global__ __launch_bounds__(64 * KS) void k_gemm_oz(const GemmProb<double>* __restrict__ probs,
                                                     const GemmTask* __restrict__ tasks) {
  constexpr int TILE = 32, NG = 64, GKS = 32;
  constexpr int LD = TILE * GKS / NG;  // elements of each operand per thread per chunk
  constexpr int STG = 2 * GKS * (TILE + 1);
  constexpr int RED = (KS / 2) * TILE * TILE;
  constexpr int SM = KS * STG > RED ? KS * STG : RED;
  static_assert(KS == 1 || KS == 2 || KS == 4, "split-k groups");
  __shared__ double sm[SM];
  const GemmTask tk = tasks[blockIdx.x];
  const GemmProb<double> P = probs[tk.prob];
  if (P.skip && *P.skip) return;
  int n = P.n;
  if (P.n_dev) n = min(n, *P.n_dev);
  const int row0 = tk.ti * TILE, col0 = tk.tj * TILE;
  if (row0 >= P.m || col0 >= n) return;
  const int grp = threadIdx.x / NG, lane = threadIdx.x % NG;
  if (((P.flags & GF_LOWER_ONLY) && row0 + TILE - 1 < col0) ||
      ((P.flags & GF_UPPER_ONLY) && col0 + TILE - 1 < row0))
    return;  // the tile lies in the triangle that is not written
  if ((P.flags & GF_LOWER_ZERO_UPPER) && row0 + TILE - 1 < col0) {
    for (int e = threadIdx.x; e < TILE * TILE; e += NG * KS) {
      const int i = row0 + (e % TILE), j = col0 + (e / TILE);
      if (i < P.m && j < n) P.C[i + (size_t)j * P.ldc] = 0.0;
    }
    return;
  }
  const bool ta = P.flags & GF_A_TRANS, tb = P.flags & GF_B_TRANS;
  const bool ua = P.flags & GF_A_UPPER, ub = P.flags & GF_B_UPPER;
  double (*As)[TILE + 1] = reinterpret_cast<double (*)[TILE + 1]>(sm + grp * STG);
  double (*Bs)[TILE + 1] = reinterpret_cast<double (*)[TILE + 1]>(sm + grp * STG + GKS * (TILE + 1));
  double acc[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[a][b][v] = 0.0;
  int klo, khi;
  gemm_k_range<TILE, GKS>(P.flags, row0, col0, P.k, klo, khi);
  const int nsteps = (khi - klo + GKS * KS - 1) / (GKS * KS);
  for (int st = 0; st < nsteps; ++st) {
    {
      double ra[LD], rb[LD];
      gemm_fetch<double, TILE, NG, LD, GKS>(P, n, row0, col0, klo + (st * KS + grp) * GKS, ta, tb, ua,
                                            ub, ra, rb, lane);
      gemm_stage<double, TILE, NG, LD, GKS, TILE + 1>(As, Bs, ra, rb, ta, tb, lane);
    }
    __syncthreads();
    oz_chunk<S, TILE + 1>(As, Bs, acc, lane);
    __syncthreads();
  }
  const int li = lane & 15, lg = lane >> 4;
#pragma unroll
  for (int h = KS / 2; h >= 1; h /= 2) {  // pairwise sum of the groups' tiles
    if (grp >= h && grp < 2 * h) {
      double* buf = sm + (size_t)(grp - h) * TILE * TILE;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int v = 0; v < 4; ++v) buf[(16 * a + li) + (16 * b + 4 * lg + v) * TILE] = acc[a][b][v];
    }
    __syncthreads();
    if (grp < h) {
      const double* buf = sm + (size_t)grp * TILE * TILE;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[a][b][v] += buf[(16 * a + li) + (16 * b + 4 * lg + v) * TILE];
    }
    __syncthreads();
  }
  if (grp == 0) {
    int ii[16], jj[16];
    double vv[16];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int q = 8 * a + 4 * b + v;
          ii[q] = row0 + 16 * a + li;
          jj[q] = col0 + 16 * b + 4 * lg + v;
          vv[q] = acc[a][b][v];
        }
    gemm_store_n<double, 16>(P, ii, jj, n, vv);
  }
}

// Tall-skinny split-precision products (A not transposed, m >= kTallMinM,
// n <= kTallMaxN, k <= kTallMaxK; the FP64 tall kernel's tasks of kTallRows x
// kTallCols): per k chunk of 32 the workgroup cuts the 32 x 32 block of op(B)
// into digits once (8 threads per column, column exponents by lane shuffles)
// and keeps them in LDS; each wavefront reads its 64 rows of A straight from
// HBM (16 consecutive rows per load) and cuts them itself.
template <int S>
__global__ __launch_bounds__(256) void k_gemm_oz_tall(const GemmProb<double>* __restrict__ probs,
                                                      const GemmTask* __restrict__ tasks) {
  constexpr int NC = kTallCols, KC = 32, RW = kTallRows / 4;  // 64 rows per wavefront
  constexpr int BDK = KC + 8;  // k stride of a digit column (bank-conflict-free 16-byte reads)
  static_assert(NC == 32 && RW == 64, "tall split-precision shape");
  __shared__ double Bs[KC][NC + 1];
  __shared__ __align__(16) __bf16 Bd[S][NC][BDK];  // rows of 80 bytes: 16-byte reads stay aligned
  __shared__ int eB[NC];
  const GemmTask tk = tasks[blockIdx.x];
  const GemmProb<double> P = probs[tk.prob];
  if (P.skip && *P.skip) return;
  int n = P.n;
  if (P.n_dev) n = min(n, *P.n_dev);
  const int row0 = tk.ti * kTallRows, col0 = tk.tj * NC;
  if (row0 >= P.m || col0 >= n) return;
  const int nc = min(NC, n - col0);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int rw = row0 + w * RW;
  const bool tb = (P.flags & GF_B_TRANS) != 0;
  const bool ua = (P.flags & GF_A_UPPER) != 0, ub = (P.flags & GF_B_UPPER) != 0;
  const bool la = (P.flags & GF_A_UNIT_LOWER) != 0, lb = (P.flags & GF_B_UNIT_LOWER) != 0;
  int klo, khi;
  gemm_k_range<kTallRows, KC>(P.flags, row0, col0, P.k, klo, khi);
  double acc[4][2][4];  // C(rw + 16 rb + li, col0 + 16 cb + 4 lg + v)
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[a][b][v] = 0.0;
  for (int k0 = klo; k0 < khi; k0 += KC) {
    for (int e = t; e < KC * NC; e += 256) {  // op(B)(k0 + kk, col0 + j)
      int kk, j;
      if (tb) { j = e % NC; kk = e / NC; }
      else    { kk = e % KC; j = e / KC; }
      double v = 0.0;
      if (k0 + kk < P.k && j < nc) {
        const int gk = k0 + kk, gj = col0 + j;
        v = tb ? gemm_masked(P.B, P.ldb, gj, gk, ub, lb) : gemm_masked(P.B, P.ldb, gk, gj, ub, lb);
      }
      Bs[kk][j] = v;
    }
    __syncthreads();
    {  // the block's digits: thread t takes column t / 8, k = 4 (t % 8) .. + 3
      const int j = t >> 3, kq = (t & 7) * 4;
      double v[4], m = 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        v[i] = Bs[kq + i][j];
        m = fmax(m, fabs(v[i]));
      }
      int e = 0;
      (void)frexp(m, &e);
      if (m == 0.0) e = -1100;
      e = max(e, __shfl_xor(e, 1));
      e = max(e, __shfl_xor(e, 2));
      e = max(e, __shfl_xor(e, 4));
      e = e == -1100 ? 0 : e;
      if ((t & 7) == 0) eB[j] = e;
      double x[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) x[i] = ldexp(v[i], -e);
#pragma unroll
      for (int sd = 0; sd < S; ++sd)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const double tt = x[i] * 256.0;
          const double d = rint(tt);
          x[i] = tt - d;
          Bd[sd][j][kq + i] = __builtin_bit_cast(__bf16, (unsigned short)(__float_as_uint((float)d) >> 16));
        }
    }
    __syncthreads();
    OzBf8 sb[2][S];
    int ej[2][4];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
#pragma unroll
      for (int sd = 0; sd < S; ++sd)
        sb[cb][sd] = *reinterpret_cast<const OzBf8*>(&Bd[sd][16 * cb + li][8 * lg]);
#pragma unroll
      for (int v = 0; v < 4; ++v) ej[cb][v] = eB[16 * cb + 4 * lg + v];
    }
#pragma unroll 1
    for (int rb = 0; rb < 4; ++rb) {
      double va[8];
      const int row = rw + 16 * rb + li;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int gk = k0 + 8 * lg + q;
        va[q] = (gk < P.k && row < P.m) ? gemm_masked(P.A, P.lda, row, gk, ua, la) : 0.0;
      }
      const int ea = oz_exponent(va);
      OzBf8 sa[S];
      oz_digits<S>(va, ea, sa);
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) {
        OzAcc lv[S];
#pragma unroll
        for (int L = 0; L < S; ++L) {
          lv[L] = OzAcc{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int sd = 0; sd <= L; ++sd)
            lv[L] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sb[cb][L - sd], sa[sd], lv[L], 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          double sum = (double)lv[S - 1][v];
#pragma unroll
          for (int L = S - 2; L >= 0; --L) sum = fma(sum, 0x1p-8, (double)lv[L][v]);
          const double c = ldexp(sum, ea + ej[cb][v] - 16);
#pragma unroll
          for (int a = 0; a < 4; ++a)
            if (a == rb) acc[a][cb][v] += c;
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)  // one value at a time, as k_gemm_tall
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        gemm_store1(P, rw + 16 * a + li, col0 + 16 * b + 4 * lg + v, n, acc[a][b][v]);
}

template <int S>
static void launch_gemm_oz_s(const GemmProb<double>* probs, const GemmTask* tasks, int ntasks,
                             int ks, hipStream_t s) {
  if (ks == 0) {
    hipLaunchKernelGGL((k_gemm_oz_tall<S>), dim3(ntasks), dim3(256), 0, s, probs, tasks);
    return;
  }
  if (ks == 4)
    hipLaunchKernelGGL((k_gemm_oz<S, 4>), dim3(ntasks), dim3(256), 0, s, probs, tasks);
  else
    hipLaunchKernelGGL((k_gemm_oz<S, 1>), dim3(ntasks), dim3(64), 0, s, probs, tasks);
}

void launch_gemm_oz(const GemmProb<double>* probs, const GemmTask* tasks, int ntasks, int slices,
                    int ks, hipStream_t s) {
  if (ntasks <= 0) return;
  switch (slices) {
    case 2: launch_gemm_oz_s<2>(probs, tasks, ntasks, ks, s); break;
    case 3: launch_gemm_oz_s<3>(probs, tasks, ntasks, ks, s); break;
    case 4: launch_gemm_oz_s<4>(probs, tasks, ntasks, ks, s); break;
    case 5: launch_gemm_oz_s<5>(probs, tasks, ntasks, ks, s); break;
    default: launch_gemm_oz_s<6>(probs, tasks, ntasks, ks, s); break;
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
  // Only the 32-tiles are split (four wavefronts): a split 64-tile workgroup
  // needs a whole CU's wave slots and LDS at once and delays the other lanes'
  // kernels more than it gains.
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
