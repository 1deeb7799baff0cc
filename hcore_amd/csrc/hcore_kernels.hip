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

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum of a group of G consecutive lanes (G a power of two, groups aligned).
template <typename T>
__device__ __forceinline__ T group_sum(T v, int g) {
  for (int off = g >> 1; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Reduce N values (N a power of two <= 64) across the 64 lanes: each exchange
// step hands half of the remaining values to the partner lane, so the cost is
// about N + 6 shuffles instead of 6N.  On return v[0] holds the wave total of
// value index reduce_scatter_index<N>(lane).  The halving is written as a
// compile-time recursion so every array index is a constant (no scratch).
template <int HALF, int OFF, int N, typename T>
__device__ __forceinline__ void rs_level(T (&v)[N], int lane) {
  if constexpr (HALF >= 1) {
    const bool upper = (lane & OFF) != 0;
#pragma unroll
    for (int i = 0; i < HALF; ++i) {
      const T send = upper ? v[i] : v[i + HALF];
      const T keep = upper ? v[i + HALF] : v[i];
      v[i] = keep + __shfl_xor(send, OFF, 64);
    }
    rs_level<HALF / 2, OFF / 2>(v, lane);
  } else {
#pragma unroll
    for (int off = OFF; off >= 1; off >>= 1) v[0] += __shfl_xor(v[0], off, 64);
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
constexpr int GT = 64;  // output tile edge
constexpr int GK = 16;  // k chunk

template <typename T>
__global__ __launch_bounds__(kThreads) void k_gemm(const GemmProb<T>* __restrict__ probs,
                                                   const GemmTask* __restrict__ tasks) {
  __shared__ T As[GK][GT];
  __shared__ T Bs[GK][GT];
  const GemmTask tk = tasks[blockIdx.x];
  const GemmProb<T> P = probs[tk.prob];
  if (P.skip && *P.skip) return;
  int n = P.n;
  if (P.n_dev) n = min(n, *P.n_dev);
  const int row0 = tk.ti * GT, col0 = tk.tj * GT;
  if (row0 >= P.m || col0 >= n) return;
  const int t = threadIdx.x;
  const int tx = t & 15, ty = t >> 4;
  const bool lzu = (P.flags & GF_LOWER_ZERO_UPPER) != 0;
  if (lzu && row0 + GT - 1 < col0) {  // tile strictly above the diagonal: zero fill
    for (int e = t; e < GT * GT; e += kThreads) {
      const int i = row0 + (e & 63), j = col0 + (e >> 6);
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

  for (int k0 = 0; k0 < P.k; k0 += GK) {
    // A tile: As[kk][i] = op(A)(row0+i, k0+kk)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int i, kk;
      if (!ta) { i = t & 63; kk = (t >> 6) + 4 * q; }
      else     { kk = t & 15; i = (t >> 4) + 16 * q; }
      const int gi = row0 + i, gk = k0 + kk;
      T v = T(0);
      if (gi < P.m && gk < P.k) {
        const int p = ta ? gk : gi, qq = ta ? gi : gk;  // stored (row, col)
        if (!(ua && p > qq)) v = P.A[p + (size_t)qq * P.lda];
      }
      As[kk][i] = v;
    }
    // B tile: Bs[kk][j] = op(B)(k0+kk, col0+j)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int j, kk;
      if (tb) { j = t & 63; kk = (t >> 6) + 4 * q; }
      else    { kk = t & 15; j = (t >> 4) + 16 * q; }
      const int gj = col0 + j, gk = k0 + kk;
      T v = T(0);
      if (gj < n && gk < P.k) {
        const int p = tb ? gj : gk, qq = tb ? gk : gj;
        if (!(ub && p > qq)) v = P.B[p + (size_t)qq * P.ldb];
      }
      Bs[kk][j] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; ++kk) {
      T av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = As[kk][tx + 16 * a];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = Bs[kk][ty + 16 * b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
    }
    __syncthreads();
  }
  const bool bz = (P.flags & GF_BETA_ZERO) != 0;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int i = row0 + tx + 16 * a, j = col0 + ty + 16 * b;
      if (i < P.m && j < n) {
        T* c = P.C + i + (size_t)j * P.ldc;
        if (lzu && i < j) { *c = T(0); continue; }
        T v = P.alpha * acc[a][b];
        if (!bz) v = fma(P.beta, *c, v);
        *c = v;
      }
    }
}

template <typename T>
void launch_gemm(const GemmProb<T>* probs, const GemmTask* tasks, int ntasks, hipStream_t s) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(k_gemm<T>, dim3(ntasks), dim3(kThreads), 0, s, probs, tasks);
}

// ------------------------------------------------------------------------------
// stacked factors (clone + scale): Us = [beta*CU | alpha*AU], Vs(:, :rc) = CV
// ------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_stack(const StackProb<T>* __restrict__ probs, int nb,
                                                    T alpha, T beta) {
  const StackProb<T> P = probs[blockIdx.x];
  const int ncol = 2 * P.rc + P.ra;
  for (int col = 0; col < ncol; ++col) {
    for (int row = threadIdx.x; row < nb; row += kThreads) {
      if (col < P.rc) {
        P.Us[row + (size_t)col * nb] = beta * P.CU[row + (size_t)col * P.ldc];
      } else if (col < P.rc + P.ra) {
        const int c = col - P.rc;
        P.Us[row + (size_t)col * nb] = alpha * P.AU[row + (size_t)c * P.lda];
      } else {
        const int c = col - P.rc - P.ra;
        P.Vs[row + (size_t)c * nb] = P.CV[row + (size_t)c * P.ldc];
      }
    }
  }
}

template <typename T>
void launch_stack(const StackProb<T>* probs, int n, int nb, T alpha, T beta, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_stack<T>, dim3(n), dim3(kThreads), 0, s, probs, nb, alpha, beta);
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
#pragma unroll
  for (int off = G / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
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
// leading dimension K), in place, reflectors discarded.  Stops at the first
// step whose largest remaining column norm^2 is <= thr2; returns that step k
// (the numerical rank).  On return rdiag[c] = R(c,c) and the strict upper
// triangle of A's first k rows holds R; piv[j] is the original column now in
// position j.  Groups of G lanes update one column each, with the column's
// rows (gl + G i, i < RPL, K <= G*RPL) in registers so that all its loads are
// in flight together; two block barriers per step.
template <typename T, int NT, int G, int RPL>
__device__ int qrcp(T* A, int K, T thr2, T* cn, T* rdiag, int* piv, T* gv, int* gi, T* red) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  constexpr int NG = NT / G;
  constexpr int NW = NT / 64;
  const int grp = tid / G, gl = tid % G;
  // exact column norms and per-group maxima
  {
    T best = T(-1);
    int bi = -1;
    for (int j = grp; j < K; j += NG) {
      const T* c = A + (size_t)j * K;
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
    T* cs_ = A + (size_t)s * K;
    T* cp_ = A + (size_t)p * K;
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
      T* cj = A + (size_t)j * K;
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
__device__ void r_transpose(T* A, int K, int k, const T* rdiag) {
  const size_t npairs = (size_t)K * (K - 1) / 2;
  for (size_t e = threadIdx.x; e < npairs; e += NT) {
    // pair e -> (lo, hi), lo < hi, enumerated column by column of the upper triangle
    int hi = (int)((1.0 + sqrt(1.0 + 8.0 * (double)e)) * 0.5);
    while ((size_t)hi * (hi - 1) / 2 > e) --hi;
    while ((size_t)(hi + 1) * hi / 2 <= e) ++hi;
    const int lo = (int)(e - (size_t)hi * (hi - 1) / 2);
    if (lo >= k) continue;
    const T v = A[lo + (size_t)hi * K];  // R(lo, hi)
    A[hi + (size_t)lo * K] = v;          // R^T(hi, lo)
    A[lo + (size_t)hi * K] = T(0);
  }
  for (int c = threadIdx.x; c < k; c += NT) A[c + (size_t)c * K] = rdiag[c];
}

// Shared epilogue: singular values (norms of the k orthogonalised columns of
// W = R^T V), descending order, truncation (hcore_dgemm.c:444-464) and the
// left singular vectors U = P W / sigma (rows permuted back by piv).
template <typename T, int NT>
__device__ void core_epilogue(const CoreProb<T>& P, const T* W, int k, const int* piv, T* sig,
                              int* perm, int* flags, int rk, int maxrk, T acc) {
  const int K = P.K;
  const int tid = threadIdx.x;
  if (tid == 0) flags[1] = K;  // flags[2] (non-finite input) was set at load time
  block_sync<NT>();
  for (int j = tid; j < K; j += NT) {
    T s = T(0);
    if (j < k) {
      const T* cj = W + (size_t)j * K;
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
    P.Zu[piv[row] + (size_t)i * K] = (sj > T(0) && j < k) ? W[row + (size_t)j * K] / sj : T(0);
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
  for (int e = tid; e < K * K; e += NT) {
    const int c = e / K, row = e - c * K;
    const T v = P.Tc[c + (size_t)row * K];
    bad |= !(fabs(v) <= T(1e300));
    A[e] = v;
  }
  if (bad) flags[2] = 1;
  block_sync<NT>();
  T thr2 = T(0);
  if (rk == 0 && acc > T(0)) {
    const T thr = T(1e-3) * acc / sqrt(T(K));
    thr2 = thr * thr;
  }
  const int k = qrcp<T, NT, G, RPL>(A, K, thr2, nrm, rdiag, piv, gv, gi, red);
  block_sync<NT>();
  r_transpose<T, NT>(A, K, k, rdiag);
  block_sync<NT>();
  const T tol = T(K) * eps_of<T>();
  int sweep = 0;
  if (k > 1) {
    for (; sweep < kMaxSweeps; ++sweep) {
      column_norms<T, NT, G, RPL>(A, K, k, K, nrm);
      if (tid == 0) flags[0] = 0;
      block_sync<NT>();
      jacobi_sweep<T, NT, G, RPL>(A, K, k, K, tol, thr2, nrm, &flags[0]);
      const int rotated = flags[0];
      block_sync<NT>();
      if (!rotated) break;
    }
  }
  if (tid == 0 && P.sweeps) *P.sweeps = sweep + 1;
  core_epilogue<T, NT>(P, A, k, piv, sig, perm, flags, rk, maxrk, acc);
}

int core_lds_max_k(size_t elem) {
  const size_t lds = 160 * 1024 - 256;
  int k = 1;
  while (core_head_bytes(elem, k + 1) + elem * (size_t)(k + 1) * (k + 1) <= lds) ++k;
  return k;
}

// variant -> (threads, group, rows per lane, storage)
#define HCORE_CORE_VARIANTS(X)                          \
  X(CORE_K16, 64, 8, 2, true)                           \
  X(CORE_K32, 64, 4, 8, true)                           \
  X(CORE_K64, 64, 2, 32, true)                          \
  X(CORE_K128, 256, 4, 32, true)                        \
  X(CORE_KLDS, 512, 4, 36, true)                        \
  X(CORE_K192, 768, 8, 24, false)                       \
  X(CORE_K256, 1024, 16, 16, false)                     \
  X(CORE_K512, 1024, 32, 16, false)

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
  if (n <= 0) return;
  const size_t head = core_head_bytes(sizeof(T), kmax);
  switch (variant) {
#define HCORE_LAUNCH(V, NT, G, RPL, LDS)                                                     \
  case V: {                                                                                  \
    const size_t sm = head + (LDS ? sizeof(T) * (size_t)kmax * kmax : 0);                    \
    hipLaunchKernelGGL((k_core<T, NT, G, RPL, LDS>), dim3(n), dim3(NT), sm, s, probs, kmax,  \
                       rk, maxrk, acc);                                                      \
    break;                                                                                   \
  }
    HCORE_CORE_VARIANTS(HCORE_LAUNCH)
#undef HCORE_LAUNCH
    default: break;
  }
}

int core_variant_for(int K, int kl) {
  if (K <= 16) return CORE_K16;
  if (K <= 32) return CORE_K32;
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
template <typename T>
void prepare_kernels() {
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
}

// ---- explicit instantiations ---------------------------------------------------
#define HCORE_INSTANTIATE(T)                                                                  \
  template void launch_gemm<T>(const GemmProb<T>*, const GemmTask*, int, hipStream_t);        \
  template void launch_stack<T>(const StackProb<T>*, int, int, T, T, hipStream_t);            \
  template void launch_geqrf<T>(const QrProb<T>*, int, int, hipStream_t);                     \
  template void launch_core<T>(const CoreProb<T>*, int, int, int, int, int, T, hipStream_t);  \
  template void launch_apply<T>(const ApplyProb<T>*, int, int, hipStream_t);                 \
  template void prepare_kernels<T>();
HCORE_INSTANTIATE(double)
HCORE_INSTANTIATE(float)

}  // namespace hcore
