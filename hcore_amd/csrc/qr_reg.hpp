// qr_reg.hpp — register-resident Householder QR kernels and their launchers.
// Included by the per-precision translation units k_qrr_*.hip and k_qrf_*.hip, so
// that the heavily unrolled column recursion compiles in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <limits>
#include <type_traits>

#include "hcore_internal.hpp"
#include "hcore_wave.hpp"

namespace hcore {
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
                                           T (*Sl)[KMAX + 1], T* taus) {
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
  // reflector products w_t (t > J) in lane t, read by every lane into scalar
  // registers KC at a time (no LDS round trip)
  const T w_l = (lane > J && lane < KMAX && j0 + lane < K) ? tau * dotv : T(0);
  if (wid == 0) {
    if (lane < J) Sl[J][lane] = dotv;
    if (lane == 0) taus[J] = tau;
  }
  // per row: the reflector entry v (rows below g), 1 (row g), 0 (rows above:
  // unchanged, the matrix being finite)
  T vr[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int row = tid + r * NT;
    const T v = x[r][J] * scal;
    vr[r] = row > g ? v : (row == g ? T(1) : T(0));
    x[r][J] = row > g ? v : (row == g ? beta : x[r][J]);
  }
  constexpr int KC = 8;
#pragma unroll
  for (int k0 = J + 1; k0 < KMAX; k0 += KC) {
    if (j0 + k0 >= K) break;  // uniform: the columns past the stack's K are empty
    T wk[KC];
#pragma unroll
    for (int c = 0; c < KC; ++c)
      if (k0 + c < KMAX) wk[c] = lane_read(w_l, k0 + c);
#pragma unroll
    for (int r = 0; r < RPT; ++r)
#pragma unroll
      for (int c = 0; c < KC; ++c)
        if (k0 + c < KMAX) x[r][k0 + c] = fma(-vr[r], wk[c], x[r][k0 + c]);
  }
  if constexpr (J + 1 < KMAX)
    qr_reg_col<J + 1, T, NT, RPT, KMAX>(x, K, j0, tid, lane, wid, part, rowj, Sl, taus);
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
__global__ __launch_bounds__(NT) void k_qr_reg(const QrProb<T>* __restrict__ probs, int nb, int j0,
                                               const int* __restrict__ q, int qc) {
  constexpr int NW = NT / 64;
  static_assert(KMAX % 16 == 0 && KMAX <= 64 && NT >= KMAX && KMAX % kPB == 0, "register QR shape");
  __shared__ T part[2][NW][KMAX];
  __shared__ T rowj[2][KMAX];
  __shared__ T Sl[KMAX][KMAX + 1];  // Sl[j][t] = y_t^T y_j, t < j (panel columns)
  __shared__ T taus[KMAX];
  __shared__ T Tf[KMAX][KMAX + 1];  // the panel's compact-WY factor
  int item = blockIdx.x;
  if (q) {  // the queue's slice of this class (uniform exit before any barrier)
    if (item >= q[3 + qc]) return;
    item = q[kQrQueueHead + q[qc] + item];
  }
  const QrProb<T> P = probs[item];
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
  qr_reg_col<0, T, NT, RPT, KMAX>(x, K, j0, tid, lane, wid, part, rowj, Sl, taus);
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
// inter-thread ordering through global memory.  MINW > 1 caps the registers
// for MINW wavefronts per SIMD (two 512-thread workgroups per CU at MINW = 4,
// at the price of a few spilled registers): for launches of many stacks.
template <typename T, int NT, int RPT, int PB, int MINW>
__global__ __launch_bounds__(NT, MINW) void k_qr_fused(const QrProb<T>* __restrict__ probs, int nb,
                                                       const int* __restrict__ q, int qc) {
  constexpr int NW = NT / 64;
  constexpr int CB = 4;
  constexpr int CH = 16;
  static_assert(PB % CH == 0 && PB <= 64 && NT >= PB * CB && PB % kPB == 0, "fused QR shape");
  __shared__ T part[2][NW][PB];
  __shared__ T rowj[2][PB];
  __shared__ T Sl[PB][PB + 1];
  __shared__ T taus[PB];
  __shared__ T Tf[PB][PB + 1];
  __shared__ T tpart[NW][CB][PB];
  __shared__ T Wm[PB][CB];
  __shared__ T W2[PB][CB];
  int item = blockIdx.x;
  if (q) {  // the queue's slice of this class (uniform exit before any barrier)
    if (item >= q[3 + qc]) return;
    item = q[kQrQueueHead + q[qc] + item];
  }
  const QrProb<T> P = probs[item];
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
    qr_reg_col<0, T, NT, RPT, PB>(x, K, j0, tid, lane, wid, part, rowj, Sl, taus);
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
// launches of at least this many stacks (two rounds of one workgroup per CU)
// take the register-capped shape that fits two workgroups per CU (measured at
// nb = 1024: 914 stacks of K 33-64 -7 %; 26 stacks of K 130-187 +5 %)
constexpr int kQrfManyStacks = 512;

template <typename T, int NT, int RPT, int KMAX>
inline void launch_qr_reg_shape(const QrProb<T>* probs, int n, int nb, hipStream_t s,
                                const int* q, int qc) {
  hipLaunchKernelGGL((k_qr_reg<T, NT, RPT, KMAX>), dim3(n), dim3(NT), 0, s, probs, nb, 0, q, qc);
}

// Only the shapes below are instantiated (nb <= 1024): K <= 16 at 256
// threads x 1, 2 or 4 rows; K <= 32 at 256 x 1, 512 x 1 or 512 x 2 rows.
template <typename T>
bool launch_qr_reg(const QrProb<T>* probs, int n, int nb, int kmax, hipStream_t s, const int* q,
                   int qc) {
  if (n <= 0) return true;
  if (nb > kQrRegMaxNb || kmax > kQrRegMaxK) return false;
  static_assert(kQr16Threads == 256 && kQr32Threads == 512, "register QR shapes");
  if (kmax <= 16) {
    if (nb <= 256) launch_qr_reg_shape<T, 256, 1, 16>(probs, n, nb, s, q, qc);
    else if (nb <= 512) launch_qr_reg_shape<T, 256, 2, 16>(probs, n, nb, s, q, qc);
    else launch_qr_reg_shape<T, 256, 4, 16>(probs, n, nb, s, q, qc);
  } else {
    if (nb <= 256) launch_qr_reg_shape<T, 256, 1, 32>(probs, n, nb, s, q, qc);
    else if (nb <= 512) launch_qr_reg_shape<T, 512, 1, 32>(probs, n, nb, s, q, qc);
    else launch_qr_reg_shape<T, 512, 2, 32>(probs, n, nb, s, q, qc);
  }
  return true;
}

template <typename T>
bool launch_qr_fused(const QrProb<T>* probs, int n, int nb, hipStream_t s, const int* q, int qc) {
  if (n <= 0) return true;
  if (nb > kQrFusedMaxNb) return false;
  constexpr int NT = kQrfThreads, PB = kQrfPanel;
  static_assert(kQrFusedMaxNb <= 4 * NT, "fused QR shapes");
  const int rpt = (nb + NT - 1) / NT;
  if (nb <= 256)
    hipLaunchKernelGGL((k_qr_fused<T, 256, 1, 32, 1>), dim3(n), dim3(256), 0, s, probs, nb, q, qc);
  else if (rpt <= 1)
    hipLaunchKernelGGL((k_qr_fused<T, NT, 1, PB, 1>), dim3(n), dim3(NT), 0, s, probs, nb, q, qc);
  else if (rpt <= 2 && n >= kQrfManyStacks)  // two workgroups per CU
    hipLaunchKernelGGL((k_qr_fused<T, NT, 2, PB, 4>), dim3(n), dim3(NT), 0, s, probs, nb, q, qc);
  else if (rpt <= 2)
    hipLaunchKernelGGL((k_qr_fused<T, NT, 2, PB, 1>), dim3(n), dim3(NT), 0, s, probs, nb, q, qc);
  else
    hipLaunchKernelGGL((k_qr_fused<T, NT, 4, PB, 1>), dim3(n), dim3(NT), 0, s, probs, nb, q, qc);
  return true;
}

}  // namespace hcore
