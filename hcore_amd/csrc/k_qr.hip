// k_qr.hip — stacking and Householder QR kernels.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <limits>
#include <type_traits>

#include "hcore_internal.hpp"
#include "hcore_wave.hpp"

namespace hcore {

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

// One workgroup: the stacks of a lane by their stacked rank K into the three
// classes of QR kernel (K <= 16, <= kQrRegMaxK, larger); the order within a
// class follows the LDS atomics and changes no result.
template <typename T>
__global__ __launch_bounds__(1024) void k_qr_queue(const QrProb<T>* __restrict__ probs, int n,
                                                   int* __restrict__ q) {
  __shared__ int cnt[3], fill[3];
  const int tid = threadIdx.x;
  if (tid < 3) {
    cnt[tid] = 0;
    fill[tid] = 0;
  }
  __syncthreads();
  auto cls = [&](int i) {
    const int K = probs[i].K;
    return K <= 16 ? 0 : (K <= kQrRegMaxK ? 1 : 2);
  };
  for (int i = tid; i < n; i += 1024) atomicAdd(&cnt[cls(i)], 1);
  __syncthreads();
  if (tid == 0) {
    q[0] = 0;
    q[1] = cnt[0];
    q[2] = cnt[0] + cnt[1];
    q[3] = cnt[0];
    q[4] = cnt[1];
    q[5] = cnt[2];
  }
  for (int i = tid; i < n; i += 1024) {
    const int c = cls(i);
    const int st = c == 0 ? 0 : (c == 1 ? cnt[0] : cnt[0] + cnt[1]);
    q[kQrQueueHead + st + atomicAdd(&fill[c], 1)] = i;
    if (probs[i].routed) *probs[i].routed = c < 2 ? 1 : 0;
  }
}

template <typename T>
void launch_qr_queue(const QrProb<T>* probs, int n, int* q, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_qr_queue<T>, dim3(1), dim3(1024), 0, s, probs, n, q);
}

// ---- explicit instantiations ---------------------------------------------------
#define HCORE_INSTANTIATE(T) \
  template void launch_stack<T>(const StackProb<T>*, int, int, int, T, T, hipStream_t); \
  template void launch_geqrf<T>(const QrProb<T>*, int, int, hipStream_t); \
  template void launch_qr_queue<T>(const QrProb<T>*, int, int*, hipStream_t);
HCORE_INSTANTIATE(double)
HCORE_INSTANTIATE(float)
#undef HCORE_INSTANTIATE

}  // namespace hcore
