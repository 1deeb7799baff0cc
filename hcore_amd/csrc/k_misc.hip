// k_misc.hip — reassembly (reflector application, compact-WY solve) and conversion kernels.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <limits>
#include <type_traits>

#include "hcore_internal.hpp"
#include "hcore_wave.hpp"

namespace hcore {

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

// ------------------------------------------------------------------------------
// compact-WY solve M = T Y1^T Z by blocked back substitution (K <= kWyMaxK):
// for the panels b from last to first,
//   M_b = T_bb (B_b - S(b, >b) M_{>b}),  S = Y^T Y,  B = Y1^T Z (in M on entry),
// which is M = T B for the forward compact-WY T (T^{-1} = diag(1/tau) +
// striu(S)).  A workgroup holds CW columns of M in LDS (64 for K <= 256, 16
// for K <= 512, where 64 columns of K rows would not fit); wave w computes
// rows 2w, 2w+1 of each panel, lane j column j % CW, the 64 / CW lanes of a
// column splitting the panel's dot products and summing them by shuffles.
// The panel's rows of S (contiguous in the stored lower triangle) and its
// kPB x kPB factor are staged in LDS by coalesced loads; the next panel's are
// fetched into registers while the current one is solved.
// ------------------------------------------------------------------------------
template <int CW>
struct WyShape {
  static constexpr int LD = CW + 1;                  // row stride of M in LDS
  static constexpr int KMAX = CW == 64 ? 256 : 512;  // rows the staging covers
  static constexpr int SPT = KMAX / kThreads;        // S entries per thread and row
};

template <typename T, int CW>
__global__ __launch_bounds__(kThreads) void k_wy_solve(const WySolveProb<T>* __restrict__ probs,
                                                       int ncb) {
  using Sh = WyShape<CW>;
  constexpr int LD = Sh::LD, KM = Sh::KMAX, SPT = Sh::SPT, SPLIT = 64 / CW;
  static_assert(kPB == 8 && kThreads == 256 && KM % kThreads == 0 && 64 % CW == 0,
                "two rows of a panel per wavefront");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const WySolveProb<T> P = probs[blockIdx.x / ncb];
  if (*P.status != ST_OK) return;
  const int r = *P.rank;
  const int j0 = (blockIdx.x % ncb) * CW;
  if (j0 >= r) return;
  const int K = P.K;
  const int nc = min(CW, r - j0);
  T* Mb = reinterpret_cast<T*>(smem_raw);  // K x CW, row c at Mb[c * LD]
  T* RH = Mb + (size_t)K * LD;              // kPB x CW right-hand sides
  T* Sb = RH + kPB * CW;                    // kPB x KM: S(hi + i, base + a) at Sb[a * KM + i]
  T* Tl = Sb + kPB * KM;                    // kPB x kPB panel factor
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int col = lane % CW, part = lane / CW;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  // columns j0 .. j0 + CW - 1 of M, read along its columns (consecutive rows per thread)
  for (int e = tid; e < K * CW; e += kThreads) {
    const int c = e % K, j = e / K;
    Mb[c * LD + j] = j < nc ? P.M[c + (size_t)(j0 + j) * K] : T(0);
  }
  const int nblk = (K + kPB - 1) / kPB;
  T pre[kPB][SPT], pret = T(0);
  auto fetch = [&](int b) {
    const int base = b * kPB, hi = min(K, base + kPB), len = K - hi;
#pragma unroll
    for (int a = 0; a < kPB; ++a)
#pragma unroll
      for (int q = 0; q < SPT; ++q) {
        const int t = tid + q * kThreads;
        pre[a][q] = (t < len && base + a < hi) ? P.S[(size_t)(base + a) * K + hi + t] : T(0);
      }
    if (tid < kPB * kPB) pret = P.Tp[(size_t)b * kPB * kPB + tid];
  };
  auto stash = [&]() {
#pragma unroll
    for (int a = 0; a < kPB; ++a)
#pragma unroll
      for (int q = 0; q < SPT; ++q) Sb[a * KM + tid + q * kThreads] = pre[a][q];
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
      const T* S0 = Sb + a0 * KM;
      const T* S1 = Sb + a1 * KM;
      T p0 = T(0), p1 = T(0), q0 = T(0), q1 = T(0);
      int i = part;
      for (; i + SPLIT < len; i += 2 * SPLIT) {
        const T m0 = Mb[(hi + i) * LD + col], m1 = Mb[(hi + i + SPLIT) * LD + col];
        p0 = fma(S0[i], m0, p0);
        q0 = fma(S1[i], m0, q0);
        p1 = fma(S0[i + SPLIT], m1, p1);
        q1 = fma(S1[i + SPLIT], m1, q1);
      }
      if (i < len) {
        const T m0 = Mb[(hi + i) * LD + col];
        p0 = fma(S0[i], m0, p0);
        q0 = fma(S1[i], m0, q0);
      }
      T ps = p0 + p1, qs = q0 + q1;
#pragma unroll
      for (int off = CW; off < 64; off <<= 1) {  // the SPLIT partial sums of a column
        ps += __shfl_xor(ps, off, 64);
        qs += __shfl_xor(qs, off, 64);
      }
      if (part == 0) {
        if (base + a0 < hi) RH[a0 * CW + col] = Mb[(base + a0) * LD + col] - ps;
        if (base + a1 < hi) RH[a1 * CW + col] = Mb[(base + a1) * LD + col] - qs;
      }
    }
    __syncthreads();
    if (part == 0) {
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int a = 2 * w + e, c = base + a;
        if (c < hi) {
          T v = T(0);
          for (int cc = a; cc < hi - base; ++cc) v = fma(Tl[a + cc * kPB], RH[cc * CW + col], v);
          Mb[c * LD + col] = v;
        }
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
    P.M[c + (size_t)(j0 + j) * K] = Mb[c * LD + j];
  }
}

template <typename T, int CW>
static size_t wy_smem(int kmax) {
  using Sh = WyShape<CW>;
  return sizeof(T) * ((size_t)kmax * Sh::LD + kPB * CW + kPB * Sh::KMAX + kPB * kPB);
}

template <typename T, int CW>
static void launch_wy_solve_cw(const WySolveProb<T>* probs, int n, int kmax, int rmax,
                               hipStream_t s) {
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_wy_solve<T, CW>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)wy_smem<T, CW>(WyShape<CW>::KMAX));
    attr = true;
  }
  if (n <= 0 || rmax <= 0) return;
  const int ncb = (rmax + CW - 1) / CW;
  const size_t sm = wy_smem<T, CW>(kmax);
  hipLaunchKernelGGL((k_wy_solve<T, CW>), dim3(n * ncb), dim3(kThreads), sm, s, probs, ncb);
}

template <typename T>
void launch_wy_solve(const WySolveProb<T>* probs, int n, int kmax, int rmax, hipStream_t s) {
  if (kmax <= WyShape<64>::KMAX) launch_wy_solve_cw<T, 64>(probs, n, kmax, rmax, s);
  else launch_wy_solve_cw<T, 16>(probs, n, kmax, rmax, s);
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

// Set the large-LDS attributes of every kernel variant up front, so that no
// attribute call happens while a stream is being captured into a graph.
template <typename T>
void prepare_kernels() {
  launch_wy_solve<T>(nullptr, 0, 1, 1, nullptr);
  launch_wy_solve<T>(nullptr, 0, kWyMaxK, 1, nullptr);
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
  launch_core<T>(nullptr, 0, CORE_R64, 1, 0, 1, T(0), nullptr);
}

// ---- explicit instantiations ---------------------------------------------------
#define HCORE_INSTANTIATE(T) \
  template void launch_apply<T>(const ApplyProb<T>*, int, int, hipStream_t); \
  template void launch_wy_solve<T>(const WySolveProb<T>*, int, int, int, hipStream_t); \
  template void prepare_kernels<T>();
HCORE_INSTANTIATE(double)
HCORE_INSTANTIATE(float)
#undef HCORE_INSTANTIATE

}  // namespace hcore
