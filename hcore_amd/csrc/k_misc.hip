// k_misc.hip — reassembly (compact-WY solve) and conversion kernels.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <limits>
#include <type_traits>

#include "hcore_internal.hpp"
#include "hcore_wave.hpp"

namespace hcore {

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
  launch_core<T>(nullptr, 0, CORE_K16, 1, 0, 1, T(0), nullptr);
  launch_core<T>(nullptr, 0, CORE_R64, 1, 0, 1, T(0), nullptr);
}

// ---- explicit instantiations ---------------------------------------------------
#define HCORE_INSTANTIATE(T) \
  template void launch_wy_solve<T>(const WySolveProb<T>*, int, int, int, hipStream_t); \
  template void prepare_kernels<T>();
HCORE_INSTANTIATE(double)
HCORE_INSTANTIATE(float)
#undef HCORE_INSTANTIATE

}  // namespace hcore
