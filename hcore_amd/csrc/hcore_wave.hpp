// hcore_wave.hpp — wavefront-level helpers shared by the CDNA4 kernels
// (lane exchanges through DPP and permlane swaps, reductions, barriers).
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>

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

// Value of the partner lane at any OFF = 1..32 (DPP up to 8; for 16 and 32
// the permlane swap of v with itself, whose two results hold the own and the
// partner's value in an order set by the lane's bit OFF).
template <int OFF>
__device__ __forceinline__ unsigned partner_u32(unsigned v, int lane) {
  if constexpr (OFF <= 8) {
    return dpp_partner<OFF>(v);
  } else {
    static_assert(OFF == 16 || OFF == 32, "partner offset");
    const auto r = OFF == 32 ? __builtin_amdgcn_permlane32_swap(v, v, false, false)
                             : __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane & OFF) ? r[0] : r[1];
  }
}
template <int OFF>
__device__ __forceinline__ int partner_any(int v, int lane) {
  return (int)partner_u32<OFF>((unsigned)v, lane);
}
template <int OFF>
__device__ __forceinline__ double partner_any(double v, int lane) {
  const unsigned long long b = __double_as_longlong(v);
  const unsigned lo = partner_u32<OFF>((unsigned)b, lane), hi = partner_u32<OFF>((unsigned)(b >> 32), lane);
  return __longlong_as_double(((unsigned long long)hi << 32) | lo);
}
template <int OFF>
__device__ __forceinline__ float partner_any(float v, int lane) {
  return __uint_as_float(partner_u32<OFF>(__float_as_uint(v), lane));
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

// Reciprocal and reciprocal square root from the hardware approximations
// (v_rcp / v_rsq) refined by Newton steps: two for double (each doubles the
// correct bits of the ~2^-26 seed), one for float.  For positive normal x
// (rcp: any normal x) they are within an ulp or two of 1/x and 1/sqrt(x),
// enough for rotation parameters and reflector scalings, in a fraction of
// the latency of IEEE division and square root.
template <typename T>
__device__ __forceinline__ T fast_rcp(T x);
template <>
__device__ __forceinline__ double fast_rcp<double>(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  return fma(fma(-x, r, 1.0), r, r);
}
template <>
__device__ __forceinline__ float fast_rcp<float>(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return fmaf(fmaf(-x, r, 1.0f), r, r);
}
template <typename T>
__device__ __forceinline__ T fast_rsq(T x);
template <>
__device__ __forceinline__ double fast_rsq<double>(double x) {
  double r = __builtin_amdgcn_rsq(x);
  r = fma(0.5 * r, fma(-x * r, r, 1.0), r);
  return fma(0.5 * r, fma(-x * r, r, 1.0), r);
}
template <>
__device__ __forceinline__ float fast_rsq<float>(float x) {
  const float r = __builtin_amdgcn_rsqf(x);
  return fmaf(0.5f * r, fmaf(-x * r, r, 1.0f), r);
}

// v of lane l (a wave-uniform lane) in scalar registers.
__device__ __forceinline__ double lane_read(double v, int l) {
  const unsigned long long b = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, l);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), l);
  return __longlong_as_double(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ float lane_read(float v, int l) {
  return __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(v), l));
}

template <int NT>
__device__ __forceinline__ void block_sync() {
  if constexpr (NT == 64) wave_sync();
  else __syncthreads();
}

// Trailing columns per update step of the blocked Householder kernels.
template <int NBW>
constexpr int qr_cb() { return NBW >= 32 ? 1 : (NBW >= 16 ? 2 : 4); }

}  // namespace hcore
