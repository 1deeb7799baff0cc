// k_core.hip — core SVD kernels (pivoted QR + one-sided Jacobi, truncation).

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <limits>
#include <type_traits>

#include "hcore_internal.hpp"
#include "hcore_wave.hpp"
#include "core_complete.hpp"

namespace hcore {

// ------------------------------------------------------------------------------
// core SVD: one-sided (Hestenes) Jacobi on the columns of T, round-robin
// ordering, then sort, truncation (hcore_dgemm.c:444-464) and U_r = G_r / sigma
// ------------------------------------------------------------------------------

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
      const T app = nrm[a], aqq = nrm[b], pq = app * aqq;
      if (fmin(app, aqq) > thr2 && pq > T(0) && fabs(apq) * fast_rsq(pq) > tol) {
        // rotation parameters through the refined hardware reciprocals
        const T zeta = (aqq - app) * (T(0.5) * fast_rcp(apq));
        T t;
        if (fabs(zeta) > T(1e30)) {
          t = T(0.5) * fast_rcp(zeta);
        } else {
          const T h2 = fma(zeta, zeta, T(1));
          t = (zeta >= T(0) ? T(1) : T(-1)) * fast_rcp(fabs(zeta) + h2 * fast_rsq(h2));
        }
        const T cs = fast_rsq(fma(t, t, T(1)));
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
      const T h = fma(alpha, alpha, xn2);
      const T nrm = h > T(1e-200) ? h * fast_rsq(h) : sqrt(h);  // the seed needs a normal h
      beta = alpha >= T(0) ? -nrm : nrm;
      tau = (beta - alpha) * fast_rcp(beta);
      scal = fast_rcp(alpha - beta);
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
                              int* perm, int* flags, int rk, int maxrk, T acc, T* nrm, T* rdiag,
                              T* red, int* gi) {
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
    if (*P.status == ST_OK) *P.status = r > maxrk ? ST_R_GT_MAXRK : ST_OK;  // keep a preset error
  }
  if (tid == 0) flags[0] = 0;
  block_sync<NT>();
  for (int e = tid; e < r * K; e += NT) {
    const int i = e / K, row = e - i * K;
    const int j = perm[i];
    const T sj = sig[j];
    const bool ok = sj > T(0) && j < k;
    P.Zu[piv[row] + (size_t)i * K] = ok ? W[row + (size_t)j * ld] / sj : T(0);
    if (!ok) flags[0] = 1;
  }
  block_sync<NT>();
  // zero singular values: complete U to an orthonormal set (scratch: the
  // pivoted-QR norms and diagonal, free by now)
  if (flags[0]) complete_basis<T, NT>(P.Zu, K, r, nrm, rdiag, red, gi);
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
                                             int rk, int maxrk, T acc, const int* __restrict__ q,
                                             int qv) {
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
  int item = blockIdx.x;
  if (q) {  // the rank queue's slice of this variant (uniform exit before any barrier)
    if (item >= q[CORE_NVARIANTS + qv]) return;
    item = q[kCoreQueueHead + q[qv] + item];
  }
  const CoreProb<T> P = probs[item];
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
  core_epilogue<T, NT>(P, A, ld, k, piv, sig, perm, flags, rk, maxrk, acc, nrm, rdiag, red, gi);
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

template <typename T>
void launch_core(const CoreProb<T>* probs, int n, int variant, int kmax, int rk, int maxrk, T acc,
                 hipStream_t s, const int* q, int qv) {
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
    launch_core_reg<T>(probs, n, variant, rk, maxrk, acc, s, q, qv);
    return;
  }
  if (n <= 0) return;
  const size_t head = core_head_bytes(sizeof(T), kmax);
  switch (variant) {
#define HCORE_LAUNCH(V, NT, G, RPL, LDS)                                                     \
  case V: {                                                                                  \
    const size_t sm = head + (LDS ? sizeof(T) * (size_t)kmax * core_ld(kmax, G) : 0);        \
    hipLaunchKernelGGL((k_core<T, NT, G, RPL, LDS>), dim3(n), dim3(NT), sm, s, probs, kmax,  \
                       rk, maxrk, acc, q, qv);                                               \
    break;                                                                                   \
  }
    HCORE_CORE_VARIANTS(HCORE_LAUNCH)
#undef HCORE_LAUNCH
    default: break;
  }
}

int core_alt_variant(int v) {
  switch (v) {
    case CORE_R64: return CORE_R64W;
    case CORE_R128: return CORE_R128W;
    case CORE_R192: return CORE_R192W;
    case CORE_R64W: return CORE_R64;
    case CORE_R128W: return CORE_R128;
    case CORE_R192W: return CORE_R192;
    default: return -1;
  }
}

int core_variant_for(int K, int kl, bool reg) { return core_variant_of(K, kl, reg); }

// One workgroup per lane: counting sort of the lane's updates by stacked rank
// K (largest first; equal K in arrival order of the LDS atomics, which does
// not change any result), then each base variant's slice of the order.
template <typename T>
__global__ __launch_bounds__(1024) void k_core_queue(const CoreProb<T>* __restrict__ probs, int n,
                                                     int kl, int reg, int* __restrict__ q) {
  __shared__ int hist[kCoreMaxK + 1];
  __shared__ int start[kCoreMaxK + 1];
  const int tid = threadIdx.x;
  auto kof = [&](int i) { return min(max(probs[i].K, 1), kCoreMaxK); };
  for (int k = tid; k <= kCoreMaxK; k += 1024) hist[k] = 0;
  __syncthreads();
  for (int i = tid; i < n; i += 1024) atomicAdd(&hist[kof(i)], 1);
  __syncthreads();
  if (tid == 0) {  // start[k]: the number of updates with a larger K
    int acc = 0;
    for (int k = kCoreMaxK; k >= 1; --k) {
      start[k] = acc;
      acc += hist[k];
    }
  }
  __syncthreads();
  if (tid < CORE_NVARIANTS) {
    int cnt = 0, st = -1;
    for (int k = kCoreMaxK; k >= 1; --k)
      if (core_variant_of(k, kl, reg != 0) == tid) {
        if (st < 0) st = start[k];
        cnt += hist[k];
      }
    q[tid] = st < 0 ? 0 : st;
    q[CORE_NVARIANTS + tid] = cnt;
  }
  __syncthreads();
  for (int k = tid; k <= kCoreMaxK; k += 1024) hist[k] = 0;  // now fill counters
  __syncthreads();
  for (int i = tid; i < n; i += 1024) {
    const int K = kof(i);
    q[kCoreQueueHead + start[K] + atomicAdd(&hist[K], 1)] = i;
  }
}

template <typename T>
void launch_core_queue(const CoreProb<T>* probs, int n, int kl, bool reg, int* q, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_core_queue<T>, dim3(1), dim3(1024), 0, s, probs, n, kl, reg ? 1 : 0, q);
}

// ---- explicit instantiations ---------------------------------------------------
#define HCORE_INSTANTIATE(T)                                                                  \
  template void launch_core<T>(const CoreProb<T>*, int, int, int, int, int, T, hipStream_t,     \
                               const int*, int);                                              \
  template void launch_core_queue<T>(const CoreProb<T>*, int, int, bool, int*, hipStream_t);
HCORE_INSTANTIATE(double)
HCORE_INSTANTIATE(float)
#undef HCORE_INSTANTIATE

}  // namespace hcore
