// k_core_reg.hip — register-resident core SVD kernels (pivoted QR + block
// one-sided Jacobi held in VGPRs) for 33 <= K <= 192.

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
  // divisions and square roots through the refined hardware approximations:
  // this chain is on the critical path of every Jacobi round
  const T zeta = (aqq - app) * (T(0.5) * fast_rcp(apq));
  if (fabs(zeta) > T(1e30)) {
    t = T(0.5) * fast_rcp(zeta);
  } else {
    const T h2 = fma(zeta, zeta, T(1));
    const T h = h2 * fast_rsq(h2);  // sqrt(1 + zeta^2)
    t = (zeta >= T(0) ? T(1) : T(-1)) * fast_rcp(fabs(zeta) + h);
  }
  cs = fast_rsq(fma(t, t, T(1)));
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
  const T pq = app * aqq;
  if (!(fmin(app, aqq) > thr2) || !(pq > T(0))) return false;
  const T rel = fabs(apq) * fast_rsq(pq);  // |apq| / sqrt(app aqq)
  rho = fmaxf(rho, (float)rel);
  if (!(rel > tol)) return false;
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
// (LDS); the rotations reach every lane through scalar registers.  The lanes that hold a pair's reduced
// inner product compute its rotation and keep the largest relative
// off-diagonal |a_pq| / sqrt(a_pp a_qq) they have seen in rho.
template <bool CROSS, typename T, int B, int RPL>
__device__ __forceinline__ void jacobi_round(T (&a)[B][RPL], T (&b)[B][RPL], int t, T* nrm, int lane,
                                             T tol, T thr2, int* rotflag, float& rho) {
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
  T cs_l = T(1), sn_l = T(0);  // pair idx's rotation, in the lane that holds its inner product
  if ((lane & ((64 / NB) - 1)) == 0 && idx < B) {
    int u, v;
    jacobi_pair<CROSS, B>(t, idx, u, v);
    const T app = nrm[u], aqq = nrm[v], apq = d[0];
    T tt;
    if (jacobi_params(app, aqq, apq, tol, thr2, cs_l, sn_l, tt, rho)) {
      nrm[u] = fma(-tt, apq, app);
      nrm[v] = fma(tt, apq, aqq);
      *rotflag = 1;
    }
  }
  // pair q's rotation read from lane q * 64 / NB into scalar registers (no LDS
  // round trip); a pair left alone carries (1, 0), applied like any other (the
  // matrix is finite here, so the identity rotation changes nothing)
  T csq[B], snq[B];
#pragma unroll
  for (int q = 0; q < B; ++q) {
    csq[q] = lane_read(cs_l, q * (64 / NB));
    snq[q] = lane_read(sn_l, q * (64 / NB));
  }
  __builtin_amdgcn_sched_barrier(0);  // the per-lane rotation registers die here
#pragma unroll
  for (int q = 0; q < B; ++q) {
    const T cs = csq[q], sn = snq[q];
    int u, v;
    jacobi_pair<CROSS, B>(t, q, u, v);
#pragma unroll
    for (int i = 0; i < RPL; ++i) {
      const T xu = jreg(a, b, u, i), xv = jreg(a, b, v, i);
      jreg(a, b, u, i) = cs * xu - sn * xv;
      jreg(a, b, v, i) = sn * xu + cs * xv;
    }
  }
  wave_sync();  // the squared norms (LDS) before the next round reads them
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
  // butterfly arg-max through DPP / permlane exchanges (no LDS crossbar)
  auto level = [&](auto off_c) {
    constexpr int OFF = decltype(off_c)::value;
    const T ov = partner_any<OFF>(bv, lane);
    const int oi = partner_any<OFF>(bi, lane);
    if (oi >= 0 && (ov > bv || (ov == bv && (bi < 0 || oi < bi)))) {
      bv = ov;
      bi = oi;
    }
  };
  level(std::integral_constant<int, 32>{});
  level(std::integral_constant<int, 16>{});
  level(std::integral_constant<int, 8>{});
  level(std::integral_constant<int, 4>{});
  level(std::integral_constant<int, 2>{});
  level(std::integral_constant<int, 1>{});
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
                                   2 * W + 2;
  static constexpr size_t n_int = 2 * W + 4 * KM + 8;
  static constexpr size_t smem = n_elem * sizeof(T) + n_int * sizeof(int);
};

template <typename T, int W, int RPL, int CPW, int B>
__global__ __launch_bounds__(64 * W) void k_core_reg(const CoreProb<T>* __restrict__ probs, int rk,
                                                      int maxrk, T acc, const int* __restrict__ q,
                                                      int qv) {
  using C = CoreReg<T, W, RPL, CPW, B>;
  constexpr int NT = C::NT, KM = C::KM;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* big = reinterpret_cast<T*>(smem_raw);  // transposition / block exchange
  T* vbuf = big + C::BIG;                   // KM: reflector
  T* sig = vbuf + KM;                       // KM
  T* xs_all = sig + KM;                     // W x XS
  T* nrm_all = xs_all + W * C::XS;          // W x 2B block norms
  T* nst = nrm_all + W * 2 * B;             // NBLK x B norms in transit
  T* cand_v = nst + C::NBLK * B;            // 2 x W pivot candidates
  T* misc = cand_v + 2 * W;                 // [0] tau
  int* cand_i = reinterpret_cast<int*>(misc + 2);  // 2 x W
  int* piv = cand_i + 2 * W;                       // KM: column at each position
  int* posv = piv + KM;                            // KM: position of each column
  int* perm = posv + KM;                           // KM: sorted order
  int* rnk = perm + KM;                            // KM: sorted position of each column
  int* flags = rnk + KM;  // [0] r, [1] nonfinite, [2..4] rotated, [5..7] max rho (sweep mod 3)

  int item = blockIdx.x;
  if (q) {  // the rank queue's slice of this variant (uniform exit before any barrier)
    if (item >= q[CORE_NVARIANTS + qv]) return;
    item = q[kCoreQueueHead + q[qv] + item];
  }
  const CoreProb<T> P = probs[item];
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
        const T h = fma(alpha, alpha, xn2);
        const T nrm = h > T(1e-200) ? h * fast_rsq(h) : sqrt(h);  // the seed needs a normal h
        beta = alpha >= T(0) ? -nrm : nrm;
        tau = (beta - alpha) * fast_rcp(beta);
        scal = fast_rcp(alpha - beta);
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
        // A column at or below the floor thr2 never rotates (jacobi_params),
        // so a block without a column above it (the slots past the pivoted
        // QR's numerical rank) makes every pair it takes part in a no-op:
        // such rounds are skipped, which changes no result.
        const T nv = lane < 2 * B ? nrm[lane] : T(0);
        const bool live_a = __any(lane < B && nv > thr2);
        const bool live_b = __any(lane >= B && lane < 2 * B && nv > thr2);
        if (o == 0 && (live_a || live_b)) {
#pragma unroll
          for (int t = 0; t < B - 1; ++t)
            jacobi_round<false>(a, b, t, nrm, lane, tol, thr2, rf, rho);
        }
        if (live_a && live_b) {
#pragma unroll
          for (int t = 0; t < B; ++t)
            jacobi_round<true>(a, b, t, nrm, lane, tol, thr2, rf, rho);
        }
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
    if (*P.status == ST_OK) *P.status = r > maxrk ? ST_R_GT_MAXRK : ST_OK;  // keep a preset error
  }
  if (tid == 0) flags[2] = 0;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 2 * B; ++q) {
    const int c = q < B ? cA + q : cB + q - B;
    if (c < K) {
      const int i = rnk[c];
      if (i < r) {
        const T s = sig[c];
        const bool ok = s > T(0) && c < k;
        const T inv = ok ? T(1) / s : T(0);
        if (!ok) flags[2] = 1;
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
  __syncthreads();
  // zero singular values: complete U to an orthonormal set (scratch: the
  // block-exchange area and the reflector buffer, free by now)
  if (flags[2]) complete_basis<T, NT>(P.Zu, K, r, big, vbuf, big + KM, cand_i);
  if (tid == 0 && P.clk) P.clk[3] = (long long)wall_clock64();
}

// register core variants: (waves, rows per lane, QRCP columns per wave, Jacobi block width)
#define HCORE_CORE_REG_VARIANTS(X) \
  X(CORE_R64, 4, 1, 16, 8)         \
  X(CORE_R128, 8, 2, 16, 8)        \
  X(CORE_R192, 8, 3, 24, 12)       \
  X(CORE_R64W, 8, 1, 8, 4)         \
  X(CORE_R128W, 16, 2, 8, 4)       \
  X(CORE_R192W, 16, 3, 12, 6)

template <typename T>
void launch_core_reg(const CoreProb<T>* probs, int n, int variant, int rk, int maxrk, T acc,
                            hipStream_t s, const int* q, int qv) {
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
                       (CoreReg<T, W, RPL, CPW, B>::smem), s, probs, rk, maxrk, acc, q, qv);  \
    break;
    HCORE_CORE_REG_VARIANTS(HCORE_LAUNCH_R)
#undef HCORE_LAUNCH_R
    default: break;
  }
}

// ---- explicit instantiations ---------------------------------------------------
template void launch_core_reg<double>(const CoreProb<double>*, int, int, int, int, double, hipStream_t,
                                      const int*, int);
template void launch_core_reg<float>(const CoreProb<float>*, int, int, int, int, float, hipStream_t,
                                     const int*, int);

}  // namespace hcore
