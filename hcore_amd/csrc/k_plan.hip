// k_plan.hip — per-execution bookkeeping of a batched plan: reading the
// device-side input ranks, rewriting the rank-dependent descriptor fields,
// and publishing the new ranks (see RankPatch in hcore_internal.hpp).
#include <hip/hip_runtime.h>

#include "hcore_internal.hpp"

namespace hcore {

// One thread per update: the ranks in use this execution (device-side ranks
// checked against the capacities; host-rank plans keep theirs), and the
// update's status reset.
__global__ __launch_bounds__(256) void k_rank_gather(RankTables t) {
  const int u = blockIdx.x * 256 + threadIdx.x;
  if (u < t.n) {
    int st = ST_OK;
    if (t.ra) {
      int r[3] = {*t.ra[u], *t.rb[u], *t.rc[u]};
      for (int q = 0; q < 3; ++q) {
        const int cap = t.cap[3 * u + q];
        if (r[q] < 1 || r[q] > cap) {
          st = ST_BAD_RANK;
          r[q] = r[q] < 1 ? 1 : cap;  // keeps every kernel inside its buffers
        }
        t.rank_in[3 * u + q] = r[q];
      }
    }
    t.status[u] = st;
  } else if (u < t.n + t.ns && t.sk) {
    const int v = u - t.n;
    const int cap = t.cap[3 * t.n + v];
    int k = *t.sk[v];
    k = k < 1 ? 1 : (k > cap ? cap : k);
    t.rank_in[3 * t.n + v] = k;
  }
}

void launch_rank_gather(const RankTables& t, hipStream_t s) {
  const int tot = t.n + t.ns;
  if (tot <= 0) return;
  hipLaunchKernelGGL(k_rank_gather, dim3((tot + 255) / 256), dim3(256), 0, s, t);
}

__global__ __launch_bounds__(256) void k_rank_patch(const RankPatch* __restrict__ patches, int np,
                                                    char* desc, const int* __restrict__ rank_in,
                                                    int n, int M, int elem) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= np) return;
  const RankPatch P = patches[i];
  char* f = desc + P.off;
  if (P.kind == PK_SK) {
    *reinterpret_cast<int*>(f) = rank_in[3 * n + P.upd];
    return;
  }
  const int ra = rank_in[3 * P.upd], rb = rank_in[3 * P.upd + 1], rc = rank_in[3 * P.upd + 2];
  const int K = ra + rc;
  int v = 0;
  switch (P.kind) {
    case PK_RA: v = ra; break;
    case PK_RB: v = rb; break;
    case PK_RC: v = rc; break;
    case PK_K: v = K; break;
    case PK_KK: v = K * K; break;
    case PK_PTR_G:
      *reinterpret_cast<long long*>(f) = P.base + (long long)rc * M * elem;
      return;
    default: return;
  }
  *reinterpret_cast<int*>(f) = v;
}

void launch_rank_patch(const RankPatch* patches, int np, char* desc, const int* rank_in, int n,
                       int M, int elem, hipStream_t s) {
  if (np <= 0) return;
  hipLaunchKernelGGL(k_rank_patch, dim3((np + 255) / 256), dim3(256), 0, s, patches, np, desc,
                     rank_in, n, M, elem);
}

// One thread per update: the new rank of every successful update replaces the
// caller's rank of C.
__global__ __launch_bounds__(256) void k_rank_publish(RankTables t) {
  const int u = blockIdx.x * 256 + threadIdx.x;
  if (u < t.n && t.status[u] == ST_OK) *t.rc[u] = t.rank_out[u];
}

void launch_rank_publish(const RankTables& t, hipStream_t s) {
  if (t.n <= 0 || !t.rc) return;
  hipLaunchKernelGGL(k_rank_publish, dim3((t.n + 255) / 256), dim3(256), 0, s, t);
}

}  // namespace hcore
