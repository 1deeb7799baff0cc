// k_tile.hip — tile generation and compression: the dense tiles of a
// covariance problem (or a caller's dense tiles) written into the padded
// K x K core layout, and the step that turns the core SVD of such a tile into
// its low-rank factors (U = left singular vectors, V = A^T U = right vectors
// scaled by sigma, the form testing_dgemm.c:119-236 builds with dgesvd).
#include <hip/hip_runtime.h>

#include <cmath>

#include "hcore_internal.hpp"

namespace hcore {

// A workgroup per kFillCols columns of one tile; the lanes run down the
// rows, so every store (and every read of a dense source) is a coalesced
// column segment.  The row points are read per entry (they stay in L2 and
// the covariance is computed once per entry); the column point is a
// per-column broadcast.
__global__ __launch_bounds__(256) void k_tile_fill(const TileFillProb* __restrict__ probs) {
  const TileFillProb P = probs[blockIdx.y];
  const int c0 = blockIdx.x * kFillCols;
  if (c0 >= P.cols) return;
  const int ce = min(c0 + kFillCols, P.cols);
  for (int j = c0; j < ce; ++j) {
    double* o = P.out + (size_t)j * P.ldo;
    if (j >= P.n) {
      for (int i = threadIdx.x; i < P.rows; i += 256) o[i] = 0.0;
      continue;
    }
    if (P.src) {
      const double* c = P.src + (size_t)j * P.lds;
      for (int i = threadIdx.x; i < P.rows; i += 256) o[i] = i < P.m ? c[i] : 0.0;
      continue;
    }
    const int gj = P.n0 + j;
    const double xj = P.px[gj], yj = P.py[gj];
    const double rb = 1.0 / P.beta;
    for (int i = threadIdx.x; i < P.rows; i += 256) {
      double v = 0.0;
      if (i < P.m) {
        const int gi = P.m0 + i;
        const double dx = P.px[gi] - xj, dy = P.py[gi] - yj;
        v = exp(-sqrt(dx * dx + dy * dy) * rb);
        if (gi == gj) v += P.nugget;
      }
      o[i] = v;
    }
  }
}

void launch_tile_fill(const TileFillProb* probs, int n, int max_cols, hipStream_t s) {
  if (n <= 0 || max_cols <= 0) return;
  hipLaunchKernelGGL(k_tile_fill, dim3((max_cols + kFillCols - 1) / kFillCols, n), dim3(256), 0,
                     s, probs);
}

__global__ __launch_bounds__(256) void k_compress_finish(const CompressProb* __restrict__ probs) {
  const CompressProb P = probs[blockIdx.y];
  int r = 0;
  if (*P.status == ST_OK && P.sigma[0] >= P.acc) r = min(min(*P.rank, P.mn), P.maxrank);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *P.rank_out = r;
    if (P.rank_out_d) *P.rank_out_d = (double)r;
  }
  const int c0 = blockIdx.x * kFinishCols;
  const int ce = min(c0 + kFinishCols, r);
  for (int c = c0; c < ce; ++c) {
    const double* z = P.Zu + (size_t)c * P.K;
    double* u = P.U + (size_t)c * P.ldu;
    for (int i = threadIdx.x; i < P.m; i += 256) u[i] = z[i];
  }
}

void launch_compress_finish(const CompressProb* probs, int n, int kmax, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_compress_finish, dim3((kmax + kFinishCols - 1) / kFinishCols, n),
                     dim3(256), 0, s, probs);
}

}  // namespace hcore
