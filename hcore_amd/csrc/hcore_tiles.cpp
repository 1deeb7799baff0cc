// hcore_tiles.cpp — tile generation and compression on the GPU: the step
// upstream of the recompression, which turns dense tiles into the U V^T
// factors and double-typed ranks that HCORE_dgemm consumes.
//
// - HCORE_dcompress: one dense tile -> U (left singular vectors), V (right
//   vectors scaled by sigma), rank = #{sigma_i >= acc} capped by maxrank:
//   the compression step of the reference's generator (generate_compress,
//   testing/testing_dgemm.c:119-236: dgesvd('A','A'), truncation at :208-212,
//   U at :219, V = (S Vt)^T at :162-166 and :224).
// - HCORE_dgytlr: generate one tile of a covariance problem and compress it,
//   the real-precision counterpart of the declared-only HCORE_zgytlr
//   (src/include/hcore_z.h:63-75): same arguments, diagonal tiles dense into
//   AD unless compress_diag.
// - HCORE_hip_gytlr_batch / HCORE_hip_compress_batch: the same for many tiles
//   of one size in few launches.
//
// Every tile is written into a K x K work matrix (K = max(m, n), zero padded)
// and goes through the core SVD kernels of the recompression (pivoted QR +
// one-sided Jacobi), then V = A^T U is one batched matrix-core GEMM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "hcore_d.h"
#include "hcore_hip.h"
#include "hcore_host.hpp"
#include "hcore_internal.hpp"

namespace hcore {
namespace {

// ---- the covariance problem ---------------------------------------------------
double g_beta = 0.1, g_nugget = 1e-4;  // exp(-d / beta), nugget on the diagonal
std::mutex g_pts_mu;
struct PointSet {
  int device, bigM;
  unsigned long long seed;
  double* d;  // x[bigM], then y[bigM]
};
std::vector<PointSet> g_pts;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// Element k >= 1 of the splitmix64 sequence of `seed`, as a double in [0, 1).
inline double u01(uint64_t seed, uint64_t k) {
  return (double)(mix64(seed + k * 0x9E3779B97F4A7C15ull) >> 11) * 0x1.0p-53;
}
// Bits of v spread to the even positions (Morton interleave).
inline uint32_t spread2(uint32_t v) {
  v &= 0xFFFF;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}
inline uint32_t grid16(double t) {
  const double q = t * 65536.0;
  return q >= 65535.0 ? 65535u : (uint32_t)q;
}

// bigM points, point i = (u_{2i+1}, u_{2i+2}) of the splitmix64 sequence of
// `seed`, sorted by their Morton code on a 2^16 x 2^16 grid (ties by index).
void covariance_points(int bigM, unsigned long long seed, std::vector<double>& xy) {
  std::vector<double> x(bigM), y(bigM);
  std::vector<uint32_t> code(bigM);
  for (int i = 0; i < bigM; ++i) {
    x[i] = u01(seed, 2 * (uint64_t)i + 1);
    y[i] = u01(seed, 2 * (uint64_t)i + 2);
    code[i] = spread2(grid16(x[i])) | (spread2(grid16(y[i])) << 1);
  }
  std::vector<int> ord(bigM);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return code[a] < code[b]; });
  xy.resize(2 * (size_t)bigM);
  for (int i = 0; i < bigM; ++i) {
    xy[i] = x[ord[i]];
    xy[bigM + i] = y[ord[i]];
  }
}

// Device copy of the problem's points, made once per (device, bigM, seed).
const double* device_points(int bigM, unsigned long long seed) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_pts_mu);
  for (const PointSet& p : g_pts)
    if (p.device == dev && p.bigM == bigM && p.seed == seed) return p.d;
  std::vector<double> xy;
  covariance_points(bigM, seed, xy);
  double* d = nullptr;
  if (hipMalloc(&d, sizeof(double) * xy.size()) != hipSuccess) return nullptr;
  if (hipMemcpy(d, xy.data(), sizeof(double) * xy.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  g_pts.push_back({dev, bigM, seed, d});
  return d;
}

// ---- batched generation / compression -----------------------------------------
struct TileJob {
  const double* src;  // dense source (device), or null: covariance tile
  int lds;
  int m0, n0;         // covariance tiles: global row / column offset
  double* U;          // compressed output (null for dense-only jobs)
  int ldu;
  double* V;
  int ldv;
  int* rank;          // device int (null: the workspace's own word)
  double* rank_d;     // device double (may be null)
  double* dense;      // dense output (the only output of a dense-only job)
  int ldd;
};

constexpr size_t kTileBudget = size_t(1) << 30;  // bytes of per-tile work matrices per chunk

size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

// Run the jobs (all m x n) on stream s, then wait for them.  *bad receives the
// index of the first job whose tile held NaN/Inf (or -1).
int run_tile_jobs(int m, int n, const std::vector<TileJob>& jobs, const double* pts, int bigM,
                  int maxrank, double acc, const ThreadRes& tr, hipStream_t s, int* bad) {
  *bad = -1;
  const int K = std::max(m, n), mn = std::min(m, n);
  const int variant = core_variant_ctx(K, sizeof(double));
  const bool scratch = core_needs_scratch(variant);
  std::vector<int> comp, dense;
  for (int i = 0; i < (int)jobs.size(); ++i) {
    if (jobs[i].U) comp.push_back(i);
    if (jobs[i].dense) dense.push_back(i);
  }
  const int nc = (int)comp.size();
  // workspace: per-job words, then the chunk's work matrices
  const size_t slot = al256(sizeof(double) * ((size_t)2 * K * K + (scratch ? (size_t)K * (K + 64) : 0) +
                                             (size_t)K));
  const int chunk = std::max(1, std::min(std::max(nc, 1), (int)(kTileBudget / slot)));
  const size_t o_words = 0;                                 // rank_core, status, rank_own per job
  const size_t o_slots = al256(sizeof(int) * 3 * (size_t)std::max(nc, 1));
  const size_t wbytes = o_slots + slot * (size_t)chunk;
  if (tr.work->reserve(wbytes) != HCORE_HIP_OK) return HCORE_HIP_ERR_RUNTIME;
  char* W = tr.work->p;
  int* w_rank = reinterpret_cast<int*>(W + o_words);
  int* w_status = w_rank + nc;
  int* w_own = w_status + nc;
  auto slot_T = [&](int c) { return reinterpret_cast<double*>(W + o_slots + slot * c); };
  auto slot_Zu = [&](int c) { return slot_T(c) + (size_t)K * K; };
  auto slot_Tj = [&](int c) { return scratch ? slot_Zu(c) + (size_t)K * K : nullptr; };
  auto slot_sg = [&](int c) {
    return slot_Zu(c) + (size_t)K * K + (scratch ? (size_t)K * (K + 64) : 0);
  };

  Packer pk;
  // dense outputs (dense-only jobs and the optional dense copy of compressed ones)
  std::vector<TileFillProb> fd;
  for (int i : dense) {
    const TileJob& j = jobs[i];
    fd.push_back({j.src, pts, pts ? pts + bigM : nullptr, j.dense, j.lds, j.ldd, j.m0, j.n0, m, n,
                  m, n, g_beta, g_nugget});
  }
  const size_t o_fd = pk.add(fd);
  struct ChunkRef {
    size_t fill, core, fin, gp;
    TaskRef tasks;
    int count;
  };
  std::vector<ChunkRef> chunks;
  for (int c0 = 0; c0 < nc; c0 += chunk) {
    const int cnt = std::min(chunk, nc - c0);
    std::vector<TileFillProb> f;
    std::vector<CoreProb<double>> cp;
    std::vector<CompressProb> fp;
    std::vector<GemmProb<double>> gp;
    TaskList tl;
    for (int q = 0; q < cnt; ++q) {
      const int u = c0 + q;
      const TileJob& j = jobs[comp[u]];
      double *T = slot_T(q), *Zu = slot_Zu(q), *sg = slot_sg(q);
      f.push_back({j.src, pts, pts ? pts + bigM : nullptr, T, j.lds, K, j.m0, j.n0, m, n, K, K,
                   g_beta, g_nugget});
      cp.push_back({T, slot_Tj(q), Zu, sg, w_rank + u, w_status + u, nullptr, nullptr, K});
      int* ro = j.rank ? j.rank : w_own + u;
      fp.push_back({Zu, sg, w_rank + u, w_status + u, ro, j.rank_d, j.U, j.ldu, K, m, mn, maxrank,
                    acc});
      // V (n x r) = A^T U: the padded work matrix transposed times Zu
      GemmProb<double> g{};
      g.A = T; g.B = Zu; g.C = j.V;
      g.m = n; g.n = K; g.k = K;
      g.lda = K; g.ldb = K; g.ldc = j.ldv;
      g.alpha = 1.0; g.beta = 0.0;
      g.flags = GF_A_TRANS | GF_BETA_ZERO;
      g.n_dev = ro;
      gp.push_back(g);
      add_tasks(tl, q, gp.back());
    }
    ChunkRef r;
    r.fill = pk.add(f);
    r.core = pk.add(cp);
    r.fin = pk.add(fp);
    r.gp = pk.add(gp);
    r.tasks = pk.add(tl);
    r.count = cnt;
    chunks.push_back(r);
  }
  if (tr.desc->reserve(pk.buf.size() + 256) != HCORE_HIP_OK) return HCORE_HIP_ERR_RUNTIME;
  char* D = tr.desc->p;
  tr.hdesc->swap(pk.buf);
  HCHECK(hipMemcpyAsync(D, tr.hdesc->data(), tr.hdesc->size(), hipMemcpyHostToDevice, s));
  if (nc > 0) HCHECK(hipMemsetAsync(w_rank, 0, sizeof(int) * 3 * (size_t)nc, s));
  launch_tile_fill(reinterpret_cast<const TileFillProb*>(D + o_fd), (int)fd.size(), n, s);
  for (const ChunkRef& c : chunks) {
    launch_tile_fill(reinterpret_cast<const TileFillProb*>(D + c.fill), c.count, K, s);
    launch_core<double>(reinterpret_cast<const CoreProb<double>*>(D + c.core), c.count, variant, K,
                        0, K, acc, s);
    launch_compress_finish(reinterpret_cast<const CompressProb*>(D + c.fin), c.count, K, s);
    launch_tasks<double>(reinterpret_cast<const GemmProb<double>*>(D + c.gp), D, c.tasks, s, true);
  }
  HCHECK(hipGetLastError());
  std::vector<int> st(nc);
  if (nc > 0)
    HCHECK(hipMemcpyAsync(st.data(), w_status, sizeof(int) * nc, hipMemcpyDeviceToHost, s));
  HCHECK(hipStreamSynchronize(s));
  for (int u = 0; u < nc; ++u)
    if (st[u] != ST_OK) {
      *bad = comp[u];
      return HCORE_HIP_ERR_NUMERIC;
    }
  return HCORE_HIP_OK;
}

int check_size(int m, int n) {
  return (m >= 1 && n >= 1 && std::max(m, n) <= kCoreMaxK) ? HCORE_HIP_OK : HCORE_HIP_ERR_ARG;
}

// One tile through run_tile_jobs with host or device operands: host
// operands are staged through the thread's pinned and device buffers.
// src: dense source (HCORE_dcompress) or null (covariance tile m0, n0).
// dense_only: write the dense tile to `dense` and nothing else.
void single_tile(const char* name, int m, int n, const double* src, int lds, int m0, int n0,
                 int bigM, unsigned long long seed, double* U, int ldu, double* V, int ldv,
                 double* rk, double* dense, int ldd, bool dense_only, int maxrank, double acc) {
  set_last_status(HCORE_HIP_OK);
  char msg[256];
  if (check_size(m, n) != HCORE_HIP_OK || (src && lds < m) || (!dense_only && (ldu < m || ldv < n)) ||
      (dense && ldd < m) || maxrank < 0) {
    snprintf(msg, sizeof msg,
             "%s: invalid arguments (m=%d n=%d, max(m, n) must be <= %d, leading dimensions >= "
             "the rows, maxrank >= 0)",
             name, m, n, kCoreMaxK);
    return report_failure(HCORE_HIP_ERR_ARG, msg);
  }
  if (!src && (bigM < 1 || m0 < 0 || n0 < 0 || m0 + m > bigM || n0 + n > bigM)) {
    snprintf(msg, sizeof msg, "%s: tile (%d, %d) of size %d x %d outside the %d-point problem",
             name, m0, n0, m, n, bigM);
    return report_failure(HCORE_HIP_ERR_ARG, msg);
  }
  if (ensure_ready() != HCORE_HIP_OK)
    return report_failure(HCORE_HIP_ERR_RUNTIME, "hcore_amd: GPU initialisation failed");
  ThreadRes tr;
  if (!thread_res(&tr)) return report_failure(HCORE_HIP_ERR_RUNTIME, "hcore_amd: per-thread GPU state failed");
  const double* pts = nullptr;
  if (!src && !(pts = device_points(bigM, seed)))
    return report_failure(HCORE_HIP_ERR_RUNTIME, "hcore_amd: covariance points failed");
  hipStream_t s = tr.stream;
  const int mn = std::min(m, n);
  const bool dev = (!src || device_pointer(src)) && (dense_only || (device_pointer(U) && device_pointer(V))) &&
                   (!dense || device_pointer(dense));
  // device images: source, U, V, dense, and the rank word
  const size_t nS = src ? (size_t)m * n : 0, nU = dense_only ? 0 : (size_t)m * mn,
               nV = dense_only ? 0 : (size_t)n * mn, nD = dense ? (size_t)m * n : 0;
  const size_t bytes = sizeof(double) * (nS + nU + nV + nD) + 256;
  if (tr.stage->reserve(bytes) != HCORE_HIP_OK || tr.pinned->reserve(bytes) != HCORE_HIP_OK)
    return report_failure(HCORE_HIP_ERR_RUNTIME, "hcore_amd: staging allocation failed");
  double* q = reinterpret_cast<double*>(tr.stage->p);
  double* h = reinterpret_cast<double*>(tr.pinned->p);
  int* d_rank = reinterpret_cast<int*>(q + nS + nU + nV + nD);
  TileJob j{};
  j.src = src;
  j.lds = lds;
  j.m0 = m0;
  j.n0 = n0;
  j.rank = d_rank;
  if (dev) {
    j.U = dense_only ? nullptr : U; j.ldu = ldu;
    j.V = V; j.ldv = ldv;
    j.dense = dense; j.ldd = ldd;
  } else {
    if (src) {
      for (int c = 0; c < n; ++c) memcpy(h + (size_t)c * m, src + (size_t)c * lds, sizeof(double) * m);
      if (hipMemcpyAsync(q, h, sizeof(double) * nS, hipMemcpyHostToDevice, s) != hipSuccess)
        return report_failure(HCORE_HIP_ERR_RUNTIME, "hcore_amd: host-to-device staging failed");
      j.src = q;
      j.lds = m;
    }
    j.U = dense_only ? nullptr : q + nS; j.ldu = m;
    j.V = q + nS + nU; j.ldv = n;
    j.dense = dense ? q + nS + nU + nV : nullptr; j.ldd = m;
  }
  std::vector<TileJob> jobs{j};
  int bad = -1;
  int st = run_tile_jobs(m, n, jobs, pts, bigM, maxrank, acc, tr, s, &bad);
  if (st == HCORE_HIP_ERR_NUMERIC) {
    snprintf(msg, sizeof msg, "%s: ERROR in the SVD of the tile (non-finite values)", name);
    return report_failure(st, msg);
  }
  if (st != HCORE_HIP_OK) return report_failure(st, "hcore_amd: GPU execution failed");
  int r = 0;
  if (!dense_only && hipMemcpy(&r, d_rank, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return report_failure(HCORE_HIP_ERR_RUNTIME, "hcore_amd: cannot read the rank");
  if (!dev) {
    // U, V (r columns) and the dense tile back through the pinned image
    const size_t back = nU + nV + nD;
    if (back && (hipMemcpy(h, q + nS, sizeof(double) * back, hipMemcpyDeviceToHost) != hipSuccess))
      return report_failure(HCORE_HIP_ERR_RUNTIME, "hcore_amd: device-to-host copy failed");
    for (int c = 0; c < r; ++c) {
      memcpy(U + (size_t)c * ldu, h + (size_t)c * m, sizeof(double) * m);
      memcpy(V + (size_t)c * ldv, h + nU + (size_t)c * n, sizeof(double) * n);
    }
    if (dense)
      for (int c = 0; c < n; ++c)
        memcpy(dense + (size_t)c * ldd, h + nU + nV + (size_t)c * m, sizeof(double) * m);
  }
  if (!dense_only && rk) {
    if (device_pointer(rk)) {
      const double rv = r;
      if (hipMemcpy(rk, &rv, sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return report_failure(HCORE_HIP_ERR_RUNTIME, "hcore_amd: cannot write the rank");
    } else {
      rk[0] = r;
    }
  }
}

// Batched entry: every pointer is device memory.
int batch_tiles(int n, int m, int nn, const double* const* src, int lds, const int* m0,
                const int* n0, int bigM, unsigned long long seed, double* const* AU, int ldu,
                double* const* AV, int ldv, int* rank, double* const* AD, int lda, int maxrank,
                double acc, void* stream) {
  set_last_status(HCORE_HIP_OK);
  if (n < 0 || check_size(m, nn) != HCORE_HIP_OK || maxrank < 0) return HCORE_HIP_ERR_ARG;
  if (n == 0) return HCORE_HIP_OK;
  if (src && lds < m) return HCORE_HIP_ERR_ARG;
  if (!src && (!m0 || !n0 || bigM < 1)) return HCORE_HIP_ERR_ARG;
  std::vector<TileJob> jobs(n);
  for (int i = 0; i < n; ++i) {
    TileJob& j = jobs[i];
    j = TileJob{};
    const bool dense_only = AD && AD[i];
    if (src) {
      if (!src[i]) return HCORE_HIP_ERR_ARG;
      j.src = src[i];
      j.lds = lds;
    } else {
      if (m0[i] < 0 || n0[i] < 0 || m0[i] + m > bigM || n0[i] + nn > bigM) return HCORE_HIP_ERR_ARG;
      j.m0 = m0[i];
      j.n0 = n0[i];
    }
    if (dense_only) {
      if (lda < m) return HCORE_HIP_ERR_ARG;
      j.dense = AD[i];
      j.ldd = lda;
    } else {
      if (!AU || !AV || !AU[i] || !AV[i] || !rank || ldu < m || ldv < nn) return HCORE_HIP_ERR_ARG;
      j.U = AU[i]; j.ldu = ldu;
      j.V = AV[i]; j.ldv = ldv;
      j.rank = rank + i;
    }
  }
  int st = ensure_ready();
  if (st != HCORE_HIP_OK) return st;
  ThreadRes tr;
  if (!thread_res(&tr)) return HCORE_HIP_ERR_RUNTIME;
  const double* pts = nullptr;
  if (!src && !(pts = device_points(bigM, seed))) return HCORE_HIP_ERR_RUNTIME;
  hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : tr.stream;
  int bad = -1;
  st = run_tile_jobs(m, nn, jobs, pts, bigM, maxrank, acc, tr, s, &bad);
  set_last_status(st);
  return st;
}

}  // namespace

void release_tile_state() {
  std::lock_guard<std::mutex> lk(g_pts_mu);
  for (const PointSet& p : g_pts) (void)hipFree(p.d);
  g_pts.clear();
}

}  // namespace hcore

using namespace hcore;

extern "C" {

void HCORE_dcompress(int m, int n, const double* A, int lda, double* U, int ldu, double* V,
                     int ldv, double* Ark, int maxrank, double acc) {
  if (!A || !U || !V) return report_failure(HCORE_HIP_ERR_ARG, "HCORE_dcompress: null operand");
  single_tile("HCORE_dcompress", m, n, A, lda, 0, 0, 0, 0, U, ldu, V, ldv, Ark, nullptr, 0, false,
              maxrank, acc);
}

void HCORE_dgytlr(int m, int n, double* AU, double* AV, double* AD, double* Ark, int lda, int ldu,
                  int ldv, int bigM, int m0, int n0, unsigned long long int seed, int maxrank,
                  double tol, int compress_diag, double* Dense) {
  const bool dense_only = m0 == n0 && !compress_diag;
  if (dense_only) {
    if (!AD) return report_failure(HCORE_HIP_ERR_ARG, "HCORE_dgytlr: diagonal tile without AD");
    single_tile("HCORE_dgytlr", m, n, nullptr, 0, m0, n0, bigM, seed, nullptr, 0, nullptr, 0,
                nullptr, AD, lda, true, maxrank, tol);
    if (Dense && HCORE_hip_last_status() == HCORE_HIP_OK)
      single_tile("HCORE_dgytlr", m, n, nullptr, 0, m0, n0, bigM, seed, nullptr, 0, nullptr, 0,
                  nullptr, Dense, lda, true, maxrank, tol);
    return;
  }
  if (!AU || !AV) return report_failure(HCORE_HIP_ERR_ARG, "HCORE_dgytlr: null AU or AV");
  single_tile("HCORE_dgytlr", m, n, nullptr, 0, m0, n0, bigM, seed, AU, ldu, AV, ldv, Ark, Dense,
              lda, false, maxrank, tol);
}

int HCORE_hip_gytlr_batch(int n, int m, int nn, const int* m0, const int* n0, int bigM,
                          unsigned long long seed, double* const* AU, int ldu, double* const* AV,
                          int ldv, int* rank, double* const* AD, int lda, int maxrank, double tol,
                          void* stream) {
  return batch_tiles(n, m, nn, nullptr, 0, m0, n0, bigM, seed, AU, ldu, AV, ldv, rank, AD, lda,
                     maxrank, tol, stream);
}

int HCORE_hip_compress_batch(int n, int m, int nn, const double* const* A, int lda,
                             double* const* AU, int ldu, double* const* AV, int ldv, int* rank,
                             int maxrank, double tol, void* stream) {
  if (!A) return HCORE_HIP_ERR_ARG;
  return batch_tiles(n, m, nn, A, lda, nullptr, nullptr, 0, 0, AU, ldu, AV, ldv, rank, nullptr, 0,
                     maxrank, tol, stream);
}

int HCORE_hip_set_covariance(double beta, double nugget) {
  if (!(beta > 0) || !(nugget >= 0)) return HCORE_HIP_ERR_ARG;
  g_beta = beta;
  g_nugget = nugget;
  return HCORE_HIP_OK;
}

int HCORE_hip_covariance_points(int bigM, unsigned long long seed, double* xy) {
  if (bigM < 1 || !xy) return HCORE_HIP_ERR_ARG;
  std::vector<double> v;
  covariance_points(bigM, seed, v);
  memcpy(xy, v.data(), sizeof(double) * v.size());
  return HCORE_HIP_OK;
}

}  // extern "C"
