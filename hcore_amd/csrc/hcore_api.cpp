// hcore_api.cpp — host side of the MI355X-native HCORE: library context,
// device workspace pool, batched plans and the drop-in C entry points
// HCORE_dgemm / HCORE_sgemm / HCORE_dsyrk / HCORE_dgemmcd.
//
// The drop-in entry points keep the reference's contracts (argument meaning,
// in-place outputs, double-typed ranks, exit(-1) on fatal conditions) and run
// every arithmetic step on the GPU through the same batched plan machinery
// that a TLR Cholesky trailing step uses (one update = a batch of one).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "flop_counts.h"
#include "hcore_d.h"
#include "hcore_hip.h"
#include "hcore_host.hpp"
#include "hcore_internal.hpp"
#include "hcore_s.h"

// ---- data symbols kept from the reference (hcore_dgemm.c:37-43 etc.) --------
extern "C" {
__attribute__((weak)) int use_scratch = 0;
int use_trmm = 0;
int gemm_print_index = 0;
int gemm_print_mat = 0;
int hc_nelm_limit = 512;
int sgemm_use_trmm = 0;
int sgemm_print_index = 0;
int sgemm_print_mat = 0;
int syrk_print_index = 0;
int gemmcd_print_index = 0;
int gemmcd_print_mat = 0;
int gemmcd_use_scratch = 0;
int gemmbdcd_print_index = 0;
int gemmbdcd_print_mat = 0;
int gemmfrk_print_index = 0;
int gemmfrk_print_mat = 0;

// Closed-form operation counts (reference src/misc/descutil/flop_counts.c:3-70),
// including its float-precision thirds in the 'q' formula.
unsigned long int flop_counts(char op, unsigned long int a, unsigned long int b,
                              unsigned long int c, unsigned long int d) {
  unsigned long int r = 0;
  switch (op) {
    case 'q':
      r = 2 * a * b * b - (unsigned long int)(2 * b * b * b / 3.0f) + 2 * a * b +
          (unsigned long int)(17 * b / 3.0f);
      break;
    case 'c': r = (unsigned long int)(a * a * a / 3 - a * a / 2.0 + a / 6); break;
    case 't': r = (c == 1) ? b * a * a : (c == 2 ? a * b * b : 0); break;
    case 'm': r = a * b * c * 2; break;
    case 's': r = 22 * a * a * a; break;
    case 'o':
      if (d == 1) r = 4 * b * a * c - 2 * b * c * c + 3 * b * c;
      else if (d == 2) r = 4 * b * a * c - 2 * a * c * c + 2 * a * c + b * c - c * c / 2 + c / 2;
      break;
    case 'r': r = (c == 1) ? b * a * a : (c == 2 ? a * b * b : 0); break;
    default: break;
  }
  return r;
}
}  // extern "C"

namespace hcore {

// ------------------------------------------------------------------------------
// context
// ------------------------------------------------------------------------------
struct Context {
  std::mutex mu;
  bool ready = false;
  int device = 0;
  int policy = HCORE_HIP_ON_ERROR_EXIT;
  int core_kl_d = 0, core_kl_f = 0;
  bool core_reg = true;  // register-resident core SVD for 33 <= K <= 192
  bool qr_reg = true;    // register-resident Householder QR for K <= kQrRegMaxK
  int generation = 0;    // bumped by HCORE_hip_finalize: per-thread state is stale
};
static Context g_ctx;
static thread_local int t_last_status = HCORE_HIP_OK;

// HCORE_ON_ERROR=return selects the recording policy for the whole process.
__attribute__((constructor)) static void policy_from_env() {
  const char* pol = getenv("HCORE_ON_ERROR");
  if (pol && strcmp(pol, "return") == 0) g_ctx.policy = HCORE_HIP_ON_ERROR_RETURN;
  // HCORE_CORE_REG=0 selects the LDS / global-memory core SVD kernels for all K
  const char* reg = getenv("HCORE_CORE_REG");
  if (reg && strcmp(reg, "0") == 0) g_ctx.core_reg = false;
  // HCORE_QR_REG=0 selects the blocked Householder QR for all stacks
  const char* qreg = getenv("HCORE_QR_REG");
  if (qreg && strcmp(qreg, "0") == 0) g_ctx.qr_reg = false;
}


static int ctx_init_locked(int device) {
  if (g_ctx.ready) return HCORE_HIP_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    fprintf(stderr, "hcore_amd: no HIP device available\n");
    return HCORE_HIP_ERR_RUNTIME;
  }
  if (device < 0) HCHECK(hipGetDevice(&device));
  HCHECK(hipSetDevice(device));
  g_ctx.device = device;
  g_ctx.core_kl_d = core_lds_max_k(sizeof(double));
  g_ctx.core_kl_f = core_lds_max_k(sizeof(float));
  prepare_kernels<double>();
  prepare_kernels<float>();
  g_ctx.ready = true;
  return HCORE_HIP_OK;
}

static int ctx_ensure() {
  if (g_ctx.ready) return HCORE_HIP_OK;
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  return ctx_init_locked(-1);
}


// Per-thread state of the single-tile entry points (HCORE_dgemm, HCORE_sgemm,
// HCORE_dsyrk, HCORE_dgemmcd): a stream of its own, and grow-only device
// buffers for the one-update plan (workspace and descriptors), the staging
// of host operands and the GEMM chains, plus pinned host words for the rank
// and status read-back.  The runtime above HCORE calls it from many worker
// threads at once, one tile each; with per-thread state those calls share
// no lock and overlap on the GPU, and a call allocates nothing once its
// thread's buffers have grown to the tile size.
struct ThreadCtx {
  hipStream_t stream = nullptr;
  DevArena plan, stage, chain;
  DevArena tile_work, tile_desc;  // tile generation and compression (hcore_tiles.cpp)
  PinnedArena hstage;      // pinned image of the staged operands and results
  std::vector<char> desc;  // host image of the last descriptors (alive until the stream syncs)
  int* hres = nullptr;     // pinned: [0] rank, [1] status
  int generation = -1;
};
static std::mutex g_tctx_mu;
static std::vector<ThreadCtx*> g_tctx;  // every thread's state, released by HCORE_hip_finalize
static thread_local ThreadCtx* t_ctx = nullptr;

static ThreadCtx* thread_ctx() {
  if (t_ctx && t_ctx->generation == g_ctx.generation) return t_ctx;
  ThreadCtx* c = new ThreadCtx();
  c->generation = g_ctx.generation;
  if (hipSetDevice(g_ctx.device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&c->hres), 4 * sizeof(int), hipHostMallocDefault) !=
          hipSuccess) {
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return nullptr;
  }
  {
    std::lock_guard<std::mutex> lk(g_tctx_mu);
    g_tctx.push_back(c);
  }
  t_ctx = c;
  return c;
}

static void release_thread_ctxs() {
  std::lock_guard<std::mutex> lk(g_tctx_mu);
  for (ThreadCtx* c : g_tctx) {
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->plan.release();
    c->stage.release();
    c->chain.release();
    c->tile_work.release();
    c->tile_desc.release();
    c->hstage.release();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->hres) (void)hipHostFree(c->hres);
    delete c;
  }
  g_tctx.clear();
}

template <typename T>
static int core_kl() { return sizeof(T) == 8 ? g_ctx.core_kl_d : g_ctx.core_kl_f; }

// ------------------------------------------------------------------------------
// autotuning cache: for a tile size M, the core's precision and the core
// variant its stacked rank selects, the variant HCORE_hip_plan_autotune found
// fastest (the same algorithm with the other wavefront count).  Plans
// finalized later start from it.  HCORE_TUNE_CACHE names a text file that is
// read at start-up and rewritten after each autotuning ("M fp64 base choice"
// per line).
// ------------------------------------------------------------------------------
struct TuneEntry {
  int M, fp64, base, choice;
};
static std::mutex g_tune_mu;
static std::vector<TuneEntry> g_tune;
static bool g_tune_loaded = false;

static void tune_load_locked() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const char* path = getenv("HCORE_TUNE_CACHE");
  if (!path) return;
  FILE* f = fopen(path, "r");
  if (!f) return;
  TuneEntry e;
  while (fscanf(f, "%d %d %d %d", &e.M, &e.fp64, &e.base, &e.choice) == 4)
    if (e.base >= 0 && e.base < CORE_NVARIANTS && e.choice >= 0 && e.choice < CORE_NVARIANTS &&
        (e.choice == e.base || e.choice == core_alt_variant(e.base)))
      g_tune.push_back(e);
  fclose(f);
}

static int tuned_variant(int M, bool fp64, int base) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_load_locked();
  for (const auto& e : g_tune)
    if (e.M == M && e.fp64 == (int)fp64 && e.base == base) return e.choice;
  return base;
}

static void tune_store(int M, bool fp64, int base, int choice) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_load_locked();
  bool found = false;
  for (auto& e : g_tune)
    if (e.M == M && e.fp64 == (int)fp64 && e.base == base) {
      e.choice = choice;
      found = true;
    }
  if (!found) g_tune.push_back({M, (int)fp64, base, choice});
  const char* path = getenv("HCORE_TUNE_CACHE");
  if (!path) return;
  FILE* f = fopen(path, "w");
  if (!f) return;
  for (const auto& e : g_tune) fprintf(f, "%d %d %d %d\n", e.M, e.fp64, e.base, e.choice);
  fclose(f);
}


}  // namespace hcore

// ------------------------------------------------------------------------------
// batched plan
// ------------------------------------------------------------------------------
struct hcore_hip_plan {
  int M = 0;
  int prec = HCORE_HIP_FP64;
  int slices_set = 0;  // split precision: digits per operand requested (0: from acc)
  bool finalized = false;
  // single-call plans: device memory from a per-thread arena (or a fixed
  // buffer), set up asynchronously on `fin_stream`, no events or lane streams
  hcore::DevArena* arena = nullptr;
  char* fixed = nullptr;          // fixed device buffer instead of an arena
  size_t fixed_cap = 0;
  hipStream_t fin_stream = nullptr;
  std::vector<char>* desc_host = nullptr;  // keeps the uploaded descriptors alive
  bool measure = false;  // finalize only computes the device bytes (dbytes)
  // element offsets in ws of update 0's Us, Vs, tau_A, tau_B, Zu, Zv
  size_t first_off[6] = {0, 0, 0, 0, 0, 0};
  // dense-B updates (HCORE_dgemm_qr_svd_b_dense): G = BD AV replaces
  // G = BU (AV^T BV)^T, and F is not formed
  std::vector<void*> BD;
  int ldbd = 0;
  // gemm group
  int ng = 0;
  std::vector<void*> AU, AV, BU, BV, CU, CV, CUo, CVo;
  std::vector<int> ra, rb, rc;
  int lda = 0, ldb = 0, ldc = 0, ldco = 0;
  double alpha = 1, beta = 1, acc = 0;
  int rk = 0, maxrk = 0;
  // syrk group
  int ns = 0;
  std::vector<void*> sAU, sAV, sCD;
  std::vector<int> sK;
  // device-side ranks (HCORE_hip_plan_add_dgemm_dev / _dsyrk_dev): the host
  // rank vectors above hold the capacities the workspace is carved for
  bool dev_ranks = false, sdev_ranks = false;
  std::vector<const int*> ra_dev, rb_dev, sk_dev;
  std::vector<int*> rc_dev;
  size_t o_patch = 0, o_tab = 0, o_cap = 0;
  int n_patch = 0;
  int* d_rank_in = nullptr;  // 3 per gemm update + 1 per dsyrk: ranks in use
  int ldau = 0, ldcd = 0;
  double salpha = -1;
  // device memory
  char* dmem = nullptr;  // owned by the plan unless single-call
  size_t dbytes = 0;
  size_t ws_bytes = 0;
  char* desc = nullptr;
  int* d_rank = nullptr;
  int* d_status = nullptr;
  std::vector<size_t> sigma_off;  // per update, element offset into ws
  bool sigma_double = true;       // sigma stored in FP64
  char* ws = nullptr;
  // the batch is split into lanes by stacked rank; each lane runs its whole
  // pipeline on its own stream so the long large-K chains overlap the rest
  struct CoreLaunch {
    size_t off;
    int count, variant, kmax;
    bool dbl;          // FP64 core of a mixed-precision plan
    int base_variant;  // the variant chosen by K (the autotuner may swap in its alternative)
    bool queue = false;  // served from the lane's device-side rank queue (count: an upper bound)
  };
  struct Lane {
    int n = 0;
    size_t o_stack = 0, o_qr = 0;
    int n_qr = 0;
    hcore::TaskRef tF, tG, tT, tZ, tB, tX;  // batched GEMM phases
    int stack_cols = 0;  // largest 2rc+ra of the lane
    int kmax = 0;        // largest stacked rank of the lane
    size_t o_wy = 0, o_qrr16 = 0, o_qrr = 0, o_qrp = 0;
    int n_qrr16 = 0, n_qrr = 0, n_qrp = 0;  // register QR for K <= 16, 17..32; panels
    bool gram = false;                      // T phase with the stacks' Gram products
    size_t o_cvin = 0, o_cvout = 0;  // mixed: FP32 <-> FP64 conversions around the core
    int n_cv = 0, cv_max = 0;
    int n_wy = 0;
    int slices = 0;  // split precision: digits per operand of the sliced products
    // device-side ranks: the lane's core SVDs routed by their stacked rank at
    // execution (queue ints at o_q, every update's descriptor at o_qprobs)
    bool queue = false, qdbl = false, qreg = false;
    size_t o_q = 0, o_qprobs = 0;
    int qkl = 0;
    // device-side ranks: the stacks' QR routed by K at execution (queue at
    // o_qrq, both stacks of every update at o_qrall)
    bool qrq = false;
    size_t o_qrq = 0, o_qrall = 0;
    int n_qrall = 0;
    std::vector<CoreLaunch> core;  // one launch per K bucket
    hipStream_t stream = nullptr;  // lanes > 0 (lane 0 uses the caller's stream)
    hipEvent_t ev_done = nullptr;
    // lanes with more than one core launch by K bucket (host ranks): in a
    // captured graph the launches after the first run on a side stream (made
    // at the first capture), forked after the lane's previous phase and joined
    // before its next (they serve disjoint updates).  Measured on the
    // nb = 1024 step: graph replay -1.5 %; executed eagerly the fifth stream
    // shares a hardware queue with another lane (+9 %), and forking the
    // rank-queue shapes of device-rank plans doubled their replay time, so
    // neither forks.
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  };
  std::vector<Lane> lanes;
  std::vector<int> lane_of;  // lane index of each update (HCORE_hip_plan_get_lanes)
  std::vector<std::vector<int>> lane_updates;  // each lane's updates in its local order
  size_t o_gp = 0, o_sp = 0;
  hcore::TaskRef tW, tXs, tC;  // dsyrk GEMM phases
  hipEvent_t ev_start = nullptr;
  int* d_sweeps = nullptr;
  long long* d_clk = nullptr;  // kCoreClk core stamps per update
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // last ranks (host) for flop accounting
  std::vector<int> last_ranks;
  // per-phase timing: a ring of event sets so timed executions do not wait
  // for each other; a set is read back when the ring wraps or on request
  static constexpr int kRing = 16;
  static constexpr int kMaxLanes = 4;  // 3 gemm lanes + the syrk stream
  bool timing = false;
  bool serial = false;  // timing mode 2: lanes one after another on the caller's stream
  hipEvent_t ev[kRing][kMaxLanes][HCORE_HIP_NPHASES][2] = {};
  bool used[kRing][kMaxLanes][HCORE_HIP_NPHASES] = {};
  bool pending[kRing] = {};
  int slot = 0;
  double phase_ms[kMaxLanes][HCORE_HIP_NPHASES] = {};
  int phase_count[kMaxLanes][HCORE_HIP_NPHASES] = {};
};

// A captured graph holds the launches as they were: capture again on the next
// replay (after the autotuners change a launch).
static void drop_graph(hcore_hip_plan* p) {
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->gexec = nullptr;
  p->graph = nullptr;
}

// FP64 storage (plain FP64 and split precision) or FP32 (FP32, mixed).
static bool plan_double(const hcore_hip_plan* p) {
  return p->prec == HCORE_HIP_FP64 || p->prec == HCORE_HIP_SPLIT;
}

namespace hcore {

// Digits per operand of a split-precision plan's products for tolerance acc:
// 8 bits each, enough for about 10 bits below acc relative to the operands'
// row and column scales (acc 1e-8: 5, 1e-6: 4, 1e-4: 3).
static int default_slices(double acc) {
  if (!(acc > 0)) return kOzMaxSlices;
  const int s = (int)std::ceil((-std::log2(acc) + 10.0) / 8.0);
  return std::min(kOzMaxSlices, std::max(kOzMinSlices, s));
}

// Streams of the lanes after the first (default priority, non-blocking), kept
// in a per-thread pool of free streams: a plan takes its streams at finalize
// and gives them back when destroyed, so a sequence of plans (one per TLR
// Cholesky step) runs on the same few streams and hardware queues instead of
// creating two or three streams per step.  A stream is handed out only on the
// device it was made on; a thread keeps at most kFreeStreamsMax free streams;
// pooled streams are not destroyed before the process ends.
constexpr size_t kFreeStreamsMax = 8;
struct FreeStream {
  int device;
  hipStream_t stream;
};
static thread_local std::vector<FreeStream> t_free_streams;
static hipError_t lane_stream_create(hipStream_t* s) {
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) {
    for (size_t i = t_free_streams.size(); i-- > 0;) {
      if (t_free_streams[i].device == dev) {
        *s = t_free_streams[i].stream;
        t_free_streams.erase(t_free_streams.begin() + (long)i);
        return hipSuccess;
      }
    }
  }
  return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}
static void lane_stream_release(hipStream_t s) {
  if (!s) return;
  hipDevice_t dev = 0;
  if (t_free_streams.size() < kFreeStreamsMax && hipStreamGetDevice(s, &dev) == hipSuccess)
    t_free_streams.push_back({(int)dev, s});
  else
    (void)hipStreamDestroy(s);
}

template <typename T>
static int plan_finalize_t(hcore_hip_plan* p) {
  const int M = p->M;
  const int n = p->ng;
  // mixed precision: FP32 everywhere except the core SVD and truncation (FP64)
  const bool mixed = p->prec == HCORE_HIP_MIXED;
  // split precision: the products through BF16 digit slices (GF_SLICED), the
  // Householder QR (with its Gram matrices), the core and the truncation FP64
  const int sl = p->prec == HCORE_HIP_SPLIT ? GF_SLICED : 0;
  const int KL = mixed ? core_kl<double>() : core_kl<T>();
  const size_t es = sizeof(T);
  auto al = [](size_t x) { return (x + 63) & ~size_t(63); };  // elements, 512 B for FP64

  // ---- stacked ranks, processing order, lanes and core variants ----
  std::vector<int> Kv(n);
  for (int u = 0; u < n; ++u) Kv[u] = p->ra[u] + p->rc[u];
  // processing order: largest stacked rank first (longest jobs start first)
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return Kv[a] > Kv[b]; });

  // lanes by stacked rank (empty lanes dropped): with the register-resident
  // core K > 128 | 65..128 | <= 64, else K > KL (core in global memory) |
  // 97..KL | <= 96
  const bool reg = g_ctx.core_reg;
  const int lane_lo[3] = {reg ? 129 : KL + 1, reg ? 65 : 97, 0};
  std::vector<std::vector<int>> lane_members(3);
  for (int idx = 0; idx < n; ++idx) {
    const int u = order[idx];
    for (int l = 0; l < 3; ++l)
      if (Kv[u] >= lane_lo[l]) { lane_members[l].push_back(u); break; }
  }
  if (n <= 64) {  // small batches: one lane
    lane_members[2].clear();
    for (int idx = 0; idx < n; ++idx) lane_members[2].push_back(order[idx]);
    lane_members[0].clear();
    lane_members[1].clear();
  }
  // core SVD variant per update: the heavy lanes use one variant for all their
  // tiles (one launch), the bulk lane buckets by K
  std::vector<int> variant(n);
  for (int l = 0; l < 3; ++l) {
    int kmax = 0;
    for (int u : lane_members[l]) kmax = std::max(kmax, Kv[u]);
    for (int u : lane_members[l])
      variant[u] = core_variant_for(l < 2 && n > 64 ? kmax : Kv[u], KL, reg);
  }
  // variants that need the K x K global scratch (work matrix, or the R^T
  // transposition of the register-resident kernels)
  auto global_core = core_needs_scratch;
  // reassembly: the compact-WY solve and GEMMs (every stacked rank the core
  // kernels accept)
  static_assert(kWyMaxK >= kCoreMaxK, "compact-WY reassembly for every stacked rank");
  // lane index as executed (empty lanes are dropped, so renumber)
  p->lane_of.assign(n, 0);
  p->lane_updates.clear();
  {
    int li = 0;
    for (int l = 0; l < 3; ++l) {
      if (lane_members[l].empty()) continue;
      for (int u : lane_members[l]) p->lane_of[u] = li;
      p->lane_updates.push_back(lane_members[l]);
      ++li;
    }
  }

  // ---- workspace layout (elements of T) ----
  std::vector<size_t> o_Us(n), o_Vs(n), o_tA(n), o_tB(n), o_TA(n), o_TB(n), o_F(n), o_Tc(n),
      o_Tj(n), o_Zu(n), o_Zv(n), o_sg(n), o_S(n), o_M(n), o_c64(n), o_rt(n);
  size_t off = 0;
  for (int u = 0; u < n; ++u) {
    const int K = Kv[u];
    o_Us[u] = off; off += al((size_t)M * K);
    o_Vs[u] = off; off += al((size_t)M * K);
    o_tA[u] = off; off += al(K);
    o_tB[u] = off; off += al(K);
    o_TA[u] = off; off += al(tp_elems(K));
    o_TB[u] = off; off += al(tp_elems(K));
    o_F[u] = off; off += al((size_t)p->ra[u] * p->rb[u]);
    o_Tc[u] = off; off += al((size_t)K * K);
    o_Tj[u] = off; off += global_core(variant[u]) ? al((size_t)K * (K + 64)) : 0;  // ld <= K + 2G - 1
    o_Zu[u] = off; off += al((size_t)K * K);
    o_Zv[u] = off; off += al((size_t)K * K);
    o_S[u] = off; off += 2 * al((size_t)K * K);   // Gram Y^T Y of both stacks
    o_M[u] = off; off += 2 * al((size_t)K * K);   // compact-WY solves of both sides
    o_sg[u] = off; off += al(K);
    if (p->dev_ranks) { o_rt[u] = off; off += al(1); }  // QR queue: routed flags of both stacks
    if (mixed) {  // FP64 core: T, scratch, left vectors and sigma, in T-sized elements
      const size_t d = (size_t)K * K + (size_t)K * (K + 64) + (size_t)K * K + K;
      o_c64[u] = off; off += al(d * sizeof(double) / es);
    }
  }
  std::vector<size_t> o_W(p->ns), o_X(p->ns);
  for (int s = 0; s < p->ns; ++s) {
    const int K = p->sK[s];
    o_W[s] = off; off += al((size_t)K * K);
    o_X[s] = off; off += al((size_t)M * K);
  }
  p->sigma_off = o_sg;
  p->sigma_double = sizeof(T) == 8;
  if (n > 0) {
    p->first_off[0] = o_Us[0]; p->first_off[1] = o_Vs[0]; p->first_off[2] = o_tA[0];
    p->first_off[3] = o_tB[0]; p->first_off[4] = o_Zu[0]; p->first_off[5] = o_Zv[0];
  }
  if (mixed) {  // sigma of the FP64 core, as an offset in doubles
    for (int u = 0; u < n; ++u) {
      const int K = Kv[u];
      p->sigma_off[u] = (o_c64[u] * es) / sizeof(double) + 2 * (size_t)K * K + (size_t)K * (K + 64);
    }
    p->sigma_double = true;
  }
  const size_t ws_elems = off;

  // ---- device memory: [rank | status | sweeps | workspace | descriptors] ----
  Packer pk;
  int* d_sweeps = nullptr;
  long long* d_clk = nullptr;
  // descriptor contents reference device addresses, so the base must be known
  // first: size the descriptors with a dry run, then allocate, then fill.
  auto build = [&](char* wsbase, int* d_rank, int* d_status) {
    T* W = reinterpret_cast<T*>(wsbase);
    pk.buf.clear();
    std::vector<GemmProb<T>> gp;
    // rank-dependent descriptor fields (device-side ranks): (vector, element,
    // field offset) resolved to byte offsets once the vectors are packed
    struct PendingPatch {
      const void* vec;
      size_t esz, idx, field;
      int kind, upd;
      long long base;
    };
    std::vector<PendingPatch> pend;
    std::vector<std::pair<const void*, size_t>> vec_off;
    const bool dvr = p->dev_ranks, dvs = p->sdev_ranks;
    auto rec = [&](const auto& vec, size_t idx, size_t field, int kind, int upd, long long base = 0) {
      pend.push_back({&vec, sizeof(vec[0]), idx, field, kind, upd, base});
    };
    auto addv = [&](const auto& vec) {
      const size_t off = pk.add(vec);
      vec_off.push_back({&vec, off});
      return off;
    };
#define HCORE_GF(f) offsetof(GemmProb<T>, f)
    struct LaneDesc {
      std::vector<StackProb<T>> stack;
      TaskList tF, tG, tT, tZ, tB, tX;
      std::vector<WySolveProb<T>> wy;
      int kmax = 0;
      std::vector<QrProb<T>> qr, qrr16, qrr, qrp;  // blocked / register single pass / register panels
      std::vector<QrProb<T>> qr_all;               // device-side ranks: every stack (the QR queue)
      bool gram = false;  // the T phase holds k = nb Gram products
      std::vector<std::vector<CoreProb<T>>> core_b;
      std::vector<std::vector<CoreProb<double>>> core_b64;  // mixed: the FP64 core
      // device-side ranks: every update's core in lane order (the rank queue)
      std::vector<CoreProb<T>> core_all;
      std::vector<CoreProb<double>> core_all64;
      std::vector<ConvProb> cv_in, cv_out;
      int cv_max = 0;
      std::vector<int> core_kmax;
    };
    constexpr int NV = CORE_NVARIANTS;
    std::vector<LaneDesc> ld;
    ld.reserve(3);  // built in place: patches refer to the lanes' vectors by address
    for (int l = 0; l < 3; ++l) {
      if (lane_members[l].empty()) continue;
      ld.emplace_back();
      LaneDesc& d = ld.back();
      d.core_b.resize(NV);
      d.core_b64.resize(NV);
      d.core_kmax.assign(NV, 1);
      for (int u : lane_members[l]) {
        const int K = Kv[u], ra = p->ra[u], rb = p->rb[u], rc = p->rc[u];
        T* Us = W + o_Us[u];
        T* Vs = W + o_Vs[u];
        d.stack.push_back({(const T*)p->CU[u], (const T*)p->CV[u], (const T*)p->AU[u], Us, Vs,
                           rc, ra, p->ldc, p->lda});
        if (dvr) {
          rec(d.stack, d.stack.size() - 1, offsetof(StackProb<T>, rc), PK_RC, u);
          rec(d.stack, d.stack.size() - 1, offsetof(StackProb<T>, ra), PK_RA, u);
        }
        GemmProb<T> g{};
        if (p->BD.empty()) {
          // F = AV^T BV   (hcore_dgemm.c:175-178)
          g.A = (const T*)p->AV[u]; g.B = (const T*)p->BV[u]; g.C = W + o_F[u];
          g.m = ra; g.n = rb; g.k = M; g.lda = p->lda; g.ldb = p->ldb; g.ldc = ra;
          g.alpha = T(1); g.beta = T(0); g.flags = GF_A_TRANS | GF_BETA_ZERO | sl;
          add_tasks(d.tF, (int)gp.size(), g); gp.push_back(g);
          if (dvr) {
            rec(gp, gp.size() - 1, HCORE_GF(m), PK_RA, u);
            rec(gp, gp.size() - 1, HCORE_GF(n), PK_RB, u);
            rec(gp, gp.size() - 1, HCORE_GF(ldc), PK_RA, u);
          }
          // G = BU F^T into Vs(:, rc:K)   (hcore_dgemm.c:199-205)
          g = GemmProb<T>{};
          g.A = (const T*)p->BU[u]; g.B = W + o_F[u]; g.C = Vs + (size_t)rc * M;
          g.m = M; g.n = ra; g.k = rb; g.lda = p->ldb; g.ldb = ra; g.ldc = M;
          g.alpha = T(1); g.beta = T(0); g.flags = GF_B_TRANS | GF_BETA_ZERO | sl;
          add_tasks(d.tG, (int)gp.size(), g); gp.push_back(g);
          if (dvr) {
            rec(gp, gp.size() - 1, HCORE_GF(n), PK_RA, u);
            rec(gp, gp.size() - 1, HCORE_GF(k), PK_RB, u);
            rec(gp, gp.size() - 1, HCORE_GF(ldb), PK_RA, u);
            rec(gp, gp.size() - 1, HCORE_GF(C), PK_PTR_G, u, (long long)(size_t)Vs);
          }
        } else {
          // G = B AV into Vs(:, rc:K) for a dense B   (hcore_dgemm.c:255-262)
          g.A = (const T*)p->BD[u]; g.B = (const T*)p->AV[u]; g.C = Vs + (size_t)rc * M;
          g.m = M; g.n = ra; g.k = M; g.lda = p->ldbd; g.ldb = p->lda; g.ldc = M;
          g.alpha = T(1); g.beta = T(0); g.flags = GF_BETA_ZERO | sl;
          add_tasks(d.tG, (int)gp.size(), g); gp.push_back(g);
        }
        // T = R_A R_B^T   (hcore_dgemm.c:384-386)
        g = GemmProb<T>{};
        g.A = Us; g.B = Vs; g.C = W + o_Tc[u];
        g.m = K; g.n = K; g.k = K; g.lda = M; g.ldb = M; g.ldc = K;
        g.alpha = T(1); g.beta = T(0);
        g.flags = GF_A_UPPER | GF_B_TRANS | GF_B_UPPER | GF_BETA_ZERO | sl;
        add_tasks(d.tT, (int)gp.size(), g); gp.push_back(g);
        auto recK = [&](std::initializer_list<size_t> fields) {
          if (dvr)
            for (size_t f : fields) rec(gp, gp.size() - 1, f, PK_K, u);
        };
        recK({HCORE_GF(m), HCORE_GF(n), HCORE_GF(k), HCORE_GF(ldc)});
        // Zv = T^T Zu = V_r S_r
        g = GemmProb<T>{};
        g.A = W + o_Tc[u]; g.B = W + o_Zu[u]; g.C = W + o_Zv[u];
        g.m = K; g.n = K; g.k = K; g.lda = K; g.ldb = K; g.ldc = K;
        g.alpha = T(1); g.beta = T(0); g.flags = GF_A_TRANS | GF_BETA_ZERO | sl;
        g.n_dev = d_rank + u; g.skip = d_status + u;
        add_tasks(d.tZ, (int)gp.size(), g); gp.push_back(g);
        recK({HCORE_GF(m), HCORE_GF(n), HCORE_GF(k), HCORE_GF(lda), HCORE_GF(ldb), HCORE_GF(ldc)});
        // register-resident QR (it also yields the Gram S and the panel factors)
        const bool qreg = g_ctx.qr_reg && K <= kQrRegMaxK && M <= kQrRegMaxNb;
        {
          T* SA = W + o_S[u];
          T* SB = SA + al((size_t)K * K);
          T* MA = W + o_M[u];
          T* MB = MA + al((size_t)K * K);
          T* Zs[2] = {W + o_Zu[u], W + o_Zv[u]};
          T* Ys[2] = {Us, Vs};
          T* Ss[2] = {SA, SB};
          T* Ms[2] = {MA, MB};
          T* Xo[2] = {(T*)(p->CUo[u] ? p->CUo[u] : p->CU[u]), (T*)(p->CVo[u] ? p->CVo[u] : p->CV[u])};
          for (int side = 0; side < 2; ++side) {
            // Gram S = Y^T Y of the unit lower trapezoidal reflector block (with T)
            g = GemmProb<T>{};
            g.A = Ys[side]; g.B = Ys[side]; g.C = Ss[side];
            g.m = K; g.n = K; g.k = M; g.lda = M; g.ldb = M; g.ldc = K;
            g.alpha = T(1); g.beta = T(0);
            // only its strict lower triangle is read (the compact-WY solve)
            g.flags = GF_A_TRANS | GF_A_UNIT_LOWER | GF_B_UNIT_LOWER | GF_BETA_ZERO | GF_LOWER_ONLY;
            if (dvr && g_ctx.qr_reg && M <= kQrRegMaxNb)  // the QR queue: a register QR wrote S
              g.skip = reinterpret_cast<const int*>(W + o_rt[u]) + side;
            if (!qreg) {
              add_tasks(d.tT, (int)gp.size(), g);
              gp.push_back(g);
              d.gram = true;
              recK({HCORE_GF(m), HCORE_GF(n), HCORE_GF(ldc)});
            }
            // B = Y1^T Z (Z = Zu with Zv's launch, Z = Zv after it)
            g = GemmProb<T>{};
            g.A = Ys[side]; g.B = Zs[side]; g.C = Ms[side];
            g.m = K; g.n = K; g.k = K; g.lda = M; g.ldb = K; g.ldc = K;
            g.alpha = T(1); g.beta = T(0); g.flags = GF_A_TRANS | GF_A_UNIT_LOWER | GF_BETA_ZERO | sl;
            g.n_dev = d_rank + u; g.skip = d_status + u;
            add_tasks(side == 0 ? d.tZ : d.tB, (int)gp.size(), g); gp.push_back(g);
            recK({HCORE_GF(m), HCORE_GF(n), HCORE_GF(k), HCORE_GF(ldb), HCORE_GF(ldc)});
            // X = [Z; 0] - Y M into the caller's factor
            g = GemmProb<T>{};
            g.A = Ys[side]; g.B = Ms[side]; g.C = Xo[side];
            g.m = M; g.n = K; g.k = K; g.lda = M; g.ldb = K; g.ldc = p->ldco;
            g.alpha = T(-1); g.beta = T(1); g.flags = GF_A_UNIT_LOWER | GF_C0 | sl;
            g.C0 = Zs[side]; g.ldc0 = K; g.m0 = K;
            g.n_dev = d_rank + u; g.skip = d_status + u;
            add_tasks(d.tX, (int)gp.size(), g); gp.push_back(g);
            recK({HCORE_GF(n), HCORE_GF(k), HCORE_GF(ldb), HCORE_GF(ldc0), HCORE_GF(m0)});
            d.wy.push_back({Ss[side], side == 0 ? W + o_TA[u] : W + o_TB[u], Ms[side], d_rank + u,
                            d_status + u, K});
            if (dvr) rec(d.wy, d.wy.size() - 1, offsetof(WySolveProb<T>, K), PK_K, u);
          }
        }
        // QR of both stacks
        auto recQ = [&](const std::vector<QrProb<T>>& q) {
          if (dvr) {
            rec(q, q.size() - 2, offsetof(QrProb<T>, K), PK_K, u);
            rec(q, q.size() - 1, offsetof(QrProb<T>, K), PK_K, u);
          }
        };
        if (dvr && g_ctx.qr_reg && M <= kQrRegMaxNb) {
          // device-side ranks: the kernel chosen by the K read at execution
          // (the register QR also writes S; the Gram products of capacities
          // above kQrRegMaxK rewrite it for every stack)
          T* SA = W + o_S[u];
          int* rt = reinterpret_cast<int*>(W + o_rt[u]);
          d.qr_all.push_back({Us, W + o_tA[u], W + o_TA[u], SA, K, rt});
          d.qr_all.push_back({Vs, W + o_tB[u], W + o_TB[u], SA + al((size_t)K * K), K, rt + 1});
          recQ(d.qr_all);
        } else if (qreg) {  // one register-resident pass (also yields S)
          T* SA = W + o_S[u];
          auto& q = K <= 16 ? d.qrr16 : d.qrr;
          q.push_back({Us, W + o_tA[u], W + o_TA[u], SA, K});
          q.push_back({Vs, W + o_tB[u], W + o_TB[u], SA + al((size_t)K * K), K});
          recQ(q);
        } else if (g_ctx.qr_reg && M <= kQrFusedMaxNb) {  // register panels + in-place trailing updates
          d.qrp.push_back({Us, W + o_tA[u], W + o_TA[u], nullptr, K});
          d.qrp.push_back({Vs, W + o_tB[u], W + o_TB[u], nullptr, K});
          recQ(d.qrp);
        } else {  // blocked Householder QR with LDS panels
          d.qr.push_back({Us, W + o_tA[u], W + o_TA[u], nullptr, K});
          d.qr.push_back({Vs, W + o_tB[u], W + o_TB[u], nullptr, K});
          recQ(d.qr);
        }
        const int b = variant[u];
        if (mixed) {
          double* c64 = reinterpret_cast<double*>(W + o_c64[u]);
          double* Tc64 = c64;
          double* Tj64 = Tc64 + (size_t)K * K;
          double* Zu64 = Tj64 + (size_t)K * (K + 64);
          double* sg64 = Zu64 + (size_t)K * K;
          CoreProb<double> c{Tc64, Tj64, Zu64, sg64, d_rank + u, d_status + u, d_sweeps + u,
                             d_clk ? d_clk + (size_t)u * kCoreClk : nullptr, K};
          d.cv_in.push_back({W + o_Tc[u], Tc64, K * K});
          d.cv_out.push_back({Zu64, W + o_Zu[u], K * K});
          d.cv_max = std::max(d.cv_max, K * K);
          if (dvr) {
            d.core_all64.push_back(c);
            rec(d.core_all64, d.core_all64.size() - 1, offsetof(CoreProb<double>, K), PK_K, u);
            rec(d.cv_in, d.cv_in.size() - 1, offsetof(ConvProb, n), PK_KK, u);
            rec(d.cv_out, d.cv_out.size() - 1, offsetof(ConvProb, n), PK_KK, u);
          } else {
            d.core_b64[b].push_back(c);
          }
        } else {
          CoreProb<T> c{W + o_Tc[u], global_core(variant[u]) ? W + o_Tj[u] : nullptr, W + o_Zu[u],
                        W + o_sg[u], d_rank + u, d_status + u, d_sweeps + u,
                        d_clk ? d_clk + (size_t)u * kCoreClk : nullptr, K};
          if (dvr) {
            d.core_all.push_back(c);
            rec(d.core_all, d.core_all.size() - 1, offsetof(CoreProb<T>, K), PK_K, u);
          } else {
            d.core_b[b].push_back(c);
          }
        }
        d.core_kmax[b] = std::max(d.core_kmax[b], K);
        d.kmax = std::max(d.kmax, K);
      }
    }
    // syrk (hcore_dsyrk.c:103, 144, 165-181)
    std::vector<GemmProb<T>> sp;
    TaskList tW, tX, tC;
    for (int s = 0; s < p->ns; ++s) {
      const int K = p->sK[s];
      GemmProb<T> g{};
      g.A = (const T*)p->sAV[s]; g.B = (const T*)p->sAV[s]; g.C = W + o_W[s];
      g.m = K; g.n = K; g.k = p->ldau; g.lda = p->ldau; g.ldb = p->ldau; g.ldc = K;
      g.alpha = T(1); g.beta = T(0); g.flags = GF_A_TRANS | GF_BETA_ZERO;
      add_tasks(tW, (int)sp.size(), g); sp.push_back(g);
      auto recS = [&](std::initializer_list<size_t> fields) {
        if (dvs)
          for (size_t f : fields) rec(sp, sp.size() - 1, f, PK_SK, s);
      };
      recS({HCORE_GF(m), HCORE_GF(n), HCORE_GF(ldc)});
      g = GemmProb<T>{};
      g.A = (const T*)p->sAU[s]; g.B = W + o_W[s]; g.C = W + o_X[s];
      g.m = M; g.n = K; g.k = K; g.lda = p->ldau; g.ldb = K; g.ldc = M;
      g.alpha = T(1); g.beta = T(0); g.flags = GF_BETA_ZERO;
      add_tasks(tX, (int)sp.size(), g); sp.push_back(g);
      recS({HCORE_GF(n), HCORE_GF(k), HCORE_GF(ldb)});
      g = GemmProb<T>{};
      g.A = W + o_X[s]; g.B = (const T*)p->sAU[s]; g.C = (T*)p->sCD[s];
      g.m = M; g.n = M; g.k = K; g.lda = M; g.ldb = p->ldau; g.ldc = p->ldcd;
      g.alpha = T(p->salpha); g.beta = T(1); g.flags = GF_B_TRANS | GF_LOWER_ZERO_UPPER;
      add_tasks(tC, (int)sp.size(), g); sp.push_back(g);
      recS({HCORE_GF(k)});
    }
    p->o_gp = addv(gp);
    if (p->lanes.size() != ld.size()) p->lanes.resize(ld.size());
    for (size_t l = 0; l < ld.size(); ++l) {
      auto& L = p->lanes[l];
      auto& d = ld[l];
      L.n = (int)d.stack.size();
      L.o_stack = addv(d.stack);
      L.stack_cols = 0;
      for (const auto& sp_ : d.stack) L.stack_cols = std::max(L.stack_cols, 2 * sp_.rc + sp_.ra);
      L.tF = pk.add(d.tF);
      L.tG = pk.add(d.tG);
      L.tT = pk.add(d.tT);
      L.tZ = pk.add(d.tZ);
      L.o_qr = addv(d.qr); L.n_qr = (int)d.qr.size();
      L.gram = d.gram;
      L.o_qrr16 = addv(d.qrr16); L.n_qrr16 = (int)d.qrr16.size();
      L.o_qrr = addv(d.qrr); L.n_qrr = (int)d.qrr.size();
      L.o_qrp = addv(d.qrp); L.n_qrp = (int)d.qrp.size();
      L.qrq = !d.qr_all.empty();
      L.n_qrall = (int)d.qr_all.size();
      L.o_qrall = addv(d.qr_all);
      L.o_qrq = pk.add(std::vector<int>(kQrQueueHead + std::max(L.n_qrall, 1), 0));
      L.tB = pk.add(d.tB);
      L.tX = pk.add(d.tX);
      L.o_wy = addv(d.wy); L.n_wy = (int)d.wy.size();
      L.kmax = d.kmax;
      L.slices = sl ? (p->slices_set ? p->slices_set : default_slices(p->acc)) : 0;
      L.core.clear();
      L.queue = dvr;
      if (dvr) {
        // one launch per base variant the lane's ranks can select (K up to the
        // lane's capacity), each served from the rank queue built at execution
        const bool dbl = mixed;
        L.qdbl = dbl;
        L.qreg = reg;
        L.qkl = KL;
        L.o_qprobs = dbl ? addv(d.core_all64) : addv(d.core_all);
        L.o_q = pk.add(std::vector<int>(kCoreQueueHead + std::max(L.n, 1), 0));
        std::vector<int> vmax(NV, 0);
        for (int K = 1; K <= d.kmax; ++K) vmax[core_variant_for(K, KL, reg)] = K;
        for (int b = 0; b < NV; ++b)
          if (vmax[b] > 0) {
            hcore_hip_plan::CoreLaunch c{L.o_qprobs, L.n, tuned_variant(M, dbl || sizeof(T) == 8, b),
                                         vmax[b], dbl, b};
            c.queue = true;
            L.core.push_back(c);
          }
      }
      for (int b = 0; b < NV; ++b) {
        if (!d.core_b[b].empty())
          L.core.push_back({addv(d.core_b[b]), (int)d.core_b[b].size(),
                            tuned_variant(M, sizeof(T) == 8, b),
                            d.core_kmax[b], false, b});
        if (!d.core_b64[b].empty())
          L.core.push_back({addv(d.core_b64[b]), (int)d.core_b64[b].size(),
                            tuned_variant(M, true, b), d.core_kmax[b], true, b});
      }
      L.o_cvin = addv(d.cv_in); L.o_cvout = addv(d.cv_out);
      L.n_cv = (int)d.cv_in.size(); L.cv_max = d.cv_max;
      // largest K first: the longest jobs start first
      std::stable_sort(L.core.begin(), L.core.end(),
                       [](const hcore_hip_plan::CoreLaunch& x, const hcore_hip_plan::CoreLaunch& y) {
                         return x.kmax > y.kmax;
                       });
    }
    p->o_sp = addv(sp);
    p->tW = pk.add(tW);
    p->tXs = pk.add(tX);
    p->tC = pk.add(tC);
#undef HCORE_GF
    // the patch list and the device rank tables
    std::vector<RankPatch> patches;
    for (const auto& q : pend) {
      size_t base_off = 0;
      for (const auto& vo : vec_off)
        if (vo.first == q.vec) base_off = vo.second;
      patches.push_back({(unsigned long long)(base_off + q.idx * q.esz + q.field), q.base, q.kind,
                         q.upd});
    }
    p->o_patch = pk.add(patches);
    p->n_patch = (int)patches.size();
    std::vector<const int*> tabs;
    std::vector<int> caps;
    if (dvr) {
      for (int u = 0; u < n; ++u) tabs.push_back(p->ra_dev[u]);
      for (int u = 0; u < n; ++u) tabs.push_back(p->rb_dev[u]);
      for (int u = 0; u < n; ++u) tabs.push_back(p->rc_dev[u]);
      for (int u = 0; u < n; ++u) {
        caps.push_back(p->ra[u]);
        caps.push_back(p->rb[u]);
        caps.push_back(p->rc[u]);
      }
    }
    if (dvs) {
      if (!dvr) caps.assign(3 * (size_t)n, 0);
      for (int v = 0; v < p->ns; ++v) tabs.push_back(p->sk_dev[v]);
      for (int v = 0; v < p->ns; ++v) caps.push_back(p->sK[v]);
    }
    p->o_tab = pk.add(tabs);
    p->o_cap = pk.add(caps);
  };
  build(nullptr, nullptr, nullptr);  // dry run: descriptor size
  // [rank | status | sweeps | ranks in use (3 per update, 1 per dsyrk)] ints
  const size_t n_in = 3 * (size_t)std::max(n, 1) + (size_t)p->ns;
  const size_t head_i =
      (((size_t)std::max(n, 1) * 3 + n_in) * sizeof(int) + 255) & ~size_t(255);
  const size_t head = head_i + (((size_t)std::max(n, 1) * kCoreClk * sizeof(long long) + 255) &
                                ~size_t(255));
  const size_t wsb = (ws_elems * es + 255) & ~size_t(255);
  const size_t total = head + wsb + pk.buf.size() + 256;
  if (p->measure) {
    p->dbytes = total;
    return HCORE_HIP_OK;
  }
  char* base = nullptr;
  const bool single = p->arena || p->fixed;
  if (p->fixed) {
    if (total > p->fixed_cap) return HCORE_HIP_ERR_ARG;
    base = p->fixed;
  } else if (p->arena) {
    int st = p->arena->reserve(total);
    if (st) return st;
    base = p->arena->p;
  } else {
    HCHECK(hipMalloc(&base, total));
    p->dmem = base;
  }
  p->dbytes = total;
  p->ws_bytes = wsb;
  p->d_rank = reinterpret_cast<int*>(base);
  p->d_status = p->d_rank + std::max(n, 1);
  p->d_sweeps = p->d_status + std::max(n, 1);
  p->d_rank_in = p->d_sweeps + std::max(n, 1);
  d_sweeps = p->d_sweeps;
  p->d_clk = reinterpret_cast<long long*>(base + head_i);
  d_clk = p->d_clk;
  p->ws = base + head;
  p->desc = base + head + wsb;
  build(p->ws, p->d_rank, p->d_status);
  if (single) {  // asynchronous on the caller's stream; the host image outlives the copy
    p->desc_host->swap(pk.buf);
    HCHECK(hipMemcpyAsync(p->desc, p->desc_host->data(), p->desc_host->size(),
                          hipMemcpyHostToDevice, p->fin_stream));
    HCHECK(hipMemsetAsync(base, 0, head, p->fin_stream));
    p->finalized = true;
    return HCORE_HIP_OK;
  }
  HCHECK(hipMemcpy(p->desc, pk.buf.data(), pk.buf.size(), hipMemcpyHostToDevice));
  HCHECK(hipMemset(base, 0, head));
  HCHECK(hipDeviceSynchronize());
  for (size_t l = 0; l < p->lanes.size(); ++l) {
    auto& L = p->lanes[l];
    if (l > 0) HCHECK(lane_stream_create(&L.stream));
    HCHECK(hipEventCreateWithFlags(&L.ev_done, hipEventDisableTiming));
  }
  HCHECK(hipEventCreateWithFlags(&p->ev_start, hipEventDisableTiming));
  p->finalized = true;
  return HCORE_HIP_OK;
}

// Accumulate the phase times of one pending timed execution.
static int collect_slot(hcore_hip_plan* p, int k) {
  if (!p->pending[k]) return HCORE_HIP_OK;
  for (int l = 0; l < hcore_hip_plan::kMaxLanes; ++l)
    for (int i = 0; i < HCORE_HIP_NPHASES; ++i) {
      if (!p->used[k][l][i]) continue;
      HCHECK(hipEventSynchronize(p->ev[k][l][i][1]));
      float ms = 0.f;
      HCHECK(hipEventElapsedTime(&ms, p->ev[k][l][i][0], p->ev[k][l][i][1]));
      p->phase_ms[l][i] += ms;
      p->phase_count[l][i] += 1;
      p->used[k][l][i] = false;
    }
  p->pending[k] = false;
  return HCORE_HIP_OK;
}

static int collect_all(hcore_hip_plan* p) {
  for (int k = 0; k < hcore_hip_plan::kRing; ++k) {
    int st = collect_slot(p, (p->slot + k) % hcore_hip_plan::kRing);
    if (st) return st;
  }
  return HCORE_HIP_OK;
}

// part: 0 the whole pipeline; 1 up to the core SVD and Zv = T^T Zu (the rank
// is known after it); 2 the rest (the reassembly of the new factors).
template <typename T>
static int plan_execute_t(hcore_hip_plan* p, hipStream_t s, bool capture = false, int part = 0) {
  const int M = p->M;
  auto D = [&](size_t off) { return p->desc + off; };
  const auto* gp = reinterpret_cast<const GemmProb<T>*>(D(p->o_gp));
  const bool tm = p->timing && !capture;
  const int k = p->slot;
  if (tm) {
    int st = collect_slot(p, k);  // the set used kRing executions ago
    if (st) return st;
    p->slot = (k + 1) % hcore_hip_plan::kRing;
  }
  // phase `id` of lane `l` on stream `st`, bracketed by events when timing
  auto phase = [&](int l, int id, hipStream_t st, bool run, auto&& fn) {
    if (!run) return;
    if (tm) (void)hipEventRecord(p->ev[k][l][id][0], st);
    fn();
    if (tm) {
      (void)hipEventRecord(p->ev[k][l][id][1], st);
      p->used[k][l][id] = true;
    }
  };
  const int nl = (int)p->lanes.size();
  // dsyrk updates (hcore_dsyrk.c:103, 144, 165-181): on the stream of one of
  // the lanes, after its recompressions, so that a plan uses at most three
  // streams (the caller's and two lanes) next to the null stream: four hardware
  // queues.  With three lanes they follow the middle one (K 65..128), which
  // finishes first; the bulk lane (K <= 64) carries the most work and the first
  // lane the longest chain.
  const int syrk_lane = nl >= 2 ? nl - 2 : 0;
  // one GEMM phase: its 64-tile and 32-tile launches
  auto gemm = [&](const GemmProb<T>* probs, const TaskRef& r, hipStream_t st, bool split = false,
                  int slices = 0) {
    launch_tasks<T>(probs, reinterpret_cast<const char*>(p->desc), r, st, split, slices);
  };
  // phases whose products reduce over the nb rows of the tiles (F = AV^T BV,
  // the stacks' Gram matrices next to T) share each tile's k range
  const bool split_k = M >= 256;
  auto syrk = [&](hipStream_t ss) {
    if (p->ns <= 0) return;
    const auto* sp = reinterpret_cast<const GemmProb<T>*>(D(p->o_sp));
    const int lt = hcore_hip_plan::kMaxLanes - 1;
    phase(lt, 9, ss, true, [&] { gemm(sp, p->tW, ss, split_k); });
    phase(lt, 10, ss, true, [&] { gemm(sp, p->tXs, ss); });
    phase(lt, 11, ss, true, [&] { gemm(sp, p->tC, ss); });
  };
  // per-execution bookkeeping of batched plans: statuses reset and, with
  // device-side ranks, the ranks read and the descriptors patched
  const bool batched = !(p->arena || p->fixed);
  RankTables rt{};
  if (batched) {
    const int* const* tab = reinterpret_cast<const int* const*>(D(p->o_tab));
    rt.ra = p->dev_ranks ? tab : nullptr;
    rt.rb = p->dev_ranks ? tab + p->ng : nullptr;
    rt.rc = p->dev_ranks ? const_cast<int* const*>(tab + 2 * p->ng) : nullptr;
    rt.sk = p->sdev_ranks ? tab + (p->dev_ranks ? 3 * p->ng : 0) : nullptr;
    rt.cap = reinterpret_cast<const int*>(D(p->o_cap));
    rt.rank_in = p->d_rank_in;
    rt.status = p->d_status;
    rt.rank_out = p->d_rank;
    rt.n = p->ng;
    rt.ns = p->sdev_ranks ? p->ns : 0;
    rt.M = M;
    rt.elem = (int)sizeof(T);
    if (part != 2) {
      launch_rank_gather(rt, s);
      launch_rank_patch(reinterpret_cast<const RankPatch*>(D(p->o_patch)), p->n_patch, p->desc,
                        p->d_rank_in, p->ng, M, (int)sizeof(T), s);
    }
  }
  if (p->ev_start) (void)hipEventRecord(p->ev_start, s);
  for (int l = 0; l < nl; ++l) {
    auto& L = p->lanes[l];
    const bool own = l > 0 && !(tm && p->serial);  // the lane runs on a stream of its own
    hipStream_t ls = own ? L.stream : s;
    if (own) (void)hipStreamWaitEvent(ls, p->ev_start, 0);
    const bool first = part != 2, second = part != 1;
    if (first) {
    phase(l, 0, ls, true, [&] {
      launch_stack<T>(reinterpret_cast<const StackProb<T>*>(D(L.o_stack)), L.n, M, L.stack_cols,
                      T(p->alpha), T(p->beta), ls);
    });
    phase(l, 1, ls, true, [&] { gemm(gp, L.tF, ls, split_k, L.slices); });
    phase(l, 2, ls, true, [&] { gemm(gp, L.tG, ls, false, L.slices); });
    bool shapes_ok = true;
    phase(l, 3, ls, true, [&] {
      if (L.qrq) {
        const auto* qa = reinterpret_cast<const QrProb<T>*>(D(L.o_qrall));
        int* qq = reinterpret_cast<int*>(p->desc + L.o_qrq);
        launch_qr_queue<T>(qa, L.n_qrall, qq, ls);
        shapes_ok &= launch_qr_fused<T>(qa, L.n_qrall, M, ls, qq, 2);
        shapes_ok &= launch_qr_reg<T>(qa, L.n_qrall, M, 32, ls, qq, 1);
        shapes_ok &= launch_qr_reg<T>(qa, L.n_qrall, M, 16, ls, qq, 0);
      }
      shapes_ok &= launch_qr_fused<T>(reinterpret_cast<const QrProb<T>*>(D(L.o_qrp)), L.n_qrp, M, ls);
      shapes_ok &= launch_qr_reg<T>(reinterpret_cast<const QrProb<T>*>(D(L.o_qrr)), L.n_qrr, M, 32, ls);
      shapes_ok &= launch_qr_reg<T>(reinterpret_cast<const QrProb<T>*>(D(L.o_qrr16)), L.n_qrr16, M, 16, ls);
      launch_geqrf<T>(reinterpret_cast<const QrProb<T>*>(D(L.o_qr)), L.n_qr, M, ls);
    });
    if (!shapes_ok) {
      fprintf(stderr, "hcore_amd: register QR routed a stack of %d rows beyond its shapes\n", M);
      return HCORE_HIP_ERR_ARG;
    }
    phase(l, 4, ls, true, [&] { gemm(gp, L.tT, ls, split_k && L.gram, L.slices); });
    // the lane's K buckets, largest first (each bucket a kernel variant with its
    // own LDS size); at most 4 streams per plan, one per hardware queue
    phase(l, 5, ls, !L.core.empty(), [&] {
      launch_convert(reinterpret_cast<const ConvProb*>(D(L.o_cvin)), L.n_cv, L.cv_max, true, ls);
      int* q = L.queue ? reinterpret_cast<int*>(p->desc + L.o_q) : nullptr;
      if (q) {
        if (L.qdbl)
          launch_core_queue<double>(reinterpret_cast<const CoreProb<double>*>(D(L.o_qprobs)), L.n,
                                    L.qkl, L.qreg, q, ls);
        else
          launch_core_queue<T>(reinterpret_cast<const CoreProb<T>*>(D(L.o_qprobs)), L.n, L.qkl,
                               L.qreg, q, ls);
      }
      // the first (largest-K) launch on the lane's stream, the others beside it
      // (one after the other on the side stream) when the lane's work is being
      // captured into a graph
      const bool fork = capture && L.aux && L.core.size() > 1;
      if (fork) {
        (void)hipEventRecord(L.ev_fork, ls);
        (void)hipStreamWaitEvent(L.aux, L.ev_fork, 0);
      }
      for (size_t ci = 0; ci < L.core.size(); ++ci) {
        const auto& c = L.core[ci];
        const int* cq = c.queue ? q : nullptr;
        hipStream_t cs = fork && ci > 0 ? L.aux : ls;
        if (c.dbl)
          launch_core<double>(reinterpret_cast<const CoreProb<double>*>(D(c.off)), c.count,
                              c.variant, c.kmax, p->rk, p->maxrk, p->acc, cs, cq, c.base_variant);
        else
          launch_core<T>(reinterpret_cast<const CoreProb<T>*>(D(c.off)), c.count, c.variant,
                         c.kmax, p->rk, p->maxrk, T(p->acc), cs, cq, c.base_variant);
      }
      if (fork) {
        (void)hipEventRecord(L.ev_join, L.aux);
        (void)hipStreamWaitEvent(ls, L.ev_join, 0);
      }
      launch_convert(reinterpret_cast<const ConvProb*>(D(L.o_cvout)), L.n_cv, L.cv_max, false, ls);
    });
    phase(l, 7, ls, true, [&] { gemm(gp, L.tZ, ls, false, L.slices); });
    }
    if (second) {
    phase(l, 6, ls, true, [&] {
      gemm(gp, L.tB, ls, false, L.slices);
      launch_wy_solve<T>(reinterpret_cast<const WySolveProb<T>*>(D(L.o_wy)), L.n_wy, L.kmax,
                         L.kmax, ls);
    });
    phase(l, 8, ls, true, [&] { gemm(gp, L.tX, ls, false, L.slices); });
    }
    if (l == syrk_lane && part == 0) syrk(ls);
    if (own) (void)hipEventRecord(L.ev_done, ls);
  }
  if (nl == 0) syrk(s);
  if (!(tm && p->serial))
    for (int l = 1; l < nl; ++l) (void)hipStreamWaitEvent(s, p->lanes[l].ev_done, 0);
  if (batched && p->dev_ranks && part != 1) launch_rank_publish(rt, s);
  if (tm) p->pending[k] = true;
  HCHECK(hipGetLastError());
  return HCORE_HIP_OK;
}

static bool is_device_ptr(const void* ptr) {
  if (!ptr) return false;
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, ptr);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged ||
         a.type == hipMemoryTypeUnified;
}

// Report a fatal condition the way the reference does, then apply the policy.
static void fail(int status, const char* msg) {
  t_last_status = status;
  fprintf(stderr, "%s\n", msg);
  if (g_ctx.policy == HCORE_HIP_ON_ERROR_EXIT) exit(-1);
}

// ------------------------------------------------------------------------------
// single-call recompression (HCORE_dgemm / HCORE_sgemm)
// ------------------------------------------------------------------------------
// Read one rank argument (a double/float on the host or on the device).
template <typename T>
static bool read_rank(const T* p, T* out) {
  if (!is_device_ptr(p)) {
    *out = *p;
    return true;
  }
  return hipMemcpy(out, p, sizeof(T), hipMemcpyDeviceToHost) == hipSuccess;
}

// Bytes of the caller's `work` under use_scratch: the layout the reference
// carves from it (hcore_dgemm.c:760-923; use_trmm = 0), clones of
// M x 2 maxrk, tau, F, newU, newV, R_A, R_B, T, sigma and superb.  A device
// `work` of that size holds this library's one-update workspace instead
// (ranks that cannot be read here give 0: the thread's own arena is used).
template <typename T>
static size_t reference_work_bytes(int M, int maxrk, const T* Ark, const T* Crk, size_t elem) {
  if (!Ark || !Crk || is_device_ptr(Ark) || is_device_ptr(Crk) || M <= 0 || maxrk <= 0) return 0;
  const size_t m = M, mk = maxrk, K = (size_t)(int)Ark[0] + (size_t)(int)Crk[0];
  const size_t rows = std::min(m, K);
  const size_t n = 2 * m * 2 * mk + 2 * m + mk * mk + std::max(m, mk) * mk + m * mk +
                   2 * rows * K + K * K + 2 * K;
  return n * elem;
}

// Where a single call's device memory comes from: the calling thread's grow-only
// arenas, or (scratch != nullptr) one fixed buffer that holds the plan and
// nothing else (the caller's factors must then be device pointers).
struct SingleMem {
  ThreadCtx* tc = nullptr;
  char* scratch = nullptr;
  size_t scratch_cap = 0;
  hipStream_t stream = nullptr;
  std::vector<char>* desc = nullptr;
  int* hres = nullptr;  // two ints the rank and status are read back into
};

template <typename T>
static void single_gemm(int M, T alpha, T* AU, T* AV, T* Ark, int LDA, T* BU, T* BV, T* Brk,
                        int LDB, T beta, T* CU, T* CV, T* Crk, int LDC, int rk, int maxrk, T acc,
                        flop_counter* flops, const char* name, void* work = nullptr,
                        size_t work_bytes = 0, SingleMem* fixed = nullptr) {
  t_last_status = HCORE_HIP_OK;
  char msg[512];
  // ranks may live on the host or the device
  T rks[3];
  if (!read_rank(Ark, &rks[0]) || !read_rank(Brk, &rks[1]) || !read_rank(Crk, &rks[2]))
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: cannot read rank");
  const int ra = (int)rks[0], rb = (int)rks[1], rc = (int)rks[2];
  if (ra == 0 || rb == 0 || rc == 0) {
    snprintf(msg, sizeof msg, "%s: _Ark=%d _Brk=%d _Crk=%d. These rank values should not be zero.",
             name, ra, rb, rc);
    return fail(HCORE_HIP_ERR_ZERO_RANK, msg);
  }
  if (ra < 0 || rb < 0 || rc < 0 || M <= 0 || LDA < M || LDB < M || LDC < M) {
    return fail(HCORE_HIP_ERR_ARG, "hcore_amd: negative rank, size or leading dimension below M");
  }
  const int K = ra + rc;
  if (K > 2 * maxrk) {
    snprintf(msg, sizeof msg, "%s: Sum of ranks (%d) is too big! _Ark:%d _Crk:%d maxrank:%d (x2: %d)",
             name, K, ra, rc, maxrk, 2 * maxrk);
    return fail(HCORE_HIP_ERR_SUM_RANKS, msg);
  }
  if (maxrk < K) {
    snprintf(msg, sizeof msg, "%s: Increase maxrank. %d is not enough ld_newV:%d CU_ncols:%d",
             name, maxrk, maxrk, K);
    return fail(HCORE_HIP_ERR_K_GT_MAXRK, msg);
  }
  if (K > M) {  // the stacks would be wider than tall
    snprintf(msg, sizeof msg,
             "%s: stacked rank %d exceeds the tile size %d (a rank above M/2 is not low rank)",
             name, K, M);
    return fail(HCORE_HIP_ERR_K_GT_MAXRK, msg);
  }
  if (K > kCoreMaxK) {
    snprintf(msg, sizeof msg, "%s: stacked rank %d exceeds the supported maximum %d", name, K,
             kCoreMaxK);
    return fail(HCORE_HIP_ERR_ARG, msg);
  }
  // argument checks above need no GPU; everything below runs on it
  if (ctx_ensure() != HCORE_HIP_OK) {
    fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: GPU initialisation failed");
    return;
  }
  SingleMem mem;
  if (fixed) {
    mem = *fixed;
  } else {
    mem.tc = thread_ctx();
    if (!mem.tc) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: per-thread GPU state failed");
    mem.stream = mem.tc->stream;
    mem.desc = &mem.tc->desc;
    mem.hres = mem.tc->hres;
  }
  if (mem.tc && work_bytes && work && is_device_ptr(work)) {  // the caller's device workspace
    mem.scratch = reinterpret_cast<char*>(work);
    mem.scratch_cap = work_bytes;
  }
  hipStream_t s = mem.stream;
  const bool dev = is_device_ptr(CU) && is_device_ptr(CV) && is_device_ptr(AU) &&
                   is_device_ptr(AV) && is_device_ptr(BU) && is_device_ptr(BV);
  if (!dev && !mem.tc)
    return fail(HCORE_HIP_ERR_ARG, "hcore_amd: this entry point takes device pointers only");
  hcore_hip_plan p;
  p.M = M;
  p.prec = sizeof(T) == 8 ? HCORE_HIP_FP64 : HCORE_HIP_FP32;
  p.ng = 1;
  p.ra = {ra}; p.rb = {rb}; p.rc = {rc};
  p.alpha = alpha; p.beta = beta; p.rk = rk; p.maxrk = maxrk; p.acc = acc;
  p.fin_stream = s;
  p.desc_host = mem.desc;
  if (mem.scratch) {
    p.fixed = mem.scratch;
    p.fixed_cap = mem.scratch_cap;
  } else {
    p.arena = &mem.tc->plan;
  }
  // host operands are staged in the thread's staging arena (M rows, packed)
  T *dAU = AU, *dAV = AV, *dBU = BU, *dBV = BV, *dCU = CU, *dCV = CV;
  T *oCU = nullptr, *oCV = nullptr;
  const size_t nA = (size_t)M * ra, nB = (size_t)M * rb, nC = (size_t)M * K;
  if (!dev) {
    // the six factors packed (ld M) into the pinned image, then one transfer;
    // the new factors come back through the same image
    const size_t in_elems = 2 * nA + 2 * nB + 2 * nC;
    const size_t bytes = sizeof(T) * (in_elems + 2 * nC) + 1024;
    if (mem.tc->stage.reserve(bytes) != HCORE_HIP_OK || mem.tc->hstage.reserve(bytes) != HCORE_HIP_OK)
      return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: staging allocation failed");
    T* q = reinterpret_cast<T*>(mem.tc->stage.p);
    dAU = q; q += nA; dAV = q; q += nA; dBU = q; q += nB; dBV = q; q += nB;
    dCU = q; q += nC; dCV = q; q += nC; oCU = q; q += nC; oCV = q;
    T* h = reinterpret_cast<T*>(mem.tc->hstage.p);
    auto pack = [&](T* dst, const T* src, int ld, int cols, int cap) {
      if (ld == M) memcpy(dst, src, sizeof(T) * (size_t)M * cols);
      else
        for (int c = 0; c < cols; ++c) memcpy(dst + (size_t)c * M, src + (size_t)c * ld, sizeof(T) * M);
      (void)cap;
    };
    pack(h, AU, LDA, ra, ra);
    pack(h + nA, AV, LDA, ra, ra);
    pack(h + 2 * nA, BU, LDB, rb, rb);
    pack(h + 2 * nA + nB, BV, LDB, rb, rb);
    pack(h + 2 * nA + 2 * nB, CU, LDC, rc, K);
    pack(h + 2 * nA + 2 * nB + nC, CV, LDC, rc, K);
    // columns rc..K-1 of the C blocks are not read; the whole image goes in one copy
    if (hipMemcpyAsync(dAU, h, sizeof(T) * in_elems, hipMemcpyHostToDevice, s) != hipSuccess)
      return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: host-to-device staging failed");
    p.lda = M; p.ldb = M; p.ldc = M; p.ldco = M;
    p.CUo = {oCU}; p.CVo = {oCV};
  } else {
    p.lda = LDA; p.ldb = LDB; p.ldc = LDC; p.ldco = LDC;
    p.CUo = {nullptr}; p.CVo = {nullptr};
  }
  p.AU = {dAU}; p.AV = {dAV}; p.BU = {dBU}; p.BV = {dBV}; p.CU = {dCU}; p.CV = {dCV};
  int st = plan_finalize_t<T>(&p);
  if (st == HCORE_HIP_ERR_ARG && mem.scratch && mem.tc) {  // caller's work too small: own arena
    p.fixed = nullptr;
    p.fixed_cap = 0;
    p.arena = &mem.tc->plan;
    st = plan_finalize_t<T>(&p);
  }
  if (st == HCORE_HIP_ERR_ARG && mem.scratch) {
    snprintf(msg, sizeof msg, "%s: M=%d, K=%d needs more than the %zu-byte workspace", name, M, K,
             mem.scratch_cap);
    return fail(HCORE_HIP_ERR_ARG, msg);
  }
  if (st == HCORE_HIP_OK) st = plan_execute_t<T>(&p, s);
  int* hres = mem.hres;
  T* hout = dev ? nullptr : reinterpret_cast<T*>(mem.tc->hstage.p) + 2 * nA + 2 * nB + 2 * nC;
  if (st == HCORE_HIP_OK) {
    if (hipMemcpyAsync(hres, p.d_rank, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(hres + 1, p.d_status, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess)
      st = HCORE_HIP_ERR_RUNTIME;
    // the new factors (at most K columns each) come back with the rank, in one sync
    if (st == HCORE_HIP_OK && !dev &&
        hipMemcpyAsync(hout, oCU, sizeof(T) * 2 * nC, hipMemcpyDeviceToHost, s) != hipSuccess)
      st = HCORE_HIP_ERR_RUNTIME;
    if (st == HCORE_HIP_OK && hipStreamSynchronize(s) != hipSuccess) st = HCORE_HIP_ERR_RUNTIME;
  }
  if (st != HCORE_HIP_OK) {
    (void)hipStreamSynchronize(s);
    return fail(st, "hcore_amd: GPU execution failed");
  }
  const int r = hres[0], dst = hres[1];
  if (dst == ST_R_GT_MAXRK) {
    snprintf(msg, sizeof msg, "%s: Rank after truncation is too big! finalrank:%d maxrank:%d", name,
             r, maxrk);
    return fail(HCORE_HIP_ERR_R_GT_MAXRK, msg);
  }
  if (dst != ST_OK) {
    snprintf(msg, sizeof msg, "%s: ERROR in the core SVD (non-finite values), K=%d", name, K);
    return fail(HCORE_HIP_ERR_NUMERIC, msg);
  }
  if (!dev) {  // oCV follows oCU in the staging arena: hout holds both
    for (int c = 0; c < r; ++c) {
      memcpy(CU + (size_t)c * LDC, hout + (size_t)c * M, sizeof(T) * M);
      memcpy(CV + (size_t)c * LDC, hout + nC + (size_t)c * M, sizeof(T) * M);
    }
  }
  if (is_device_ptr(Crk)) {
    T rv = (T)r;
    if (hipMemcpyAsync(Crk, &rv, sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: cannot write rank");
  } else {
    Crk[0] = (T)r;
  }
  if (flops) {
    // the reference's increments (hcore_dgemm.c:126, 179, 208, 224, 434, 493, 595, 666)
    unsigned long f = flop_counts('q', M, K, 0, 0) + flop_counts('m', ra, rb, M, 0) +
                      flop_counts('m', M, ra, rb, 0) + flop_counts('q', M, K, 0, 0) +
                      flop_counts('s', K, 0, 0, 0) + (unsigned long)r * (unsigned long)K +
                      flop_counts('o', M, r, K, 1) + flop_counts('o', r, M, K, 2);
    flops->update += f;
  }
}

// Run a short list of dependent GEMMs on the calling thread's stream (dsyrk,
// dgemmcd); the descriptors go to the thread's chain arena.
static int run_gemm_chain(ThreadCtx* tc, const std::vector<GemmProb<double>>& gp) {
  std::vector<TaskList> tl(gp.size());
  Packer pk;
  size_t o_gp = pk.add(gp);
  std::vector<TaskRef> o_t(gp.size());
  for (size_t i = 0; i < gp.size(); ++i) {
    add_tasks(tl[i], (int)i, gp[i]);
    o_t[i] = pk.add(tl[i]);
  }
  int st = tc->chain.reserve(pk.buf.size() + 256);
  if (st) return st;
  char* d = tc->chain.p;
  hipStream_t s = tc->stream;
  tc->desc.swap(pk.buf);
  HCHECK(hipMemcpyAsync(d, tc->desc.data(), tc->desc.size(), hipMemcpyHostToDevice, s));
  for (size_t i = 0; i < gp.size(); ++i) {
    const auto* probs = reinterpret_cast<const GemmProb<double>*>(d + o_gp);
    launch_tasks<double>(probs, d, o_t[i], s, true);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "gemm chain");
  return HCORE_HIP_OK;
}

// Device copies of host matrices (rows x cols, leading dimension ld) for the
// dense-output routines, carved from the thread's staging arena; `dev` true
// means the caller's pointers are used as they are.
struct Staged {
  struct Req {
    const double* h;
    int rows, ld, cols;
    bool upload;
  };
  std::vector<Req> req;
  int add(const double* h, int rows, int ld, int cols, bool upload) {
    req.push_back({h, rows, ld, std::max(cols, 1), upload});
    return (int)req.size() - 1;
  }
  // stage every request (dev: only the scratch ones); out[i] receives the
  // device address of request i
  int place(ThreadCtx* tc, bool dev, std::vector<double*>& out) {
    size_t total = 0;
    for (const Req& r : req)
      if (!(dev && r.h)) total += (((size_t)r.rows * r.cols * sizeof(double)) + 255) & ~size_t(255);
    int st = tc->stage.reserve(total + 256);
    if (st) return st;
    char* q = tc->stage.p;
    out.assign(req.size(), nullptr);
    for (size_t i = 0; i < req.size(); ++i) {
      const Req& r = req[i];
      if (dev && r.h) {
        out[i] = const_cast<double*>(r.h);
        continue;
      }
      out[i] = reinterpret_cast<double*>(q);
      q += (((size_t)r.rows * r.cols * sizeof(double)) + 255) & ~size_t(255);
      if (r.upload &&
          hipMemcpy2DAsync(out[i], sizeof(double) * r.rows, r.h, sizeof(double) * r.ld,
                           sizeof(double) * r.rows, r.cols, hipMemcpyHostToDevice,
                           tc->stream) != hipSuccess)
        return HCORE_HIP_ERR_RUNTIME;
    }
    return HCORE_HIP_OK;
  }
};

// One HCORE_dgemm on device pointers whose one-update workspace, descriptors
// and status words go to a fixed device buffer, on stream s: nothing is
// allocated on the device.  Fails (policy, status) when the update does not
// fit in ws_bytes.
void dgemm_fixed_workspace(int M, double alpha, double* AU, double* AV, double* Ark, int LDA,
                           double* BU, double* BV, double* Brk, int LDB, double beta, double* CU,
                           double* CV, double* Crk, int LDC, int rk, int maxrk, double acc,
                           flop_counter* flops, const char* name, char* ws, size_t ws_bytes,
                           hipStream_t s) {
  SingleMem mem;
  mem.scratch = ws;
  mem.scratch_cap = ws_bytes;
  mem.stream = s;
  std::vector<char> desc;  // host image of the descriptors, alive until the call returns
  int hres[2] = {0, 0};
  mem.desc = &desc;
  mem.hres = hres;
  single_gemm<double>(M, alpha, AU, AV, Ark, LDA, BU, BV, Brk, LDB, beta, CU, CV, Crk, LDC, rk,
                      maxrk, acc, flops, name, nullptr, 0, &mem);
}

void report_failure(int status, const char* msg) { fail(status, msg); }

int ensure_ready() { return ctx_ensure(); }
bool device_pointer(const void* ptr) { return is_device_ptr(ptr); }
void set_last_status(int status) { t_last_status = status; }
int core_variant_ctx(int K, size_t elem) {
  return core_variant_for(K, elem == 8 ? g_ctx.core_kl_d : g_ctx.core_kl_f, g_ctx.core_reg);
}
bool thread_res(ThreadRes* out) {
  ThreadCtx* tc = thread_ctx();
  if (!tc) return false;
  out->stream = tc->stream;
  out->work = &tc->tile_work;
  out->stage = &tc->stage;
  out->desc = &tc->tile_desc;
  out->pinned = &tc->hstage;
  out->hdesc = &tc->desc;
  return true;
}

}  // namespace hcore

using namespace hcore;

// ==============================================================================
// C API
// ==============================================================================
extern "C" {

void HCORE_dgemm(HCORE_enum transA, int transB, int M, int N, double alpha, double* AU,
                 double* AV, double* Ark, int LDA, double* BU, double* BV, double* Brk, int LDB,
                 double beta, double* CU, double* CV, double* Crk, int LDC, int rk, int maxrk,
                 double acc, double* work, flop_counter* flops) {
  (void)transA; (void)transB; (void)N;
  single_gemm<double>(M, alpha, AU, AV, Ark, LDA, BU, BV, Brk, LDB, beta, CU, CV, Crk, LDC, rk,
                      maxrk, acc, flops, "HCORE_dgemm", use_scratch ? work : nullptr,
                      reference_work_bytes(M, maxrk, Ark, Crk, sizeof(double)));
}

void HCORE_sgemm(HCORE_enum transA, HCORE_enum transB, int M, int N, float alpha, float* AU,
                 float* AV, float* Ark, int LDA, float* BU, float* BV, float* Brk, int LDB,
                 float beta, float* CU, float* CV, float* Crk, int LDC, int rk, int maxrk,
                 float acc, float* work, flop_counter* flops) {
  (void)transA; (void)transB; (void)N;
  single_gemm<float>(M, alpha, AU, AV, Ark, LDA, BU, BV, Brk, LDB, beta, CU, CV, Crk, LDC, rk,
                     maxrk, acc, flops, "HCORE_sgemm", use_scratch ? work : nullptr,
                     reference_work_bytes(M, maxrk, Ark, Crk, sizeof(float)));
}

void HCORE_dsyrk(HCORE_enum uplo, HCORE_enum trans, int M, int K, double alpha, const double* AU,
                 int LDAU, const double* AV, int LDAV, double beta, double* CD, int LDCD,
                 double* work, flop_counter* flops) {
  (void)uplo; (void)trans; (void)LDAV; (void)beta; (void)work;
  t_last_status = HCORE_HIP_OK;
  if (ctx_ensure() != HCORE_HIP_OK) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: GPU init failed");
  ThreadCtx* tc = thread_ctx();
  if (!tc) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: per-thread GPU state failed");
  if (M <= 0 || K <= 0 || LDAU < M || LDCD < M)
    return fail(HCORE_HIP_ERR_ARG, "hcore_amd: dsyrk with an empty tile or leading dimension below M");
  const int N = LDAU;  // rows of AV, as in the reference (hcore_dsyrk.c:57)
  const bool dev = is_device_ptr(AU) && is_device_ptr(AV) && is_device_ptr(CD);
  Staged sg;
  const int iAU = sg.add(AU, LDAU, LDAU, K, true), iAV = sg.add(AV, LDAU, LDAU, K, true);
  const int iCD = sg.add(CD, LDCD, LDCD, M, true);
  const int iW = sg.add(nullptr, std::max(K, 1), 0, K, false), iX = sg.add(nullptr, M, 0, K, false);
  std::vector<double*> d;
  if (sg.place(tc, dev, d) != HCORE_HIP_OK)
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: staging failed");
  std::vector<GemmProb<double>> gp(3);
  gp[0] = GemmProb<double>{d[iAV], d[iAV], d[iW], K, K, N, LDAU, LDAU, K, 1.0, 0.0,
                           GF_A_TRANS | GF_BETA_ZERO, nullptr, nullptr};
  gp[1] = GemmProb<double>{d[iAU], d[iW], d[iX], M, K, K, LDAU, K, M, 1.0, 0.0, GF_BETA_ZERO,
                           nullptr, nullptr};
  gp[2] = GemmProb<double>{d[iX], d[iAU], d[iCD], M, M, K, M, LDAU, LDCD, alpha, 1.0,
                           GF_B_TRANS | GF_LOWER_ZERO_UPPER, nullptr, nullptr};
  if (run_gemm_chain(tc, gp) != HCORE_HIP_OK)
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: dsyrk failed");
  if (!dev && (hipMemcpy2DAsync(CD, sizeof(double) * LDCD, d[iCD], sizeof(double) * LDCD,
                                sizeof(double) * LDCD, M, hipMemcpyDeviceToHost,
                                tc->stream) != hipSuccess ||
               hipStreamSynchronize(tc->stream) != hipSuccess))
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: dsyrk copy-back failed");
  if (flops)
    flops->syrk += flop_counts('m', K, K, N, 0) + flop_counts('m', M, K, K, 0) +
                   flop_counts('m', M, M, K, 0);
}

void HCORE_dgemmcd(HCORE_enum transA, HCORE_enum transB, int M, double alpha, double* Au,
                   double* Av, int Arank, int LDA, double* Bu, double* Bv, int Brank, int LDB,
                   double beta, double* CD, int LDC, int maxrank, double* work,
                   flop_counter* flops) {
  (void)transA; (void)transB; (void)alpha; (void)LDA; (void)LDB; (void)beta; (void)LDC;
  (void)maxrank; (void)work; (void)flops;
  t_last_status = HCORE_HIP_OK;
  if (ctx_ensure() != HCORE_HIP_OK) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: GPU init failed");
  ThreadCtx* tc = thread_ctx();
  if (!tc) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: per-thread GPU state failed");
  if (M <= 0 || Arank <= 0 || Brank <= 0)
    return fail(HCORE_HIP_ERR_ARG, "hcore_amd: dgemmcd with an empty tile or a rank of 0");
  const bool dev = is_device_ptr(Au) && is_device_ptr(Av) && is_device_ptr(Bu) &&
                   is_device_ptr(Bv) && is_device_ptr(CD);
  // every leading dimension is M (hcore_dgemmcd.c:66-100)
  Staged sg;
  const int iAu = sg.add(Au, M, M, Arank, true), iAv = sg.add(Av, M, M, Arank, true);
  const int iBu = sg.add(Bu, M, M, Brank, true), iBv = sg.add(Bv, M, M, Brank, true);
  const int iCD = sg.add(CD, M, M, M, true);
  const int iW = sg.add(nullptr, std::max(Arank, 1), 0, Brank, false);
  const int iX = sg.add(nullptr, M, 0, std::max(Arank, Brank), false);
  std::vector<double*> d;
  if (sg.place(tc, dev, d) != HCORE_HIP_OK)
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: staging failed");
  std::vector<GemmProb<double>> gp(3);
  gp[0] = GemmProb<double>{d[iAv], d[iBv], d[iW], Arank, Brank, M, M, M, Arank, 1.0, 0.0,
                           GF_A_TRANS | GF_BETA_ZERO, nullptr, nullptr};
  if (Arank > Brank) {  // X = Au W (M x Brank); CD -= X Bu^T
    gp[1] = GemmProb<double>{d[iAu], d[iW], d[iX], M, Brank, Arank, M, Arank, M, 1.0, 0.0,
                             GF_BETA_ZERO, nullptr, nullptr};
    gp[2] = GemmProb<double>{d[iX], d[iBu], d[iCD], M, M, Brank, M, M, M, -1.0, 1.0, GF_B_TRANS,
                             nullptr, nullptr};
  } else {  // X = W Bu^T (Arank x M); CD -= Au X
    gp[1] = GemmProb<double>{d[iW], d[iBu], d[iX], Arank, M, Brank, Arank, M, Arank, 1.0, 0.0,
                             GF_B_TRANS | GF_BETA_ZERO, nullptr, nullptr};
    gp[2] = GemmProb<double>{d[iAu], d[iX], d[iCD], M, M, Arank, M, Arank, M, -1.0, 1.0, 0,
                             nullptr, nullptr};
  }
  if (run_gemm_chain(tc, gp) != HCORE_HIP_OK)
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: dgemmcd failed");
  if (!dev && (hipMemcpyAsync(CD, d[iCD], sizeof(double) * (size_t)M * M, hipMemcpyDeviceToHost,
                              tc->stream) != hipSuccess ||
               hipStreamSynchronize(tc->stream) != hipSuccess))
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: dgemmcd copy-back failed");
}

// ---- dense neighbours of the recompression: potrf, trsm, dgemmbdcd, uncompress ----
}  // extern "C"

namespace hcore {

// HCORE_enum values are the CBLAS ones (reference hcore.h:30-44).
constexpr int kNoTrans = 111, kTrans = 112, kConjTrans = 113, kUpper = 121, kLower = 122,
              kNonUnit = 131, kUnit = 132, kLeft = 141, kRight = 142;

// A list of GEMM steps packed once, uploaded once to the thread's chain
// arena, then launched one step at a time between other kernels.
struct GemmSteps {
  std::vector<GemmProb<double>> gp;
  std::vector<TaskList> tl;
  std::vector<TaskRef> ref;
  size_t o_gp = 0, o_extra = 0;
  char* dev = nullptr;
  int add(const GemmProb<double>& g) {
    gp.push_back(g);
    tl.emplace_back();
    add_tasks(tl.back(), (int)gp.size() - 1, g);
    return (int)gp.size() - 1;
  }
  void pack(Packer& pk) {
    o_gp = pk.add(gp);
    ref.resize(gp.size());
    for (size_t i = 0; i < gp.size(); ++i) ref[i] = pk.add(tl[i]);
    o_extra = (pk.buf.size() + 255) & ~size_t(255);
  }
  // Reserve the thread's chain arena for the descriptors plus `extra` bytes
  // of device scratch after them (256-aligned); the addresses are final
  // from here on, so descriptors may point into the scratch.
  int reserve(ThreadCtx* tc, size_t extra) {
    Packer pk;
    pack(pk);
    int st = tc->chain.reserve(o_extra + extra + 256);
    if (st) return st;
    dev = tc->chain.p;
    return HCORE_HIP_OK;
  }
  // Pack and copy the descriptors (after reserve) on the thread's stream; the
  // host image stays alive in the thread state until the stream syncs.
  int upload(ThreadCtx* tc) {
    Packer pk;
    pack(pk);
    tc->desc.swap(pk.buf);
    if (!tc->desc.empty())
      HCHECK(hipMemcpyAsync(dev, tc->desc.data(), tc->desc.size(), hipMemcpyHostToDevice,
                            tc->stream));
    return HCORE_HIP_OK;
  }
  char* extra() const { return dev + o_extra; }
  void launch(int i, hipStream_t s) const {
    launch_tasks<double>(reinterpret_cast<const GemmProb<double>*>(dev + o_gp), dev, ref[i], s,
                         true);
  }
};

// Blocked triangular solve on device memory: op(A) X = B (left) or
// X op(A) = B (right), B (M x N) overwritten by X, after B = alpha B.  The
// diagonal kTrsmNB blocks are solved by k_trsm_diag, the rest of B is updated
// by GEMMs (hcore_dtrsm.c:24-48 calls cblas_dtrsm).  *skip (device) != 0
// skips every kernel.  The GEMM steps go through `steps` (already populated
// steps are kept; the caller uploads and the solve appends its own).
struct TrsmPlan {
  TrsmDiagProb<double> d{};
  std::vector<int> order;   // diagonal blocks in solve order
  std::vector<int> gstep;   // GEMM step after each block (-1: none)
};

static TrsmPlan trsm_plan(GemmSteps& steps, bool left, bool lowerA, bool transA, bool unit, int M,
                          int N, const double* A, int lda, double* B, int ldb, const int* skip) {
  TrsmPlan t;
  const bool Tlow = lowerA != transA;  // op(A) lower triangular
  const int n = left ? M : N;          // order of op(A)
  t.d.A = A;
  t.d.lda = lda;
  t.d.B = B;
  t.d.alpha = 1.0;
  t.d.skip = skip;
  t.d.unit = unit;
  t.d.nrhs = left ? N : M;
  if (left) {
    t.d.es = 1; t.d.cs = ldb; t.d.trans = transA; t.d.lower = Tlow;
  } else {  // x^T T = b^T  <=>  T^T x = b for every row of B
    t.d.es = ldb; t.d.cs = 1; t.d.trans = !transA; t.d.lower = !Tlow;
  }
  const bool fwd = t.d.lower;
  const int nblk = (n + kTrsmNB - 1) / kTrsmNB;
  for (int idx = 0; idx < nblk; ++idx) {
    const int b = fwd ? idx : nblk - 1 - idx;
    const int b0 = b * kTrsmNB, nbk = std::min(kTrsmNB, n - b0);
    const int lo = fwd ? b0 + nbk : 0, hi = fwd ? n : b0;  // the part still to update
    t.order.push_back(b);
    if (hi <= lo) {
      t.gstep.push_back(-1);
      continue;
    }
    GemmProb<double> g{};
    g.alpha = -1.0;
    g.beta = 1.0;
    g.skip = skip;
    if (left) {  // B[lo:hi, :] -= op(A)[lo:hi, b] X_b
      g.C = B + lo;
      g.ldc = ldb;
      g.m = hi - lo; g.n = N; g.k = nbk;
      g.A = transA ? A + b0 + (size_t)lo * lda : A + lo + (size_t)b0 * lda;
      g.lda = lda;
      g.flags = transA ? GF_A_TRANS : 0;
      g.B = B + b0;
      g.ldb = ldb;
    } else {  // B[:, lo:hi] -= X_b op(A)[b, lo:hi]
      g.C = B + (size_t)lo * ldb;
      g.ldc = ldb;
      g.m = M; g.n = hi - lo; g.k = nbk;
      g.A = B + (size_t)b0 * ldb;
      g.lda = ldb;
      g.B = transA ? A + lo + (size_t)b0 * lda : A + b0 + (size_t)lo * lda;
      g.ldb = lda;
      g.flags = transA ? GF_B_TRANS : 0;
    }
    t.gstep.push_back(steps.add(g));
  }
  return t;
}

static void trsm_run(const TrsmPlan& t, const GemmSteps& steps, int n, hipStream_t s) {
  for (size_t i = 0; i < t.order.size(); ++i) {
    TrsmDiagProb<double> d = t.d;
    d.b0 = t.order[i] * kTrsmNB;
    d.nbk = std::min(kTrsmNB, n - d.b0);
    launch_trsm_diag<double>(d, s);
    if (t.gstep[i] >= 0) steps.launch(t.gstep[i], s);
  }
}

}  // namespace hcore

extern "C" {

void HCORE_dtrsm(HCORE_enum side, HCORE_enum uplo, HCORE_enum transA, HCORE_enum diag, int M,
                 int Brank, double alpha, double* A, int LDA, double* BV, int LDBV,
                 flop_counter* flops) {
  t_last_status = HCORE_HIP_OK;
  if (side != kLeft && side != kRight)  // the reference asserts (hcore_dtrsm.c:45-47)
    return fail(HCORE_HIP_ERR_ARG, "HCORE_dtrsm: side is not CblasLeft or CblasRight");
  if ((uplo != kUpper && uplo != kLower) ||
      (transA != kNoTrans && transA != kTrans && transA != kConjTrans) ||
      (diag != kNonUnit && diag != kUnit) || M < 0 || Brank < 0)
    return fail(HCORE_HIP_ERR_ARG, "HCORE_dtrsm: invalid argument");
  const int n = side == kLeft ? M : Brank;  // order of A
  if (LDA < std::max(1, n) || LDBV < std::max(1, M))
    return fail(HCORE_HIP_ERR_ARG, "HCORE_dtrsm: leading dimension too small");
  if (flops) flops->trsm += flop_counts('t', M, Brank, side == kLeft ? 1 : 2, 0);
  if (M == 0 || Brank == 0) return;
  if (ctx_ensure() != HCORE_HIP_OK) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: GPU init failed");
  ThreadCtx* tc = thread_ctx();
  if (!tc) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: per-thread GPU state failed");
  const bool dev = is_device_ptr(A) && is_device_ptr(BV);
  Staged sg;
  const int iA = sg.add(A, LDA, LDA, n, true), iB = sg.add(BV, LDBV, LDBV, Brank, true);
  std::vector<double*> d;
  if (sg.place(tc, dev, d) != HCORE_HIP_OK)
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: staging failed");
  hipStream_t s = tc->stream;
  if (alpha != 1.0) launch_scale<double>(d[iB], LDBV, M, Brank, alpha, s);
  if (alpha != 0.0) {
    GemmSteps steps;
    TrsmPlan t = trsm_plan(steps, side == kLeft, uplo == kLower, transA != kNoTrans,
                           diag == kUnit, M, Brank, d[iA], LDA, d[iB], LDBV, nullptr);
    if (steps.reserve(tc, 0) != HCORE_HIP_OK || steps.upload(tc) != HCORE_HIP_OK)
      return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: dtrsm descriptors failed");
    trsm_run(t, steps, n, s);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && !dev)
    e = hipMemcpy2DAsync(BV, sizeof(double) * LDBV, d[iB], sizeof(double) * LDBV,
                         sizeof(double) * M, Brank, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: dtrsm failed");
}

void HCORE_dpotrf(HCORE_enum uplo, int N, double* A, int LDA, int* info, flop_counter* flops) {
  t_last_status = HCORE_HIP_OK;
  // the reference passes the enum to LAPACKE as the uplo character
  // (hcore_dpotrf.c:29-32), which LAPACK rejects; the intended triangle is
  // taken here, from the enum or from the LAPACK character
  const bool upper = uplo == kUpper || uplo == 'U' || uplo == 'u';
  const bool lower = uplo == kLower || uplo == 'L' || uplo == 'l';
  if (!upper && !lower) {
    if (info) *info = -1;
    return;
  }
  if (N < 0 || LDA < std::max(1, N)) {
    if (info) *info = N < 0 ? -2 : -4;
    return;
  }
  if (flops) flops->potrf += flop_counts('c', N, 0, 0, 0);
  if (info) *info = 0;
  if (N == 0) return;
  if (ctx_ensure() != HCORE_HIP_OK) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: GPU init failed");
  ThreadCtx* tc = thread_ctx();
  if (!tc) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: per-thread GPU state failed");
  const bool dev = is_device_ptr(A);
  Staged sg;
  const int iA = sg.add(A, LDA, LDA, N, true);
  std::vector<double*> d;
  if (sg.place(tc, dev, d) != HCORE_HIP_OK)
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: staging failed");
  double* dA = d[iA];
  hipStream_t s = tc->stream;
  // right-looking blocked Cholesky: factor the diagonal block, solve the
  // panel against it, update the trailing triangle (one GEMM per block)
  GemmSteps steps;
  std::vector<int> tstep;
  const int nblk = (N + kTrsmNB - 1) / kTrsmNB;
  for (int b = 0; b < nblk; ++b) {
    const int b0 = b * kTrsmNB, nbk = std::min(kTrsmNB, N - b0), r1 = b0 + nbk;
    if (r1 >= N) {
      tstep.push_back(-1);
      continue;
    }
    GemmProb<double> g{};
    g.alpha = -1.0;
    g.beta = 1.0;
    g.m = g.n = N - r1;
    g.k = nbk;
    g.C = dA + r1 + (size_t)r1 * LDA;
    g.ldc = g.lda = g.ldb = LDA;
    if (lower) {  // A[r1:, r1:] -= L[r1:, b] L[r1:, b]^T, lower triangle
      g.A = g.B = dA + r1 + (size_t)b0 * LDA;
      g.flags = GF_B_TRANS | GF_LOWER_ONLY;
    } else {  // A[r1:, r1:] -= U[b, r1:]^T U[b, r1:], upper triangle
      g.A = g.B = dA + b0 + (size_t)r1 * LDA;
      g.flags = GF_A_TRANS | GF_UPPER_ONLY;
    }
    tstep.push_back(steps.add(g));
  }
  if (steps.reserve(tc, sizeof(int)) != HCORE_HIP_OK)
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: dpotrf descriptors failed");
  int* dinfo = reinterpret_cast<int*>(steps.extra());  // device info word after the descriptors
  for (auto& g : steps.gp) g.skip = dinfo;
  if (steps.upload(tc) != HCORE_HIP_OK)
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: dpotrf descriptors failed");
  hipError_t e = hipMemsetAsync(dinfo, 0, sizeof(int), s);
  for (int b = 0; b < nblk && e == hipSuccess; ++b) {
    const int b0 = b * kTrsmNB, nbk = std::min(kTrsmNB, N - b0), r1 = b0 + nbk;
    launch_potf2<double>(dA, LDA, b0, nbk, upper ? 1 : 0, dinfo, s);
    if (r1 >= N) break;
    TrsmDiagProb<double> pd{};
    pd.A = dA; pd.lda = LDA; pd.alpha = 1.0; pd.skip = dinfo; pd.unit = 0;
    pd.b0 = b0; pd.nbk = nbk; pd.nrhs = N - r1; pd.lower = 1;
    if (lower) {  // X L_bb^T = A[r1:, b]: every row of the panel is a right-hand side
      pd.B = dA + r1; pd.es = LDA; pd.cs = 1; pd.trans = 0;
    } else {      // U_bb^T X = A[b, r1:]: every column of the panel
      pd.B = dA + (size_t)r1 * LDA; pd.es = 1; pd.cs = LDA; pd.trans = 1;
    }
    launch_trsm_diag<double>(pd, s);
    steps.launch(tstep[b], s);
    e = hipGetLastError();
  }
  int* hres = tc->hres;
  if (e == hipSuccess) e = hipMemcpyAsync(hres, dinfo, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && !dev)
    e = hipMemcpy2DAsync(A, sizeof(double) * LDA, dA, sizeof(double) * LDA, sizeof(double) * N, N,
                         hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: dpotrf failed");
  if (info) *info = hres[0];
}

void HCORE_dgemmbdcd(HCORE_enum transA, HCORE_enum transB, int M, int N, double alpha,
                     double* AU, double* AV, double* Ark, int LDA, double* BD, int LDB, double beta,
                     double* CD, int LDC, double* work, flop_counter* flops) {
  (void)transB; (void)work;
  t_last_status = HCORE_HIP_OK;
  double rk = 0;
  if (!read_rank(Ark, &rk)) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: cannot read rank");
  const int ark = (int)rk;
  // the reference's two branches (hcore_dgemmbdcd.c:76-116): inner size of
  // AVBD = AV^T BD and row count of CD = alpha AU AVBD + beta CD
  const int k1 = transA == kTrans ? M : LDB;
  const int m2 = transA == kTrans ? LDA : M;
  if (ark <= 0 || M <= 0 || N <= 0 || LDA < std::max(k1, m2) || LDB < k1 || LDC < m2)
    return fail(HCORE_HIP_ERR_ARG, "HCORE_dgemmbdcd: invalid rank or dimensions");
  if (ctx_ensure() != HCORE_HIP_OK) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: GPU init failed");
  ThreadCtx* tc = thread_ctx();
  if (!tc) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: per-thread GPU state failed");
  const bool dev = is_device_ptr(AU) && is_device_ptr(AV) && is_device_ptr(BD) && is_device_ptr(CD);
  Staged sg;
  const int iAU = sg.add(AU, LDA, LDA, ark, true), iAV = sg.add(AV, LDA, LDA, ark, true);
  const int iBD = sg.add(BD, LDB, LDB, N, true), iCD = sg.add(CD, LDC, LDC, N, beta != 0.0);
  const int iW = sg.add(nullptr, ark, 0, N, false);
  std::vector<double*> d;
  if (sg.place(tc, dev, d) != HCORE_HIP_OK)
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: staging failed");
  std::vector<GemmProb<double>> gp(2);
  gp[0] = GemmProb<double>{d[iAV], d[iBD], d[iW], ark, N, k1, LDA, LDB, ark, 1.0, 0.0,
                           GF_A_TRANS | GF_BETA_ZERO, nullptr, nullptr};
  gp[1] = GemmProb<double>{d[iAU], d[iW], d[iCD], m2, N, ark, LDA, ark, LDC, alpha, beta,
                           beta == 0.0 ? GF_BETA_ZERO : 0, nullptr, nullptr};
  if (run_gemm_chain(tc, gp) != HCORE_HIP_OK)
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: dgemmbdcd failed");
  if (!dev && (hipMemcpy2DAsync(CD, sizeof(double) * LDC, d[iCD], sizeof(double) * LDC,
                                sizeof(double) * m2, N, hipMemcpyDeviceToHost,
                                tc->stream) != hipSuccess ||
               hipStreamSynchronize(tc->stream) != hipSuccess))
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: dgemmbdcd copy-back failed");
  if (flops)
    flops->update += flop_counts('m', ark, N, k1, 0) + flop_counts('m', m2, N, ark, 0);
}

void HCORE_duncompress(HCORE_enum transA, HCORE_enum transB, int M, int N, double alpha,
                       double* AU, double* Ark, int LDA, double* BV, double* Brk, int LDB,
                       double beta, double* CD, int LDC) {
  (void)transA; (void)transB; (void)alpha; (void)beta;
  t_last_status = HCORE_HIP_OK;
  double ra = 0, rb = 0;
  if (!read_rank(Ark, &ra) || !read_rank(Brk, &rb))
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: cannot read rank");
  const int ark = (int)ra;
  if ((int)ra != (int)rb)  // the reference asserts (hcore_duncompress.c:67)
    return fail(HCORE_HIP_ERR_ARG, "HCORE_duncompress: the two ranks must be equal");
  if (ark <= 0 || M <= 0 || N <= 0 || LDA < M || LDB < N || LDC < M)
    return fail(HCORE_HIP_ERR_ARG, "HCORE_duncompress: invalid rank or dimensions");
  if (ctx_ensure() != HCORE_HIP_OK) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: GPU init failed");
  ThreadCtx* tc = thread_ctx();
  if (!tc) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: per-thread GPU state failed");
  const bool dev = is_device_ptr(AU) && is_device_ptr(BV) && is_device_ptr(CD);
  Staged sg;
  const int iAU = sg.add(AU, LDA, LDA, ark, true), iBV = sg.add(BV, LDB, LDB, ark, true);
  const int iCD = sg.add(CD, LDC, LDC, N, false);
  std::vector<double*> d;
  if (sg.place(tc, dev, d) != HCORE_HIP_OK)
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: staging failed");
  std::vector<GemmProb<double>> gp(1);
  gp[0] = GemmProb<double>{d[iAU], d[iBV], d[iCD], M, N, ark, LDA, LDB, LDC, 1.0, 0.0,
                           GF_B_TRANS | GF_BETA_ZERO, nullptr, nullptr};
  if (run_gemm_chain(tc, gp) != HCORE_HIP_OK)
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: duncompress failed");
  if (!dev && (hipMemcpy2DAsync(CD, sizeof(double) * LDC, d[iCD], sizeof(double) * LDC,
                                sizeof(double) * M, N, hipMemcpyDeviceToHost,
                                tc->stream) != hipSuccess ||
               hipStreamSynchronize(tc->stream) != hipSuccess))
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: duncompress copy-back failed");
}

// ---- two-phase recompression (hcore_dgemm.c:1022-1699) --------------------------
}  // extern "C"

namespace hcore {

// One update split at the rank reveal: phase 1 (stacking, both QRs, the core
// SVD and truncation, Zv) leaves its state on the device; phase 2 (the
// reassembly) finds it again through the address it handed out as newU.  The
// state owns its device memory (from a pool shared by all threads), so the
// caller may run other HCORE calls, on any thread, between the two phases.
struct PendingGemm {
  hcore_hip_plan plan;
  DevArena* mem = nullptr;
  std::vector<char> desc;
  hipStream_t stream = nullptr;
  int M = 0, ra = 0, rb = 0, rc = 0, K = 0, r = 0;
  double* oCU = nullptr;  // the new factors, M x K (ld M), written by phase 2
  double* oCV = nullptr;
  bool bdense = false;
};
static std::mutex g_pending_mu;
static std::vector<std::pair<const void*, PendingGemm*>> g_pending;
static std::vector<DevArena*> g_arena_pool;

static DevArena* pool_take() {
  std::lock_guard<std::mutex> lk(g_pending_mu);
  if (g_arena_pool.empty()) return new DevArena();
  DevArena* a = g_arena_pool.back();
  g_arena_pool.pop_back();
  return a;
}
static void pool_give(DevArena* a) {
  std::lock_guard<std::mutex> lk(g_pending_mu);
  g_arena_pool.push_back(a);
}
static void pending_release(PendingGemm* pg) {
  if (pg->mem) pool_give(pg->mem);
  delete pg;
}

// Phase 1 of the split recompression; B_dense != nullptr selects G = B AV.
static void split_phase1(int M, double alpha, double* AU, double* AV, int* Ark, int LDA,
                         double* BU, double* BV, int* Brk, double* B_dense, int LDB, double beta,
                         double* CU, double* CV, int* Crk, int LDC, int rk, int maxrk, double acc,
                         double* work, flop_counter* flops, double** p_work_new, double** p_CU,
                         double** p_CV, int* p_CU_ncols, int* p_new_UVrk, double** p_newU,
                         int* p_ld_newU, double** p_qrtauA, int* p_CV_ncols, double** p_newV,
                         int* p_ld_newV, double** p_qrtauB, int* p_use_clone, double** p_CUclone,
                         int* p_ld_CUclone, double** p_CU_save, double** p_CVclone,
                         int* p_ld_CVclone, double** p_CV_save, const char* name) {
  t_last_status = HCORE_HIP_OK;
  char msg[512];
  const bool bd = B_dense != nullptr;
  int rks[3] = {0, 0, 0};
  auto rd = [](const int* p, int* out) {
    if (!is_device_ptr(p)) {
      *out = *p;
      return true;
    }
    return hipMemcpy(out, p, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess;
  };
  if (!rd(Ark, &rks[0]) || (!bd && !rd(Brk, &rks[1])) || !rd(Crk, &rks[2]))
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: cannot read rank");
  const int ra = rks[0], rb = bd ? 0 : rks[1], rc = rks[2];
  if (ra <= 0 || rc <= 0 || (!bd && rb <= 0)) {
    snprintf(msg, sizeof msg, "%s: _Ark=%d _Crk=%d. These rank values should not be zero.", name,
             ra, rc);
    return fail(HCORE_HIP_ERR_ZERO_RANK, msg);
  }
  const int K = ra + rc;
  if (M <= 0 || LDA < M || LDB < M || LDC < M)
    return fail(HCORE_HIP_ERR_ARG, "hcore_amd: size or leading dimension below M");
  if (K > 2 * maxrk) {
    snprintf(msg, sizeof msg, "%s: Sum of ranks (%d) is too big! _Ark:%d _Crk:%d maxrank:%d (x2: %d)",
             name, K, ra, rc, maxrk, 2 * maxrk);
    return fail(HCORE_HIP_ERR_SUM_RANKS, msg);
  }
  if (K > M || K > kCoreMaxK) {
    snprintf(msg, sizeof msg, "%s: stacked rank %d exceeds the tile size %d or the maximum %d", name,
             K, M, kCoreMaxK);
    return fail(HCORE_HIP_ERR_K_GT_MAXRK, msg);
  }
  if (ctx_ensure() != HCORE_HIP_OK) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: GPU init failed");
  ThreadCtx* tc = thread_ctx();
  if (!tc) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: per-thread GPU state failed");
  hipStream_t s = tc->stream;
  PendingGemm* pg = new PendingGemm();
  pg->M = M; pg->ra = ra; pg->rb = rb; pg->rc = rc; pg->K = K; pg->stream = s; pg->bdense = bd;
  hcore_hip_plan& p = pg->plan;
  p.M = M;
  p.prec = HCORE_HIP_FP64;
  p.ng = 1;
  p.ra = {ra}; p.rb = {bd ? 1 : rb}; p.rc = {rc};
  p.alpha = alpha; p.beta = beta; p.rk = rk; p.maxrk = maxrk; p.acc = acc;
  p.fin_stream = s;
  p.desc_host = &pg->desc;
  p.lda = M; p.ldb = M; p.ldc = M; p.ldco = M; p.ldbd = M;
  // inputs are copied (device or host) into the state's own memory, ld = M:
  // the caller's buffers are free to change between the phases
  const size_t nA = (size_t)M * ra, nB = bd ? (size_t)M * M : (size_t)M * rb, nC = (size_t)M * K;
  const size_t stage = sizeof(double) * (2 * nA + 2 * nB + 4 * nC) + 2048;
  p.measure = true;
  p.AU = {nullptr}; p.AV = {nullptr}; p.BU = {nullptr}; p.BV = {nullptr};
  p.CU = {nullptr}; p.CV = {nullptr}; p.CUo = {nullptr}; p.CVo = {nullptr};
  if (bd) p.BD = {nullptr};
  if (plan_finalize_t<double>(&p) != HCORE_HIP_OK) {
    pending_release(pg);
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: split plan sizing failed");
  }
  const size_t plan_bytes = (p.dbytes + 255) & ~size_t(255);
  pg->mem = pool_take();
  if (pg->mem->reserve(plan_bytes + stage) != HCORE_HIP_OK) {
    pending_release(pg);
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: split workspace allocation failed");
  }
  double* q = reinterpret_cast<double*>(pg->mem->p + plan_bytes);
  double *dAU = q, *dAV = q + nA; q += 2 * nA;
  double *dBU = q, *dBV = bd ? q : q + nB; q += bd ? nB : 2 * nB;
  double *dCU = q, *dCV = q + nC; q += 2 * nC;
  pg->oCU = q; pg->oCV = q + nC;
  auto up = [&](double* d, const double* h, int ld, int cols) {
    return hipMemcpy2DAsync(d, sizeof(double) * M, h, sizeof(double) * ld, sizeof(double) * M,
                            cols, hipMemcpyDefault, s);
  };
  hipError_t e = up(dAU, AU, LDA, ra);
  if (!e) e = up(dAV, AV, LDA, ra);
  if (!e) e = bd ? up(dBU, B_dense, LDB, M) : up(dBU, BU, LDB, rb);
  if (!e && !bd) e = up(dBV, BV, LDB, rb);
  if (!e) e = up(dCU, CU, LDC, rc);
  if (!e) e = up(dCV, CV, LDC, rc);
  if (e) {
    (void)hipStreamSynchronize(s);
    pending_release(pg);
    return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: split staging failed");
  }
  p.measure = false;
  p.fixed = pg->mem->p;
  p.fixed_cap = plan_bytes;
  p.AU = {dAU}; p.AV = {dAV}; p.CU = {dCU}; p.CV = {dCV};
  p.CUo = {pg->oCU}; p.CVo = {pg->oCV};
  if (bd) {
    p.BD = {dBU};
    p.BU = {dBU}; p.BV = {dBU};  // not read
  } else {
    p.BU = {dBU}; p.BV = {dBV};
  }
  int st = plan_finalize_t<double>(&p);
  if (st == HCORE_HIP_OK) st = plan_execute_t<double>(&p, s, false, 1);
  int* hres = tc->hres;
  if (st == HCORE_HIP_OK &&
      (hipMemcpyAsync(hres, p.d_rank, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
       hipMemcpyAsync(hres + 1, p.d_status, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess))
    st = HCORE_HIP_ERR_RUNTIME;
  if (st != HCORE_HIP_OK) {
    (void)hipStreamSynchronize(s);
    pending_release(pg);
    return fail(st, "hcore_amd: GPU execution failed");
  }
  const int r = hres[0], dst = hres[1];
  if (dst == ST_R_GT_MAXRK) {
    pending_release(pg);
    snprintf(msg, sizeof msg, "%s: Rank after truncation is too big! finalrank:%d maxrank:%d", name,
             r, maxrk);
    return fail(HCORE_HIP_ERR_R_GT_MAXRK, msg);
  }
  if (dst != ST_OK) {
    pending_release(pg);
    snprintf(msg, sizeof msg, "%s: ERROR in the core SVD (non-finite values), K=%d", name, K);
    return fail(HCORE_HIP_ERR_NUMERIC, msg);
  }
  pg->r = r;
  // the workspace layout puts Zu, Zv and the taus at fixed offsets; hand out
  // their device addresses in the reference's out-parameters
  double* ws = reinterpret_cast<double*>(p.ws);
  double* Zu = ws + p.first_off[4];
  double* Zv = ws + p.first_off[5];
  double *Us = ws + p.first_off[0], *Vs = ws + p.first_off[1];
  double *tauA = ws + p.first_off[2], *tauB = ws + p.first_off[3];
  {
    std::lock_guard<std::mutex> lk(g_pending_mu);
    g_pending.push_back({Zu, pg});
  }
  if (flops) {  // hcore_dgemm.c:126, 179, 208, 224, 434, 493 (b_dense: 263, 287)
    unsigned long f = flop_counts('q', M, K, 0, 0) + flop_counts('q', M, K, 0, 0) +
                      flop_counts('s', K, 0, 0, 0) + (unsigned long)r * (unsigned long)K;
    f += bd ? flop_counts('m', M, ra, M, 0)
            : flop_counts('m', ra, rb, M, 0) + flop_counts('m', M, ra, rb, 0);
    flops->update += f;
  }
  *p_work_new = work;
  *p_CU = Us; *p_CV = Vs;
  *p_CU_ncols = K; *p_CV_ncols = K;
  *p_new_UVrk = r;
  *p_newU = Zu; *p_ld_newU = K;
  *p_newV = Zv; *p_ld_newV = K;
  *p_qrtauA = tauA; *p_qrtauB = tauB;
  *p_use_clone = 1;
  *p_CUclone = Us; *p_ld_CUclone = M;
  *p_CVclone = Vs; *p_ld_CVclone = M;
  *p_CU_save = CU; *p_CV_save = CV;
}

}  // namespace hcore

extern "C" {

void HCORE_dgemm_qr_svd(int transA, int transB, int M, int N, double alpha, double* AU, double* AV,
                        int* Ark, int LDA, double* BU, double* BV, int* Brk, int LDB, double beta,
                        double* CU, double* CV, int* Crk, int LDC, int rk, int storage_maxrank,
                        int maxrk, double acc, double* work, flop_counter* flops,
                        double** _p_work_new, double** _p__CU, double** _p__CV, int* _p_CU_ncols,
                        int* _p_new_UVrk, double** _p_newU, int* _p_ld_newU, double** _p_qrtauA,
                        int* _p_CV_ncols, double** _p_newV, int* _p_ld_newV, double** _p_qrtauB,
                        int* _p_use_CUV_clone, double** _p_CUclone, int* _p_ld_CUclone,
                        double** _p__CU_save, double** _p_CVclone, int* _p_ld_CVclone,
                        double** _p__CV_save) {
  (void)transA; (void)transB; (void)N; (void)storage_maxrank;
  split_phase1(M, alpha, AU, AV, Ark, LDA, BU, BV, Brk, nullptr, LDB, beta, CU, CV, Crk, LDC, rk,
               maxrk, acc, work, flops, _p_work_new, _p__CU, _p__CV, _p_CU_ncols, _p_new_UVrk,
               _p_newU, _p_ld_newU, _p_qrtauA, _p_CV_ncols, _p_newV, _p_ld_newV, _p_qrtauB,
               _p_use_CUV_clone, _p_CUclone, _p_ld_CUclone, _p__CU_save, _p_CVclone,
               _p_ld_CVclone, _p__CV_save, "HCORE_dgemm_qr_svd");
}

void HCORE_dgemm_qr_svd_b_dense(int transA, int transB, int M, int N, double alpha, double* AU,
                                double* AV, int* Ark, int LDA, double* B, int LDB, double beta,
                                double* CU, double* CV, int* Crk, int LDC, int rk,
                                int storage_maxrank, int maxrk, double acc, double* work,
                                flop_counter* flops, double** _p_work_new, double** _p__CU,
                                double** _p__CV, int* _p_CU_ncols, int* _p_new_UVrk,
                                double** _p_newU, int* _p_ld_newU, double** _p_qrtauA,
                                int* _p_CV_ncols, double** _p_newV, int* _p_ld_newV,
                                double** _p_qrtauB, int* _p_use_CUV_clone, double** _p_CUclone,
                                int* _p_ld_CUclone, double** _p__CU_save, double** _p_CVclone,
                                int* _p_ld_CVclone, double** _p__CV_save) {
  (void)transA; (void)transB; (void)N; (void)storage_maxrank;
  if (!B) return fail(HCORE_HIP_ERR_ARG, "HCORE_dgemm_qr_svd_b_dense: B is NULL");
  split_phase1(M, alpha, AU, AV, Ark, LDA, nullptr, nullptr, nullptr, B, LDB, beta, CU, CV, Crk,
               LDC, rk, maxrk, acc, work, flops, _p_work_new, _p__CU, _p__CV, _p_CU_ncols,
               _p_new_UVrk, _p_newU, _p_ld_newU, _p_qrtauA, _p_CV_ncols, _p_newV, _p_ld_newV,
               _p_qrtauB, _p_use_CUV_clone, _p_CUclone, _p_ld_CUclone, _p__CU_save, _p_CVclone,
               _p_ld_CVclone, _p__CV_save, "HCORE_dgemm_qr_svd_b_dense");
}

void HCORE_dgemm_ormqr(int transA, int transB, int M, int N, double alpha, double* AU, double* AV,
                       int* Ark, int LDA, double beta, double* CU, double* CV, int* Crk, int LDC,
                       int rk, int storage_maxrank, int maxrk, double acc, double* work,
                       flop_counter* flops, double* _CU, double* _CV, int CU_ncols, int new_UVrk,
                       double* newU, int ld_newU, double* qrtauA, int CV_ncols, double* newV,
                       int ld_newV, double* qrtauB, int use_CUV_clone, double* CUclone,
                       int ld_CUclone, double* _CU_save, double* CVclone, int ld_CVclone,
                       double* _CV_save) {
  (void)transA; (void)transB; (void)N; (void)alpha; (void)AU; (void)AV; (void)Ark; (void)LDA;
  (void)beta; (void)CU; (void)CV; (void)rk; (void)maxrk; (void)acc; (void)work; (void)_CU;
  (void)_CV; (void)CU_ncols; (void)ld_newU; (void)qrtauA; (void)CV_ncols; (void)newV;
  (void)ld_newV; (void)qrtauB; (void)use_CUV_clone; (void)CUclone; (void)ld_CUclone;
  (void)CVclone; (void)ld_CVclone;
  t_last_status = HCORE_HIP_OK;
  char msg[512];
  PendingGemm* pg = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_pending_mu);
    for (size_t i = 0; i < g_pending.size(); ++i)
      if (g_pending[i].first == newU) {
        pg = g_pending[i].second;
        g_pending.erase(g_pending.begin() + i);
        break;
      }
  }
  if (!pg)
    return fail(HCORE_HIP_ERR_ARG, "HCORE_dgemm_ormqr: newU does not come from HCORE_dgemm_qr_svd");
  if (new_UVrk > storage_maxrank) {  // hcore_dgemm.c:1648-1651
    snprintf(msg, sizeof msg,
             "HCORE_dgemm_ormqr: Increase storage maxrank %d. Observed rank for U and V is %d",
             storage_maxrank, new_UVrk);
    pending_release(pg);
    return fail(HCORE_HIP_ERR_R_GT_MAXRK, msg);
  }
  const int Mp = pg->M, r = pg->r, K = pg->K;
  if (M != Mp || LDC < Mp || new_UVrk != r) {
    pending_release(pg);
    return fail(HCORE_HIP_ERR_ARG, "HCORE_dgemm_ormqr: arguments differ from the first phase");
  }
  hipStream_t s = pg->stream;
  int st = plan_execute_t<double>(&pg->plan, s, false, 2);
  hipError_t e = st == HCORE_HIP_OK ? hipSuccess : hipErrorUnknown;
  if (!e) e = hipMemcpy2DAsync(_CU_save, sizeof(double) * LDC, pg->oCU, sizeof(double) * Mp,
                               sizeof(double) * Mp, r, hipMemcpyDefault, s);
  if (!e) e = hipMemcpy2DAsync(_CV_save, sizeof(double) * LDC, pg->oCV, sizeof(double) * Mp,
                               sizeof(double) * Mp, r, hipMemcpyDefault, s);
  if (!e) e = hipStreamSynchronize(s);
  pending_release(pg);
  if (e) return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: HCORE_dgemm_ormqr failed");
  if (is_device_ptr(Crk)) {
    if (hipMemcpy(Crk, &r, sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      return fail(HCORE_HIP_ERR_RUNTIME, "hcore_amd: cannot write rank");
  } else {
    Crk[0] = r;
  }
  if (flops)  // hcore_dgemm.c:595, 666
    flops->update += flop_counts('o', Mp, r, K, 1) + flop_counts('o', r, Mp, K, 2);
}

// ---- the GEMM engine on its own ---------------------------------------------------
int HCORE_hip_gemm(int precision, int transA, int transB, int m, int n, int k, double alpha,
                   const void* A, int lda, const void* B, int ldb, double beta, void* C, int ldc,
                   int path, void* stream) {
  if (precision != HCORE_HIP_FP64 && precision != HCORE_HIP_FP32) return HCORE_HIP_ERR_ARG;
  if (m < 0 || n < 0 || k < 0 || path < 0 || path > 4) return HCORE_HIP_ERR_ARG;
  const bool ta = transA == kTrans || transA == kConjTrans, tb = transB == kTrans || transB == kConjTrans;
  if (lda < std::max(1, ta ? k : m) || ldb < std::max(1, tb ? n : k) || ldc < std::max(1, m))
    return HCORE_HIP_ERR_ARG;
  if (m == 0 || n == 0) return HCORE_HIP_OK;
  if (path == 4 && (ta || m < 1 || n > kTallMaxN || k > kTallMaxK)) return HCORE_HIP_ERR_ARG;
  int st = ctx_ensure();
  if (st) return st;
  ThreadCtx* tc = thread_ctx();
  if (!tc) return HCORE_HIP_ERR_RUNTIME;
  hipStream_t s = stream ? (hipStream_t)stream : tc->stream;
  const int flags = (ta ? GF_A_TRANS : 0) | (tb ? GF_B_TRANS : 0) | (beta == 0.0 ? GF_BETA_ZERO : 0);
  auto run = [&](auto zero) -> int {
    using T = decltype(zero);
    GemmProb<T> g{};
    g.A = (const T*)A; g.B = (const T*)B; g.C = (T*)C;
    g.m = m; g.n = n; g.k = k; g.lda = lda; g.ldb = ldb; g.ldc = ldc;
    g.alpha = T(alpha); g.beta = T(beta); g.flags = flags;
    TaskList tl;
    const int tile = path == 1 ? 64 : 32;
    if (path == 0) {
      add_tasks(tl, 0, g);
    } else if (path == 4) {
      for (int tj = 0; tj < (n + kTallCols - 1) / kTallCols; ++tj)
        for (int ti = 0; ti < (m + kTallRows - 1) / kTallRows; ++ti) tl.tall.push_back({0, ti, tj});
    } else {
      auto& v = tile == 64 ? tl.big : tl.small;
      for (int tj = 0; tj < (n + tile - 1) / tile; ++tj)
        for (int ti = 0; ti < (m + tile - 1) / tile; ++ti) v.push_back({0, ti, tj});
    }
    Packer pk;
    const size_t o_gp = pk.add(std::vector<GemmProb<T>>{g});
    const TaskRef ref = pk.add(tl);
    int st2 = tc->chain.reserve(pk.buf.size() + 256);
    if (st2) return st2;
    char* d = tc->chain.p;
    tc->desc.swap(pk.buf);
    HCHECK(hipMemcpyAsync(d, tc->desc.data(), tc->desc.size(), hipMemcpyHostToDevice, s));
    launch_tasks<T>(reinterpret_cast<const GemmProb<T>*>(d + o_gp), d, ref, s, path == 0 || path == 3);
    HCHECK(hipGetLastError());
    HCHECK(hipStreamSynchronize(s));
    return HCORE_HIP_OK;
  };
  return precision == HCORE_HIP_FP64 ? run(0.0) : run(0.0f);
}

// The split-precision product on its own: C = alpha op(A) op(B) + beta C for
// FP64 operands through `slices` BF16 digits (0: the same tiles on the FP64
// matrix cores).
int HCORE_hip_gemm_split(int slices, int transA, int transB, int m, int n, int k, double alpha,
                         const double* A, int lda, const double* B, int ldb, double beta, double* C,
                         int ldc, void* stream) {
  if (slices != 0 && (slices < kOzMinSlices || slices > kOzMaxSlices)) return HCORE_HIP_ERR_ARG;
  if (m < 0 || n < 0 || k < 0) return HCORE_HIP_ERR_ARG;
  const bool ta = transA == kTrans || transA == kConjTrans, tb = transB == kTrans || transB == kConjTrans;
  if (lda < std::max(1, ta ? k : m) || ldb < std::max(1, tb ? n : k) || ldc < std::max(1, m))
    return HCORE_HIP_ERR_ARG;
  if (m == 0 || n == 0) return HCORE_HIP_OK;
  int st = ctx_ensure();
  if (st) return st;
  ThreadCtx* tc = thread_ctx();
  if (!tc) return HCORE_HIP_ERR_RUNTIME;
  hipStream_t s = stream ? (hipStream_t)stream : tc->stream;
  GemmProb<double> g{};
  g.A = A; g.B = B; g.C = C;
  g.m = m; g.n = n; g.k = k; g.lda = lda; g.ldb = ldb; g.ldc = ldc;
  g.alpha = alpha; g.beta = beta;
  g.flags = (ta ? GF_A_TRANS : 0) | (tb ? GF_B_TRANS : 0) | (beta == 0.0 ? GF_BETA_ZERO : 0) | GF_SLICED;
  TaskList tl;
  add_tasks(tl, 0, g);
  Packer pk;
  const size_t o_gp = pk.add(std::vector<GemmProb<double>>{g});
  const TaskRef ref = pk.add(tl);
  st = tc->chain.reserve(pk.buf.size() + 256);
  if (st) return st;
  char* d = tc->chain.p;
  tc->desc.swap(pk.buf);
  HCHECK(hipMemcpyAsync(d, tc->desc.data(), tc->desc.size(), hipMemcpyHostToDevice, s));
  launch_tasks<double>(reinterpret_cast<const GemmProb<double>*>(d + o_gp), d, ref, s, false, slices);
  HCHECK(hipGetLastError());
  HCHECK(hipStreamSynchronize(s));
  return HCORE_HIP_OK;
}

// ---- context -----------------------------------------------------------------
int HCORE_hip_init(int device) {
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  return ctx_init_locked(device);
}

int HCORE_hip_finalize(void) {
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  if (!g_ctx.ready) return HCORE_HIP_OK;
  release_thread_ctxs();
  release_tile_state();
  ++g_ctx.generation;
  g_ctx.ready = false;
  return HCORE_HIP_OK;
}

int HCORE_hip_set_error_policy(int policy) {
  if (policy != HCORE_HIP_ON_ERROR_EXIT && policy != HCORE_HIP_ON_ERROR_RETURN)
    return HCORE_HIP_ERR_ARG;
  g_ctx.policy = policy;
  return HCORE_HIP_OK;
}

int HCORE_hip_last_status(void) { return t_last_status; }

// ---- plans -----------------------------------------------------------------------
int HCORE_hip_plan_create(hcore_hip_plan** plan, int M, int precision) {
  if (!plan || M <= 0) return HCORE_HIP_ERR_ARG;
  if (precision != HCORE_HIP_FP64 && precision != HCORE_HIP_FP32 && precision != HCORE_HIP_MIXED &&
      precision != HCORE_HIP_SPLIT)
    return HCORE_HIP_ERR_ARG;
  int st = ctx_ensure();
  if (st) return st;
  hcore_hip_plan* p = new hcore_hip_plan();
  p->M = M;
  p->prec = precision;
  *plan = p;
  return HCORE_HIP_OK;
}

int HCORE_hip_plan_add_dgemm(hcore_hip_plan* p, int n, void* const* AU, void* const* AV,
                             const int* Ark, int LDA, void* const* BU, void* const* BV,
                             const int* Brk, int LDB, void* const* CU, void* const* CV,
                             const int* Crk, int LDC, void* const* CUout, void* const* CVout,
                             int LDCout, double alpha, double beta, int rk, int maxrk,
                             double acc) {
  if (!p || p->finalized || p->ng > 0 || n < 0) return HCORE_HIP_ERR_ARG;
  // outputs: both in place (leading dimension LDC) or both out of place
  // (leading dimension LDCout), every entry given
  if ((CUout == nullptr) != (CVout == nullptr)) return HCORE_HIP_ERR_ARG;
  const int M = p->M;
  if (LDA < M || LDB < M || LDC < M || (CUout && LDCout < M)) return HCORE_HIP_ERR_ARG;
  if (n > 0 && (!AU || !AV || !BU || !BV || !CU || !CV || !Ark || !Brk || !Crk))
    return HCORE_HIP_ERR_ARG;
  for (int u = 0; u < n; ++u) {
    if (!AU[u] || !AV[u] || !BU[u] || !BV[u] || !CU[u] || !CV[u]) return HCORE_HIP_ERR_ARG;
    if (CUout && (!CUout[u] || !CVout[u])) return HCORE_HIP_ERR_ARG;
    const int K = Ark[u] + Crk[u];
    if (Ark[u] <= 0 || Brk[u] <= 0 || Crk[u] <= 0) return HCORE_HIP_ERR_ZERO_RANK;
    if (K > 2 * maxrk) return HCORE_HIP_ERR_SUM_RANKS;
    if (K > maxrk || K > p->M) return HCORE_HIP_ERR_K_GT_MAXRK;
    if (K > kCoreMaxK) return HCORE_HIP_ERR_ARG;  // stacked rank beyond the core SVD kernels
  }
  p->ng = n;
  p->AU.assign(AU, AU + n); p->AV.assign(AV, AV + n);
  p->BU.assign(BU, BU + n); p->BV.assign(BV, BV + n);
  p->CU.assign(CU, CU + n); p->CV.assign(CV, CV + n);
  p->CUo.assign(n, nullptr); p->CVo.assign(n, nullptr);
  if (CUout) p->CUo.assign(CUout, CUout + n);
  if (CVout) p->CVo.assign(CVout, CVout + n);
  p->ra.assign(Ark, Ark + n); p->rb.assign(Brk, Brk + n); p->rc.assign(Crk, Crk + n);
  p->lda = LDA; p->ldb = LDB; p->ldc = LDC; p->ldco = CUout ? LDCout : LDC;
  p->alpha = alpha; p->beta = beta; p->rk = rk; p->maxrk = maxrk; p->acc = acc;
  return HCORE_HIP_OK;
}

int HCORE_hip_plan_add_dgemm_dev(hcore_hip_plan* p, int n, void* const* AU, void* const* AV,
                                 const int* const* Ark, const int* Ark_cap, int LDA,
                                 void* const* BU, void* const* BV, const int* const* Brk,
                                 const int* Brk_cap, int LDB, void* const* CU, void* const* CV,
                                 int* const* Crk, const int* Crk_cap, int LDC, void* const* CUout,
                                 void* const* CVout, int LDCout, double alpha, double beta, int rk,
                                 int maxrk, double acc) {
  if (!p || n < 0 || (n > 0 && (!Ark || !Brk || !Crk))) return HCORE_HIP_ERR_ARG;
  for (int u = 0; u < n; ++u)
    if (!Ark[u] || !Brk[u] || !Crk[u]) return HCORE_HIP_ERR_ARG;
  int st = HCORE_hip_plan_add_dgemm(p, n, AU, AV, Ark_cap, LDA, BU, BV, Brk_cap, LDB, CU, CV,
                                    Crk_cap, LDC, CUout, CVout, LDCout, alpha, beta, rk, maxrk, acc);
  if (st) return st;
  p->dev_ranks = true;
  p->ra_dev.assign(Ark, Ark + n);
  p->rb_dev.assign(Brk, Brk + n);
  p->rc_dev.assign(Crk, Crk + n);
  return HCORE_HIP_OK;
}

int HCORE_hip_plan_add_dsyrk_dev(hcore_hip_plan* p, int n, void* const* AU, void* const* AV,
                                 const int* const* K, const int* K_cap, int LDAU,
                                 void* const* CD, int LDCD, double alpha) {
  if (!p || n < 0 || (n > 0 && !K)) return HCORE_HIP_ERR_ARG;
  for (int u = 0; u < n; ++u)
    if (!K[u]) return HCORE_HIP_ERR_ARG;
  int st = HCORE_hip_plan_add_dsyrk(p, n, AU, AV, K_cap, LDAU, CD, LDCD, alpha);
  if (st) return st;
  p->sdev_ranks = true;
  p->sk_dev.assign(K, K + n);
  return HCORE_HIP_OK;
}

int HCORE_hip_plan_add_dsyrk(hcore_hip_plan* p, int n, void* const* AU, void* const* AV,
                             const int* K, int LDAU, void* const* CD, int LDCD, double alpha) {
  if (!p || p->finalized || p->ns > 0 || n < 0) return HCORE_HIP_ERR_ARG;
  if (LDAU < p->M || LDCD < p->M) return HCORE_HIP_ERR_ARG;
  if (n > 0 && (!AU || !AV || !K || !CD)) return HCORE_HIP_ERR_ARG;
  for (int u = 0; u < n; ++u) {
    if (!AU[u] || !AV[u] || !CD[u]) return HCORE_HIP_ERR_ARG;
    if (K[u] <= 0 || K[u] > LDAU) return HCORE_HIP_ERR_ARG;
  }
  p->ns = n;
  p->sAU.assign(AU, AU + n); p->sAV.assign(AV, AV + n); p->sCD.assign(CD, CD + n);
  p->sK.assign(K, K + n);
  p->ldau = LDAU; p->ldcd = LDCD; p->salpha = alpha;
  return HCORE_HIP_OK;
}

int HCORE_hip_plan_finalize(hcore_hip_plan* p) {
  if (!p || p->finalized) return HCORE_HIP_ERR_ARG;
  return plan_double(p) ? plan_finalize_t<double>(p) : plan_finalize_t<float>(p);
}

int HCORE_hip_plan_execute(hcore_hip_plan* p, void* stream) {
  if (!p || !p->finalized) return HCORE_HIP_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  return plan_double(p) ? plan_execute_t<double>(p, s) : plan_execute_t<float>(p, s);
}

int HCORE_hip_plan_execute_graph(hcore_hip_plan* p, void* stream) {
  if (!p || !p->finalized) return HCORE_HIP_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (!p->gexec) {
    // side streams of the lanes with several core launches by K bucket, made
    // at the first capture (plans executed eagerly never pay for them)
    for (auto& L : p->lanes)
      if (L.core.size() > 1 && !L.queue && !L.aux) {
        HCHECK(lane_stream_create(&L.aux));
        HCHECK(hipEventCreateWithFlags(&L.ev_fork, hipEventDisableTiming));
        HCHECK(hipEventCreateWithFlags(&L.ev_join, hipEventDisableTiming));
      }
    hipStream_t cs = nullptr;
    HCHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    HCHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    int st = plan_double(p) ? plan_execute_t<double>(p, cs, true)
                            : plan_execute_t<float>(p, cs, true);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (st) return st;
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    p->graph = g;
    HCHECK(hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0));
  }
  HCHECK(hipGraphLaunch(p->gexec, s));
  return HCORE_HIP_OK;
}

static int plan_ranks_in(hcore_hip_plan* p, std::vector<int>& r3);

int HCORE_hip_plan_set_slices(hcore_hip_plan* p, int slices) {
  if (!p || p->prec != HCORE_HIP_SPLIT) return HCORE_HIP_ERR_ARG;
  if (slices != 0 && (slices < kOzMinSlices || slices > kOzMaxSlices)) return HCORE_HIP_ERR_ARG;
  p->slices_set = slices;
  if (p->finalized) {
    const int s = slices ? slices : default_slices(p->acc);
    for (auto& L : p->lanes) L.slices = s;
    drop_graph(p);
  }
  return HCORE_HIP_OK;
}

int HCORE_hip_plan_get_slices(hcore_hip_plan* p, int* out, int nlanes) {
  if (!p || !p->finalized) return -HCORE_HIP_ERR_ARG;
  const int nl = (int)p->lanes.size();
  if (out)
    for (int l = 0; l < std::min(nl, nlanes); ++l) out[l] = p->lanes[l].slices;
  return nl;
}

// Split-precision slice autotuner: the plan's first part (up to the core SVD;
// it reads the inputs and writes only workspace, ranks and statuses) with the
// sliced products through the FP64 matrix cores, then with 2..6 digits; per
// lane the smallest count whose core singular values stay within tol * acc of
// the FP64 ones.  Ranks, statuses and sweep counts are restored afterwards.
int HCORE_hip_plan_autotune_slices(hcore_hip_plan* p, void* stream, double tol, double* report,
                                   int nrep) {
  if (!p || !p->finalized || p->prec != HCORE_HIP_SPLIT || p->arena || p->fixed)
    return -HCORE_HIP_ERR_ARG;
  if (!(tol > 0)) tol = 0.1;
  hipStream_t s = (hipStream_t)stream;
  const int n = p->ng, nl = (int)p->lanes.size();
  if (n == 0 || nl == 0) return 0;
  std::vector<int> r3;
  int st = plan_ranks_in(p, r3);
  if (st) return -st;
  // the host copy of every update's singular values after one part-1 run
  std::vector<size_t> so(n + 1, 0);
  for (int u = 0; u < n; ++u) so[u + 1] = so[u] + (size_t)(r3[3 * u] + r3[3 * u + 2]);
  auto sigmas = [&](std::vector<double>& out) -> int {
    int e = plan_execute_t<double>(p, s, false, 1);
    if (e) return e;
    HCHECK(hipStreamSynchronize(s));
    out.assign(so[n], 0.0);
    for (int u = 0; u < n; ++u)
      if (so[u + 1] > so[u])
        HCHECK(hipMemcpy(out.data() + so[u], reinterpret_cast<double*>(p->ws) + p->sigma_off[u],
                         sizeof(double) * (so[u + 1] - so[u]), hipMemcpyDeviceToHost));
    return HCORE_HIP_OK;
  };
  int* save = nullptr;
  HCHECK(hipDeviceSynchronize());
  HCHECK(hipMalloc(&save, sizeof(int) * 3 * n));
  HCHECK(hipMemcpy(save, p->d_rank, sizeof(int) * 3 * n, hipMemcpyDeviceToDevice));
  std::vector<int> keep(nl);
  for (int l = 0; l < nl; ++l) keep[l] = p->lanes[l].slices;
  const double thr = tol * (p->acc > 0 ? p->acc : 1e-12);
  std::vector<double> ref, got;
  // dev[l][c]: the lane's largest deviation with c digits
  std::vector<std::vector<double>> dev(nl, std::vector<double>(kOzMaxSlices + 1, -1.0));
  for (auto& L : p->lanes) L.slices = 0;
  st = sigmas(ref);
  for (int c = kOzMinSlices; c <= kOzMaxSlices && !st; ++c) {
    for (auto& L : p->lanes) L.slices = c;
    st = sigmas(got);
    if (st) break;
    for (int u = 0; u < n; ++u) {
      double d = 0.0;
      for (size_t i = so[u]; i < so[u + 1]; ++i) d = std::max(d, std::fabs(got[i] - ref[i]));
      if (!(d == d)) d = INFINITY;  // NaN from a non-finite input: never within tolerance
      double& e = dev[p->lane_of[u]][c];
      e = std::max(e, d);
    }
  }
  int nr = 0;
  for (int l = 0; l < nl; ++l) {
    int pick = kOzMaxSlices;
    for (int c = kOzMinSlices; c <= kOzMaxSlices; ++c)
      if (dev[l][c] >= 0 && dev[l][c] <= thr) { pick = c; break; }
    p->lanes[l].slices = st ? keep[l] : pick;
    if (report && nr < nrep) {
      double* q = report + 4 * nr;
      q[0] = l;
      q[1] = p->lanes[l].slices;
      q[2] = dev[l][p->lanes[l].slices] / (p->acc > 0 ? p->acc : 1.0);
      q[3] = p->lanes[l].slices > kOzMinSlices
                 ? dev[l][p->lanes[l].slices - 1] / (p->acc > 0 ? p->acc : 1.0)
                 : -1.0;
    }
    ++nr;
  }
  HCHECK(hipMemcpy(p->d_rank, save, sizeof(int) * 3 * n, hipMemcpyDeviceToDevice));
  HCHECK(hipDeviceSynchronize());
  (void)hipFree(save);
  drop_graph(p);
  return st ? -st : nr;
}

}  // extern "C"

namespace hcore {

// Time every register-resident core launch of an executed plan with its
// variant and with the alternative wavefront count; keep the faster.  The
// core only reads T (kept intact) and rewrites Zu, sigma, rank and status, so
// the re-runs leave the plan's results as they were (rank, status and sweep
// words are saved and restored).
template <typename T>
static int plan_autotune_t(hcore_hip_plan* p, hipStream_t s, int reps, double* report, int nrep) {
  const int n = std::max(p->ng, 1);
  int* save = nullptr;
  HCHECK(hipDeviceSynchronize());
  HCHECK(hipMalloc(&save, sizeof(int) * 3 * n));
  HCHECK(hipMemcpy(save, p->d_rank, sizeof(int) * 3 * n, hipMemcpyDeviceToDevice));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HCHECK(hipEventCreate(&e0));
  HCHECK(hipEventCreate(&e1));
  int nr = 0;
  for (size_t l = 0; l < p->lanes.size(); ++l) {
    for (auto& c : p->lanes[l].core) {
      const int alt = core_alt_variant(c.base_variant);
      if (alt < 0) continue;
      const int cand[2] = {c.base_variant, alt};
      float best[2] = {1e30f, 1e30f};
      for (int r = 0; r < reps; ++r)
        for (int k = 0; k < 2; ++k) {
          HCHECK(hipEventRecord(e0, s));
          const int* cq =
              c.queue ? reinterpret_cast<const int*>(p->desc + p->lanes[l].o_q) : nullptr;
          if (c.dbl)
            launch_core<double>(reinterpret_cast<const CoreProb<double>*>(p->desc + c.off), c.count,
                                cand[k], c.kmax, p->rk, p->maxrk, p->acc, s, cq, c.base_variant);
          else
            launch_core<T>(reinterpret_cast<const CoreProb<T>*>(p->desc + c.off), c.count, cand[k],
                           c.kmax, p->rk, p->maxrk, T(p->acc), s, cq, c.base_variant);
          HCHECK(hipEventRecord(e1, s));
          HCHECK(hipEventSynchronize(e1));
          float ms = 0.f;
          HCHECK(hipEventElapsedTime(&ms, e0, e1));
          best[k] = std::min(best[k], ms);
        }
      c.variant = best[1] < best[0] ? alt : c.base_variant;
      tune_store(p->M, c.dbl || sizeof(T) == 8, c.base_variant, c.variant);
      if (report && nr < nrep) {
        double* q = report + 6 * nr;
        q[0] = (double)l; q[1] = c.base_variant; q[2] = c.variant; q[3] = best[0]; q[4] = best[1];
        q[5] = c.count;
      }
      ++nr;
    }
  }
  HCHECK(hipMemcpy(p->d_rank, save, sizeof(int) * 3 * n, hipMemcpyDeviceToDevice));
  HCHECK(hipDeviceSynchronize());
  (void)hipFree(save);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  drop_graph(p);
  return nr;
}

}  // namespace hcore

extern "C" {

int HCORE_hip_plan_autotune(hcore_hip_plan* p, void* stream, int reps, double* report, int nrep) {
  if (!p || !p->finalized || p->arena || p->fixed || reps < 1) return -HCORE_HIP_ERR_ARG;
  const int r = plan_double(p)
                    ? plan_autotune_t<double>(p, (hipStream_t)stream, reps, report, nrep)
                    : plan_autotune_t<float>(p, (hipStream_t)stream, reps, report, nrep);
  return r;
}

const int* HCORE_hip_plan_ranks_device(const hcore_hip_plan* p) { return p ? p->d_rank : nullptr; }

int HCORE_hip_plan_get_ranks(hcore_hip_plan* p, int* out) {
  if (!p || !p->finalized) return HCORE_HIP_ERR_ARG;
  HCHECK(hipDeviceSynchronize());
  p->last_ranks.resize(p->ng);
  if (p->ng) HCHECK(hipMemcpy(p->last_ranks.data(), p->d_rank, sizeof(int) * p->ng, hipMemcpyDeviceToHost));
  if (out) std::copy(p->last_ranks.begin(), p->last_ranks.end(), out);
  return HCORE_HIP_OK;
}

int HCORE_hip_plan_get_status(hcore_hip_plan* p, int* out) {
  if (!p || !p->finalized || !out) return HCORE_HIP_ERR_ARG;
  HCHECK(hipDeviceSynchronize());
  if (p->ng) HCHECK(hipMemcpy(out, p->d_status, sizeof(int) * p->ng, hipMemcpyDeviceToHost));
  return HCORE_HIP_OK;
}

int HCORE_hip_plan_get_core_queue(hcore_hip_plan* p, int lane, int* out, int n) {
  if (!p || !p->finalized || lane < 0 || lane >= (int)p->lanes.size()) return -HCORE_HIP_ERR_ARG;
  const auto& L = p->lanes[lane];
  if (!L.queue) return 0;
  std::vector<int> q(kCoreQueueHead + L.n);
  HCHECK(hipDeviceSynchronize());
  HCHECK(hipMemcpy(q.data(), p->desc + L.o_q, sizeof(int) * q.size(), hipMemcpyDeviceToHost));
  const auto& mem = p->lane_updates[lane];
  if (out)
    for (int i = 0; i < std::min(n, L.n); ++i) out[i] = mem[q[kCoreQueueHead + i]];
  return L.n;
}

int HCORE_hip_plan_get_lanes(hcore_hip_plan* p, int* out) {
  if (!p || !p->finalized || !out) return HCORE_HIP_ERR_ARG;
  std::copy(p->lane_of.begin(), p->lane_of.end(), out);
  return HCORE_HIP_OK;
}

int HCORE_hip_plan_get_sweeps(hcore_hip_plan* p, int* out) {
  if (!p || !p->finalized || !out) return HCORE_HIP_ERR_ARG;
  HCHECK(hipDeviceSynchronize());
  if (p->ng) HCHECK(hipMemcpy(out, p->d_sweeps, sizeof(int) * p->ng, hipMemcpyDeviceToHost));
  return HCORE_HIP_OK;
}

int HCORE_hip_plan_get_core_profile(hcore_hip_plan* p, long long* out, double* tick_ns) {
  if (!p || !p->finalized) return HCORE_HIP_ERR_ARG;
  if (tick_ns) {
    int khz = 0;
    HCHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g_ctx.device));
    *tick_ns = khz > 0 ? 1e6 / khz : 0.0;
  }
  if (!out) return HCORE_HIP_OK;
  HCHECK(hipDeviceSynchronize());
  if (p->ng)
    HCHECK(hipMemcpy(out, p->d_clk, sizeof(long long) * kCoreClk * p->ng, hipMemcpyDeviceToHost));
  return HCORE_HIP_OK;
}

// The ranks in use by the last execution (device-side ranks) or the host ranks.
static int plan_ranks_in(hcore_hip_plan* p, std::vector<int>& r3) {
  r3.resize(3 * (size_t)p->ng);
  if (p->dev_ranks) {
    HCHECK(hipDeviceSynchronize());
    if (p->ng)
      HCHECK(hipMemcpy(r3.data(), p->d_rank_in, sizeof(int) * 3 * p->ng, hipMemcpyDeviceToHost));
  } else {
    for (int u = 0; u < p->ng; ++u) {
      r3[3 * u] = p->ra[u];
      r3[3 * u + 1] = p->rb[u];
      r3[3 * u + 2] = p->rc[u];
    }
  }
  return HCORE_HIP_OK;
}

int HCORE_hip_plan_get_sigma(hcore_hip_plan* p, int u, double* out, int* K_out) {
  if (!p || !p->finalized || u < 0 || u >= p->ng) return HCORE_HIP_ERR_ARG;
  std::vector<int> r3;
  int st = plan_ranks_in(p, r3);
  if (st) return st;
  const int K = r3[3 * u] + r3[3 * u + 2];
  if (K_out) *K_out = K;
  if (!out) return HCORE_HIP_OK;
  HCHECK(hipDeviceSynchronize());
  if (p->sigma_double) {
    HCHECK(hipMemcpy(out, reinterpret_cast<double*>(p->ws) + p->sigma_off[u], sizeof(double) * K,
                     hipMemcpyDeviceToHost));
  } else {
    std::vector<float> f(K);
    HCHECK(hipMemcpy(f.data(), reinterpret_cast<float*>(p->ws) + p->sigma_off[u], sizeof(float) * K,
                     hipMemcpyDeviceToHost));
    for (int i = 0; i < K; ++i) out[i] = f[i];
  }
  return HCORE_HIP_OK;
}

long long HCORE_hip_plan_workspace_bytes(const hcore_hip_plan* p) {
  return p ? (long long)p->dbytes : 0;
}

int HCORE_hip_plan_flops(hcore_hip_plan* p, unsigned long* update, unsigned long* syrk) {
  if (!p || !p->finalized) return HCORE_HIP_ERR_ARG;
  if ((int)p->last_ranks.size() != p->ng) return HCORE_HIP_ERR_ARG;
  unsigned long fu = 0, fs = 0;
  const unsigned long M = p->M;
  std::vector<int> r3;
  int st = plan_ranks_in(p, r3);
  if (st) return st;
  for (int u = 0; u < p->ng; ++u) {
    const unsigned long ra = r3[3 * u], rb = r3[3 * u + 1], K = r3[3 * u] + r3[3 * u + 2];
    const unsigned long r = p->last_ranks[u];
    fu += flop_counts('q', M, K, 0, 0) * 2 + flop_counts('m', ra, rb, M, 0) +
          flop_counts('m', M, ra, rb, 0) + flop_counts('s', K, 0, 0, 0) + r * K +
          flop_counts('o', M, r, K, 1) + flop_counts('o', r, M, K, 2);
  }
  for (int s = 0; s < p->ns; ++s) {
    const unsigned long K = p->sK[s], N = p->ldau;
    fs += flop_counts('m', K, K, N, 0) + flop_counts('m', M, K, K, 0) + flop_counts('m', M, M, K, 0);
  }
  if (update) *update = fu;
  if (syrk) *syrk = fs;
  return HCORE_HIP_OK;
}

void HCORE_hip_plan_destroy(hcore_hip_plan* p) {
  if (!p) return;
  for (int k = 0; k < hcore_hip_plan::kRing; ++k)
    for (int l = 0; l < hcore_hip_plan::kMaxLanes; ++l)
      for (int i = 0; i < HCORE_HIP_NPHASES; ++i)
        for (int j = 0; j < 2; ++j)
          if (p->ev[k][l][i][j]) (void)hipEventDestroy(p->ev[k][l][i][j]);
  for (auto& L : p->lanes) {
    lane_stream_release(L.stream);
    if (L.ev_done) (void)hipEventDestroy(L.ev_done);
    lane_stream_release(L.aux);
    if (L.ev_fork) (void)hipEventDestroy(L.ev_fork);
    if (L.ev_join) (void)hipEventDestroy(L.ev_join);
  }
  if (p->ev_start) (void)hipEventDestroy(p->ev_start);
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->dmem) {
    (void)hipDeviceSynchronize();
    (void)hipFree(p->dmem);
  }
  delete p;
}

int HCORE_hip_plan_set_timing(hcore_hip_plan* p, int enable) {
  if (!p) return HCORE_HIP_ERR_ARG;
  if (enable && !p->ev[0][0][0][0])
    for (int k = 0; k < hcore_hip_plan::kRing; ++k)
      for (int l = 0; l < hcore_hip_plan::kMaxLanes; ++l)
        for (int i = 0; i < HCORE_HIP_NPHASES; ++i)
          for (int j = 0; j < 2; ++j) HCHECK(hipEventCreate(&p->ev[k][l][i][j]));
  p->timing = enable != 0;
  p->serial = enable == 2;
  return HCORE_HIP_OK;
}

int HCORE_hip_plan_phase_times(hcore_hip_plan* p, double* ms, int* count) {
  if (!p) return HCORE_HIP_ERR_ARG;
  int st = collect_all(p);
  if (st) return st;
  for (int l = 0; l < hcore_hip_plan::kMaxLanes; ++l)
    for (int i = 0; i < HCORE_HIP_NPHASES; ++i) {
      if (ms) ms[l * HCORE_HIP_NPHASES + i] = p->phase_ms[l][i];
      if (count) count[l * HCORE_HIP_NPHASES + i] = p->phase_count[l][i];
    }
  return HCORE_HIP_OK;
}

// ---- device memory / streams / events ------------------------------------------
int HCORE_hip_malloc(void** ptr, size_t bytes) {
  int st = ctx_ensure();
  if (st) return st;
  HCHECK(hipMalloc(ptr, bytes));
  return HCORE_HIP_OK;
}
int HCORE_hip_free(void* ptr) {
  HCHECK(hipFree(ptr));
  return HCORE_HIP_OK;
}
int HCORE_hip_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  HCHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return HCORE_HIP_OK;
}
int HCORE_hip_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  HCHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return HCORE_HIP_OK;
}
int HCORE_hip_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  HCHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return HCORE_HIP_OK;
}
int HCORE_hip_memset(void* ptr, int value, size_t bytes) {
  HCHECK(hipMemset(ptr, value, bytes));
  return HCORE_HIP_OK;
}
int HCORE_hip_device_synchronize(void) {
  HCHECK(hipDeviceSynchronize());
  return HCORE_HIP_OK;
}
int HCORE_hip_stream_create(void** stream) {
  int st = ctx_ensure();
  if (st) return st;
  hipStream_t s = nullptr;
  HCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return HCORE_HIP_OK;
}
int HCORE_hip_stream_destroy(void* stream) {
  HCHECK(hipStreamDestroy((hipStream_t)stream));
  return HCORE_HIP_OK;
}
int HCORE_hip_stream_synchronize(void* stream) {
  HCHECK(hipStreamSynchronize((hipStream_t)stream));
  return HCORE_HIP_OK;
}
int HCORE_hip_event_create(void** event) {
  int st = ctx_ensure();
  if (st) return st;
  hipEvent_t e = nullptr;
  HCHECK(hipEventCreate(&e));
  *event = e;
  return HCORE_HIP_OK;
}
int HCORE_hip_event_destroy(void* event) {
  HCHECK(hipEventDestroy((hipEvent_t)event));
  return HCORE_HIP_OK;
}
int HCORE_hip_event_record(void* event, void* stream) {
  HCHECK(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
  return HCORE_HIP_OK;
}
int HCORE_hip_event_elapsed_ms(void* start, void* stop, float* ms) {
  HCHECK(hipEventSynchronize((hipEvent_t)stop));
  HCHECK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return HCORE_HIP_OK;
}
int HCORE_hip_mem_info(size_t* free_bytes, size_t* total_bytes) {
  int st = ctx_ensure();
  if (st) return st;
  size_t f = 0, t = 0;
  HCHECK(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return HCORE_HIP_OK;
}

int HCORE_hip_device_info(char* name, int len, int* num_cu) {
  int st = ctx_ensure();
  if (st) return st;
  hipDeviceProp_t prop;
  HCHECK(hipGetDeviceProperties(&prop, g_ctx.device));
  if (name && len > 0) {
    strncpy(name, prop.name[0] ? prop.name : prop.gcnArchName, len - 1);
    name[len - 1] = 0;
  }
  if (num_cu) *num_cu = prop.multiProcessorCount;
  return HCORE_HIP_OK;
}

}  // extern "C"
