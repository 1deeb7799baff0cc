// hcore_internal.hpp — device work descriptors and kernel launchers of the
// MI355X-native HCORE.  Internal to the library (not installed).
//
// Every kernel is batched: it takes a device array of per-item descriptors
// and processes one item per workgroup, so a single tile (the HCORE_* C API)
// and a whole TLR Cholesky trailing step use the same code.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "flop_util_structs.h"

namespace hcore {

// ---- status codes written per update by the device --------------------------
enum : int {
  ST_OK = 0,
  ST_R_GT_MAXRK = 4,  // truncated rank above maxrk (reference hcore_dgemm.c:465-468)
  ST_NONFINITE = 5,   // NaN/Inf reached the core SVD (the reference's LAPACK info!=0 class)
  ST_BAD_RANK = 7,    // a device-side input rank outside 1..its capacity (HCORE_HIP_ERR_ARG)
};

// ---- generic batched GEMM on 64x64 output tiles ------------------------------
// C(m x n) = alpha * op(A) * op(B) + beta * C, column-major.
enum : int {
  GF_A_TRANS = 1,           // op(A) = A^T
  GF_B_TRANS = 2,           // op(B) = B^T
  GF_A_UPPER = 4,           // stored A(p,q) read as 0 when p > q
  GF_B_UPPER = 8,           // stored B(p,q) read as 0 when p > q
  GF_LOWER_ZERO_UPPER = 16, // write i >= j only; strict upper triangle of C set to 0
  GF_BETA_ZERO = 32,        // C is not read
  GF_A_UNIT_LOWER = 64,     // stored A(p,q) read as 0 when p < q and 1 when p == q
  GF_B_UNIT_LOWER = 128,    // same for B
  GF_C0 = 256,              // the beta term reads C0 (rows < m0; 0 below) instead of C
  GF_LOWER_ONLY = 512,      // write i >= j only; the strict upper triangle of C is not touched
  GF_UPPER_ONLY = 1024,     // write i <= j only; the strict lower triangle of C is not touched
  GF_SLICED = 2048,         // FP64 product through BF16 digit slices (k_gemm_oz)
};

template <typename T>
struct GemmProb {
  const T* A;
  const T* B;
  T* C;
  int m, n, k;
  int lda, ldb, ldc;
  T alpha, beta;
  int flags;
  const int* n_dev;   // if set, n is clamped to *n_dev (device-side rank)
  const int* skip;    // if set and *skip != 0 the problem is skipped
  const T* C0;        // GF_C0: source of the beta term, leading dimension ldc0
  int ldc0, m0;
};

struct GemmTask {
  int prob;
  int ti, tj;  // tile coordinates (64 x 64, or 32 x 32 in the small-tile lists)
};

// ---- stacked factors: Us = [beta*CU | alpha*AU], Vs[:, :rc] = CV --------------
template <typename T>
struct StackProb {
  const T* CU;
  const T* CV;
  const T* AU;
  T* Us;
  T* Vs;
  int rc, ra, ldc, lda;
};

// ---- Householder QR of one nb x K stack (in place, ld = nb) ------------------
template <typename T>
struct QrProb {
  T* A;    // nb x K: R on and above the diagonal, reflectors below
  T* tau;  // K
  T* Tp;   // ceil(K/PB) blocks of PB x PB compact-WY factors
  T* S;    // K x K Gram Y^T Y, lower triangle (single-panel register QR only)
  int K;
  int* routed;  // QR queue: set to 1 when a register QR (which writes S) takes the stack
};

// ---- core SVD: one-sided Jacobi on T = R_A R_B^T -----------------------------
template <typename T>
struct CoreProb {
  const T* Tc;  // K x K, ld K (kept intact)
  T* Tj;        // K x K scratch used when the matrix does not fit in LDS
  T* Zu;        // K x K, ld K: first r columns = left singular vectors
  T* sigma;     // K singular values, descending
  int* rank;    // new rank r
  int* status;  // ST_*
  int* sweeps;  // Jacobi sweeps used (diagnostic, may be null)
  long long* clk;  // kCoreClk wall-clock stamps (diagnostic, may be null)
  int K;
};
// core stamps: [0] start, [1] after the pivoted QR, [2] after the Jacobi
// sweeps, [3] end, [4] numerical rank k of the pivoted QR
constexpr int kCoreClk = 8;

// ---- compact-WY solve: M = T Y1^T Z for Q [Z; 0] = [Z; 0] - Y M --------------
// With S = Y^T Y and the per-panel factors T_bb, M solves
// (diag(1/tau) + striu(S)) M = Y1^T Z; on entry M holds Y1^T Z.
template <typename T>
struct WySolveProb {
  const T* S;    // K x K Gram Y^T Y (ld K; its lower triangle is read)
  const T* Tp;   // per-panel compact-WY factors (kPB x kPB each)
  T* M;          // K x r, ld K: Y1^T Z on entry, M on exit
  const int* rank;
  const int* status;
  int K;
};

// ---- precision conversion of one K x K core (mixed-precision plans) ---------
struct ConvProb {
  const void* src;
  void* dst;
  int n;  // elements
};

constexpr int kThreads = 256;
constexpr int kPB = 8;  // Householder panel width

// Leading dimension of the core SVD's work matrix: the smallest odd multiple
// of the group width G that is >= K.  Groups of G lanes read G consecutive
// rows of different columns at the same time; an odd multiple of G between
// columns puts the groups of a wavefront on distinct LDS banks.
__host__ __device__ constexpr int core_ld(int K, int G) {
  const int q = (K + G - 1) / G;
  return G * (q | 1);
}
constexpr int kCoreLdsG = 4;  // group width of the largest LDS-resident core variant

// Largest K whose K x K FP64 core fits in LDS next to the Jacobi bookkeeping.
int core_lds_max_k(size_t elem_size);

// ---- launchers (k_*.hip) ------------------------------------------------------
// tile: 64 (256 threads) or 32 (64 threads, for problems with m or n <= 32);
// split: the k range of each tile is shared by several thread groups (for
// k = nb reductions into small outputs)
template <typename T>
void launch_gemm(const GemmProb<T>* probs, const GemmTask* tasks, int ntasks, int tile,
                 hipStream_t s, bool split = false);
// tall-skinny products: A not transposed, m >= kTallMinM, n <= kTallMaxN,
// k <= kTallMaxK; tasks are (row block of kTallRows, column block of kTallCols)
constexpr int kTallRows = 256, kTallCols = 32, kTallK = 64;
constexpr int kTallMinM = 256, kTallMaxN = 128, kTallMaxK = 128;
template <typename T>
void launch_gemm_tall(const GemmProb<T>* probs, const GemmTask* tasks, int ntasks, hipStream_t s);
// split-precision FP64 products on the BF16 matrix cores: 32 x 32 output tiles
// (tasks as for tile 32), `slices` digits per operand (kOzMinSlices..
// kOzMaxSlices), ks = 1 or 4 wavefronts sharing a tile's k range; ks = 0: the
// tall-skinny kernel on kTallRows x kTallCols tasks (as launch_gemm_tall)
constexpr int kOzMinSlices = 2, kOzMaxSlices = 6;
void launch_gemm_oz(const GemmProb<double>* probs, const GemmTask* tasks, int ntasks, int slices,
                    int ks, hipStream_t s);
constexpr int kStackCols = 4;  // stacked columns per workgroup of k_stack
template <typename T>
void launch_stack(const StackProb<T>* probs, int n, int nb, int maxcols, T alpha, T beta,
                  hipStream_t s);
template <typename T>
void launch_geqrf(const QrProb<T>* probs, int n, int nb, hipStream_t s);
// register-resident Householder QR in one pass (nb <= kQrRegMaxNb, K <= kQrRegMaxK):
// kmax <= 16 selects the 16-column kernel, else the 32-column one.  Returns
// false (and launches nothing) for a shape beyond those limits.
constexpr int kQrRegMaxK = 32;
constexpr int kQrRegMaxNb = 1024;  // rows held by the largest register QR shape (512 x 2)
constexpr int kQrFusedMaxNb = 2048;  // rows of the largest fused-QR shape (512 x 4)
// q (optional): a device-side QR queue (launch_qr_queue), the launch serving
// its class qc (n an upper bound on the class's size).
template <typename T>
bool launch_qr_reg(const QrProb<T>* probs, int n, int nb, int kmax, hipStream_t s,
                   const int* q = nullptr, int qc = -1);
// whole-stack QR in one workgroup (nb <= kQrFusedMaxNb): register panels of 16
// columns (32 at nb <= 256), each applied to the trailing columns in place;
// false for nb beyond the limit
template <typename T>
bool launch_qr_fused(const QrProb<T>* probs, int n, int nb, hipStream_t s, const int* q = nullptr,
                     int qc = -1);
// ---- device-side QR queue (plans with device ranks) --------------------------
// q = [start (3) | count (3) | items (n)]: the stacks by the stacked rank K in
// their (patched) descriptors, class 0: K <= 16, 1: 17..kQrRegMaxK (the
// register QR shapes), 2: larger (the fused QR).
constexpr int kQrQueueHead = 6;
template <typename T>
void launch_qr_queue(const QrProb<T>* probs, int n, int* q, hipStream_t s);
// core SVD variants by stacked rank K: one wavefront for K <= 64, one
// workgroup with the core in LDS up to the LDS limit, 1024 threads with the
// core in global memory beyond it
// (LDS/global variants), then the register-resident variants (k_core_reg)
enum : int {
  CORE_K16, CORE_K32, CORE_K64, CORE_K128, CORE_KLDS, CORE_K192, CORE_K256, CORE_K512,
  CORE_R64, CORE_R128, CORE_R192,
  // the register-resident variants with twice the wavefronts per tile (half
  // the Jacobi block width): more parallelism per tile, more registers per CU
  CORE_R64W, CORE_R128W, CORE_R192W, CORE_NVARIANTS
};
int core_variant_for(int K, int lds_max_k, bool reg);
// The same choice on the device (the rank queue below).
__host__ __device__ inline int core_variant_of(int K, int kl, bool reg) {
  if (K <= 16) return CORE_K16;
  if (K <= 32) return CORE_K32;
  if (reg) {
    if (K <= 64) return CORE_R64;
    if (K <= 128) return CORE_R128;
    if (K <= 192) return CORE_R192;
  }
  if (K <= 64) return CORE_K64;
  if (K <= 128) return CORE_K128;
  if (K <= kl && K <= 144) return CORE_KLDS;
  if (K <= 192) return CORE_K192;
  if (K <= 256) return CORE_K256;
  return CORE_K512;
}
// Variants whose work matrix (or R^T transposition) lives in the K x (K + 64)
// global scratch CoreProb::Tj.
inline bool core_needs_scratch(int v) {
  return v == CORE_K192 || v == CORE_K256 || v == CORE_K512 || v >= CORE_R64;
}
// The other wavefront count of a register-resident variant (CORE_R64 <->
// CORE_R64W, ...), or -1: the alternative an autotuned plan may pick.
int core_alt_variant(int variant);
// the register-resident variants (CORE_R64 .. CORE_R192); launch_core forwards to it
template <typename T>
void launch_core_reg(const CoreProb<T>* probs, int n, int variant, int rk, int maxrk, T acc,
                     hipStream_t s, const int* q = nullptr, int qv = -1);
constexpr int kCoreMaxK = 512;
constexpr int kMaxSweeps = 40;  // Jacobi sweep limit of every core variant
// q (optional): a lane's device-side rank queue (launch_core_queue); the launch
// then serves the updates of base variant qv in it, largest K first, n
// being an upper bound on their number (workgroups beyond it exit at once).
template <typename T>
void launch_core(const CoreProb<T>* probs, int n, int variant, int kmax, int rk, int maxrk,
                 T acc, hipStream_t s, const int* q = nullptr, int qv = -1);
// ---- device-side rank queue of a lane's core SVDs (plans with device ranks) --
// Built each execution from the stacked ranks K in the (patched) descriptors:
// q = [vstart (CORE_NVARIANTS) | vcount (CORE_NVARIANTS) | items (n)], items
// the lane's updates ordered by K, largest first, and [vstart[v], + vcount[v])
// the slice whose K selects base variant v (core_variant_of).
constexpr int kCoreQueueHead = 2 * CORE_NVARIANTS;
template <typename T>
void launch_core_queue(const CoreProb<T>* probs, int n, int kl, bool reg, int* q, hipStream_t s);
// blocked back substitution of the compact-WY solve (64 columns per workgroup
// up to K = 256, 16 beyond)
constexpr int kWyMaxK = 512;
template <typename T>
void launch_wy_solve(const WySolveProb<T>* probs, int n, int kmax, int rmax, hipStream_t s);

// batched element-wise conversion FP32 -> FP64 (to_double) or FP64 -> FP32
void launch_convert(const ConvProb* probs, int n, int nmax, bool to_double, hipStream_t s);

template <typename T>
void prepare_kernels();

// ---- single-update entry with a caller-provided workspace (hcore_api.cpp) ----
// One HCORE_dgemm on device pointers whose one-update workspace, descriptors
// and status words go to the fixed device buffer ws (ws_bytes), on stream s:
// nothing is allocated on the device.  Fails through the library's error
// policy when the update does not fit.
void dgemm_fixed_workspace(int M, double alpha, double* AU, double* AV, double* Ark, int LDA,
                           double* BU, double* BV, double* Brk, int LDB, double beta, double* CU,
                           double* CV, double* Crk, int LDC, int rk, int maxrk, double acc,
                           flop_counter* flops, const char* name, char* ws, size_t ws_bytes,
                           hipStream_t s);
// Print msg, record status for HCORE_hip_last_status and apply the error policy.
void report_failure(int status, const char* msg);

// ---- device-side ranks of a plan (k_plan.hip) --------------------------------
// Plans whose input ranks live in device memory keep every rank-dependent
// descriptor field in a patch list; at the start of each execution the
// ranks are read (and checked against the capacities the workspace was
// carved for), then every patch rewrites its field, so one captured graph
// serves successive steps whose ranks changed on the device.
enum : int {
  PK_RA,     // int field = rank of A
  PK_RB,     // int field = rank of B
  PK_RC,     // int field = rank of C
  PK_K,      // int field = ra + rc (stacked rank)
  PK_KK,     // int field = K * K
  PK_PTR_G,  // pointer field = base + rc * M elements (G's place in the V stack)
  PK_SK,     // int field = rank of a dsyrk update
};
struct RankPatch {
  unsigned long long off;  // byte offset of the field in the descriptor buffer
  long long base;          // PK_PTR_G: device address of the V stack
  int kind, upd;
};
struct RankTables {
  const int* const* ra;  // per gemm update: device address of its rank of A (null: host ranks)
  const int* const* rb;
  int* const* rc;        // also receives the new rank after a successful update
  const int* const* sk;  // per dsyrk update
  const int* cap;        // 3 per gemm update (ra, rb, rc capacities), then 1 per dsyrk
  int* rank_in;          // out: 3 per gemm update, then 1 per dsyrk (the ranks in use)
  int* status;           // per gemm update: reset to ST_OK, or ST_BAD_RANK
  const int* rank_out;   // per gemm update: the new rank (publish)
  int n, ns, M, elem;
};
void launch_rank_gather(const RankTables& t, hipStream_t s);
void launch_rank_patch(const RankPatch* patches, int np, char* desc, const int* rank_in, int n,
                       int M, int elem, hipStream_t s);
void launch_rank_publish(const RankTables& t, hipStream_t s);

// ---- dense triangular kernels (k_dense.hip) ----------------------------------
// One diagonal block of a triangular solve: for every right-hand side c <
// nrhs, solve S y = alpha z in place, where z_k = B[(b0 + k) es + c cs]
// (k < nbk <= kTrsmNB) and S(k, l) is A(b0 + k, b0 + l) (A(b0 + l, b0 + k)
// when trans), lower (forward substitution) or upper (backward), with a unit
// diagonal when unit.  Skipped when *skip != 0.
constexpr int kTrsmNB = 64;
template <typename T>
struct TrsmDiagProb {
  const T* A;
  T* B;
  long long es, cs;
  const int* skip;
  T alpha;
  int lda, b0, nbk, nrhs;
  int trans, lower, unit;
};
template <typename T>
void launch_trsm_diag(const TrsmDiagProb<T>& p, hipStream_t s);
// Unblocked Cholesky of the diagonal block A(b0.., b0..) of size nbk <=
// kTrsmNB (lower: L L^T in the lower triangle; upper: U^T U in the upper);
// the other triangle is not touched.  The first non-positive (or NaN) pivot
// k sets *info = b0 + k + 1 and stops; nothing runs when *info != 0.
template <typename T>
void launch_potf2(T* A, int lda, int b0, int nbk, int upper, int* info, hipStream_t s);
// B = alpha * B for an m x n block (alpha == 0: zero fill).
template <typename T>
void launch_scale(T* B, int ldb, int m, int n, T alpha, hipStream_t s);

// ---- tile generation and compression (k_tile.hip) ----------------------------
// One tile written into out (leading dimension ldo): entries (i, j) with i < m
// and j < n come from the dense source src (leading dimension lds) or, when
// src is null, from the covariance of points p_(m0+i) and p_(n0+j),
// exp(-|p - q| / beta), plus nugget where the global row equals the global
// column; entries with i < rows, j < cols outside m x n are written as 0.
struct TileFillProb {
  const double* src;
  const double* px;  // sorted point coordinates: x[bigM], then y[bigM]
  const double* py;
  double* out;
  int lds, ldo;
  int m0, n0, m, n, rows, cols;
  double beta, nugget;
};
constexpr int kFillCols = 8;  // tile columns per workgroup of k_tile_fill
void launch_tile_fill(const TileFillProb* probs, int n, int max_cols, hipStream_t s);

// After the core SVD of a K x K padded tile: the truncated rank (0 when the
// core failed or sigma_0 < acc, else the core's rank capped by min(m, n) and
// maxrank) to rank_out (and rank_out_d as a double when set), and the first
// rank columns of Zu (rows < m) to U.
struct CompressProb {
  const double* Zu;
  const double* sigma;
  const int* rank;
  const int* status;
  int* rank_out;
  double* rank_out_d;
  double* U;
  int ldu, K, m, mn, maxrank;
  double acc;
};
constexpr int kFinishCols = 16;  // columns of U per workgroup of k_compress_finish
void launch_compress_finish(const CompressProb* probs, int n, int kmax, hipStream_t s);

// Size in elements of the compact-WY storage for a K-column stack.
inline int tp_elems(int K) { return ((K + kPB - 1) / kPB) * kPB * kPB; }

}  // namespace hcore
