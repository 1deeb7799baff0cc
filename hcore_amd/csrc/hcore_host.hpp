// hcore_host.hpp — host-side building blocks shared by the library's
// translation units: HIP error checks, grow-only device and pinned buffers,
// batched-GEMM task lists and descriptor packing, and the services of the
// single-call entry points (context, per-thread stream and buffers, error
// policy).  Internal to the library (not installed).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

#include "hcore_hip.h"
#include "hcore_internal.hpp"

namespace hcore {

inline int hip_fail(hipError_t e, const char* what) {
  fprintf(stderr, "hcore_amd: HIP error in %s: %s\n", what, hipGetErrorString(e));
  return HCORE_HIP_ERR_RUNTIME;
}
#define HCHECK(call)                                          \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) return hip_fail(e_, #call);         \
  } while (0)

// Grow-only device buffer.  Growing drops the contents; callers grow it only
// when no work of theirs on it is in flight.
struct DevArena {
  char* p = nullptr;
  size_t cap = 0;
  int reserve(size_t bytes) {
    if (cap >= bytes) return HCORE_HIP_OK;
    if (p) HCHECK(hipFree(p));
    p = nullptr;
    cap = 0;
    const size_t mib = size_t(1) << 20;
    const size_t c = ((std::max(bytes, cap * 2) + mib - 1) / mib) * mib;
    HCHECK(hipMalloc(&p, c));
    cap = c;
    return HCORE_HIP_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Grow-only pinned host buffer (hipHostMalloc): host operands are packed into
// it on the CPU and moved by one DMA transfer each way, instead of one
// pageable transfer per column.
struct PinnedArena {
  char* p = nullptr;
  size_t cap = 0;
  int reserve(size_t bytes) {
    if (cap >= bytes) return HCORE_HIP_OK;
    if (p) HCHECK(hipHostFree(p));
    p = nullptr;
    cap = 0;
    const size_t mib = size_t(1) << 20;
    const size_t c = ((bytes + mib - 1) / mib) * mib;
    HCHECK(hipHostMalloc(reinterpret_cast<void**>(&p), c, hipHostMallocDefault));
    cap = c;
    return HCORE_HIP_OK;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

// ------------------------------------------------------------------------------
// host-side descriptor packing
// ------------------------------------------------------------------------------
// Output tiles of one batched GEMM phase: 64 x 64 tiles, or 32 x 32 tiles
// (one wavefront each) for problems with m or n <= 32, which would otherwise
// compute mostly padding; the split-precision products (GF_SLICED) as 32 x 32
// tiles with one wavefront (k <= 64) or four sharing the k range.
struct TaskList {
  std::vector<GemmTask> big, small, tall, oz1, oz4, ozt;
  size_t size() const {
    return big.size() + small.size() + tall.size() + oz1.size() + oz4.size() + ozt.size();
  }
};
struct TaskRef {
  size_t ob = 0, os = 0, ot = 0, o1 = 0, o4 = 0, oT = 0;
  int nb = 0, ns = 0, nt = 0, n1 = 0, n4 = 0, nT = 0;
};

struct Packer {
  std::vector<char> buf;
  template <typename X>
  size_t add(const std::vector<X>& v) {
    size_t off = (buf.size() + 255) & ~size_t(255);
    buf.resize(off + v.size() * sizeof(X));
    if (!v.empty()) memcpy(buf.data() + off, v.data(), v.size() * sizeof(X));
    return off;
  }
  TaskRef add(const TaskList& tl) {
    TaskRef r;
    r.ob = add(tl.big);
    r.nb = (int)tl.big.size();
    r.os = add(tl.small);
    r.ns = (int)tl.small.size();
    r.ot = add(tl.tall);
    r.nt = (int)tl.tall.size();
    r.o1 = add(tl.oz1);
    r.n1 = (int)tl.oz1.size();
    r.o4 = add(tl.oz4);
    r.n4 = (int)tl.oz4.size();
    r.oT = add(tl.ozt);
    r.nT = (int)tl.ozt.size();
    return r;
  }
};

// Output tiles of one product: tall-skinny blocks (kTallRows x kTallCols) for
// products of the nb rows of a tile with a few dozen columns, else 64 x 64
// tiles, or 32 x 32 ones when m or n <= 32.
template <typename T>
void add_tasks(TaskList& tl, int prob, const GemmProb<T>& g) {
  const int m = g.m, n = g.n;
  const bool tall = !(g.flags & (GF_A_TRANS | GF_LOWER_ZERO_UPPER | GF_LOWER_ONLY | GF_UPPER_ONLY)) &&
                    m >= kTallMinM &&
                    n <= kTallMaxN && g.k <= kTallMaxK;
  if ((g.flags & GF_SLICED) && tall) {  // the tall split-precision kernel (same tiles)
    for (int tj = 0; tj < (n + kTallCols - 1) / kTallCols; ++tj)
      for (int ti = 0; ti < (m + kTallRows - 1) / kTallRows; ++ti) tl.ozt.push_back({prob, ti, tj});
    return;
  }
  if (g.flags & GF_SLICED) {
    auto& v = g.k > 64 ? tl.oz4 : tl.oz1;
    for (int tj = 0; tj < (n + 31) / 32; ++tj)
      for (int ti = 0; ti < (m + 31) / 32; ++ti) v.push_back({prob, ti, tj});
    return;
  }
  if (tall) {
    for (int tj = 0; tj < (n + kTallCols - 1) / kTallCols; ++tj)
      for (int ti = 0; ti < (m + kTallRows - 1) / kTallRows; ++ti) tl.tall.push_back({prob, ti, tj});
    return;
  }
  // 32-tiles for thin outputs, and for long reductions (k = nb) into outputs
  // of at most 128 x 128, whose phases share each tile's k range among four
  // wavefronts (split-k)
  const bool sm = m <= 32 || n <= 32 || (g.k >= 256 && m <= 128 && n <= 128);
  const int t = sm ? 32 : 64;
  auto& v = sm ? tl.small : tl.big;
  for (int tj = 0; tj < (n + t - 1) / t; ++tj)
    for (int ti = 0; ti < (m + t - 1) / t; ++ti) v.push_back({prob, ti, tj});
}

// The launches of one task list: 64-tiles, 32-tiles, tall-skinny blocks, and
// the split-precision tiles with `slices` digits (0: the same tiles through
// the FP64 matrix-core kernels, a plan's FP64 reference for the autotuner).
template <typename T>
void launch_tasks(const GemmProb<T>* probs, const char* base, const TaskRef& r,
                         hipStream_t s, bool split, int slices = 0) {
  launch_gemm<T>(probs, reinterpret_cast<const GemmTask*>(base + r.ob), r.nb, 64, s, split);
  launch_gemm<T>(probs, reinterpret_cast<const GemmTask*>(base + r.os), r.ns, 32, s, split);
  launch_gemm_tall<T>(probs, reinterpret_cast<const GemmTask*>(base + r.ot), r.nt, s);
  if constexpr (std::is_same<T, double>::value) {
    const auto* t1 = reinterpret_cast<const GemmTask*>(base + r.o1);
    const auto* t4 = reinterpret_cast<const GemmTask*>(base + r.o4);
    const auto* tT = reinterpret_cast<const GemmTask*>(base + r.oT);
    if (slices > 0) {
      launch_gemm_oz(probs, t1, r.n1, slices, 1, s);
      launch_gemm_oz(probs, t4, r.n4, slices, 4, s);
      launch_gemm_oz(probs, tT, r.nT, slices, 0, s);
    } else {
      launch_gemm<T>(probs, t1, r.n1, 32, s, false);
      launch_gemm<T>(probs, t4, r.n4, 32, s, true);
      launch_gemm_tall<T>(probs, tT, r.nT, s);
    }
  }
}

// ---- services of the single-call entry points (hcore_api.cpp) ----------------
// GPU context, initialised on first use (HCORE_HIP_OK or an error status).
int ensure_ready();
// True when ptr is device (or managed) memory of the HIP runtime.
bool device_pointer(const void* ptr);
// Record the calling thread's status for HCORE_hip_last_status.
void set_last_status(int status);
// Core SVD variant for a K x K core with the context's settings.
int core_variant_ctx(int K, size_t elem);
// The calling thread's stream and grow-only buffers (null on failure).
struct ThreadRes {
  hipStream_t stream;
  DevArena* work;           // workspace of the tile routines
  DevArena* stage;          // device staging of host operands
  DevArena* desc;           // device descriptors
  PinnedArena* pinned;      // host staging
  std::vector<char>* hdesc; // host image of the descriptors
};
bool thread_res(ThreadRes* out);
// Release the tile routines' process-wide state (hcore_tiles.cpp).
void release_tile_state();

}  // namespace hcore
