// hcore_noalloc.cpp — HCORE_dgemm_noalloc: HCORE_dgemm on device pointers
// without any device allocation, for callers that cannot allocate (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The one-update workspace lives in a static device array, `scratch`, of
// kScratchBytes; every GPU holds its own copy until the process ends.  Its
// device address is taken once per device with that device current
// (hipGetSymbolAddress) and only that address is used afterwards.  Calls on
// one GPU share the array and keep the last call's contents (the status
// words are cleared by every call), so calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "hcore_d.h"
#include "hcore_hip.h"
#include "hcore_internal.hpp"

namespace hcore {

constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];
constexpr int kMaxScratchDevices = 64;
static char* g_scratch_ptr[kMaxScratchDevices] = {};
static std::mutex g_scratch_mu;

// Device address of the current GPU's scratch array, looked up once per
// device with that device current.
static char* scratch_for_current_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_scratch_ptr[dev] = reinterpret_cast<char*>(ptr);
  }
  return g_scratch_ptr[dev];
}

}  // namespace hcore

extern "C" {

void HCORE_dgemm_noalloc(HCORE_enum transA, int transB, int M, int N, double alpha, double* AU,
                         double* AV, double* Ark, int LDA, double* BU, double* BV, double* Brk,
                         int LDB, double beta, double* CU, double* CV, double* Crk, int LDC, int rk,
                         int maxrk, double acc, double* work, flop_counter* flops) {
  (void)transA; (void)transB; (void)N; (void)work;
  char* ws = hcore::scratch_for_current_device();
  if (!ws) {
    hcore::report_failure(HCORE_HIP_ERR_RUNTIME,
                          "HCORE_dgemm_noalloc: no device address for the static workspace");
    return;
  }
  // the null stream: no stream is created
  hcore::dgemm_fixed_workspace(M, alpha, AU, AV, Ark, LDA, BU, BV, Brk, LDB, beta, CU, CV, Crk,
                               LDC, rk, maxrk, acc, flops, "HCORE_dgemm_noalloc", ws,
                               hcore::kScratchBytes, nullptr);
}

size_t HCORE_dgemm_noalloc_workspace_bytes(void) { return hcore::kScratchBytes; }

}  // extern "C"
