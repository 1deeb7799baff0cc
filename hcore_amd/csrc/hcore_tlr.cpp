// hcore_tlr.cpp — the TLR Cholesky update stream: the right-looking tile
// low-rank Cholesky that HiCMA drives on top of HCORE (per step k: potrf of
// the diagonal tile, trsm of the panel's V factors, dsyrk of the diagonal
// tiles below, HCORE_dgemm recompression of every lower tile of the
// trailing matrix), run here as one batched plan per step on the GPU.
//
// HCORE_hip_tlr_dpotrf_form runs the recompressions as FP64 or as
// split-precision plans (BF16-digit products); potrf, trsm and dsyrk stay FP64.
//
// Storage (all device memory): dense diagonal tiles D[k] (nb x nb, ld nb);
// off-diagonal tiles (i > j) as U[i + j nt], V[i + j nt] (nb x cap, ld nb)
// with host int ranks rank[i + j nt].  Panel tiles whose V factors lie cap
// columns apart in one buffer are solved by one trsm per step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "hcore_d.h"
#include "hcore_hip.h"
#include "hcore_host.hpp"

extern "C" int HCORE_hip_tlr_dpotrf_form(int precision, int slices, int nt, int nb,
                                         double* const* D, double* const* U, double* const* V,
                                         int* rank, int cap, int maxrk, double acc, void* stream,
                                         int* info, flop_counter* flops, int* step_ranks_max) {
  using namespace hcore;
  if (info) *info = 0;
  if (nt < 1 || nb < 1 || !D || !rank || cap < 1 || maxrk < 1 || maxrk > cap ||
      (nt > 1 && (!U || !V)))
    return HCORE_HIP_ERR_ARG;
  if (precision != HCORE_HIP_FP64 && precision != HCORE_HIP_SPLIT) return HCORE_HIP_ERR_ARG;
  if (precision == HCORE_HIP_SPLIT && slices != 0 && (slices < kOzMinSlices || slices > kOzMaxSlices))
    return HCORE_HIP_ERR_ARG;
  for (int j = 0; j < nt; ++j) {
    if (!D[j]) return HCORE_HIP_ERR_ARG;
    for (int i = j + 1; i < nt; ++i) {
      const int t = i + j * nt;
      if (!U[t] || !V[t] || rank[t] < 1 || rank[t] > cap) return HCORE_HIP_ERR_ARG;
    }
  }
  int st = ensure_ready();
  if (st != HCORE_HIP_OK) return st;
  flop_counter fl{};
  // HCORE_TLR_PROFILE=1: host-side seconds per part of the stream on stderr
  static const bool prof = getenv("HCORE_TLR_PROFILE") != nullptr;
  using clk = std::chrono::steady_clock;
  double t_potrf = 0, t_trsm = 0, t_build = 0, t_run = 0, t_free = 0;
  auto since = [](clk::time_point t0) {
    return std::chrono::duration<double>(clk::now() - t0).count();
  };
  for (int k = 0; k < nt; ++k) {
    int pinfo = 0;
    set_last_status(HCORE_HIP_OK);
    auto t0 = clk::now();
    HCORE_dpotrf(HCORE_Lower, nb, D[k], nb, &pinfo, &fl);
    t_potrf += since(t0);
    if ((st = HCORE_hip_last_status()) != HCORE_HIP_OK) break;
    if (pinfo) {
      if (info) *info = k * nb + pinfo;
      break;
    }
    if (k == nt - 1) break;
    // panel: V_ik <- L_kk^{-1} V_ik for i > k (A_ik L_kk^{-T} = U_ik (L_kk^{-1} V_ik)^T)
    t0 = clk::now();
    bool contiguous = true;
    for (int i = k + 2; i < nt; ++i)
      contiguous &= V[i + k * nt] == V[k + 1 + k * nt] + (size_t)(i - k - 1) * cap * nb;
    if (contiguous) {
      HCORE_dtrsm(HCORE_Left, HCORE_Lower, HCORE_NoTrans, HCORE_NonUnit, nb,
                  cap * (nt - k - 1), 1.0, D[k], nb, V[k + 1 + k * nt], nb, &fl);
      st = HCORE_hip_last_status();
    } else {
      for (int i = k + 1; i < nt && st == HCORE_HIP_OK; ++i) {
        HCORE_dtrsm(HCORE_Left, HCORE_Lower, HCORE_NoTrans, HCORE_NonUnit, nb, rank[i + k * nt],
                    1.0, D[k], nb, V[i + k * nt], nb, &fl);
        st = HCORE_hip_last_status();
      }
    }
    t_trsm += since(t0);
    if (st != HCORE_HIP_OK) break;
    // trailing matrix: dsyrk on D_i (i > k), recompression of A_ij (i > j > k)
    t0 = clk::now();
    std::vector<void*> sAU, sAV, sCD;
    std::vector<int> sK;
    for (int i = k + 1; i < nt; ++i) {
      sAU.push_back(U[i + k * nt]);
      sAV.push_back(V[i + k * nt]);
      sK.push_back(rank[i + k * nt]);
      sCD.push_back(D[i]);
    }
    std::vector<void*> AU, AV, BU, BV, CU, CV;
    std::vector<int> ra, rb, rc, tile;
    for (int j = k + 1; j < nt; ++j)
      for (int i = j + 1; i < nt; ++i) {
        AU.push_back(U[i + k * nt]); AV.push_back(V[i + k * nt]); ra.push_back(rank[i + k * nt]);
        BU.push_back(U[j + k * nt]); BV.push_back(V[j + k * nt]); rb.push_back(rank[j + k * nt]);
        CU.push_back(U[i + j * nt]); CV.push_back(V[i + j * nt]); rc.push_back(rank[i + j * nt]);
        tile.push_back(i + j * nt);
      }
    hcore_hip_plan* p = nullptr;
    st = HCORE_hip_plan_create(&p, nb, precision);
    if (st == HCORE_HIP_OK && precision == HCORE_HIP_SPLIT && slices)
      st = HCORE_hip_plan_set_slices(p, slices);
    const int n = (int)AU.size();
    if (st == HCORE_HIP_OK && n)
      st = HCORE_hip_plan_add_dgemm(p, n, AU.data(), AV.data(), ra.data(), nb, BU.data(),
                                    BV.data(), rb.data(), nb, CU.data(), CV.data(), rc.data(), nb,
                                    nullptr, nullptr, nb, -1.0, 1.0, 0, maxrk, acc);
    if (st == HCORE_HIP_OK)
      st = HCORE_hip_plan_add_dsyrk(p, (int)sAU.size(), sAU.data(), sAV.data(), sK.data(), nb,
                                    sCD.data(), nb, -1.0);
    if (st == HCORE_HIP_OK) st = HCORE_hip_plan_finalize(p);
    t_build += since(t0);
    t0 = clk::now();
    if (st == HCORE_HIP_OK) st = HCORE_hip_plan_execute(p, stream);
    std::vector<int> nr(std::max(n, 1)), ns(std::max(n, 1));
    if (st == HCORE_HIP_OK && n) st = HCORE_hip_plan_get_ranks(p, nr.data());
    if (st == HCORE_HIP_OK && n) st = HCORE_hip_plan_get_status(p, ns.data());
    if (st == HCORE_HIP_OK && !n) st = HCORE_hip_stream_synchronize(stream);
    if (st == HCORE_HIP_OK) {
      unsigned long upd = 0, syrk = 0;
      HCORE_hip_plan_flops(p, &upd, &syrk);
      fl.update += upd;
      fl.syrk += syrk;
    }
    t_run += since(t0);
    t0 = clk::now();
    if (p) HCORE_hip_plan_destroy(p);
    t_free += since(t0);
    if (st != HCORE_HIP_OK) break;
    int rmax = 0;
    for (int u = 0; u < n; ++u) {
      if (ns[u] != HCORE_HIP_OK) {
        st = ns[u];
        break;
      }
      rank[tile[u]] = nr[u];
      rmax = std::max(rmax, nr[u]);
    }
    if (step_ranks_max) step_ranks_max[k] = rmax;
    if (st != HCORE_HIP_OK) break;
  }
  if (prof)
    fprintf(stderr,
            "HCORE_hip_tlr_dpotrf: potrf %.3f s, trsm %.3f s, plan build %.3f s, execute %.3f s, "
            "plan release %.3f s\n",
            t_potrf, t_trsm, t_build, t_run, t_free);
  if (flops) {
    flops->potrf += fl.potrf;
    flops->trsm += fl.trsm;
    flops->syrk += fl.syrk;
    flops->update += fl.update;
  }
  set_last_status(st);
  return st;
}

extern "C" int HCORE_hip_tlr_dpotrf(int nt, int nb, double* const* D, double* const* U,
                                    double* const* V, int* rank, int cap, int maxrk, double acc,
                                    void* stream, int* info, flop_counter* flops,
                                    int* step_ranks_max) {
  return HCORE_hip_tlr_dpotrf_form(HCORE_HIP_FP64, 0, nt, nb, D, U, V, rank, cap, maxrk, acc,
                                   stream, info, flops, step_ranks_max);
}
