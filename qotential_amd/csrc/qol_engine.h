// Host engine for the batched close-evaluation path (see engine.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>

#include "qol_host.h"
#include "qol_kernels.h"

#define HIPCHK(x)                                                                              \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess)                                                                      \
      throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + ":" + \
                               std::to_string(__LINE__) + ")");                                \
  } while (0)

namespace qol {

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
  void alloc(size_t b);
};

struct DeviceTables {
  int h = 0, ntot = 0, nsubtot = 0;
  LevelArgs LA{};
  RRQTables RT{};
  DevBuf d_interp, d_imp, d_wt, d_suboff, d_subtyp, d_chunks, d_tgl, d_wgl, d_Leg, d_umatr, d_uvn, d_ktab;
  DeviceTables(i64 p, hipStream_t s);
};

struct Engine {
  int p = 0, h = 0, nq = 0, ncoef = 0, npan = 0, ichart = 0;
  size_t nsrc = 0;
  int nt = -1;
  i64 nnz = -1;
  bool levels_built = false, evaluated = false, have_omega = false, rrq_alloc = false;
  bool density_set = false;   // ub, ubn resident (eval then also makes the corrections)
  bool val_current = false;   // d_val holds the corrections of the last eval
  i64 tlists_nnz = -1;
  hipStream_t stream = nullptr, s1 = nullptr, s2 = nullptr, s3 = nullptr;
  hipEvent_t tev[10][2] = {};
  hipEvent_t ev_start = nullptr, ev_lists = nullptr, ev_direct = nullptr, ev_setup = nullptr, ev_system = nullptr, ev_solve = nullptr,
            ev_edge = nullptr;
  std::unique_ptr<DeviceTables> tab;
  DevBuf d_sx, d_snx, d_sw, d_rts, d_rps, d_sxbd, d_rv;
  DevBuf d_tx, d_nball, d_cnt, d_offs, d_tidx;
  DevBuf d_label, d_ppanel, d_nrrq, d_lcount, d_lstart, d_ltot, d_lpairs, d_wout, d_val, d_omega, d_ias;
  DevBuf d_lgeo, d_balls, d_pq, d_w1, d_oe, d_root, d_eorder, d_vt;
  DevBuf d_frame, d_edge, d_bX, d_cx, d_basis, d_Fq, d_Msl, d_Fn, d_G, d_Binv, d_scr;
  DevBuf d_ub, d_ubn, d_u, d_tcount, d_toffs, d_tcur, d_tpairs;
  RRQPanels RP{};

  Engine(i64 p, i64 npan, const double* sx, const double* snx, const double* sw, const double* rts,
         const double* rps, i64 ichart, const double* sxbd, const double* rv, bool device_ptrs);
  ~Engine();
  void set_targets(i64 nt, const double* tx, bool device_ptr);
  i64 nearlist();
  void set_pairs(const i64* offs, const i64* idx1);
  void get_pairs(i64* offs, i64* idx1);
  void build_levels();
  void alloc_pairs();
  void alloc_rrq();
  void eval(int exterior, double side_tol, bool want_omega);
  void eval_levels_only();
  void classify_lists();
  void direct(hipStream_t s);
  void stage_ms(double* ms);
  void get_weights(double* As, double* Ad, i64* label, double* Omega, double* Ias);
  void apply(const double* ub, const double* ubn, double* u, double* val, bool device_ptrs);
  void apply_device();
  void launch_correction(int first, int last, hipStream_t s);
  void build_target_lists();
  void far(const double* ub, const double* ubn, double* u);
  void get_values(double* val, double* u);
  void level_data(double* lgeo, double* balls);
  void triplets(const i64* trow, i64* si, i64* sj, double* vs, double* vd);
};

}  // namespace qol
