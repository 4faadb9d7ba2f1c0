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

// Where device buffers come from: hipMalloc by default; while the calling
// thread has an arena set, DevBuf::alloc takes 256-byte aligned slices of it
// instead and the buffer is never freed (the arena's owner keeps the memory).
struct Arena {
  char* base = nullptr;
  size_t cap = 0, used = 0;
};
Arena*& thread_arena();

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  bool owned = false;  // from hipMalloc (freed here) or from an arena (not)
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
  DevBuf d_interp, d_imp, d_impf, d_wt, d_suboff, d_subtyp, d_chunks, d_tgl, d_wgl, d_Leg, d_umatr, d_uvn, d_ktab;
  DevBuf d_corner;  // 3 x h rows interpolating a panel at its corners (selector features)
  DevBuf d_aseg;    // triangle levels: the adaptive tree cuts' leaf cells
  DevBuf d_l2m;     // Legendre to power-basis coefficients, nq x nq
  DeviceTables(i64 p, bool square, hipStream_t s);
};

// Streams and events an engine borrows instead of creating (and destroying) its
// own; all four stream roles may be the same stream.
struct StreamSet {
  hipStream_t stream = nullptr, s1 = nullptr, s2 = nullptr, s3 = nullptr;
  hipEvent_t tev[10][2] = {};
  hipEvent_t ev_start = nullptr, ev_lists = nullptr, ev_direct = nullptr, ev_setup = nullptr, ev_system = nullptr,
             ev_solve = nullptr, ev_edge = nullptr, ev_proj = nullptr;
  void create_one_stream();  // one non-blocking stream in every role, events for every stage
};

struct Engine {
  int p = 0, h = 0, nq = 0, ncoef = 0, npan = 0, ichart = 0;
  size_t nsrc = 0;
  int nt = -1;
  i64 nnz = -1;
  bool levels_built = false, evaluated = false, have_omega = false, rrq_alloc = false;
  bool borrowed = false;      // streams and events belong to a StreamSet
  bool density_set = false;   // ub, ubn resident (eval then also makes the corrections)
  bool val_current = false;   // d_val holds the corrections of the last eval
  i64 tlists_nnz = -1;
  int split_levels = 0;       // levels 1..split_levels use the split-precision (FP32) direct kernel
  SelectorArgs sel{};         // how pairs are labelled: ball rule, forced label, or learned selector
  bool timing = true;         // record the stage timing events (off while a step is captured into a graph)
  bool one_stream_run = false;  // run every stage on the main stream (a scheduling choice; same results)
  int solve_variant = 0;      // RRQ solve kernels: 0 default, 1 fused elimination, 2 matrix-core elimination
  bool square = true;         // level form: square regroup (PR_SQUARE) or triangle sub-patches
  bool adaptive = false;      // triangle levels served by the adaptive walk (PR_ADAPTIVE)
  double near_diam = 0.0;     // near list: 0 the reference ball, f > 0 radius f x (largest corner distance)
  hipStream_t stream = nullptr, s1 = nullptr, s2 = nullptr, s3 = nullptr;
  hipEvent_t tev[10][2] = {};
  hipEvent_t ev_start = nullptr, ev_lists = nullptr, ev_direct = nullptr, ev_setup = nullptr, ev_system = nullptr, ev_solve = nullptr,
            ev_edge = nullptr, ev_proj = nullptr;
  std::unique_ptr<DeviceTables> tab;
  DevBuf d_sx, d_snx, d_sw, d_rts, d_rps, d_sxbd, d_rv;
  DevBuf d_tx, d_nball, d_cnt, d_offs, d_tidx;
  DevBuf d_label, d_ppanel, d_nrrq, d_lcount, d_lstart, d_ltot, d_lpairs, d_wout, d_val, d_omega, d_ias;
  DevBuf d_abucket, d_apairs, d_acnt;  // adaptive levels: tree cut per pair, pairs sorted by cut, cut counts
  DevBuf d_lgeo, d_balls, d_pq, d_w1, d_oe, d_root, d_eorder, d_vt;
  DevBuf d_pfail;              // projections ended at the step cap: count, largest last step
  bool projected = false;      // the last eval ran the projection (it had RRQ pairs to run)
  DevBuf d_frame, d_edge, d_bX, d_cx, d_basis, d_Fq, d_Msl, d_Fn, d_G, d_Binv, d_scr;
  DevBuf d_ub, d_ubn, d_u, d_tcount, d_toffs, d_tcur, d_tpairs;
  RRQPanels RP{};

  Engine(i64 p, i64 npan, const double* sx, const double* snx, const double* sw, const double* rts,
         const double* rps, i64 ichart, const double* sxbd, const double* rv, bool device_ptrs,
         const StreamSet* borrow = nullptr, bool square_levels = true);
  // Device bytes one engine of one panel takes for m targets (constructor,
  // set_targets, set_pairs and eval with Omega), each buffer rounded up to 256
  // bytes as an arena slices it.
  static size_t one_panel_bytes(i64 p, i64 m, i64 ichart, bool square = true);
  // new panel data of the same order, panel count and ichart; targets, pairs
  // and everything evaluated are dropped
  void set_surface(const double* sx, const double* snx, const double* sw, const double* rts, const double* rps,
                   const double* sxbd, const double* rv, bool device_ptrs);
  ~Engine();
  void set_level_mode(bool adaptive, bool square);
  void set_targets(i64 nt, const double* tx, bool device_ptr);
  i64 nearlist();
  i64 nearlist_grid();
  void set_pairs(const i64* offs, const i64* idx1);
  void get_pairs(i64* offs, i64* idx1);
  void build_levels();
  void alloc_levels();
  void alloc_pairs();
  void alloc_rrq();
  void eval(int exterior, double side_tol, bool want_omega);
  void eval_levels_only();
  void classify_lists();
  void direct(hipStream_t s);
  void stage_ms(double* ms);
  // RRQ pairs of the last eval whose Gauss-Newton projection had not stopped by its
  // last (40th) step, and the largest such last step |du| + |dv|; their foot
  // point is the last iterate (waits for the stream)
  i64 projection_unconverged(double* max_step = nullptr);
  void get_weights(double* As, double* Ad, i64* label, double* Omega, double* Ias);
  void pair_features(float* feat);
  void apply(const double* ub, const double* ubn, double* u, double* val, bool device_ptrs);
  void apply_device();
  void launch_correction(int first, int last, hipStream_t s);
  void build_target_lists();
  void far(const double* ub, const double* ubn, double* u);
  void get_values(double* val, double* u);
  void level_data(double* lgeo, double* balls);
  void triplets(const i64* trow, i64* si, i64* sj, double* vs, double* vd);
};

// engine stage times (ms: levels, classify, direct, rrq setup, rrq solve,
// projection, root + SSQ, moments+recursion, solid angle, weights) added into
// the reference's timeinfo slots (seconds, accumulated; lap3d_close_mod.f90:
// 714-871): 1 levels, 2 boundary precompute, 3 level direct, 4 SSQ, 5 solid
// angle (with the projection that orients it), 6 moments and weights, 8 solve
void add_timeinfo(const double* ms, double* timeinfo);

// the level form the drop-in entry points use (qol_set_level_mode: the
// reference's module variables PR_ADAPTIVE and PR_SQUARE)
void level_mode(bool* adaptive, bool* square);

// the selector arguments of a model made by qol_selector_create (mode 2), or
// the ball rule (mode 0) for a null model
SelectorArgs selector_args(void* mlp, double threshold);

}  // namespace qol
