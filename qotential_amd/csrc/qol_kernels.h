// Kernel interfaces shared by the HIP translation units and the host engine.
#pragma once
#include <vector>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qol_device.h"

namespace qol {

// per-order level tables on the device
struct LevelArgs {
  int h, ntot, nsubtot;
  int lev_n[4], lev_off[4], lev_nsub[4], sub_base[4];
  const double* interp;  // level l at interp + h*lev_off[l], column-major lev_n[l] x h
  const double* imp;     // level l at imp + pad*lev_off[l], row-major lev_n[l] x pad, zero-padded
  const float* impf;     // imp rounded to FP32 (the split-precision direct kernel's B operand)
  int pad;               // 16 * ceil(h / 16)
  const double* wt;      // ntot
  const int* suboff;     // per level (nsub+1) node offsets, concatenated
  const int* subtyp;     // nsubtot
  double dist2;
  const int* chunks;     // nchunk x {level, first node, end node, first sub-patch, end sub-patch}
  int nchunk;            // (node and sub-patch indices within the level)
  // the reference's adaptive quadtree (PR_ADAPTIVE, triangle levels only): n1
  // nodes per sub-patch and, per tree cut, the node offsets of its leaf cells
  int adaptive = 0;
  int n1 = 0;
  const int* aseg = nullptr;  // kAdaptBuckets x (1 + kAdaptSegs): count, then offsets into the levels' nodes
};

// Tree cuts of the adaptive walk over levels 1-3: cut 0 is level 1 whole, cut
// 1 + m (m a 4-bit mask) has the level-2 cells not in m as leaves and the four
// level-3 children of each cell in m.
constexpr int kAdaptBuckets = 17, kAdaptSegs = 16;
std::vector<int> adaptive_segments(const LevelArgs& A);
__global__ void zero_ints_kernel(int* p, int n);

// per-order RRQ tables on the device
struct RRQTables {
  int p, h, nq, ncoef;
  const double* tgl;     // nq
  const double* wgl;     // nq
  const double* Legmat;  // nq x nq (coef = Legmat * vals)
  const double* leg2mon; // nq x nq: [k * nq + j] = coefficient of s^j in P_k(s)
  const double* umatr;   // h x h
  const double* uvn;     // 2 x h VR nodes
  const double* ktab;    // Koornwinder evaluation constants (koorn_eval_table(p))
};
std::vector<double> koorn_eval_table(int N);

// per-panel RRQ data on the device (one slot per panel)
struct RRQPanels {
  double* frame;   // [npan][16]: R(9) c(3) alpha sigma L pad
  double* edge;    // [npan][3][3][nq] power-basis coefficients of the frame boundary (s in [-1, 1])
  double* bX;      // [npan][3*nq][6] boundary X, dX (frame)
  double* cx;      // [npan][3][h] Koornwinder coefficients of the geometry (original coords)
  double* basis;   // [npan][3*nq][p(p+1)/2] complex: H'_n^m (n < p) at the boundary nodes
  double* Fq;      // [npan][h][h][3] pure quaternions grad F_k(y_i)
  double* Msl;     // [npan][h][h] Msl^T (column-major)
  double* Fn;      // [npan][h][h] F_k at the nodes, F(i,k)
  double* G;       // [npan][5h][2h] weights operator (transposed)
  double* Binv;    // [npan][h][h] (Msl^T)^{-1}, row-major
  double* scr;     // [npan][4h^2] G1 when it does not fit in LDS, else null
};

__global__ void level_geom_kernel(LevelArgs A, int npan, const double* sx, const double* rts, const double* rps,
                                  double* lgeo, double* balls);
size_t level_geom_smem(int h);
std::vector<int> level_chunks(const LevelArgs& A, const int* suboff_host);
int level_geom_blocks(int npan);
void launch_level_geom(const LevelArgs& A, int npan, const double* sx, const double* rts, const double* rps,
                       double* lgeo, double* balls, hipStream_t s);
// uniform grid over the panel balls for the near-list search
struct GridArgs {
  double lo[3];   // grid corner
  double inv_c;   // 1 / cell side
  int n[3];       // cells per dimension
};
__global__ void grid_bin_kernel(int nt, const double* tx, GridArgs G, int* cell_of, int* cell_count);
__global__ void grid_scatter_kernel(int nt, const int* cell_of, const int* cell_start, int* cursor, int* cell_targets);
__global__ void grid_panel_kernel(int npan, const double* ball, GridArgs G, const int* cell_start,
                                  const int* cell_targets, const double* tx, int fill, int* count, const int* offs,
                                  int* cursor, int* idx);
__global__ void segment_sort_kernel(int npan, const int* offs, int* idx);

// Learned level selector: an MLP 10 -> 64 -> 64 -> 7 (ReLU) on per-pair geometry
// features, BF16 weights with FP32 biases and accumulation.  Classes: 0 the
// panel's own nodes suffice (no correction), 1..4 direct on level 1..4,
// 5 RRQ, 6 direct on level 4 although the ball rule asks for RRQ.
constexpr int kSelIn = 10, kSelHid = 64, kSelOut = 7;
constexpr int kSelW1 = kSelHid * kSelIn, kSelW2 = kSelHid * kSelHid, kSelW3 = kSelOut * kSelHid;
constexpr int kSelNW = kSelW1 + kSelW2 + kSelW3;  // weights: W1 (64x10), W2 (64x64), W3 (7x64), row-major
constexpr int kSelNB = 2 * kSelHid + kSelOut;     // biases b1, b2, b3
constexpr int kSelNP = kSelNW + kSelNB;
struct SelectorArgs {
  int mode = 0;             // 0 ball rule, 1 every pair gets `forced`, 2 learned (rule below threshold),
                            // 3 ball rule with level pairs served `forced` levels finer (at most level 4)
  int forced = 0;           // label 0..5 for mode 1, level shift for mode 3
  float threshold = 0.9f;   // softmax confidence below which the ball rule decides
  const uint16_t* w = nullptr;   // kSelNW BF16 weights
  const float* b = nullptr;      // kSelNB biases
  const float* norm = nullptr;   // feature mean (10) then 1/std (10)
  const double* corner = nullptr;  // 3 x h rows interpolating the panel at its corners
};
// the label a selector class serves: 0 none, 1..4 levels, 5 RRQ (class 6 -> level 4)
__host__ __device__ inline int selector_label(int cls) { return cls == 6 ? 4 : cls; }

// abucket (adaptive levels only): the tree cut of every pair the tree serves (label 2)
void launch_classify(const LevelArgs& A, const SelectorArgs& S, int npan, const int* offs, const int* tidx,
                     const double* tx, const double* sx, const double* snx, const double* sw, const double* balls,
                     int8_t* label, int* ppanel, int* nrrq, int* lcount, int8_t* abucket, hipStream_t s);
size_t classify_smem(const LevelArgs& A, const SelectorArgs& S);
// per-pair selector features (raw, before normalisation), feat[nnz][10]
__global__ void pair_features_kernel(LevelArgs A, const double* corner, const int* offs, const int* tidx,
                                     const double* tx, const double* sx, const double* snx, const double* sw,
                                     const double* balls, float* feat);
// naive weights for the pairs labelled 0 (their correction is then exactly 0)
__global__ void naive_fill_kernel(int nnz, int h, const int8_t* label, const int* tidx, const int* ppanel,
                                  const double* tx, const double* sx, const double* snx, const double* sw,
                                  double* wout, double* val);

__global__ void level_scan_kernel(const int* lcount, int npan, int* lstart, int* ltot);
__global__ void compact_kernel(const int* offs, const int8_t* label, const int* lstart, const int* ltot, int npan,
                               int* lpairs);
// direct S/D of the pairs of levels lo..hi (0-based, inclusive) on FP64 matrix cores
// the pairs of list 1 (label 2) served by an adaptive tree cut: sorted by cut
// (acnt: 3 x kAdaptBuckets ints of counts, starts and cursors; apairs: nnz),
// then direct S/D on the cut's leaf nodes contracted with their interpolation rows
void launch_direct_adaptive(LevelArgs A, const int* ltot, const int* lpairs, const int8_t* abucket, int* acnt,
                            int* apairs, const int* ppanel, const int* tidx, const double* tx, const double* lgeo,
                            double* wout, int64_t nnz, hipStream_t s);
void launch_direct_mfma(LevelArgs A, const int* lpairs, const int* ltot, const int* ppanel, const int* tidx,
                        const double* tx, const double* lgeo, double* wout, int64_t nnz, hipStream_t s, int lo = 0,
                        int hi = 3);
// the same on FP32 matrix cores: x - y in FP64, kernel entries in FP32, FP32
// accumulation over chunks of 16 level nodes added into FP64
void launch_direct_split(LevelArgs A, const int* lpairs, const int* ltot, const int* ppanel, const int* tidx,
                         const double* tx, const double* lgeo, double* wout, int64_t nnz, hipStream_t s, int lo,
                         int hi);

__global__ void nearlist_prep_kernel(int npan, int h, const double* sx, const double* sw, const double* corner,
                                     double diam, double* ball);
__global__ void nearlist_count_kernel(int npan, int nt, const double* tx, const double* ball, int* count);
__global__ void nearlist_fill_kernel(int npan, int nt, const double* tx, const double* ball, const int* offs,
                                     int* idx);
__global__ void scan_kernel(const int* in, int n, int* out);

__global__ void naive_apply_kernel(int nnz, int h, const int* tidx, const int* ppanel, const double* tx,
                                   const double* sx, const double* snx, const double* sw, const double* ub,
                                   const double* ubn, const double* wout, double* val);
__global__ void naive_apply_list_kernel(const int* lpairs, const int* ltot, int first, int last, int h,
                                        const int* tidx, const int* ppanel, const double* tx, const double* sx,
                                        const double* snx, const double* sw, const double* ub, const double* ubn,
                                        const double* wout, double* val);
__global__ void target_count_kernel(int nnz, const int* tidx, int* tcount);
__global__ void target_fill_kernel(int nnz, const int* tidx, const int* toffs, int* tcur, int* tpairs);
__global__ void target_sum_kernel(int nt, const int* toffs, int* tpairs, const double* val, double* u);
__global__ void hbm_copy_kernel(size_t n, const double2* src, double2* dst);
__global__ void far_direct_kernel(int nt, const double* tx, int ns, const double* sx, const double* snx,
                                  const double* sw, const double* ub, const double* ubn, double* u);
__global__ void triplet_kernel(int nnz, int h, const int* tidx, const int* ppanel, const int* trow,
                               const int* offs, const double* tx, const double* sx, const double* snx,
                               const double* sw, const double* wout, int64_t* si, int64_t* sj, double* vs,
                               double* vd);

// RRQ
__global__ void rrq_panel_frame_kernel(RRQTables T, RRQPanels P, int npan, const int* nrrq, const double* sx,
                                       const double* snx, const double* sw, int ichart, const double* sxbd_chart,
                                       const double* rv_chart);
__global__ void rrq_panel_system_kernel(RRQTables T, RRQPanels P, int npan, const int* nrrq, const double* sx,
                                        const double* snx);
// variant 0: register-resident inverses up to order 8 (QOL_SOLVE_FUSED=1 / QOL_SOLVE_MFMA=1
// select the others), 1: the fused vector-unit elimination, 2: the matrix-core
// quaternion elimination (orders up to 8; the default above)
void launch_rrq_panel_solve(RRQTables T, RRQPanels P, int npan, const int* nrrq, hipStream_t s, int variant = 0);
bool rrq_solve_g1_in_lds(int h);
void launch_rrq_project(RRQTables T, RRQPanels P, const int* lpairs, const int* ltot, const int* ppanel,
                        const int* tidx, const double* tx, const double* sx, int exterior, double side_tol,
                        double* pq, int* order, int* nfail, int64_t nnz, hipStream_t s);
// root + SSQ kernel, then (after event `mid`, if given) the solid-angle kernel
size_t rrq_edge_order_ints(int64_t nnz);
void launch_rrq_edge(RRQTables T, RRQPanels P, const int* lpairs, const int* ltot, const int* ppanel,
                     const int* tidx, const double* tx, const double* pq, double* w1, double* root, double* oe,
                     int* order, int64_t nnz, hipStream_t s, hipEvent_t mid);
__global__ void rrq_target_kernel(RRQTables T, RRQPanels P, const int* lpairs, const int* lstart, const int* ltot,
                                  int npan, const int* tidx, const double* tx, const int* nrrq, const double* w1g,
                                  const double* oeg, double* vt, double* omega_out, double* ias_out);
void launch_rrq_weights(RRQTables T, RRQPanels P, const int* lpairs, const int* lstart, const int* ltot, int npan,
                        const int* nrrq, const double* vt, double* wout, hipStream_t s);
size_t rrq_frame_smem(int h, int nq);
size_t rrq_system_smem(int h, int nq, int ncoef);
size_t rrq_target_smem(int p, int h, int nq, int ncoef);

}  // namespace qol
