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
  int pad;               // 16 * ceil(h / 16)
  const double* wt;      // ntot
  const int* suboff;     // per level (nsub+1) node offsets, concatenated
  const int* subtyp;     // nsubtot
  double dist2;
  const int* chunks;     // nchunk x {level, first node, end node, first sub-patch, end sub-patch}
  int nchunk;            // (node and sub-patch indices within the level)
};

// per-order RRQ tables on the device
struct RRQTables {
  int p, h, nq, ncoef;
  const double* tgl;     // nq
  const double* wgl;     // nq
  const double* Legmat;  // nq x nq (coef = Legmat * vals)
  const double* umatr;   // h x h
  const double* uvn;     // 2 x h VR nodes
  const double* ktab;    // Koornwinder evaluation constants (koorn_eval_table(p))
};
std::vector<double> koorn_eval_table(int N);

// per-panel RRQ data on the device (one slot per panel)
struct RRQPanels {
  double* frame;   // [npan][16]: R(9) c(3) alpha sigma L pad
  double* edge;    // [npan][3][3][nq] Legendre coefficients of the frame boundary
  double* bX;      // [npan][3*nq][6] boundary X, dX (frame)
  double* cx;      // [npan][3][h] Koornwinder coefficients of the geometry (original coords)
  double* basis;   // [npan][3*nq][ncoef][12] grad H (3 complex) and V (3 complex)
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
__global__ void classify_kernel(LevelArgs A, const int* offs, const int* tidx, const double* tx,
                                const double* balls, int8_t* label, int* ppanel, int* nrrq, int* lcount);

__global__ void level_scan_kernel(const int* lcount, int npan, int* lstart, int* ltot);
__global__ void compact_kernel(const int* offs, const int8_t* label, const int* lstart, const int* ltot, int npan,
                               int* lpairs);
void launch_direct_mfma(LevelArgs A, const int* lpairs, const int* ltot, const int* ppanel, const int* tidx,
                        const double* tx, const double* lgeo, double* wout, int64_t nnz, hipStream_t s);

__global__ void nearlist_prep_kernel(int npan, int h, const double* sx, const double* sw, double* ball);
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
void launch_rrq_panel_solve(RRQTables T, RRQPanels P, int npan, const int* nrrq, hipStream_t s);
bool rrq_solve_g1_in_lds(int h);
void launch_rrq_project(RRQTables T, RRQPanels P, const int* lpairs, const int* ltot, const int* ppanel,
                        const int* tidx, const double* tx, const double* sx, int exterior, double side_tol,
                        double* pq, int* order, int64_t nnz, hipStream_t s);
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
size_t rrq_target_smem(int h, int nq, int ncoef);

}  // namespace qol
