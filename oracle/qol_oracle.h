/* Oracle: a plain CPU FP64 restatement of qotential's RRQ close-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (qotential_amd/) may call,
 * link or load this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker or the timed CPU baseline.
 *
 * Array conventions follow the reference Fortran: column-major, x(3,n) is
 * x[3*j + c], A(m,n) is A[i + m*j]; index outputs (idxs, CSR) are 1-based.
 * Compiled with -O2 -ffp-contract=off and no fast-math (reference FP semantics,
 * Makefile:95-99), so every sum is a left-to-right IEEE sum.
 *
 * How each part is pinned (tests/test_oracle_*.py); the reference itself cannot
 * be built (its QuatApproximation / LineQuaaadrature submodules are absent):
 *   gauss, legendre_vals2coefs, bclag_weights   GL-8 of test/test_lap3d_close.f90:35-42,
 *                                               numpy leggauss / legfit, barycentric exactness
 *   vr_nodes                                    table of utils/get_vioreanu_nodes.m, exactness
 *   koorn_pols, lu_inverse, simplex_precomp     orthonormality, vmatr*umatr = I, interpolation
 *   patch_ref                                   level counts, exact composite rules
 *   build_patch_levels, lap3dsdlpmat_levels     brute-force panel quadrature
 *   lap3dsdlpmat                                closed form, coincident zero, Gauss's law
 *   ssq_w1                                      adaptive quadrature; exact on straight edges
 *   solid angle                                 Van Oosterom-Strackee on flat triangles
 *   rrq                                         brute force on smooth densities (spectral in p)
 *   nearlist                                    137,259 pairs on the order-4 fixture surface
 *   close_apply, far_direct, getnearquad        fixture GRF, Gauss's law, COO consistency
 * Parity partially pinned: the level label rule (which level serves a pair) is
 * the restatement of patch_refine_mod.f90:516-615 with no reference vector of
 * labels; the results each label produces are pinned by brute force.
 */
#pragma once
#include <cstdint>
#include <vector>
#include <complex>

typedef int64_t i64;
typedef std::complex<double> cplx;

namespace orc {

/* ---- 1-D and simplex primitives (tables.cpp) ---- */
void gauss(i64 n, double* t, double* w, double* D);
void bclag_weights(i64 n, const double* t, double* wb);
void legendre_vals2coefs(i64 n, const double* t, const double* w, double* Lmat);
void koorn_pols(double u, double v, i64 nmax, double* pols, double* du = nullptr,
                double* dv = nullptr);
void vr_nodes(i64 order, double* uv, double* w);
i64 vr_count(i64 order);
void lu_inverse(i64 n, double* A);  /* in place, partial pivoting */
void simplex_precomp(i64 nquad, i64 korder, i64 kpols, double* tgl, double* wgl,
                     double* Dgl, double* w_bclag, double* Legmat, double* umatr,
                     double* vmatr);

/* ---- per-order level tables (build_patch_ref_r64) ---- */
struct RefLevel {
  i64 n = 0, nsub = 0;
  std::vector<double> uv, wt, interpmat; /* uv(2,n), wt(n), interpmat(n,npols) */
  std::vector<i64> suboff, subtyp;       /* suboff(nsub+1) 0-based offsets */
};
struct PatchRef {
  i64 korder = -1, npols = -1, order = -1, nlevel = -1;
  bool square = true;
  std::vector<RefLevel> lev;
};
/* PR_ADAPTIVE / PR_SQUARE (patch_refine_mod.f90:51-55); both true is refused
   (returns false, nothing changes), as the reference stops at :116-117 */
bool set_level_mode(bool adaptive, bool square);
bool level_adaptive();
bool level_square();
/* the tables of the current PR_SQUARE */
const PatchRef& patch_ref(i64 korder, i64 npols, i64 order, i64 nlevel);

/* ---- per-panel levels (build_patch_levels_r64) ---- */
struct PanelLevel {
  i64 n = 0, nsub = 0;
  std::vector<double> sx, snx, sw, qpoint, qradii2;
};
struct PanelLevels {
  std::vector<PanelLevel> lev;
};
void build_patch_levels(i64 korder, i64 npols, const double* sx, const double* rts,
                        const double* rps, i64 order, double dist, i64 nlevel,
                        PanelLevels& lv);
/* levels classification + direct + contraction; label(j)=level 1..nlevel or
 * nlevel+1 for targets left to RRQ.  idxs are 1-based.  With PR_ADAPTIVE the
 * adaptive walk serves first: label 1 when level 1 is the whole cut, 2 for a
 * cut over levels 2-3 (cut(j) = 1 + the mask of opened level-2 cells), then
 * level nlevel (label nlevel) and RRQ (nlevel + 1). */
void lap3dsdlpmat_levels(i64 m, const double* tx, i64 npols, const PanelLevels& lv,
                         const PatchRef& ref, double* As, double* Ad, i64* idxs, i64* ms,
                         i64* label, i64* cut = nullptr);
/* the walk of one target (patch_refine_mod.f90:740-771): leaf cells (level
 * 0-based, cell) in walk order, ok = false when a level-ntree cell holds x */
void adaptive_walk(const PanelLevels& lv, i64 ntree, const double* x, std::vector<i64>& cl,
                   std::vector<i64>& cs, bool& ok);
void lap3dsdlpmat(i64 m, const double* tx, i64 n, const double* sx, const double* snx,
                  const double* sw, double* As, double* Ad);
/* the ball test of one level: x inside some sub-patch ball (strict <) */
bool inside_any_ball(const PanelLevel& L, const double* x);

/* ---- tensor-patch close DLP (closepanel.cpp): Ac(m, p^2), column-major ---- */
void closepanel_dlp(i64 m, const double* tx, i64 p, const double* sx, i64 ref, double* Ac);

/* ---- RRQ (rrq.cpp) ---- */
struct RRQStats {
  double t[20] = {0};
};
void rrq(i64 m, const double* tx, i64 n, const double* sx, const double* snx,
         const double* sw, const double* rts, const double* rps, i64 order, i64 nquad,
         i64 orderff, double distff, bool exterior, i64 ichart, const double* sxbd_chart,
         const double* rv_chart, const double* tgl, const double* wgl, const double* Dgl,
         const double* w_bclag, const double* Legmat, const double* umatr,
         const double* vmatr, double* As, double* Ad, double* Omega, double* Ias,
         double* timeinfo, i64* label);
/* projections since the last reset that had not stopped by their 40th step,
   and (max_step) the largest last step |du| + |dv| among them */
i64 projection_unconverged(bool reset, double* max_step);

}  // namespace orc
