/* qotential_hip.h -- C-ABI of the MI355X RRQ close-evaluation library
 * (libqotential_hip.so).
 *
 * Conventions (same as the reference's Fortran/MEX boundary):
 *   - every integer is int64_t; arrays are caller-allocated, column-major:
 *     x(3,n) is x[3*j+c], A(m,n) is A[i+m*j];
 *   - index outputs (idxs, CSR target indices, COO rows/cols) are 1-based;
 *   - the *_mex_ symbols keep gfortran's trailing-underscore mangling and pass
 *     every argument by reference, so a Fortran or MEX caller binds them unchanged.
 * Status: functions returning int return 0 on success; qol_last_error() then
 * holds a message for a nonzero return.  The drop-in *_mex_ entry points keep
 * the reference's void signatures and report failure through qol_last_status().
 */
#ifndef QOTENTIAL_HIP_H
#define QOTENTIAL_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* qol_last_error(void);
int64_t qol_last_status(void);
/* RRQ pairs of this thread's last qol_rrq_mex_, qol_rrq_noalloc_ or
 * qol_getnearquad_lap_rrq_mex_ call whose Gauss-Newton projection onto the patch
 * had not stopped by its 40th and last step.  It stops when the step |du| + |dv|
 * falls below 1e-15 or below eight times the step that rounding of the residual
 * makes (8 eps |x| / the smallest singular value of the surface Jacobian, bounded
 * above); such a pair keeps its last iterate as its foot point.  *max_step (if not
 * NULL) receives the largest of those last steps, in units of the reference
 * triangle.  No counterpart in the reference, whose drop-ins report nothing of
 * the kind. */
int64_t qol_last_projection_unconverged(double* max_step);

/* ---------------- workload generation (host; qol_geo_build_gpu on the GPU) ----------------
 * Replaces stellarator_geo_ntri / stellarator_geo / stellarator_geo_charts
 * (utils/stellarator_geo_mex.c:548-552, 758-770, 787-877).  nmode == 0 selects
 * the built-in 7-term surface; otherwise mnrz holds nmode rows (m, n, Rc, Zs)
 * of the VMEC form.  iterate_aspect: 0 = single aspect pass (built-in surface),
 * 1 = aspect splits iterated to convergence (build_charts_adaptive, restol<=0). */
int64_t qol_geo_count(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode,
                      const double* mnrz, int64_t iterate_aspect);
int qol_geo_build(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode,
                  const double* mnrz, int64_t iterate_aspect, double* sx, double* snx,
                  double* sw, double* rts, double* rps, double* charts);
/* qol_geo_build on the GPU: the same arguments and outputs, with the per-chart
 * work of stellarator_geo / stellarator_geo_charts (utils/stellarator_geo_mex.c:
 * 636-749: tensor grid, spectral derivatives, node interpolation, normals and
 * weights) in one GPU wave per chart; the chart list and diagonals are built on
 * the host as in qol_geo_build.  Nodes agree with qol_geo_build to the rounding
 * of the device sin/cos.  Synchronous; host output arrays. */
int qol_geo_build_gpu(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode,
                      const double* mnrz, int64_t iterate_aspect, double* sx, double* snx,
                      double* sw, double* rts, double* rps, double* charts);
/* exact chart boundary per triangle (ichart = 1 input of rrq): replaces the
 * uvbd table + stellarator_geo_uv2x loop of stellarator_grf_core_mod.f90:613-629,
 * 819-828.  sxbd(3, 3*nquad, ntri), rv(3, 3, ntri). */
int qol_geo_chart_boundary(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode,
                           const double* mnrz, int64_t iterate_aspect, int64_t nquad,
                           double* sxbd, double* rv);
/* The adaptive chart list of stellarator_charts (utils/stellarator_geo_mex.c:
 * 779-808, build_charts_adaptive :427-546): aspect splits iterated, then for
 * restol > 0 four-way splits until the mean curvature's tensor-Legendre tail
 * (panel_resolution, :359-411) is at most restol, then 2:1 balance between face
 * neighbours.  Returns the chart count (charts(6, n): a1 b1 a2 b2 jj ii; may be
 * NULL to count only), -1 when the list would pass cap (cap <= 0: 200000), -2 on
 * error.  A triangle list has 2 n triangles. */
int64_t qol_geo_charts(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode,
                       const double* mnrz, double restol, int64_t cap, double* charts);
/* qol_geo_charts with each curvature pass's tails evaluated on the GPU (one wave
 * per chart); the same list up to the rounding of the device sin/cos near restol */
int64_t qol_geo_charts_gpu(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode,
                           const double* mnrz, double restol, int64_t cap, double* charts);
/* The chart list a driver's mesh initialisation builds (the reference's
 * stellarator_mesh_init_r64, called at test/stellarator_bench.f90:60 and
 * web/stellarator/fortran/stellarator_grf_core_mod.f90:472): restol > 0 as
 * qol_geo_charts; otherwise the aspect-split list of qol_geo_build (one pass
 * when iterate_aspect is 0).  Returns the chart count, -1 past cap (cap > 0),
 * -2 on error; charts may be NULL to count only. */
int64_t qol_geo_plan_charts(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode,
                            const double* mnrz, int64_t iterate_aspect, double restol, int64_t cap,
                            double* charts);
/* panel_resolution (utils/stellarator_geo_mex.c:359-411) of nchart charts on the
 * GPU: tails(nchart) */
int qol_geo_curvature_tails_gpu(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode,
                                const double* mnrz, int64_t nchart, const double* charts, double* tails);
/* panel_resolution of one chart (6 values as above); -1 on error */
double qol_geo_curvature_tail(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode,
                              const double* mnrz, const double* chart);
/* the geometry and the chart boundary of a given chart list, as qol_geo_build,
 * qol_geo_build_gpu and qol_geo_chart_boundary (stellarator_geo_charts,
 * utils/stellarator_geo_mex.c:837-877) */
int qol_geo_build_charts(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode,
                         const double* mnrz, int64_t nchart, const double* charts, double* sx,
                         double* snx, double* sw, double* rts, double* rps);
int qol_geo_build_charts_gpu(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode,
                             const double* mnrz, int64_t nchart, const double* charts, double* sx,
                             double* snx, double* sw, double* rts, double* rps);
int qol_geo_chart_boundary_charts(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode,
                                  const double* mnrz, int64_t nchart, const double* charts,
                                  int64_t nquad, double* sxbd, double* rv);
/* near-surface targets of the benchmark workloads (no reference counterpart:
 * the reference drivers use the surface nodes themselves as targets) */
int qol_make_targets(int64_t p, int64_t ntri, const double* sx, const double* snx,
                     int64_t nt, double umin, double umax, int64_t sides, uint64_t seed,
                     double* tx, int64_t* tri, double* dist);

/* ---------------- per-order tables ----------------
 * qol_simplex_precomp_mex_ replaces simplex_precomp_r64 / qol_simplex_precomp_mex
 * (src/lap3d_close_mod.f90:115-133, src/lap3d_close_mex.f90:48-59): Gauss-Legendre
 * nodes/weights/differentiation matrix, barycentric weights, Legendre
 * values->coefficients, Koornwinder values<->coefficients at the VR nodes. */
void qol_simplex_precomp_mex_(const int64_t* nquad, const int64_t* korder, const int64_t* kpols,
                              double* tgl, double* wgl, double* Dgl, double* w_bclag,
                              double* Legmat, double* umatr, double* vmatr);
/* The level form every drop-in entry point uses: the reference's module
 * variables PR_ADAPTIVE and PR_SQUARE (src/patch_refine_mod.f90:51-55), which
 * its drivers set from iadap (test/stellarator_bench.f90:46-47).  square = 1
 * (the default) regroups levels 2-4 into Gauss-Legendre squares; square = 0
 * keeps triangle sub-patches (:229-240); adaptive = 1 (triangle levels only)
 * serves pairs by the adaptive walk over levels 1-3, then level 4, then RRQ
 * (:532-539, 617-842).  Both set is refused at the next evaluation, as the
 * reference stops ("PR_SQUARE and PR_ADAPTIVE cannot be combined", :116-117).
 * Process-wide, like the module variables; the order-table queries below
 * follow square. */
int qol_set_level_mode(int64_t adaptive, int64_t square);
int qol_get_level_mode(int64_t* adaptive, int64_t* square);
/* level sizes of patch order p: info[0..12] = h, nquad, ncoef, ntot, nsubtot,
 * nl(1..4), nsub(1..4).  No reference counterpart (the reference keeps them in
 * the module variable REF, src/patch_refine_mod.f90:51). */
int qol_order_info(int64_t p, int64_t* info);
/* one refinement level of patch order p (lev 1..4): uv(2,n), wt(n), interp(n,h)
 * column-major, suboff(nsub+1) 0-based, subtyp(nsub) -- the per-order table
 * that build_patch_ref_r64 keeps in REF (src/patch_refine_mod.f90:100-262). */
int qol_order_level(int64_t p, int64_t lev, double* uv, double* wt, double* interp,
                    int64_t* suboff, int64_t* subtyp);

/* ---------------- drop-in close evaluation ----------------
 * qol_rrq_mex_ replaces qol_rrq_mex -> rrq_r64 (src/lap3d_close_mex.f90:18-45,
 * src/lap3d_close_mod.f90:622-881): As, Ad (m,n), Omega (4h,m),
 * IalphaAsvestas (m) for m targets against one panel; timeinfo(20) accumulated. */
void qol_rrq_mex_(const int64_t* m, const double* tx, const int64_t* n, const double* sx,
                  const double* snx, const double* sw, const double* rts, const double* rps,
                  const int64_t* order, const int64_t* nquad, const int64_t* orderff,
                  const double* distff, const int64_t* exterior, const int64_t* isimd,
                  const int64_t* ichart, const double* sxbd_chart, const double* rv_chart,
                  const double* tgl, const double* wgl, const double* Dgl, const double* w_bclag,
                  const double* Legmat, const double* umatr, const double* vmatr, double* As,
                  double* Ad, double* Omega, double* IalphaAsvestas, double* timeinfo);
/* qol_rrq_noalloc_: no reference counterpart.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc): the arguments, outputs and error reporting of
 * qol_rrq_mex_, with every device buffer of the call taken from a fixed 64 MiB
 * device array per GPU (set up on the first call on that device) and no
 * hipMalloc / hipFree per call.  Calls needing more than the array -- about
 * 17,800 targets at order 8, 7,700 at order 12 -- fail with qol_last_status() = 1
 * and a message, before any device work.  Calls on one GPU share the array and
 * are serialised by the library. */
void qol_rrq_noalloc_(const int64_t* m, const double* tx, const int64_t* n, const double* sx,
                      const double* snx, const double* sw, const double* rts, const double* rps,
                      const int64_t* order, const int64_t* nquad, const int64_t* orderff,
                      const double* distff, const int64_t* exterior, const int64_t* isimd,
                      const int64_t* ichart, const double* sxbd_chart, const double* rv_chart,
                      const double* tgl, const double* wgl, const double* Dgl, const double* w_bclag,
                      const double* Legmat, const double* umatr, const double* vmatr, double* As,
                      double* Ad, double* Omega, double* IalphaAsvestas, double* timeinfo);
/* qol_getnearquad_lap_rrq_mex_ replaces qol_getnearquad_lap_rrq_mex ->
 * getnearquad_lap_rrq_r64 (src/lap3d_close_mex.f90:62-81,
 * src/lap3d_close_mod.f90:884-960): COO triplets of (S - S_naive, K - K_naive),
 * out index (tcxi(k)-1)*hdim + (j-1)*m + i, 1-based rows/cols. */
void qol_getnearquad_lap_rrq_mex_(const int64_t* npan, const int64_t* nterms, const int64_t* hdim,
                                  const int64_t* nquad, const int64_t* nsrc, const double* sx,
                                  const double* snx, const double* sw, const double* rts,
                                  const double* rps, const int64_t* ntcx, const double* tcx,
                                  const int64_t* tcxrow, const int64_t* tcxi,
                                  const int64_t* orderff, const double* distff,
                                  const int64_t* iside, const int64_t* isimd, const int64_t* nnz,
                                  int64_t* sparsei, int64_t* sparsej, double* sparsevs,
                                  double* sparsevd, double* timeinfo);
/* qol_lap3ddlp_closepanel_mex_ replaces qol_lap3ddlp_closepanel_mex ->
 * Lap3dDLP_closepanel_r64 (src/lap3d_close_mex.f90:2-15, src/lap3d_close_mod.f90:
 * 135-270): the close double-layer matrix Ac(m_tgt, npat) of one tensor-product
 * patch, npat = order^2 nodes on the Gauss-Legendre tensor grid (node b*order+a,
 * a the fast first-parameter index): sum_ip Ac(k, ip) sigma_ip is the DLP of the
 * interpolated density at t_x(:, k), kernel (x - y).n/(4 pi |x - y|^3), n along
 * y_u x y_v.  ref raises the leaf quadrature order (2 order + 2 ref); if_adapt_d
 * must be nonzero (the reference stops for the non-adaptive form).  order <= 16. */
void qol_lap3ddlp_closepanel_mex_(const int64_t* m_tgt, const double* t_x, const int64_t* npat,
                                  const double* s_x, const int64_t* order, const int64_t* ref,
                                  const double* if_adapt_d, double* Ac);
/* handle-based refinement levels, replacing src/patch_refine_mex.f90:198-242
 * (qol_build_patch_levels_mex, qol_lap3dsdlpmat_levels_mex,
 * qol_clear_patch_levels_mex, qol_clear_patch_levels_all_mex).  One slot,
 * generation-tagged handle; ier 0 ok, 1 no free slot, 2 bad handle,
 * 3 npols mismatch, 5 unsupported level set, 6 device failure. */
void qol_build_patch_levels_mex_(const int64_t* korder, const int64_t* npols, const double* sx,
                                 const double* rts, const double* rps, const int64_t* order,
                                 const double* dist, const int64_t* nlevel, int64_t* lv,
                                 int64_t* ier);
void qol_lap3dsdlpmat_levels_mex_(const int64_t* m, const double* tx, const int64_t* npols,
                                  const int64_t* lv, const int64_t* isimd, double* As, double* Ad,
                                  int64_t* idxs, int64_t* ms, int64_t* ier);
void qol_clear_patch_levels_mex_(const int64_t* lv, int64_t* ier);
void qol_clear_patch_levels_all_mex_(int64_t* ier);

/* ---------------- batched engine (GPU-resident surface) ----------------
 * No single reference counterpart: one engine does for a whole surface what
 * stellarator_grf_core_mod.f90:645-899 does panel by panel (ball count, CSR,
 * rrq_r64 + naive block per panel, apply, fixed-order scatter), with geometry,
 * targets, near list and weights resident in HBM.  device_ptrs != 0 means the
 * array arguments are device pointers. */
int qol_engine_create(int64_t p, int64_t npan, const double* sx, const double* snx,
                      const double* sw, const double* rts, const double* rps, int64_t ichart,
                      const double* sxbd, const double* rv, int64_t device_ptrs, void** engine);
int qol_engine_destroy(void* engine);
int qol_engine_set_targets(void* engine, int64_t nt, const double* tx, int64_t device_ptr);
/* near list by the ball rule (stellarator_grf_core_mod.f90:652-695) */
int qol_engine_nearlist(void* engine, int64_t* nnz);
/* the same CSR, bit for bit, through a uniform grid of the targets: each panel
 * tests only the cells its ball covers (the spatial search of the MATLAB
 * caller's k-d tree, test/torus_fmm3dbie_rrq_comparison.m:81-99), then sorts
 * its list into ascending target order */
int qol_engine_nearlist_grid(void* engine, int64_t* nnz);
/* the radius the next qol_engine_nearlist / _grid uses: diam_factor = 0 (the
 * default) is the reference's ball, 1.75 sqrt(sum sw) about the mean of the
 * panel's nodes; f > 0 is f h about the same centre, h the largest distance
 * between the corners of the panel's interpolant (f = 2: the "kd-tree radius 2h"
 * near list of the benchmark's order-8 configuration).  No counterpart in the
 * reference's entry points, whose callers build the list themselves. */
int qol_engine_set_near_radius(void* engine, double diam_factor);
/* caller's CSR by panel: offs(npan+1) 0-based, idx(nnz) 1-based target indices */
int qol_engine_set_pairs(void* engine, const int64_t* offs, const int64_t* idx);
int qol_engine_get_pairs(void* engine, int64_t* offs, int64_t* idx);
/* levels + RRQ weights for every pair.  Targets within side_tol*patch size of a
 * patch take the exterior (1) / interior (0) limit; others their own side. */
int qol_engine_eval(void* engine, int64_t exterior, double side_tol, int64_t want_omega);
int qol_engine_sync(void* engine);
/* stage times (ms) of the last eval, ms[10]: levels, classify (+ per-level pair
 * lists), direct, rrq setup, rrq solve, rrq projection, rrq root + SSQ,
 * rrq moments + recursion, rrq solid angle, rrq weights.  Stages on different
 * streams overlap. */
int qol_engine_stage_ms(void* engine, double* ms);
/* RRQ pairs of the last eval whose Gauss-Newton projection had not stopped by
 * its 40th and last step (they keep that last iterate), and in *max_step (if not
 * NULL) the largest such last step, as for qol_last_projection_unconverged;
 * waits for the engine's stream.  No counterpart in the reference. */
int qol_engine_projection_unconverged(void* engine, int64_t* count, double* max_step);
/* per-pair weights As, Ad (nnz x h, column-major), labels (1-4 the level, 5
 * RRQ, 0 the panel's own nodes), and, when the eval asked for them, Omega and
 * IalphaAsvestas; any output pointer may be NULL, and with As and Ad both NULL
 * the weights are not copied at all */
int qol_engine_get_weights(void* engine, double* As, double* Ad, int64_t* label, double* Omega,
                           double* Ias);
/* near-correction apply + fixed-order scatter into u (in: far field) */
int qol_engine_apply(void* engine, const double* ub, const double* ubn, double* u, double* val);
/* with ub, ubn already resident (from qol_engine_apply or qol_engine_far): the
 * corrections (made during eval when the density was resident before it) and the
 * per-target sum into the resident u */
int qol_engine_apply_resident(void* engine);
/* copy back the per-pair corrections val(nnz) and the resident u(nt) (no compute) */
int qol_engine_get_values(void* engine, double* val, double* u);
/* the resident potential u(nt) that the per-target sums add into (no compute) */
int qol_engine_set_potential(void* engine, const double* u);
/* far field by direct summation (stellarator_grf_core_mod.f90:560-611) */
int qol_engine_far(void* engine, const double* ub, const double* ubn, double* u);
/* this engine's level sizes: info[0..11] = ntot, nsubtot, nl(1..4), nsub(1..4),
 * square, adaptive (they follow its level form); size qol_engine_level_data's
 * buffers from them */
int qol_engine_level_info(void* engine, int64_t* info);
/* the level nodes lgeo(npan, 7, ntot) (x, y, z, nx, ny, nz, w) and balls
 * (npan, nsubtot, 4) of every panel */
int qol_engine_level_data(void* engine, double* lgeo, double* balls);
int qol_engine_triplets(void* engine, const int64_t* trow, int64_t* si, int64_t* sj, double* vs,
                        double* vd);
int qol_engine_stream(void* engine, void** stream);
/* The level form of one engine (see qol_set_level_mode); new level tables are
 * uploaded when square changes.  Adaptive levels take the ball-rule walk only
 * (no selector, forced labels or split precision).  Labels then read: 1 level 1
 * whole, 2 a tree cut of levels 2-3, 4 level 4, 5 RRQ. */
int qol_engine_set_level_mode(void* engine, int64_t adaptive, int64_t square);
/* Split precision (no reference counterpart; the reference computes every
 * entry in FP64, patch_refine_mod.f90:569-602): pairs served at levels
 * 1..split_levels get their direct S/D weights from the FP32 matrix cores
 * (x - y in FP64, entries and chunked accumulation in FP32, totals in FP64);
 * 0 (the default) keeps every level in FP64.  Takes effect at the next eval. */
int qol_engine_set_precision(void* engine, int64_t split_levels);

/* ---------------- learned level selector ----------------
 * No reference counterpart: the reference labels every pair by the ball rule
 * (src/patch_refine_mod.f90:541-613).  An MLP 10 -> 64 -> 64 -> 7 (ReLU; BF16
 * weights, FP32 biases and accumulation) on per-pair geometry features picks
 * class 0 (the panel's own nodes suffice, no correction), 1..4 (direct on that
 * level), 5 (RRQ) or 6 (direct on level 4 where the rule asks for RRQ); below
 * the confidence threshold the ball rule labels the pair.  params holds the
 * kSelNP = 5319 FP32 values W1 (64x10), W2 (64x64), W3 (7x64) row-major, then
 * b1 (64), b2 (64), b3 (7); norm holds the feature means (10) then 1/std (10).
 * Features: feat(10, n), see qol_engine_pair_features. */
int qol_selector_create(void** selector);
int qol_selector_destroy(void* selector);
int qol_selector_set_params(void* selector, const float* params, const float* norm);
int qol_selector_get_params(void* selector, float* params, float* norm);
/* one batch of all n samples: loss (class-weighted cross-entropy, mean) and its
 * gradient with respect to params (weights as rounded to BF16) */
int qol_selector_grad(void* selector, int64_t n, const float* feat, const int64_t* labels,
                      const float* class_w, float* grad, double* loss);
/* Adam (beta 0.9, 0.999; eps 1e-8) on FP32 master parameters, BF16 weights
 * rounded after every step; loss_epoch(epochs) gets each epoch's mean loss */
int qol_selector_train(void* selector, int64_t n, const float* feat, const int64_t* labels,
                       const float* class_w, int64_t epochs, int64_t batch, double lr, uint64_t seed,
                       double* loss_epoch);
/* class probabilities prob(7, n) */
int qol_selector_predict(void* selector, int64_t n, const float* feat, float* prob);
/* the engine labels pairs with the selector (NULL: the ball rule); the model
 * must outlive the engine's use of it */
int qol_engine_set_selector(void* engine, void* selector, double threshold);
/* every pair gets one label (0..5; -1 returns to the ball rule): evaluates all
 * pairs at one level, by RRQ, or with the panel's own nodes (the selector's
 * training labels come from such evaluations) */
int qol_engine_force_label(void* engine, int64_t label);
/* the ball rule with every level-served pair served `shift` (0..3) levels
 * finer, at most level 4 (a uniform-refinement comparison for the selector) */
int qol_engine_refine_levels(void* engine, int64_t shift);
/* the selector's 10 features of every pair, feat(10, nnz) */
int qol_engine_pair_features(void* engine, float* feat);

/* ---------------- mixed-order batches ----------------
 * No single reference counterpart: the reference evaluates a mixed-order
 * surface panel by panel with each panel's own order (lap3d_close_mod.f90:
 * 622-649 in the loop of stellarator_grf_core_mod.f90:759-888).  A batch holds
 * one engine per order (each already given targets, pairs and a resident
 * density, with one eager eval + apply done); qol_batch_capture records one
 * step of all of them -- eval, the near corrections and the per-target sums --
 * into one HIP graph, and qol_batch_launch replays it as one launch; capture
 * is switched off for now (its first run ended the process with a host-side
 * segmentation fault whose cause is not yet found), so qol_batch_capture
 * returns an error and qol_batch_launch launches the step directly.  The
 * engines must outlive the batch and keep their sizes.
 * qol_batch_autotune times, per engine, the four- and one-stream schedules and
 * (orders up to 8) the three RRQ solve variants, keeps the fastest, then the
 * engines' order in the graph; ms(6 n + 2) and choice(2 n + 1) report the
 * candidates' ms per launch (-1 where not applicable) and the choices
 * (one-stream flag, solve variant per engine; 1 = descending order). */
int qol_batch_create(int64_t n, void** engines, void** batch);
int qol_batch_destroy(void* batch);
int qol_batch_capture(void* batch, int64_t exterior, double side_tol);
int qol_batch_launch(void* batch);
int qol_batch_sync(void* batch);
int qol_batch_autotune(void* batch, int64_t reps, double* ms, int64_t* choice);

/* No reference counterpart: one plain device-to-device copy of nbytes (a
 * multiple of 16) through hbm_copy_kernel, on buffers it allocates and frees,
 * synchronous; *ms = its event time.  A workload of known HBM traffic
 * (nbytes read + nbytes written) that scales the byte counters of a profiler. */
int qol_hbm_copy(int64_t nbytes, double* ms);

#ifdef __cplusplus
}
#endif
#endif
