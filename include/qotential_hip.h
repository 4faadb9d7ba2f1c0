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

/* ---------------- workload generation (host only, no GPU) ----------------
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
/* exact chart boundary per triangle (ichart = 1 input of rrq): replaces the
 * uvbd table + stellarator_geo_uv2x loop of stellarator_grf_core_mod.f90:613-629,
 * 819-828.  sxbd(3, 3*nquad, ntri), rv(3, 3, ntri). */
int qol_geo_chart_boundary(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode,
                           const double* mnrz, int64_t iterate_aspect, int64_t nquad,
                           double* sxbd, double* rv);
/* near-surface targets of the benchmark workloads (no reference counterpart:
 * the reference drivers use the surface nodes themselves as targets) */
int qol_make_targets(int64_t p, int64_t ntri, const double* sx, const double* snx,
                     int64_t nt, double umin, double umax, int64_t sides, uint64_t seed,
                     double* tx, int64_t* tri, double* dist);

#ifdef __cplusplus
}
#endif
#endif
