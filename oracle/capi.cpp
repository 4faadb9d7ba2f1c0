/* Oracle C entry points (ctypes-callable).  TEST INFRASTRUCTURE ONLY.
 *
 * orc_getnearquad   restates getnearquad_lap_rrq_r64 (lap3d_close_mod.f90:884-960)
 * orc_nearlist      restates the ball rule (stellarator_grf_core_mod.f90:652-695)
 * orc_close_apply   restates the per-panel rrq + naive block + apply
 *                   (stellarator_grf_core_mod.f90:759-888) and the fixed-order
 *                   scatter (:895-899)
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include "qol_oracle.h"

namespace orc { void set_side_tol(double tol); }

using namespace orc;

extern "C" {

void orc_set_side_tol(double tol) { set_side_tol(tol); }

/* Gauss-Newton projections that had not stopped by the step cap since the last reset,
   and the largest last step among them */
i64 orc_projection_unconverged(i64 reset, double* max_step) { return projection_unconverged(reset != 0, max_step); }

/* PR_ADAPTIVE / PR_SQUARE; returns 0, or 1 (nothing changed) for both set */
i64 orc_set_level_mode(i64 adaptive, i64 square) { return set_level_mode(adaptive != 0, square != 0) ? 0 : 1; }

void orc_gauss(i64 n, double* t, double* w, double* D) { gauss(n, t, w, D); }

void orc_koorn_pols(double u, double v, i64 nmax, double* pols, double* du, double* dv) {
  koorn_pols(u, v, nmax, pols, du, dv);
}

void orc_vr_nodes(i64 order, double* uv, double* w) { vr_nodes(order, uv, w); }

void orc_simplex_precomp(i64 nquad, i64 korder, i64 kpols, double* tgl, double* wgl, double* Dgl,
                         double* w_bclag, double* Legmat, double* umatr, double* vmatr) {
  simplex_precomp(nquad, korder, kpols, tgl, wgl, Dgl, w_bclag, Legmat, umatr, vmatr);
}

/* level table sizes: n(l), nsub(l) for l=1..4 */
void orc_patch_ref_info(i64 korder, i64 npols, i64 order, i64* nl, i64* nsub) {
  const PatchRef& R = patch_ref(korder, npols, order, 4);
  for (int l = 0; l < 4; ++l) { nl[l] = R.lev[l].n; nsub[l] = R.lev[l].nsub; }
}

/* copy one level table: uv(2,n), wt(n), interpmat(n,npols), suboff(nsub+1), subtyp(nsub) */
void orc_patch_ref_level(i64 korder, i64 npols, i64 order, i64 lev, double* uv, double* wt,
                         double* interpmat, i64* suboff, i64* subtyp) {
  const PatchRef& R = patch_ref(korder, npols, order, 4);
  const RefLevel& L = R.lev[lev - 1];
  std::memcpy(uv, L.uv.data(), sizeof(double) * 2 * L.n);
  std::memcpy(wt, L.wt.data(), sizeof(double) * L.n);
  std::memcpy(interpmat, L.interpmat.data(), sizeof(double) * L.n * npols);
  std::memcpy(suboff, L.suboff.data(), sizeof(i64) * (L.nsub + 1));
  std::memcpy(subtyp, L.subtyp.data(), sizeof(i64) * L.nsub);
}

/* per-panel level geometry: concatenated over levels, sx/snx (3,ntot), sw(ntot),
   qpoint (3,nsubtot), qradii2(nsubtot) */
void orc_patch_levels(i64 korder, i64 npols, const double* sx, const double* rts, const double* rps,
                      i64 order, double dist, double* lsx, double* lsnx, double* lsw, double* qpoint,
                      double* qradii2) {
  PanelLevels lv;
  build_patch_levels(korder, npols, sx, rts, rps, order, dist, 4, lv);
  i64 o = 0, os = 0;
  for (auto& L : lv.lev) {
    std::memcpy(lsx + 3 * o, L.sx.data(), sizeof(double) * 3 * L.n);
    std::memcpy(lsnx + 3 * o, L.snx.data(), sizeof(double) * 3 * L.n);
    std::memcpy(lsw + o, L.sw.data(), sizeof(double) * L.n);
    std::memcpy(qpoint + 3 * os, L.qpoint.data(), sizeof(double) * 3 * L.nsub);
    std::memcpy(qradii2 + os, L.qradii2.data(), sizeof(double) * L.nsub);
    o += L.n;
    os += L.nsub;
  }
}

void orc_lap3dsdlpmat(i64 m, const double* tx, i64 n, const double* sx, const double* snx,
                      const double* sw, double* As, double* Ad) {
  lap3dsdlpmat(m, tx, n, sx, snx, sw, As, Ad);
}

/* qol_lap3dsdlpmat_levels: classify + direct on levels; idxs 1-based, label per target */
void orc_levels_eval(i64 m, const double* tx, i64 n, const double* sx, const double* rts,
                     const double* rps, i64 korder, i64 orderff, double distff, double* As,
                     double* Ad, i64* idxs, i64* ms, i64* label) {
  PanelLevels lv;
  build_patch_levels(korder, n, sx, rts, rps, orderff, distff, 4, lv);
  const PatchRef& ref = patch_ref(korder, n, orderff, 4);
  for (i64 i = 0; i < m * n; ++i) { As[i] = 0.0; Ad[i] = 0.0; }
  lap3dsdlpmat_levels(m, tx, n, lv, ref, As, Ad, idxs, ms, label);
}

void orc_rrq(i64 m, const double* tx, i64 n, const double* sx, const double* snx, const double* sw,
             const double* rts, const double* rps, i64 order, i64 nquad, i64 orderff, double distff,
             i64 exterior, i64 ichart, const double* sxbd_chart, const double* rv_chart,
             const double* tgl, const double* wgl, const double* Dgl, const double* w_bclag,
             const double* Legmat, const double* umatr, const double* vmatr, double* As, double* Ad,
             double* Omega, double* Ias, double* timeinfo, i64* label) {
  rrq(m, tx, n, sx, snx, sw, rts, rps, order, nquad, orderff, distff, exterior != 0, ichart,
      sxbd_chart, rv_chart, tgl, wgl, Dgl, w_bclag, Legmat, umatr, vmatr, As, Ad, Omega, Ias,
      timeinfo, label);
}

/* CSR near list by the ball rule: for each panel k (hdim nodes each),
   offs(npan+1) 0-based, idx 1-based target indices ascending.  Two passes;
   panels are independent, so each pass runs over panels in nthreads threads
   (the result does not depend on the thread count).  With a diameter factor
   f > 0 (orc_set_near_radius) the radius is f h, h the largest distance between
   the corners (0,0), (1,0), (0,1) of the panel's degree-(p-1) interpolant. */
static double g_near_diam = 0.0;

/* rows interpolating a panel of hdim nodes at its three corners: Koornwinder
   values at the corner times umatr (the interpolant's coefficients from values) */
static const std::vector<double>& corner_rows(i64 hdim) {
  static std::vector<double> rows[16];
  i64 p = 1;
  while (p * (p + 1) / 2 < hdim) ++p;
  std::vector<double>& r = rows[p];
  if (r.empty()) {
    const i64 korder = p - 1, nq = p + 2;
    std::vector<double> t(nq), w(nq), D(nq * nq), wb(nq), L(nq * nq), um(hdim * hdim), vm(hdim * hdim), pol(hdim);
    simplex_precomp(nq, korder, hdim, t.data(), w.data(), D.data(), wb.data(), L.data(), um.data(), vm.data());
    const double cuv[3][2] = {{0.0, 0.0}, {1.0, 0.0}, {0.0, 1.0}};
    std::vector<double> out(3 * hdim);
    for (int c = 0; c < 3; ++c) {
      koorn_pols(cuv[c][0], cuv[c][1], korder, pol.data());
      for (i64 j = 0; j < hdim; ++j) {
        double a = 0.0;
        for (i64 q = 0; q < hdim; ++q) a += pol[q] * um[q + hdim * j];
        out[c * hdim + j] = a;
      }
    }
    r = out;
  }
  return r;
}

static void panel_ball(i64 k, i64 hdim, const double* sx, const double* sw, double qp[3], double* qr2) {
  double s = 0.0;
  qp[0] = qp[1] = qp[2] = 0.0;
  for (i64 i = 0; i < hdim; ++i)
    for (int c = 0; c < 3; ++c) qp[c] += sx[3 * (k * hdim + i) + c];
  for (i64 i = 0; i < hdim; ++i) s += sw[k * hdim + i];
  double qr = 1.75 * std::sqrt(s);
  if (g_near_diam > 0.0) {
    const std::vector<double>& cr = corner_rows(hdim);
    double v[3][3];
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < 3; ++d) {
        double a = 0.0;
        for (i64 i = 0; i < hdim; ++i) a += cr[c * hdim + i] * sx[3 * (k * hdim + i) + d];
        v[c][d] = a;
      }
    double hd = 0.0;
    for (int a = 0; a < 3; ++a)
      for (int b = a + 1; b < 3; ++b) {
        const double e0 = v[a][0] - v[b][0], e1 = v[a][1] - v[b][1], e2 = v[a][2] - v[b][2];
        hd = std::max(hd, std::sqrt(e0 * e0 + e1 * e1 + e2 * e2));
      }
    qr = g_near_diam * hd;
  }
  for (int c = 0; c < 3; ++c) qp[c] /= (double)hdim;
  *qr2 = qr * qr;
}

void orc_set_near_radius(double diam_factor) { g_near_diam = diam_factor; }

i64 orc_nearlist_count_mt(i64 npan, i64 hdim, const double* sx, const double* sw, i64 nt,
                          const double* tx, i64* offs, i64 nthreads) {
  std::vector<i64> cnt(npan, 0);
  if (g_near_diam > 0.0) (void)corner_rows(hdim);  // filled before the threads read it
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? (int)nthreads : 1)
  for (i64 k = 0; k < npan; ++k) {
    double qp[3], qr2;
    panel_ball(k, hdim, sx, sw, qp, &qr2);
    i64 c = 0;
    for (i64 j = 0; j < nt; ++j) {
      double d0 = tx[3 * j] - qp[0], d1 = tx[3 * j + 1] - qp[1], d2 = tx[3 * j + 2] - qp[2];
      if (d0 * d0 + d1 * d1 + d2 * d2 < qr2) ++c;
    }
    cnt[k] = c;
  }
  offs[0] = 0;
  for (i64 k = 0; k < npan; ++k) offs[k + 1] = offs[k] + cnt[k];
  return offs[npan];
}

void orc_nearlist_fill_mt(i64 npan, i64 hdim, const double* sx, const double* sw, i64 nt,
                          const double* tx, const i64* offs, i64* idx, i64 nthreads) {
  if (g_near_diam > 0.0) (void)corner_rows(hdim);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? (int)nthreads : 1)
  for (i64 k = 0; k < npan; ++k) {
    double qp[3], qr2;
    panel_ball(k, hdim, sx, sw, qp, &qr2);
    i64 o = offs[k];
    for (i64 j = 0; j < nt; ++j) {
      double d0 = tx[3 * j] - qp[0], d1 = tx[3 * j + 1] - qp[1], d2 = tx[3 * j + 2] - qp[2];
      if (d0 * d0 + d1 * d1 + d2 * d2 < qr2) idx[o++] = j + 1;
    }
  }
}

i64 orc_nearlist_count(i64 npan, i64 hdim, const double* sx, const double* sw, i64 nt,
                       const double* tx, i64* offs) {
  return orc_nearlist_count_mt(npan, hdim, sx, sw, nt, tx, offs, 1);
}

void orc_nearlist_fill(i64 npan, i64 hdim, const double* sx, const double* sw, i64 nt,
                       const double* tx, const i64* offs, i64* idx) {
  orc_nearlist_fill_mt(npan, hdim, sx, sw, nt, tx, offs, idx, 1);
}

/* Labels only (patch_refine_mod.f90:423-502 levels, :516-615 classification):
   for every pair of the CSR (offs 0-based, idx 1-based), the level 1..4 that
   serves it or 5 for RRQ; with PR_ADAPTIVE the walk's labels (1 level 1 whole,
   2 a cut of levels 2-3, with cut = 1 + opened level-2 cells when cut is not
   null), then 4 and 5.  Panels in nthreads threads. */
void orc_labels(i64 npan, i64 order, const double* sx, const double* rts, const double* rps,
                const double* tx, const i64* offs, const i64* idx, i64* label, i64* cut, i64 nthreads) {
  const bool adaptive = level_adaptive();
  const i64 h = order * (order + 1) / 2, korder = order - 1, orderff = order + 2;
  patch_ref(korder, h, orderff, 4);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? (int)nthreads : 1)
  for (i64 k = 0; k < npan; ++k) {
    const i64 m = offs[k + 1] - offs[k];
    if (m <= 0) continue;
    PanelLevels lv;
    build_patch_levels(korder, h, sx + 3 * k * h, rts + 3 * k * h, rps + 3 * k * h, orderff, 1.4, 4, lv);
    std::vector<i64> cl, cs;
    for (i64 j = 0; j < m; ++j) {
      const double* x = tx + 3 * (idx[offs[k] + j] - 1);
      i64 lab = 5, c = 0;
      bool ok = false;
      if (adaptive) adaptive_walk(lv, 3, x, cl, cs, ok);
      if (ok) {
        const bool whole = cl.size() == 1 && cl[0] == 0;
        lab = whole ? 1 : 2;
        i64 open = 0;
        for (size_t q = 0; q < cl.size(); ++q)
          if (cl[q] == 2) open |= (i64)1 << (cs[q] / 4);
        c = whole ? 0 : 1 + open;
      } else {
        for (i64 il = adaptive ? 3 : 0; il < 4; ++il)
          if (!inside_any_ball(lv.lev[il], x)) { lab = il + 1; break; }
      }
      label[offs[k] + j] = lab;
      if (cut) cut[offs[k] + j] = c;
    }
  }
}

/* qol_getnearquad_lap_rrq: COO triplets for the near correction (As-As_naive, Ad-Ad_naive);
   tcxrow(ntcx) holds the source index of each target when it is a source node (or 0). */
void orc_getnearquad(i64 npan, i64 nterms, i64 hdim, i64 nquad, i64 nsrc, const double* sx,
                     const double* snx, const double* sw, const double* rts, const double* rps,
                     i64 ntcx, const double* tcx, const i64* tcxrow, const i64* tcxi, i64 orderff,
                     double distff, i64 iside, i64 nnz, i64* sparsei, i64* sparsej, double* sparsevs,
                     double* sparsevd, double* timeinfo) {
  (void)nsrc; (void)ntcx; (void)nnz;
  std::vector<double> tgl(nquad), wgl(nquad), Dgl(nquad * nquad), wb(nquad), Leg(nquad * nquad),
      um(hdim * hdim), vm(hdim * hdim);
  simplex_precomp(nquad, nterms - 1, hdim, tgl.data(), wgl.data(), Dgl.data(), wb.data(), Leg.data(),
                  um.data(), vm.data());
  bool exterior = (iside % 10) == 0;
  std::vector<double> sxbd(3 * 3 * nquad, 0.0), rv(9, 0.0);
  for (i64 k = 0; k < npan; ++k) {
    i64 idxlo = k * hdim, jlo = tcxi[k] - 1, m = tcxi[k + 1] - tcxi[k];
    if (m <= 0) continue;
    std::vector<double> As(m * hdim), Ad(m * hdim), An(m * hdim), Dn(m * hdim), Om(4 * hdim * m), Ia(m);
    rrq(m, tcx + 3 * jlo, hdim, sx + 3 * idxlo, snx + 3 * idxlo, sw + idxlo, rts + 3 * idxlo,
        rps + 3 * idxlo, nterms, nquad, orderff, distff, exterior, 0, sxbd.data(), rv.data(),
        tgl.data(), wgl.data(), Dgl.data(), wb.data(), Leg.data(), um.data(), vm.data(), As.data(),
        Ad.data(), Om.data(), Ia.data(), timeinfo, nullptr);
    lap3dsdlpmat(m, tcx + 3 * jlo, hdim, sx + 3 * idxlo, snx + 3 * idxlo, sw + idxlo, An.data(), Dn.data());
    for (i64 i = 0; i < m; ++i) {
      i64 row = tcxrow[jlo + i];
      if (row >= idxlo + 1 && row <= idxlo + hdim) {
        i64 sc = row - idxlo - 1;
        An[i + m * sc] = 0.0;
        Dn[i + m * sc] = 0.0;
      }
    }
    for (i64 j = 0; j < hdim; ++j)
      for (i64 i = 0; i < m; ++i) {
        i64 out = (tcxi[k] - 1) * hdim + j * m + i;
        sparsei[out] = tcxrow[jlo + i];
        sparsej[out] = idxlo + j + 1;
        sparsevs[out] = As[i + m * j] - An[i + m * j];
        sparsevd[out] = Ad[i + m * j] - Dn[i + m * j];
      }
  }
}

/* Close correction for the GRF: for panel k and near targets idx(offs(k)+1..),
   val(pair) = sum_i (S-S_naive)_{pair,i} ubn_i - (K-K_naive)_{pair,i} ub_i, then
   u(t) += val in panel order.  label(pair) gets the level (1..4) or 5 = RRQ. */
void orc_close_apply(i64 npan, i64 nterms, i64 hdim, i64 nquad, const double* sx, const double* snx,
                     const double* sw, const double* rts, const double* rps, const double* tx,
                     const i64* offs, const i64* idx, const double* ub, const double* ubn, i64 exterior,
                     i64 ichart, const double* sxbd_chart_all, const double* rv_chart_all,
                     double* csr_val, double* u, i64* label, double* timeinfo, i64 nthreads) {
  std::vector<double> tgl(nquad), wgl(nquad), Dgl(nquad * nquad), wb(nquad), Leg(nquad * nquad),
      um(hdim * hdim), vm(hdim * hdim);
  simplex_precomp(nquad, nterms - 1, hdim, tgl.data(), wgl.data(), Dgl.data(), wb.data(), Leg.data(),
                  um.data(), vm.data());
  i64 orderff = nterms + 2;
  double distff = 1.4;
  /* warm the per-order cache before threading */
  patch_ref(nterms - 1, hdim, orderff, 4);
  std::vector<double> tinfo_acc(20, 0.0);
#pragma omp parallel num_threads(nthreads > 0 ? (int)nthreads : 1)
  {
    std::vector<double> tloc(20, 0.0);
#pragma omp for schedule(dynamic, 4)
    for (i64 k = 0; k < npan; ++k) {
      i64 m = offs[k + 1] - offs[k];
      if (m <= 0) continue;
      i64 o0 = offs[k], i0 = k * hdim;
      std::vector<double> t3(3 * m), As(m * hdim), Ad(m * hdim), An(m * hdim), Dn(m * hdim),
          Om(4 * hdim * m), Ia(m);
      std::vector<i64> lab(m);
      for (i64 j = 0; j < m; ++j)
        for (int c = 0; c < 3; ++c) t3[3 * j + c] = tx[3 * (idx[o0 + j] - 1) + c];
      std::vector<double> sxbd(3 * 3 * nquad, 0.0), rv(9, 0.0);
      if (ichart == 1) {
        for (i64 i = 0; i < 9 * nquad; ++i) sxbd[i] = sxbd_chart_all[9 * nquad * k + i];
        for (int i = 0; i < 9; ++i) rv[i] = rv_chart_all[9 * k + i];
      }
      rrq(m, t3.data(), hdim, sx + 3 * i0, snx + 3 * i0, sw + i0, rts + 3 * i0, rps + 3 * i0, nterms,
          nquad, orderff, distff, exterior != 0, ichart, sxbd.data(), rv.data(), tgl.data(), wgl.data(),
          Dgl.data(), wb.data(), Leg.data(), um.data(), vm.data(), As.data(), Ad.data(), Om.data(),
          Ia.data(), tloc.data(), lab.data());
      lap3dsdlpmat(m, t3.data(), hdim, sx + 3 * i0, snx + 3 * i0, sw + i0, An.data(), Dn.data());
      for (i64 j = 0; j < m; ++j) csr_val[o0 + j] = 0.0;
      for (i64 i = 0; i < hdim; ++i)
        for (i64 j = 0; j < m; ++j)
          csr_val[o0 + j] = csr_val[o0 + j] + (As[j + m * i] - An[j + m * i]) * ubn[i0 + i] -
                            (Ad[j + m * i] - Dn[j + m * i]) * ub[i0 + i];
      if (label)
        for (i64 j = 0; j < m; ++j) label[o0 + j] = lab[j];
    }
#pragma omp critical
    for (int i = 0; i < 20; ++i) tinfo_acc[i] += tloc[i];
  }
  if (timeinfo)
    for (int i = 0; i < 20; ++i) timeinfo[i] += tinfo_acc[i];
  if (u)
    for (i64 k = 0; k < npan; ++k)
      for (i64 j = offs[k]; j < offs[k + 1]; ++j) u[idx[j] - 1] = u[idx[j] - 1] + csr_val[j];
}

void orc_closepanel_dlp(i64 m, const double* tx, i64 p, const double* sx, i64 ref, double* Ac) {
  closepanel_dlp(m, tx, p, sx, ref, Ac);
}

/* far field by direct summation with the coincident term dropped:
   u(t) = sum_j As(t,j) ubn_j - Ad(t,j) ub_j */
void orc_far_direct(i64 nt, const double* tx, i64 ns, const double* sx, const double* snx,
                    const double* sw, const double* ub, const double* ubn, double* u, i64 nthreads) {
  const double pi4inv = 1.0 / (16.0 * std::atan(1.0));
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? (int)nthreads : 1)
  for (i64 i = 0; i < nt; ++i) {
    double acc = 0.0;
    for (i64 j = 0; j < ns; ++j) {
      double dx = tx[3 * i] - sx[3 * j], dy = tx[3 * i + 1] - sx[3 * j + 1], dz = tx[3 * i + 2] - sx[3 * j + 2];
      double rr = dx * dx + dy * dy + dz * dz;
      if (rr > 1.0e-28) {
        double r = std::sqrt(rr);
        double a = pi4inv * sw[j] / r;
        double d = pi4inv * sw[j] * (dx * snx[3 * j] + dy * snx[3 * j + 1] + dz * snx[3 * j + 2]) / (r * rr);
        acc = acc + a * ubn[j] - d * ub[j];
      }
    }
    u[i] = acc;
  }
}

}  // extern "C"

extern "C" {
/* Brute-force reference for one panel: the reference triangle split into nsub^2
   sub-triangles, each with the order-qord Vioreanu-Rokhlin rule; the patch is the
   degree-(p-1) interpolant of sx through the VR nodes and the density basis is
   its Lagrange basis (pols * umatr), exactly the object rrq_r64 integrates.
   Returns As, Ad (m, h) and the solid-angle moment int z.snx/|z|^3 dS per target. */
void orc_brute(i64 m, const double* tx, i64 order, const double* sx, const double* snx, i64 nsub, i64 qord,
               double* As, double* Ad, double* omega) {
  i64 h = order * (order + 1) / 2, korder = order - 1;
  std::vector<double> um(h * h), vm(h * h), uvn(2 * h), pp(h), pu(h), pv(h);
  vr_nodes(korder, uvn.data(), nullptr);
  for (i64 i = 0; i < h; ++i) {
    koorn_pols(uvn[2 * i], uvn[2 * i + 1], korder, pp.data());
    for (i64 j = 0; j < h; ++j) vm[i + h * j] = pp[j];
  }
  um = vm;
  lu_inverse(h, um.data());
  std::vector<double> cx(3 * h);
  for (int c = 0; c < 3; ++c)
    for (i64 k = 0; k < h; ++k) {
      double s = 0.0;
      for (i64 i = 0; i < h; ++i) s += um[k + h * i] * sx[3 * i + c];
      cx[k + h * c] = s;
    }
  double mn[3] = {0, 0, 0};
  for (i64 i = 0; i < h; ++i) for (int c = 0; c < 3; ++c) mn[c] += snx[3 * i + c];
  i64 nqn = vr_count(qord);
  std::vector<double> quv(2 * nqn), qw(nqn);
  vr_nodes(qord, quv.data(), qw.data());
  for (i64 i = 0; i < m * h; ++i) { As[i] = 0.0; Ad[i] = 0.0; }
  for (i64 t = 0; t < m; ++t) omega[t] = 0.0;
  const double pi4 = 1.0 / (16.0 * std::atan(1.0));
  double hs = 1.0 / (double)nsub;
  std::vector<double> L(h);
  for (i64 a = 0; a < nsub; ++a)
    for (i64 b = 0; b < nsub - a; ++b)
      for (int up = 0; up < 2; ++up) {
        if (up == 1 && a + b >= nsub - 1) continue;
        for (i64 q = 0; q < nqn; ++q) {
          double s = quv[2 * q], r = quv[2 * q + 1], u, v;
          if (up == 0) { u = (a + s) * hs; v = (b + r) * hs; }
          else { u = (a + 1 - s) * hs; v = (b + 1 - r) * hs; }
          double jw = qw[q] * hs * hs;
          koorn_pols(u, v, korder, pp.data(), pu.data(), pv.data());
          double y[3], yu[3], yv[3];
          for (int c = 0; c < 3; ++c) {
            double s0 = 0, s1 = 0, s2 = 0;
            for (i64 k = 0; k < h; ++k) { s0 += pp[k] * cx[k + h * c]; s1 += pu[k] * cx[k + h * c]; s2 += pv[k] * cx[k + h * c]; }
            y[c] = s0; yu[c] = s1; yv[c] = s2;
          }
          double nn[3] = {yu[1] * yv[2] - yu[2] * yv[1], yu[2] * yv[0] - yu[0] * yv[2], yu[0] * yv[1] - yu[1] * yv[0]};
          double J = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
          if (nn[0] * mn[0] + nn[1] * mn[1] + nn[2] * mn[2] < 0) for (int c = 0; c < 3; ++c) nn[c] = -nn[c];
          for (int c = 0; c < 3; ++c) nn[c] /= J;
          for (i64 j = 0; j < h; ++j) {
            double s0 = 0;
            for (i64 k = 0; k < h; ++k) s0 += pp[k] * um[k + h * j];
            L[j] = s0;
          }
          double w = jw * J;
          for (i64 t = 0; t < m; ++t) {
            double d[3] = {tx[3 * t] - y[0], tx[3 * t + 1] - y[1], tx[3 * t + 2] - y[2]};
            double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2], r1 = std::sqrt(rr);
            double kS = pi4 * w / r1, kD = pi4 * w * (d[0] * nn[0] + d[1] * nn[1] + d[2] * nn[2]) / (r1 * rr);
            omega[t] += -w * (d[0] * nn[0] + d[1] * nn[1] + d[2] * nn[2]) / (r1 * rr);
            for (i64 j = 0; j < h; ++j) { As[t + m * j] += kS * L[j]; Ad[t + m * j] += kD * L[j]; }
          }
        }
      }
}
}  // extern "C"
