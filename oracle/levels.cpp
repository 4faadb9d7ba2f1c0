/* Oracle: per-panel refinement levels, level classification, direct S/D and the
 * near list.  TEST INFRASTRUCTURE ONLY (see qol_oracle.h).
 *
 * build_patch_levels   restates patch_refine_mod.f90:423-502
 * lap3dsdlpmat_levels  restates patch_refine_mod.f90:516-615, and with PR_ADAPTIVE
 *                      adaptive_pass / cut_gI / walk / one_level (:617-842)
 * lap3dsdlpmat         restates the direct kernel (legacy scalar form
 *                      utils/f/ckernels_expts.c:52-124: zero when r^2 <= 1e-28)
 * nearlist             restates the ball rule of stellarator_grf_core_mod.f90:652-695
 *                      and test/stellarator_bench.f90:88-94
 * All sums run left to right, matching the dgemm-free reading of the reference.
 */
#include "qol_oracle.h"
#include <cmath>

namespace orc {

void build_patch_levels(i64 korder, i64 npols, const double* sx, const double* rts,
                        const double* rps, i64 order, double dist, i64 nlevel,
                        PanelLevels& lv) {
  const PatchRef& R = patch_ref(korder, npols, order, nlevel);
  double dist2 = dist * dist;
  lv.lev.assign(nlevel, PanelLevel());
  std::vector<double> srcT(npols * 9);
  for (i64 k = 0; k < npols; ++k)
    for (int c = 0; c < 3; ++c) {
      srcT[k + npols * c] = sx[3 * k + c];
      srcT[k + npols * (3 + c)] = rts[3 * k + c];
      srcT[k + npols * (6 + c)] = rps[3 * k + c];
    }
  for (i64 il = 0; il < nlevel; ++il) {
    const RefLevel& RL = R.lev[il];
    PanelLevel& L = lv.lev[il];
    i64 n = RL.n, nsub = RL.nsub;
    L.n = n; L.nsub = nsub;
    L.sx.assign(3 * n, 0.0); L.snx.assign(3 * n, 0.0); L.sw.assign(n, 0.0);
    L.qpoint.assign(3 * nsub, 0.0); L.qradii2.assign(nsub, 0.0);
    std::vector<double> geo(9);
    for (i64 i = 0; i < n; ++i) {
      for (int c = 0; c < 9; ++c) {
        double s = 0.0;
        for (i64 k = 0; k < npols; ++k) s = std::fma(RL.interpmat[i + n * k], srcT[k + npols * c], s);
        geo[c] = s;
      }
      L.sx[3 * i] = geo[0]; L.sx[3 * i + 1] = geo[1]; L.sx[3 * i + 2] = geo[2];
      double n0 = geo[7] * geo[5] - geo[8] * geo[4];
      double n1 = geo[8] * geo[3] - geo[6] * geo[5];
      double n2 = geo[6] * geo[4] - geo[7] * geo[3];
      double sp = std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
      L.snx[3 * i] = n0 / sp; L.snx[3 * i + 1] = n1 / sp; L.snx[3 * i + 2] = n2 / sp;
      L.sw[i] = sp * RL.wt[i];
    }
    for (i64 k = 0; k < nsub; ++k) {
      i64 i0 = RL.suboff[k], nk = RL.suboff[k + 1] - i0;
      double swsum = 0.0;
      for (i64 i = i0; i < i0 + nk; ++i) swsum += L.sw[i];
      double qp[3];
      for (int c = 0; c < 3; ++c) {
        double s = 0.0;
        for (i64 i = i0; i < i0 + nk; ++i) s += L.sx[3 * i + c] * L.sw[i];
        qp[c] = s / swsum;
      }
      double rmax2 = 0.0;
      for (i64 i = i0; i < i0 + nk; ++i) {
        double d0 = L.sx[3 * i] - qp[0], d1 = L.sx[3 * i + 1] - qp[1], d2 = L.sx[3 * i + 2] - qp[2];
        double d2s = d0 * d0 + d1 * d1 + d2 * d2;
        rmax2 = std::fmax(rmax2, d2s);
      }
      L.qpoint[3 * k] = qp[0]; L.qpoint[3 * k + 1] = qp[1]; L.qpoint[3 * k + 2] = qp[2];
      if (RL.subtyp[k] == 1) L.qradii2[k] = dist2 * std::fmax(rmax2, 0.5 * swsum);
      else L.qradii2[k] = dist2 * std::fmax(rmax2, swsum);
    }
  }
}

void lap3dsdlpmat(i64 m, const double* tx, i64 n, const double* sx, const double* snx,
                  const double* sw, double* As, double* Ad) {
  const double pi4inv = 1.0 / (16.0 * std::atan(1.0));
  for (i64 j = 0; j < n; ++j)
    for (i64 i = 0; i < m; ++i) {
      double dx = tx[3 * i] - sx[3 * j], dy = tx[3 * i + 1] - sx[3 * j + 1],
             dz = tx[3 * i + 2] - sx[3 * j + 2];
      double rr = dx * dx + dy * dy + dz * dz;
      if (rr > 1.0e-28) {
        double r = std::sqrt(rr);
        As[i + m * j] = pi4inv * sw[j] / r;
        Ad[i + m * j] = pi4inv * sw[j] * (dx * snx[3 * j] + dy * snx[3 * j + 1] + dz * snx[3 * j + 2]) / (r * rr);
      } else {
        As[i + m * j] = 0.0;
        Ad[i + m * j] = 0.0;
      }
    }
}

bool inside_any_ball(const PanelLevel& L, const double* x) {
  for (i64 k = 0; k < L.nsub; ++k) {
    double d0 = L.qpoint[3 * k] - x[0], d1 = L.qpoint[3 * k + 1] - x[1], d2 = L.qpoint[3 * k + 2] - x[2];
    if (d0 * d0 + d1 * d1 + d2 * d2 < L.qradii2[k]) return true;
  }
  return false;
}

/* walk (patch_refine_mod.f90:740-771): from the level-1 cell, a cell is a leaf
   when the target is outside its ball (strict >); otherwise its four children
   (cells 4q..4q+3 of the next level, 0-based) are visited, and a visited cell
   of level ntree holding the target ends the walk with ok = false */
void adaptive_walk(const PanelLevels& lv, i64 ntree, const double* x, std::vector<i64>& cl,
                   std::vector<i64>& cs, bool& ok) {
  cl.clear(); cs.clear(); ok = true;
  std::vector<i64> act = {0}, nxt;
  for (i64 il = 0; il < ntree; ++il) {
    nxt.clear();
    const PanelLevel& L = lv.lev[il];
    for (i64 q : act) {
      double d0 = x[0] - L.qpoint[3 * q], d1 = x[1] - L.qpoint[3 * q + 1], d2 = x[2] - L.qpoint[3 * q + 2];
      if (d0 * d0 + d1 * d1 + d2 * d2 > L.qradii2[q]) {
        cl.push_back(il); cs.push_back(q);
      } else if (il == ntree - 1) {
        ok = false;
        return;
      } else {
        for (i64 b = 0; b < 4; ++b) nxt.push_back(4 * q + b);
      }
    }
    if (nxt.empty()) break;
    act = nxt;
  }
}

/* adaptive_pass + cut_gI (:617-738): every target the walk serves gets direct
   S/D on its leaf cells' nodes contracted with those nodes' interpolation rows
   (the reference's per-mask gI), leaves in walk order; the rest go on to
   one_level at the last level (:773-842), then RRQ */
static void adaptive_levels(i64 m, const double* tx, i64 npols, const PanelLevels& lv, const PatchRef& ref,
                            double* As, double* Ad, std::vector<i64>& irem, i64& nrem, i64* label, i64* cut) {
  const i64 nlevel = (i64)lv.lev.size(), ntree = nlevel - 1, n1 = ref.lev[0].n;
  std::vector<i64> inext, cl, cs;
  std::vector<double> a(n1), b(n1);
  for (i64 jj = 0; jj < nrem; ++jj) {
    const i64 j = irem[jj];
    bool ok = false;
    adaptive_walk(lv, ntree, tx + 3 * j, cl, cs, ok);
    if (!ok) { inext.push_back(j); continue; }
    std::vector<double> s(npols, 0.0), d(npols, 0.0);
    i64 open = 0;
    for (size_t c = 0; c < cl.size(); ++c) {
      const PanelLevel& L = lv.lev[cl[c]];
      const RefLevel& R = ref.lev[cl[c]];
      const i64 i0 = cs[c] * n1;
      lap3dsdlpmat(1, tx + 3 * j, n1, L.sx.data() + 3 * i0, L.snx.data() + 3 * i0, L.sw.data() + i0, a.data(),
                   b.data());
      for (i64 k = 0; k < n1; ++k)
        for (i64 col = 0; col < npols; ++col) {
          s[col] += a[k] * R.interpmat[i0 + k + R.n * col];
          d[col] += b[k] * R.interpmat[i0 + k + R.n * col];
        }
      if (cl[c] == 2) open |= (i64)1 << (cs[c] / 4);
    }
    for (i64 col = 0; col < npols; ++col) { As[j + m * col] = s[col]; Ad[j + m * col] = d[col]; }
    const bool whole = cl.size() == 1 && cl[0] == 0;
    if (label) label[j] = whole ? 1 : 2;
    if (cut) cut[j] = whole ? 0 : 1 + open;
  }
  nrem = (i64)inext.size();
  irem.assign(inext.begin(), inext.end());
}

void lap3dsdlpmat_levels(i64 m, const double* tx, i64 npols, const PanelLevels& lv,
                         const PatchRef& ref, double* As, double* Ad, i64* idxs, i64* ms,
                         i64* label, i64* cut) {
  std::vector<i64> irem(m), ibat, inext;
  for (i64 j = 0; j < m; ++j) irem[j] = j;
  i64 nrem = m;
  i64 nlevel = (i64)lv.lev.size();
  i64 il0 = 0;
  if (level_adaptive() && nlevel >= 3) {
    adaptive_levels(m, tx, npols, lv, ref, As, Ad, irem, nrem, label, cut);
    il0 = nlevel - 1;  /* one_level at the last level */
  }
  for (i64 il = il0; il < nlevel && nrem > 0; ++il) {
    const PanelLevel& L = lv.lev[il];
    ibat.clear(); inext.clear();
    for (i64 jj = 0; jj < nrem; ++jj) {
      i64 j = irem[jj];
      if (inside_any_ball(L, tx + 3 * j)) inext.push_back(j);
      else ibat.push_back(j);
    }
    i64 nbat = (i64)ibat.size(), nl = L.n;
    if (nbat > 0) {
      std::vector<double> txb(3 * nbat), Ab(nbat * nl), Bb(nbat * nl);
      for (i64 i = 0; i < nbat; ++i)
        for (int c = 0; c < 3; ++c) txb[3 * i + c] = tx[3 * ibat[i] + c];
      lap3dsdlpmat(nbat, txb.data(), nl, L.sx.data(), L.snx.data(), L.sw.data(), Ab.data(), Bb.data());
      const std::vector<double>& IM = ref.lev[il].interpmat;
      for (i64 i = 0; i < nbat; ++i) {
        for (i64 c = 0; c < npols; ++c) {
          double s = 0.0, d = 0.0;
          for (i64 k = 0; k < nl; ++k) {
            s += Ab[i + nbat * k] * IM[k + nl * c];
            d += Bb[i + nbat * k] * IM[k + nl * c];
          }
          As[ibat[i] + m * c] = s;
          Ad[ibat[i] + m * c] = d;
        }
        if (label) label[ibat[i]] = il + 1;
      }
    }
    nrem = (i64)inext.size();
    irem.assign(inext.begin(), inext.end());
  }
  *ms = nrem;
  for (i64 j = 0; j < nrem; ++j) {
    idxs[j] = irem[j] + 1;
    if (label) label[irem[j]] = nlevel + 1;
  }
}

}  // namespace orc
