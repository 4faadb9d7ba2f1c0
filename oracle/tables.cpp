/* Oracle: per-order tables.  TEST INFRASTRUCTURE ONLY (see qol_oracle.h).
 *
 * gauss/bclag/legendre:  textbook Gauss-Legendre (Newton on P_n), second-form
 *   barycentric weights, Legendre values->coefficients; the reference calls them
 *   from the absent LQ/QA modules at lap3d_close_mod.f90:127-129.
 * koorn_pols:            an orthogonal (Koornwinder/Dubiner) basis on the simplex;
 *   umatr = V^{-1} at the VR nodes (koorn_vals2coefs_coefs2vals,
 *   lap3d_close_mod.f90:132), so interpmat = pols(uv)*umatr is the Lagrange
 *   interpolation matrix and does not depend on the basis normalisation.
 * patch_ref:             restates build_patch_ref_r64, cc_subidx_at and
 *   ref_square_regroup (patch_refine_mod.f90:100-421) line by line.
 */
#include "qol_oracle.h"
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

#include "../qotential_amd/csrc/vr_tables.inc"

namespace orc {

void gauss(i64 n, double* t, double* w, double* D) {
  const double pi = 4.0 * std::atan(1.0);
  for (i64 i = 0; i < n; ++i) {
    double x = -std::cos(pi * ((double)i + 0.75) / ((double)n + 0.5));
    double dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = x;
      for (i64 k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / (double)k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) { p1 = x; p0 = 1.0; }
      dp = (double)n * (x * p1 - p0) / (x * x - 1.0);
      double dx = p1 / dp;
      x -= dx;
      if (std::fabs(dx) < 1e-17) break;
    }
    double p0 = 1.0, p1 = x;
    for (i64 k = 2; k <= n; ++k) {
      double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / (double)k;
      p0 = p1;
      p1 = p2;
    }
    if (n == 1) { p1 = x; p0 = 1.0; }
    dp = (double)n * (x * p1 - p0) / (x * x - 1.0);
    t[i] = x;
    w[i] = 2.0 / ((1.0 - x * x) * dp * dp);
  }
  for (i64 i = 0; i < n / 2; ++i) { /* exact symmetry */
    double a = 0.5 * (t[n - 1 - i] - t[i]);
    t[i] = -a;
    t[n - 1 - i] = a;
    double ww = 0.5 * (w[i] + w[n - 1 - i]);
    w[i] = ww;
    w[n - 1 - i] = ww;
  }
  if (n % 2 == 1) t[n / 2] = 0.0;
  if (D) {
    std::vector<double> a(n);
    for (i64 k = 0; k < n; ++k) {
      double p = 1.0;
      for (i64 j = 0; j < n; ++j)
        if (j != k) p *= (t[k] - t[j]);
      a[k] = p;
    }
    for (i64 k = 0; k < n; ++k) {
      double s = 0.0;
      for (i64 i = 0; i < n; ++i) {
        if (i == k) continue;
        D[i + n * k] = (a[i] / a[k]) / (t[i] - t[k]);
        s += 1.0 / (t[k] - t[i]);
      }
      D[k + n * k] = s;
    }
  }
}

void bclag_weights(i64 n, const double* t, double* wb) {
  for (i64 j = 0; j < n; ++j) {
    double p = 1.0;
    for (i64 k = 0; k < n; ++k)
      if (k != j) p *= (t[j] - t[k]);
    wb[j] = 1.0 / p;
  }
}

/* Lmat(k,i) = P_k(t_i) w_i (2k+1)/2, so coefs = Lmat * vals (exact for the
   degree-(n-1) interpolant on Gauss-Legendre nodes). */
void legendre_vals2coefs(i64 n, const double* t, const double* w, double* Lmat) {
  for (i64 i = 0; i < n; ++i) {
    double p0 = 1.0, p1 = t[i];
    for (i64 k = 0; k < n; ++k) {
      double pk;
      if (k == 0) pk = 1.0;
      else if (k == 1) pk = t[i];
      else {
        double p2 = ((2.0 * k - 1.0) * t[i] * p1 - (k - 1.0) * p0) / (double)k;
        p0 = p1;
        p1 = p2;
        pk = p2;
      }
      Lmat[k + n * i] = pk * w[i] * (2.0 * k + 1.0) / 2.0;
    }
  }
}

/* Koornwinder basis K_{a,b}(u,v) = Q_a(u,v) J_{a,b}(v), total degree a+b <= nmax,
   Q_a = (1-v)^a P_a((2u+v-1)/(1-v)), J_{a,b} = P_b^{(2a+1,0)}(2v-1).  Ordering:
   degree d outer, b = 0..d inner (a = d-b).  Optional d/du, d/dv. */
void koorn_pols(double u, double v, i64 nmax, double* pols, double* du, double* dv) {
  const i64 N = nmax + 1;
  std::vector<double> Q(N), Qu(N), Qv(N);
  double a = 2.0 * u + v - 1.0, b = 1.0 - v;
  Q[0] = 1.0; Qu[0] = 0.0; Qv[0] = 0.0;
  if (N > 1) { Q[1] = a; Qu[1] = 2.0; Qv[1] = 1.0; }
  for (i64 m = 1; m + 1 < N; ++m) {
    const double c1 = 2.0 * m + 1.0, c2 = (double)m, inv = 1.0 / (double)(m + 1), bb = b * b;
    Q[m + 1] = (c1 * a * Q[m] - c2 * bb * Q[m - 1]) * inv;
    Qu[m + 1] = (c1 * (2.0 * Q[m] + a * Qu[m]) - c2 * bb * Qu[m - 1]) * inv;
    Qv[m + 1] = (c1 * (Q[m] + a * Qv[m]) - c2 * (bb * Qv[m - 1] - 2.0 * b * Q[m - 1])) * inv;
  }
  double x = 2.0 * v - 1.0;
  std::vector<double> J(N), Jv(N);
  i64 idx = 0;
  /* collect per a the Jacobi sequence, then place by degree */
  std::vector<double> JJ(N * N), JJv(N * N);
  for (i64 am = 0; am < N; ++am) {
    double al = 2.0 * am + 1.0;
    i64 nb = N - am;
    J[0] = 1.0; Jv[0] = 0.0;
    if (nb > 1) { J[1] = ((al + 2.0) * x + al) / 2.0; Jv[1] = (al + 2.0); }
    for (i64 n = 1; n + 1 < nb; ++n) {
      double dn = (double)n;
      double a1 = 2.0 * (dn + 1.0) * (dn + al + 1.0) * (2.0 * dn + al);
      double a2 = (2.0 * dn + al + 1.0) * al * al;
      double a3 = (2.0 * dn + al) * (2.0 * dn + al + 1.0) * (2.0 * dn + al + 2.0);
      double a4 = 2.0 * (dn + al) * dn * (2.0 * dn + al + 2.0);
      J[n + 1] = ((a2 + a3 * x) * J[n] - a4 * J[n - 1]) / a1;
      Jv[n + 1] = ((a2 + a3 * x) * Jv[n] + 2.0 * a3 * J[n] - a4 * Jv[n - 1]) / a1;
    }
    for (i64 n = 0; n < nb; ++n) { JJ[am * N + n] = J[n]; JJv[am * N + n] = Jv[n]; }
  }
  for (i64 d = 0; d < N; ++d) {
    for (i64 bb = 0; bb <= d; ++bb) {
      i64 am = d - bb;
      double sc = std::sqrt((2.0 * am + 1.0) * 2.0 * (double)(d + 1));
      double q = Q[am], j = JJ[am * N + bb];
      pols[idx] = sc * q * j;
      if (du) du[idx] = sc * Qu[am] * j;
      if (dv) dv[idx] = sc * (Qv[am] * j + q * JJv[am * N + bb]);
      ++idx;
    }
  }
}

i64 vr_count(i64 order) { return (order + 1) * (order + 2) / 2; }

void vr_nodes(i64 order, double* uv, double* w) {
  if (order < 0 || order > VR_MAXORD) {
    std::fprintf(stderr, "oracle: VR order %ld out of range\n", (long)order);
    std::abort();
  }
  i64 o = VR_OFF[order], n = VR_OFF[order + 1] - o;
  for (i64 i = 0; i < n; ++i) {
    if (uv) { uv[2 * i] = VR_UV[2 * (o + i)]; uv[2 * i + 1] = VR_UV[2 * (o + i) + 1]; }
    if (w) w[i] = VR_W[o + i];
  }
}

void lu_inverse(i64 n, double* A) {
  std::vector<double> B(n * n, 0.0);
  for (i64 i = 0; i < n; ++i) B[i + n * i] = 1.0;
  for (i64 k = 0; k < n; ++k) {
    i64 ip = k;
    double pm = std::fabs(A[k + n * k]);
    for (i64 i = k + 1; i < n; ++i)
      if (std::fabs(A[i + n * k]) > pm) { pm = std::fabs(A[i + n * k]); ip = i; }
    if (ip != k) {
      for (i64 j = 0; j < n; ++j) { std::swap(A[k + n * j], A[ip + n * j]); std::swap(B[k + n * j], B[ip + n * j]); }
    }
    for (i64 i = k + 1; i < n; ++i) {
      double f = A[i + n * k] / A[k + n * k];
      A[i + n * k] = f;
      for (i64 j = k + 1; j < n; ++j) A[i + n * j] -= f * A[k + n * j];
      for (i64 j = 0; j < n; ++j) B[i + n * j] -= f * B[k + n * j];
    }
  }
  for (i64 j = 0; j < n; ++j)
    for (i64 i = n - 1; i >= 0; --i) {
      double s = B[i + n * j];
      for (i64 k = i + 1; k < n; ++k) s -= A[i + n * k] * B[k + n * j];
      B[i + n * j] = s / A[i + n * i];
    }
  for (i64 i = 0; i < n * n; ++i) A[i] = B[i];
}

void simplex_precomp(i64 nquad, i64 korder, i64 kpols, double* tgl, double* wgl,
                     double* Dgl, double* w_bclag, double* Legmat, double* umatr,
                     double* vmatr) {
  gauss(nquad, tgl, wgl, Dgl);
  bclag_weights(nquad, tgl, w_bclag);
  legendre_vals2coefs(nquad, tgl, wgl, Legmat);
  std::vector<double> uv(2 * kpols), p(kpols);
  vr_nodes(korder, uv.data(), nullptr);
  for (i64 i = 0; i < kpols; ++i) {
    koorn_pols(uv[2 * i], uv[2 * i + 1], korder, p.data());
    for (i64 j = 0; j < kpols; ++j) vmatr[i + kpols * j] = p[j];
  }
  for (i64 i = 0; i < kpols * kpols; ++i) umatr[i] = vmatr[i];
  lu_inverse(kpols, umatr);
}

/* ---------------- build_patch_ref_r64 ---------------- */
static const i64 CC_NSUB = 39;
static const double CC_EXTRA_SCALE = 0.625, CC_EXTRA_OFFSET = 0.4375;
static const i64 CC_SUBIDX[39] = {1, 2, 3, 4, 7, 8, 9, 10, 11, 13, 14, 16, 17,
                                  27, 32, 33, 34, 35, 37, 40, 41, 42, 43, 44, 46, 47,
                                  48, 49, 50, 52, 53, 55, 58, 59, 60, 61, 62, 63, 64};

/* patch_refine_mod.f90:313-421 */
static void square_regroup(PatchRef& R, i64 ilev, i64 n1, i64 korder, bool is_cc) {
  RefLevel& L = R.lev[ilev - 1];
  i64 nsub_old = L.nsub, npairable = nsub_old, ncell;
  if (is_cc) { npairable = nsub_old - 4; ncell = 8; }
  else ncell = (i64)1 << (ilev - 1);
  double side = 1.0 / (double)ncell;
  i64 qg = korder + 1;
  if (qg < 8) qg = korder + 3;
  double fac = 0.0;
  for (double w : R.lev[0].wt) fac += w;
  fac = 2.0 * fac;
  std::vector<double> uv_old = L.uv, wt_old = L.wt;
  std::vector<i64> cellcnt(ncell * ncell, 0), singles;
  auto cellof = [&](i64 q, i64& ic, i64& jc) {
    i64 i0 = q * n1;
    double cu = 0.0, cv = 0.0;
    for (i64 i = 0; i < n1; ++i) cu += uv_old[2 * (i0 + i)];
    for (i64 i = 0; i < n1; ++i) cv += uv_old[2 * (i0 + i) + 1];
    cu /= (double)n1;
    cv /= (double)n1;
    ic = std::min(std::max((i64)(cu / side), (i64)0), ncell - 1);
    jc = std::min(std::max((i64)(cv / side), (i64)0), ncell - 1);
  };
  for (i64 q = 0; q < npairable; ++q) {
    i64 ic, jc;
    cellof(q, ic, jc);
    if (++cellcnt[ic + ncell * jc] > 2) {
      std::fprintf(stderr, "oracle: square regroup: >2 in a cell\n");
      std::abort();
    }
  }
  for (i64 q = 0; q < npairable; ++q) {
    i64 ic, jc;
    cellof(q, ic, jc);
    if (cellcnt[ic + ncell * jc] == 1) singles.push_back(q);
  }
  for (i64 q = npairable; q < nsub_old; ++q) singles.push_back(q);
  i64 npair = 0;
  for (i64 k = 0; k < ncell * ncell; ++k)
    if (cellcnt[k] == 2) ++npair;
  i64 nsingle = (i64)singles.size();
  i64 n = npair * qg * qg + nsingle * n1;
  L.n = n;
  L.nsub = npair + nsingle;
  L.uv.assign(2 * n, 0.0);
  L.wt.assign(n, 0.0);
  L.suboff.assign(L.nsub + 1, 0);
  L.subtyp.assign(L.nsub, 0);
  for (i64 k = 0; k < npair; ++k) L.subtyp[k] = 1;
  std::vector<double> tq(qg), wq(qg);
  gauss(qg, tq.data(), wq.data(), nullptr);
  std::vector<double> u01(qg), w01(qg);
  for (i64 i = 0; i < qg; ++i) { u01[i] = 0.5 * (tq[i] + 1.0); w01[i] = 0.5 * wq[i]; }
  i64 row = 0, k = 0;
  for (i64 jc = 0; jc < ncell; ++jc)
    for (i64 ic = 0; ic < ncell; ++ic) {
      if (cellcnt[ic + ncell * jc] != 2) continue;
      for (i64 j = 0; j < qg; ++j)
        for (i64 i = 0; i < qg; ++i) {
          L.uv[2 * row] = ((double)ic + u01[i]) * side;
          L.uv[2 * row + 1] = ((double)jc + u01[j]) * side;
          L.wt[row] = w01[i] * w01[j] * side * side * fac;
          ++row;
        }
      ++k;
      L.suboff[k] = row;
    }
  for (i64 s = 0; s < nsingle; ++s) {
    i64 i0 = singles[s] * n1;
    for (i64 j = 0; j < n1; ++j) {
      L.uv[2 * row] = uv_old[2 * (i0 + j)];
      L.uv[2 * row + 1] = uv_old[2 * (i0 + j) + 1];
      L.wt[row] = wt_old[i0 + j];
      ++row;
    }
    ++k;
    L.suboff[k] = row;
  }
}

static void split4(const RefLevel& P, std::vector<double>& uv, std::vector<double>& wt) {
  i64 np = P.n;
  uv.assign(2 * 4 * np, 0.0);
  wt.assign(4 * np, 0.0);
  for (i64 i = 0; i < np; ++i) {
    double u0 = 2.0 * P.uv[2 * i] - 1.0, v0 = 2.0 * P.uv[2 * i + 1] - 1.0;
    uv[2 * i] = u0;               uv[2 * i + 1] = v0;
    uv[2 * (np + i)] = -u0;       uv[2 * (np + i) + 1] = -v0;
    uv[2 * (2 * np + i)] = u0;    uv[2 * (2 * np + i) + 1] = v0 + 2.0;
    uv[2 * (3 * np + i)] = u0 + 2.0; uv[2 * (3 * np + i) + 1] = v0;
    for (int b = 0; b < 4; ++b) wt[b * np + i] = 0.25 * P.wt[i];
  }
  for (i64 i = 0; i < 4 * np; ++i) {
    uv[2 * i] = ((uv[2 * i] - 1.0) * 0.5 + 1.0) * 0.5;
    uv[2 * i + 1] = ((uv[2 * i + 1] - 1.0) * 0.5 + 1.0) * 0.5;
  }
}

static PatchRef build_ref(i64 korder, i64 npols, i64 order, i64 nlevel, bool square) {
  PatchRef R;
  R.korder = korder; R.npols = npols; R.order = order; R.nlevel = nlevel;
  R.square = square;
  R.lev.resize(nlevel);
  std::vector<double> umatr(npols * npols), vmatr(npols * npols), uvv(2 * npols), pp(npols);
  vr_nodes(korder, uvv.data(), nullptr);
  for (i64 i = 0; i < npols; ++i) {
    koorn_pols(uvv[2 * i], uvv[2 * i + 1], korder, pp.data());
    for (i64 j = 0; j < npols; ++j) vmatr[i + npols * j] = pp[j];
  }
  umatr = vmatr;
  lu_inverse(npols, umatr.data());
  i64 n1 = order * (order + 1) / 2, ncc = CC_NSUB + 4;
  for (i64 ilev = 1; ilev <= nlevel; ++ilev) {
    RefLevel& L = R.lev[ilev - 1];
    if (ilev == 1) {
      L.n = n1; L.nsub = 1;
      L.uv.resize(2 * n1); L.wt.resize(n1);
      vr_nodes(order - 1, L.uv.data(), L.wt.data());
    } else if (ilev == nlevel && nlevel == 4) {
      std::vector<double> uvt, wtt;
      split4(R.lev[ilev - 2], uvt, wtt);
      L.n = ncc * n1; L.nsub = ncc;
      L.uv.assign(2 * L.n, 0.0); L.wt.assign(L.n, 0.0);
      for (i64 q = 0; q < CC_NSUB; ++q)
        for (i64 i = 0; i < n1; ++i) {
          i64 isrc = (CC_SUBIDX[q] - 1) * n1 + i, idst = q * n1 + i;
          L.uv[2 * idst] = uvt[2 * isrc];
          L.uv[2 * idst + 1] = uvt[2 * isrc + 1];
          L.wt[idst] = wtt[isrc];
        }
      /* the 4 extras use level 2 as stored before square regrouping */
      for (i64 q = 0; q < 4; ++q)
        for (i64 i = 0; i < n1; ++i) {
          i64 isrc = q * n1 + i, idst = (CC_NSUB + q) * n1 + i;
          L.uv[2 * idst] = (R.lev[1].uv[2 * isrc] - 0.5) * CC_EXTRA_SCALE + CC_EXTRA_OFFSET;
          L.uv[2 * idst + 1] = (R.lev[1].uv[2 * isrc + 1] - 0.5) * CC_EXTRA_SCALE + CC_EXTRA_OFFSET;
          L.wt[idst] = CC_EXTRA_SCALE * CC_EXTRA_SCALE * R.lev[1].wt[isrc];
        }
    } else {
      const RefLevel& P = R.lev[ilev - 2];
      split4(P, L.uv, L.wt);
      L.n = 4 * P.n; L.nsub = 4 * P.nsub;
    }
  }
  for (i64 ilev = 1; ilev <= nlevel; ++ilev) {
    RefLevel& L = R.lev[ilev - 1];
    if (square && ilev >= 2) {
      square_regroup(R, ilev, n1, korder, ilev == nlevel && nlevel == 4);
    } else {
      /* PR_SQUARE off (patch_refine_mod.f90:234-239): n1 nodes per sub-patch */
      L.suboff.resize(L.nsub + 1);
      L.subtyp.assign(L.nsub, 0);
      for (i64 k = 0; k <= L.nsub; ++k) L.suboff[k] = k * n1;
    }
    /* interpmat = pols(uv) * umatr, sequential k sum */
    L.interpmat.assign(L.n * npols, 0.0);
    for (i64 i = 0; i < L.n; ++i) {
      koorn_pols(L.uv[2 * i], L.uv[2 * i + 1], korder, pp.data());
      for (i64 j = 0; j < npols; ++j) {
        double s = 0.0;
        for (i64 k = 0; k < npols; ++k) s += pp[k] * umatr[k + npols * j];
        L.interpmat[i + L.n * j] = s;
      }
    }
  }
  return R;
}

/* PR_ADAPTIVE / PR_SQUARE (patch_refine_mod.f90:51-55): process-wide, as the
   reference's module variables are */
static std::atomic<int> g_adaptive{0}, g_square{1};
bool set_level_mode(bool adaptive, bool square) {
  if (adaptive && square) return false; /* patch_refine_mod.f90:116-117 stops */
  g_adaptive.store(adaptive ? 1 : 0);
  g_square.store(square ? 1 : 0);
  return true;
}
bool level_adaptive() { return g_adaptive.load() != 0; }
bool level_square() { return g_square.load() != 0; }

const PatchRef& patch_ref(i64 korder, i64 npols, i64 order, i64 nlevel) {
  static std::mutex mu;
  static std::map<std::tuple<i64, i64, i64, i64, bool>, PatchRef> cache;
  const bool square = level_square();
  std::lock_guard<std::mutex> g(mu);
  auto key = std::make_tuple(korder, npols, order, nlevel, square);
  auto it = cache.find(key);
  if (it == cache.end()) it = cache.emplace(key, build_ref(korder, npols, order, nlevel, square)).first;
  return it->second;
}

}  // namespace orc
