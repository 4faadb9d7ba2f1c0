// Per-order tables for the RRQ path (host side, built once per patch order).
//
// Reference interfaces these tables stand in for:
//   simplex_precomp_r64           lap3d_close_mod.f90:115-133
//   build_patch_ref_r64           patch_refine_mod.f90:100-262
//   ref_square_regroup            patch_refine_mod.f90:313-421
//   cc_subidx_at (the 39 kept cells of the 8x8 split)   patch_refine_mod.f90:264-311
#include "qol_host.h"

#include <cmath>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>

#include "vr_tables.inc"

namespace qol {

void gl_nodes(i64 n, double* t, double* w, double* D) {
  const double pi = 3.14159265358979323846;
  for (i64 i = 0; i < n; ++i) {
    // ascending nodes: Newton on P_n from the Chebyshev-like guess
    double x = -std::cos(pi * ((double)i + 0.75) / ((double)n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = x;
      for (i64 k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / (double)k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) { p0 = 1.0; p1 = x; }
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
    if (n == 1) { p0 = 1.0; p1 = x; }
    dp = (double)n * (x * p1 - p0) / (x * x - 1.0);
    t[i] = x;
    w[i] = 2.0 / ((1.0 - x * x) * dp * dp);
  }
  for (i64 i = 0; i < n / 2; ++i) {
    double a = 0.5 * (t[n - 1 - i] - t[i]), ww = 0.5 * (w[i] + w[n - 1 - i]);
    t[i] = -a; t[n - 1 - i] = a; w[i] = ww; w[n - 1 - i] = ww;
  }
  if (n % 2) t[n / 2] = 0.0;
  if (D) {
    std::vector<double> a(n);
    for (i64 k = 0; k < n; ++k) {
      double pr = 1.0;
      for (i64 j = 0; j < n; ++j) if (j != k) pr *= t[k] - t[j];
      a[k] = pr;
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

// Orthogonal polynomials on the triangle (Koornwinder/Dubiner), ordered by total
// degree d, then b = 0..d with a = d-b: K = c_ab Q_a(u,v) P_b^{(2a+1,0)}(2v-1),
// Q_a = (1-v)^a P_a((2u+v-1)/(1-v)).  Only the spanned space matters to the path
// (interpolation matrices are V^{-1}-normalised), so the scaling c_ab is a choice.
void koornwinder(double u, double v, i64 nmax, double* f, double* fu, double* fv) {
  const i64 N = nmax + 1;
  double Q[32], Qu[32], Qv[32];
  const double a = 2.0 * u + v - 1.0, b = 1.0 - v, bb = b * b;
  Q[0] = 1.0; Qu[0] = 0.0; Qv[0] = 0.0;
  if (N > 1) { Q[1] = a; Qu[1] = 2.0; Qv[1] = 1.0; }
  for (i64 m = 1; m + 1 < N; ++m) {
    const double c1 = 2.0 * m + 1.0, c2 = (double)m, inv = 1.0 / (double)(m + 1);
    Q[m + 1] = (c1 * a * Q[m] - c2 * bb * Q[m - 1]) * inv;
    Qu[m + 1] = (c1 * (2.0 * Q[m] + a * Qu[m]) - c2 * bb * Qu[m - 1]) * inv;
    Qv[m + 1] = (c1 * (Q[m] + a * Qv[m]) - c2 * (bb * Qv[m - 1] - 2.0 * b * Q[m - 1])) * inv;
  }
  const double x = 2.0 * v - 1.0;
  i64 idx = 0;
  double J[32], Jv[32];
  // per degree d, per b: need P_b^{(2a+1,0)} with a = d-b
  for (i64 d = 0; d < N; ++d) {
    for (i64 bdeg = 0; bdeg <= d; ++bdeg) {
      const i64 am = d - bdeg;
      const double al = 2.0 * am + 1.0;
      J[0] = 1.0; Jv[0] = 0.0;
      if (bdeg >= 1) { J[1] = 0.5 * ((al + 2.0) * x + al); Jv[1] = al + 2.0; }
      for (i64 k = 1; k < bdeg; ++k) {
        const double dk = (double)k;
        const double a1 = 2.0 * (dk + 1.0) * (dk + al + 1.0) * (2.0 * dk + al);
        const double a2 = (2.0 * dk + al + 1.0) * al * al;
        const double a3 = (2.0 * dk + al) * (2.0 * dk + al + 1.0) * (2.0 * dk + al + 2.0);
        const double a4 = 2.0 * (dk + al) * dk * (2.0 * dk + al + 2.0);
        J[k + 1] = ((a2 + a3 * x) * J[k] - a4 * J[k - 1]) / a1;
        Jv[k + 1] = ((a2 + a3 * x) * Jv[k] + 2.0 * a3 * J[k] - a4 * Jv[k - 1]) / a1;
      }
      const double sc = std::sqrt((2.0 * am + 1.0) * 2.0 * (double)(d + 1));
      f[idx] = sc * Q[am] * J[bdeg];
      if (fu) fu[idx] = sc * Qu[am] * J[bdeg];
      if (fv) fv[idx] = sc * (Qv[am] * J[bdeg] + Q[am] * Jv[bdeg]);
      ++idx;
    }
  }
}

void vr_table(i64 order, double* uv, double* w) {
  if (order < 0 || order > VR_MAXORD) throw std::runtime_error("VR order out of range");
  const i64 o = VR_OFF[order], n = VR_OFF[order + 1] - o;
  for (i64 i = 0; i < n; ++i) {
    if (uv) { uv[2 * i] = VR_UV[2 * (o + i)]; uv[2 * i + 1] = VR_UV[2 * (o + i) + 1]; }
    if (w) w[i] = VR_W[o + i];
  }
}

void invert(i64 n, double* A) {
  std::vector<double> B(n * n, 0.0);
  for (i64 i = 0; i < n; ++i) B[i + n * i] = 1.0;
  for (i64 k = 0; k < n; ++k) {
    i64 ip = k;
    for (i64 i = k + 1; i < n; ++i)
      if (std::fabs(A[i + n * k]) > std::fabs(A[ip + n * k])) ip = i;
    if (ip != k)
      for (i64 j = 0; j < n; ++j) { std::swap(A[k + n * j], A[ip + n * j]); std::swap(B[k + n * j], B[ip + n * j]); }
    for (i64 i = k + 1; i < n; ++i) {
      const double f = A[i + n * k] / A[k + n * k];
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

namespace {

const i64 kKeptCells[39] = {1, 2, 3, 4, 7, 8, 9, 10, 11, 13, 14, 16, 17,
                            27, 32, 33, 34, 35, 37, 40, 41, 42, 43, 44, 46, 47,
                            48, 49, 50, 52, 53, 55, 58, 59, 60, 61, 62, 63, 64};

// four-way split of every sub-triangle: corner, centre (point reflected), top, right
void split_level(const LevelTable& prev, std::vector<double>& uv, std::vector<double>& wt) {
  const i64 np = prev.n;
  uv.assign(8 * np, 0.0);
  wt.assign(4 * np, 0.0);
  for (i64 i = 0; i < np; ++i) {
    const double u0 = 2.0 * prev.uv[2 * i] - 1.0, v0 = 2.0 * prev.uv[2 * i + 1] - 1.0;
    const double cu[4] = {u0, -u0, u0, u0 + 2.0}, cv[4] = {v0, -v0, v0 + 2.0, v0};
    for (int b = 0; b < 4; ++b) {
      uv[2 * (b * np + i)] = ((cu[b] - 1.0) * 0.5 + 1.0) * 0.5;
      uv[2 * (b * np + i) + 1] = ((cv[b] - 1.0) * 0.5 + 1.0) * 0.5;
      wt[b * np + i] = 0.25 * prev.wt[i];
    }
  }
}

// cells of a uniform grid that hold two sub-triangles become one qg x qg
// Gauss-Legendre square; cells with one keep their triangle nodes
void regroup_squares(LevelTable& L, i64 n1, i64 p, bool cc, i64 ilev, double wsum1) {
  const i64 npairable = cc ? L.nsub - 4 : L.nsub;
  const i64 ncell = cc ? 8 : ((i64)1 << (ilev - 1));
  const double side = 1.0 / (double)ncell;
  i64 qg = p;  // korder + 1
  if (qg < 8) qg = p + 2;
  const double fac = 2.0 * wsum1;
  const std::vector<double> uv0 = L.uv, wt0 = L.wt;
  std::vector<i64> cnt(ncell * ncell, 0), cellid(npairable);
  for (i64 q = 0; q < npairable; ++q) {
    double cu = 0.0, cv = 0.0;
    for (i64 i = 0; i < n1; ++i) cu += uv0[2 * (q * n1 + i)];
    for (i64 i = 0; i < n1; ++i) cv += uv0[2 * (q * n1 + i) + 1];
    cu /= (double)n1;
    cv /= (double)n1;
    i64 ic = (i64)(cu / side), jc = (i64)(cv / side);
    ic = ic < 0 ? 0 : (ic > ncell - 1 ? ncell - 1 : ic);
    jc = jc < 0 ? 0 : (jc > ncell - 1 ? ncell - 1 : jc);
    cellid[q] = ic + ncell * jc;
    if (++cnt[cellid[q]] > 2) throw std::runtime_error("square regroup: more than two sub-triangles in a cell");
  }
  std::vector<i64> singles;
  for (i64 q = 0; q < npairable; ++q)
    if (cnt[cellid[q]] == 1) singles.push_back(q);
  for (i64 q = npairable; q < L.nsub; ++q) singles.push_back(q);
  i64 npair = 0;
  for (i64 c = 0; c < ncell * ncell; ++c) npair += (cnt[c] == 2);
  const i64 n = npair * qg * qg + (i64)singles.size() * n1;
  L.n = n;
  L.nsub = npair + (i64)singles.size();
  L.uv.assign(2 * n, 0.0);
  L.wt.assign(n, 0.0);
  L.suboff.assign(L.nsub + 1, 0);
  L.subtyp.assign(L.nsub, 0);
  std::vector<double> tq(qg), wq(qg);
  gl_nodes(qg, tq.data(), wq.data(), nullptr);
  i64 row = 0, k = 0;
  for (i64 jc = 0; jc < ncell; ++jc)
    for (i64 ic = 0; ic < ncell; ++ic) {
      if (cnt[ic + ncell * jc] != 2) continue;
      for (i64 j = 0; j < qg; ++j)
        for (i64 i = 0; i < qg; ++i, ++row) {
          L.uv[2 * row] = ((double)ic + 0.5 * (tq[i] + 1.0)) * side;
          L.uv[2 * row + 1] = ((double)jc + 0.5 * (tq[j] + 1.0)) * side;
          L.wt[row] = (0.5 * wq[i]) * (0.5 * wq[j]) * side * side * fac;
        }
      L.subtyp[k] = 1;
      L.suboff[++k] = row;
    }
  for (i64 q : singles) {
    for (i64 j = 0; j < n1; ++j, ++row) {
      L.uv[2 * row] = uv0[2 * (q * n1 + j)];
      L.uv[2 * row + 1] = uv0[2 * (q * n1 + j) + 1];
      L.wt[row] = wt0[q * n1 + j];
    }
    L.suboff[++k] = row;
  }
}

std::unique_ptr<OrderTables> make_tables(i64 p, bool square) {
  auto T = std::make_unique<OrderTables>();
  T->p = p;
  T->square = square;
  T->h = p * (p + 1) / 2;
  T->nq = p + 2;
  T->korder = p - 1;
  T->orderff = p + 2;
  T->ncoef = (p + 1) * (p + 2) / 2;
  const i64 nq = T->nq, h = T->h;
  T->tgl.resize(nq); T->wgl.resize(nq); T->Dgl.resize(nq * nq); T->wbc.resize(nq); T->Legmat.resize(nq * nq);
  gl_nodes(nq, T->tgl.data(), T->wgl.data(), T->Dgl.data());
  for (i64 j = 0; j < nq; ++j) {
    double pr = 1.0;
    for (i64 k = 0; k < nq; ++k) if (k != j) pr *= T->tgl[j] - T->tgl[k];
    T->wbc[j] = 1.0 / pr;
  }
  for (i64 i = 0; i < nq; ++i) {
    double pkm1 = 0.0, pk = 1.0;
    for (i64 k = 0; k < nq; ++k) {
      T->Legmat[k + nq * i] = pk * T->wgl[i] * (2.0 * k + 1.0) * 0.5;
      const double pn = ((2.0 * k + 1.0) * T->tgl[i] * pk - (double)k * pkm1) / (double)(k + 1);
      pkm1 = pk;
      pk = pn;
    }
  }
  T->uvn.resize(2 * h); T->wn.resize(h);
  vr_table(p - 1, T->uvn.data(), T->wn.data());
  T->vmatr.resize(h * h);
  std::vector<double> f(h);
  for (i64 i = 0; i < h; ++i) {
    koornwinder(T->uvn[2 * i], T->uvn[2 * i + 1], p - 1, f.data(), nullptr, nullptr);
    for (i64 j = 0; j < h; ++j) T->vmatr[i + h * j] = f[j];
  }
  T->umatr = T->vmatr;
  invert(h, T->umatr.data());
  // refinement levels: VR(orderff-1), two 4-way splits, the graded "cc" level
  const i64 of = T->orderff, n1 = of * (of + 1) / 2;
  LevelTable* L = T->lev;
  L[0].n = n1; L[0].nsub = 1;
  L[0].uv.resize(2 * n1); L[0].wt.resize(n1);
  vr_table(of - 1, L[0].uv.data(), L[0].wt.data());
  for (int l = 1; l <= 2; ++l) {
    split_level(L[l - 1], L[l].uv, L[l].wt);
    L[l].n = 4 * L[l - 1].n;
    L[l].nsub = 4 * L[l - 1].nsub;
  }
  {
    std::vector<double> uvt, wtt;
    split_level(L[2], uvt, wtt);  // the 64-cell split; keep 39 cells
    L[3].nsub = 43;
    L[3].n = 43 * n1;
    L[3].uv.assign(2 * L[3].n, 0.0);
    L[3].wt.assign(L[3].n, 0.0);
    for (i64 q = 0; q < 39; ++q)
      for (i64 i = 0; i < n1; ++i) {
        const i64 s = (kKeptCells[q] - 1) * n1 + i, d = q * n1 + i;
        L[3].uv[2 * d] = uvt[2 * s]; L[3].uv[2 * d + 1] = uvt[2 * s + 1]; L[3].wt[d] = wtt[s];
      }
    // the interior is covered by level 2 shrunk by 0.625 about (0.5,0.5)+offset
    for (i64 q = 0; q < 4; ++q)
      for (i64 i = 0; i < n1; ++i) {
        const i64 s = q * n1 + i, d = (39 + q) * n1 + i;
        L[3].uv[2 * d] = (L[1].uv[2 * s] - 0.5) * 0.625 + 0.4375;
        L[3].uv[2 * d + 1] = (L[1].uv[2 * s + 1] - 0.5) * 0.625 + 0.4375;
        L[3].wt[d] = 0.625 * 0.625 * L[1].wt[s];
      }
  }
  double wsum1 = 0.0;
  for (double w : L[0].wt) wsum1 += w;
  L[0].suboff = {0, n1};
  L[0].subtyp = {0};
  for (int l = 1; l < 4; ++l) {
    if (square) {
      regroup_squares(L[l], n1, p, l == 3, l + 1, wsum1);
    } else {
      // PR_SQUARE off: every sub-patch keeps its n1 triangle nodes
      L[l].suboff.assign(L[l].nsub + 1, 0);
      for (i64 q = 0; q <= L[l].nsub; ++q) L[l].suboff[q] = q * n1;
      L[l].subtyp.assign(L[l].nsub, 0);
    }
  }
  T->ntot = 0;
  T->nsubtot = 0;
  std::vector<double> pol(h);
  for (int l = 0; l < 4; ++l) {
    LevelTable& Lv = L[l];
    Lv.interpmat.assign(Lv.n * h, 0.0);
    for (i64 i = 0; i < Lv.n; ++i) {
      koornwinder(Lv.uv[2 * i], Lv.uv[2 * i + 1], p - 1, pol.data(), nullptr, nullptr);
      for (i64 j = 0; j < h; ++j) {
        double s = 0.0;
        for (i64 k = 0; k < h; ++k) s += pol[k] * T->umatr[k + h * j];
        Lv.interpmat[i + Lv.n * j] = s;
      }
    }
    T->ntot += Lv.n;
    T->nsubtot += Lv.nsub;
  }
  return T;
}

}  // namespace

const OrderTables& order_tables(i64 p, bool square) {
  static std::mutex mu;
  static std::map<std::pair<i64, bool>, std::unique_ptr<OrderTables>> cache;
  if (p < 2 || p > 20) throw std::runtime_error("patch order must be in 2..20, got " + std::to_string(p));
  std::lock_guard<std::mutex> g(mu);
  const auto key = std::make_pair(p, square);
  auto it = cache.find(key);
  if (it == cache.end()) it = cache.emplace(key, make_tables(p, square)).first;
  return *it->second;
}

}  // namespace qol
