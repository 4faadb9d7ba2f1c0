/* Oracle: close DLP matrix of a tensor-product patch.  TEST INFRASTRUCTURE ONLY
 * (see qol_oracle.h).
 *
 * Lap3dDLP_closepanel_r64 (src/lap3d_close_mod.f90:135-270) returns, for m
 * targets near one patch given by its nodes on a p x p Gauss-Legendre tensor
 * grid, the matrix Ac(m, p^2) with D[sigma](x_k) = sum_ip Ac(k, ip) sigma(y_ip)
 * for sigma interpolated on the grid.  The reference builds it from boundary
 * moments in its absent QuatApproximation / LineQuaaadrature submodules; this
 * restatement computes the same integrals,
 *   Ac(k, ip) = 1/(4 pi) int (x_k - y).(y_u x y_v) / |x_k - y|^3 L_ip(u, v) du dv,
 * by adaptive quadrature: a quadtree on [-1,1]^2 whose cells are split while
 * the target is within ETA cell diameters of the cell's centre image (up to
 * depth MAXDEPTH), each leaf integrated with a q x q Gauss rule, q = 2p + 2 ref.
 * Node index ip = b p + a with a the first (fast) parameter index, as the
 * reference's patch loops (test/test_lap3d_close.f90:56-66).
 */
#include "qol_oracle.h"
#include <cmath>
#include <vector>

namespace orc {

static const double CP_ETA = 2.5;
static const int CP_MAXDEPTH = 48;

/* Lagrange basis on nodes t (n) and its derivative at u, by products */
static void lagrange1d(i64 n, const double* t, double u, double* l, double* dl) {
  for (i64 j = 0; j < n; ++j) {
    double v = 1.0;
    for (i64 k = 0; k < n; ++k)
      if (k != j) v *= (u - t[k]) / (t[j] - t[k]);
    l[j] = v;
    double d = 0.0;
    for (i64 q = 0; q < n; ++q) {
      if (q == j) continue;
      double w = 1.0 / (t[j] - t[q]);
      for (i64 k = 0; k < n; ++k)
        if (k != j && k != q) w *= (u - t[k]) / (t[j] - t[k]);
      d += w;
    }
    dl[j] = d;
  }
}

struct CpCell {
  double u0, v0, s;  /* lower-left corner and side in parameter space */
  int depth;
};

void closepanel_dlp(i64 m, const double* tx, i64 p, const double* sx, i64 ref, double* Ac) {
  const i64 h = p * p;
  std::vector<double> t(p), w(p);
  gauss(p, t.data(), w.data(), nullptr);
  const i64 q = 2 * p + 2 * (ref > 0 ? ref : 1);
  std::vector<double> gq(q), gw(q);
  gauss(q, gq.data(), gw.data(), nullptr);
  std::vector<double> lu(p), dlu(p), lv(p), dlv(p);
  const double pi4 = 1.0 / (16.0 * std::atan(1.0));
  auto geom = [&](double u, double v, double* y, double* yu, double* yv) {
    lagrange1d(p, t.data(), u, lu.data(), dlu.data());
    lagrange1d(p, t.data(), v, lv.data(), dlv.data());
    for (int c = 0; c < 3; ++c) { y[c] = 0.0; yu[c] = 0.0; yv[c] = 0.0; }
    for (i64 b = 0; b < p; ++b)
      for (i64 a = 0; a < p; ++a) {
        const double* s = sx + 3 * (b * p + a);
        for (int c = 0; c < 3; ++c) {
          y[c] += lu[a] * lv[b] * s[c];
          yu[c] += dlu[a] * lv[b] * s[c];
          yv[c] += lu[a] * dlv[b] * s[c];
        }
      }
  };
  for (i64 k = 0; k < m; ++k) {
    const double* x = tx + 3 * k;
    std::vector<double> acc(h, 0.0);
    std::vector<CpCell> cur = {{-1.0, -1.0, 2.0, 0}}, next, leaves;
    while (!cur.empty()) {
      next.clear();
      for (const CpCell& c : cur) {
        double yc[3], a1[3], a2[3];
        geom(c.u0 + 0.5 * c.s, c.v0 + 0.5 * c.s, yc, a1, a2);
        /* physical diameter: the largest distance between corner images */
        double cy[4][3];
        const double cu[4] = {c.u0, c.u0 + c.s, c.u0, c.u0 + c.s}, cv[4] = {c.v0, c.v0, c.v0 + c.s, c.v0 + c.s};
        for (int i = 0; i < 4; ++i) geom(cu[i], cv[i], cy[i], a1, a2);
        double diam2 = 0.0;
        for (int i = 0; i < 4; ++i)
          for (int j = i + 1; j < 4; ++j) {
            const double d0 = cy[i][0] - cy[j][0], d1 = cy[i][1] - cy[j][1], d2 = cy[i][2] - cy[j][2];
            diam2 = std::fmax(diam2, d0 * d0 + d1 * d1 + d2 * d2);
          }
        const double e0 = x[0] - yc[0], e1 = x[1] - yc[1], e2 = x[2] - yc[2];
        const double dist2 = e0 * e0 + e1 * e1 + e2 * e2;
        if (c.depth < CP_MAXDEPTH && dist2 < CP_ETA * CP_ETA * diam2) {
          const double hs = 0.5 * c.s;
          next.push_back({c.u0, c.v0, hs, c.depth + 1});
          next.push_back({c.u0 + hs, c.v0, hs, c.depth + 1});
          next.push_back({c.u0, c.v0 + hs, hs, c.depth + 1});
          next.push_back({c.u0 + hs, c.v0 + hs, hs, c.depth + 1});
        } else {
          leaves.push_back(c);
        }
      }
      cur.swap(next);
    }
    for (const CpCell& c : leaves) {
      const double hs = 0.5 * c.s;
      for (i64 jv = 0; jv < q; ++jv)
        for (i64 ju = 0; ju < q; ++ju) {
          const double u = c.u0 + hs * (gq[ju] + 1.0), v = c.v0 + hs * (gq[jv] + 1.0);
          double y[3], yu[3], yv[3];
          geom(u, v, y, yu, yv);
          const double nn[3] = {yu[1] * yv[2] - yu[2] * yv[1], yu[2] * yv[0] - yu[0] * yv[2],
                                yu[0] * yv[1] - yu[1] * yv[0]};
          const double d0 = x[0] - y[0], d1 = x[1] - y[1], d2 = x[2] - y[2];
          const double rr = d0 * d0 + d1 * d1 + d2 * d2, r = std::sqrt(rr);
          const double kd = pi4 * gw[ju] * gw[jv] * hs * hs * (d0 * nn[0] + d1 * nn[1] + d2 * nn[2]) / (r * rr);
          for (i64 b = 0; b < p; ++b)
            for (i64 a = 0; a < p; ++a) acc[b * p + a] += kd * lu[a] * lv[b];
        }
    }
    for (i64 j = 0; j < h; ++j) Ac[k + m * j] = acc[j];
  }
}

}  // namespace orc
