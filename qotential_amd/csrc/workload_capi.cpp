// C-ABI for the workload generator (host only; no GPU calls).
//
// These entry points build the inputs the reference's drivers build before they
// call the close-evaluation path: the stellarator geometry of
// utils/stellarator_geo_mex.c (stellarator_geo / stellarator_geo_charts /
// stellarator_geo_uv2x, :758-877), the chart-boundary table of
// stellarator_grf_core_mod.f90:613-629, and the synthetic near-surface targets
// and point-charge densities of the benchmark workloads.
#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>

#include "qol_host.h"
#include "qotential_hip.h"

using namespace qol;

namespace {

Surface make_surface(int64_t nfp, int64_t nmode, const double* mnrz) {
  Surface s;
  s.nfp = nfp;
  for (int64_t k = 0; k < nmode; ++k) {
    s.m.push_back(mnrz[4 * k]);
    s.n.push_back(mnrz[4 * k + 1]);
    s.rc.push_back(mnrz[4 * k + 2]);
    s.zs.push_back(mnrz[4 * k + 3]);
  }
  return s;
}

std::vector<Chart> unpack_charts(int64_t nchart, const double* cl) {
  if (nchart < 0 || (nchart > 0 && !cl)) throw std::runtime_error("chart list: bad count or null pointer");
  std::vector<Chart> L(nchart);
  for (int64_t k = 0; k < nchart; ++k)
    L[k] = Chart{cl[6 * k], cl[6 * k + 1], cl[6 * k + 2], cl[6 * k + 3], cl[6 * k + 4], cl[6 * k + 5]};
  return L;
}

void copy_geometry(const Geometry& g, double* sx, double* snx, double* sw, double* rts, double* rps) {
  std::memcpy(sx, g.sx.data(), sizeof(double) * g.sx.size());
  std::memcpy(snx, g.snx.data(), sizeof(double) * g.snx.size());
  std::memcpy(sw, g.sw.data(), sizeof(double) * g.sw.size());
  std::memcpy(rts, g.rts.data(), sizeof(double) * g.rts.size());
  std::memcpy(rps, g.rps.data(), sizeof(double) * g.rps.size());
}

void chart_boundary(const Surface& s, int64_t mp, int64_t np, int64_t p, std::vector<Chart> charts, int64_t nquad,
                    double* sxbd, double* rv) {
  Geometry g;
  g.p = p; g.mp = mp; g.np = np;
  g.charts = std::move(charts);
  g.ntri = 2 * (int64_t)g.charts.size();
  std::vector<double> t(nquad), w(nquad);
  gl_nodes(nquad, t.data(), w.data(), nullptr);
  std::vector<double> uv(2 * (3 * nquad + 3));
  const double pi = 3.14159265358979323846, th1 = 2.0 * pi / 3.0, th2 = 4.0 * pi / 3.0, ti1 = 1.0 / th1;
  for (int64_t k = 0; k < 3; ++k)
    for (int64_t i = 0; i < nquad; ++i) {
      const double t0 = (double)k * 2.0 * pi / 3.0, t1 = (double)(k + 1) * 2.0 * pi / 3.0;
      const double tt = t0 + 0.5 * (1.0 + t[i]) * (t1 - t0);
      double u = 0.0, v = 0.0;
      if (tt < th1) v = 1.0 - ti1 * tt;
      else if (tt < th2) u = ti1 * (tt - th1);
      else { u = 1.0 - ti1 * (tt - th2); v = ti1 * (tt - th2); }
      uv[2 * (k * nquad + i)] = u;
      uv[2 * (k * nquad + i) + 1] = v;
    }
  const double vert[6] = {0, 0, 1, 0, 0, 1};
  for (int i = 0; i < 6; ++i) uv[2 * 3 * nquad + i] = vert[i];
  std::vector<double> x(3 * (3 * nquad + 3));
  for (int64_t it = 0; it < g.ntri; ++it) {
    chart_uv2x(s, g, it, 3 * nquad + 3, uv.data(), x.data());
    std::memcpy(sxbd + 9 * nquad * it, x.data(), sizeof(double) * 9 * nquad);
    std::memcpy(rv + 9 * it, x.data() + 9 * nquad, sizeof(double) * 9);
  }
}

}  // namespace

namespace qol {
void set_error(const std::string& s);
const char* error_text();
}

extern "C" {

const char* qol_last_error(void) { return qol::error_text(); }

int64_t qol_geo_count(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode, const double* mnrz,
                      int64_t iterate_aspect) {
  try {
    Surface s = make_surface(nfp, nmode, mnrz);
    return 2 * (int64_t)build_charts(s, mp, np, p, iterate_aspect != 0).size();
  } catch (const std::exception& e) {
    qol::set_error(e.what());
    return -1;
  }
}

int qol_geo_build(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode, const double* mnrz,
                  int64_t iterate_aspect, double* sx, double* snx, double* sw, double* rts, double* rps,
                  double* charts) {
  try {
    Surface s = make_surface(nfp, nmode, mnrz);
    Geometry g;
    build_geometry(s, mp, np, p, iterate_aspect != 0, g);
    copy_geometry(g, sx, snx, sw, rts, rps);
    if (charts)
      for (size_t k = 0; k < g.charts.size(); ++k) {
        const Chart& c = g.charts[k];
        const double v[6] = {c.a1, c.b1, c.a2, c.b2, c.jj, c.ii};
        std::memcpy(charts + 6 * k, v, sizeof(v));
      }
    return 0;
  } catch (const std::exception& e) {
    qol::set_error(e.what());
    return 1;
  }
}

// exact chart boundary for every triangle: sxbd(3, 3*nquad, ntri) at the
// Gauss-Legendre nodes of the three edges (0,1)->(0,0)->(1,0)->(0,1), and the
// vertices rv(3,3,ntri) at (0,0),(1,0),(0,1)
int qol_geo_chart_boundary(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode, const double* mnrz,
                           int64_t iterate_aspect, int64_t nquad, double* sxbd, double* rv) {
  try {
    Surface s = make_surface(nfp, nmode, mnrz);
    chart_boundary(s, mp, np, p, build_charts(s, mp, np, p, iterate_aspect != 0), nquad, sxbd, rv);
    return 0;
  } catch (const std::exception& e) {
    qol::set_error(e.what());
    return 1;
  }
}

int64_t qol_geo_charts(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode, const double* mnrz,
                       double restol, int64_t cap, double* charts) {
  try {
    Surface s = make_surface(nfp, nmode, mnrz);
    bool over = false;
    const std::vector<Chart> L = build_charts_refined(s, mp, np, p, restol, cap, &over);
    if (over) return -1;
    if (charts)
      for (size_t k = 0; k < L.size(); ++k) {
        const Chart& c = L[k];
        const double v[6] = {c.a1, c.b1, c.a2, c.b2, c.jj, c.ii};
        std::memcpy(charts + 6 * k, v, sizeof(v));
      }
    return (int64_t)L.size();
  } catch (const std::exception& e) {
    qol::set_error(e.what());
    return -2;
  }
}

int64_t qol_geo_plan_charts(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode, const double* mnrz,
                            int64_t iterate_aspect, double restol, int64_t cap, double* charts) {
  if (restol > 0.0) return qol_geo_charts(mp, np, p, nfp, nmode, mnrz, restol, cap, charts);
  try {
    Surface s = make_surface(nfp, nmode, mnrz);
    GeometryPlan P;
    plan_geometry(s, mp, np, p, iterate_aspect != 0, P);
    if (cap > 0 && (int64_t)P.charts.size() > cap) return -1;
    if (charts)
      for (size_t k = 0; k < P.charts.size(); ++k) {
        const Chart& c = P.charts[k];
        const double v[6] = {c.a1, c.b1, c.a2, c.b2, c.jj, c.ii};
        std::memcpy(charts + 6 * k, v, sizeof(v));
      }
    return (int64_t)P.charts.size();
  } catch (const std::exception& e) {
    qol::set_error(e.what());
    return -2;
  }
}

double qol_geo_curvature_tail(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode, const double* mnrz,
                              const double* chart) {
  try {
    Surface s = make_surface(nfp, nmode, mnrz);
    const Chart c{chart[0], chart[1], chart[2], chart[3], chart[4], chart[5]};
    return chart_curvature_tail(s, c, mp, np, p);
  } catch (const std::exception& e) {
    qol::set_error(e.what());
    return -1.0;
  }
}

int qol_geo_build_charts(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode, const double* mnrz,
                         int64_t nchart, const double* charts, double* sx, double* snx, double* sw, double* rts,
                         double* rps) {
  try {
    Surface s = make_surface(nfp, nmode, mnrz);
    Geometry g;
    build_geometry(s, mp, np, p, unpack_charts(nchart, charts), g);
    copy_geometry(g, sx, snx, sw, rts, rps);
    return 0;
  } catch (const std::exception& e) {
    qol::set_error(e.what());
    return 1;
  }
}

int qol_geo_chart_boundary_charts(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode,
                                  const double* mnrz, int64_t nchart, const double* charts, int64_t nquad,
                                  double* sxbd, double* rv) {
  try {
    Surface s = make_surface(nfp, nmode, mnrz);
    chart_boundary(s, mp, np, p, unpack_charts(nchart, charts), nquad, sxbd, rv);
    return 0;
  } catch (const std::exception& e) {
    qol::set_error(e.what());
    return 1;
  }
}

// Near-surface targets: triangle k uniform, (u,v) uniform on the reference
// triangle, point and unit normal of the triangle's degree-(p-1) interpolant,
// offset side * h * 10^U along the normal, U ~ U[umin, umax], h = the largest
// vertex distance of the triangle.  The normal is oriented like the panel
// normals snx; side = -1, +1, or random per target (sides == 0).
// Writes tx(3,nt), tri(nt) (0-based), dist(nt) (signed, along snx).
int qol_make_targets(int64_t p, int64_t ntri, const double* sx, const double* snx, int64_t nt, double umin,
                     double umax, int64_t sides, uint64_t seed, double* tx, int64_t* tri, double* dist) {
  try {
    const OrderTables& T = order_tables(p);
    const int64_t h = T.h;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U01(0.0, 1.0);
    std::vector<double> f(h), fu(h), fv(h), cx(3 * h);
    for (int64_t j = 0; j < nt; ++j) {
      const int64_t k = std::min<int64_t>(ntri - 1, (int64_t)(U01(rng) * (double)ntri));
      double u = U01(rng), v = U01(rng);
      if (u + v > 1.0) { u = 1.0 - u; v = 1.0 - v; }
      const double ex = umin + (umax - umin) * U01(rng);
      const double side = sides < 0 ? -1.0 : (sides > 0 ? 1.0 : (U01(rng) < 0.5 ? -1.0 : 1.0));
      double mn[3] = {0.0, 0.0, 0.0};
      for (int c = 0; c < 3; ++c) {
        for (int64_t i = 0; i < h; ++i) mn[c] += snx[3 * (k * h + i) + c];
        for (int64_t a = 0; a < h; ++a) {
          double s = 0.0;
          for (int64_t i = 0; i < h; ++i) s += T.umatr[a + h * i] * sx[3 * (k * h + i) + c];
          cx[a + h * c] = s;
        }
      }
      auto eval = [&](double uu, double vv, double* x, double* xu, double* xv) {
        koornwinder(uu, vv, p - 1, f.data(), fu.data(), fv.data());
        for (int c = 0; c < 3; ++c) {
          double s = 0.0, su = 0.0, sv = 0.0;
          for (int64_t a = 0; a < h; ++a) {
            s += f[a] * cx[a + h * c];
            su += fu[a] * cx[a + h * c];
            sv += fv[a] * cx[a + h * c];
          }
          x[c] = s;
          if (xu) xu[c] = su;
          if (xv) xv[c] = sv;
        }
      };
      double x[3], xu[3], xv[3], vtx[9];
      eval(u, v, x, xu, xv);
      eval(0.0, 0.0, vtx, nullptr, nullptr);
      eval(1.0, 0.0, vtx + 3, nullptr, nullptr);
      eval(0.0, 1.0, vtx + 6, nullptr, nullptr);
      double hd = 0.0;
      for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b) {
          double d = 0.0;
          for (int c = 0; c < 3; ++c) d += (vtx[3 * a + c] - vtx[3 * b + c]) * (vtx[3 * a + c] - vtx[3 * b + c]);
          hd = std::max(hd, std::sqrt(d));
        }
      double n[3] = {xu[1] * xv[2] - xu[2] * xv[1], xu[2] * xv[0] - xu[0] * xv[2], xu[0] * xv[1] - xu[1] * xv[0]};
      double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      if (n[0] * mn[0] + n[1] * mn[1] + n[2] * mn[2] < 0.0) nn = -nn;
      const double d = side * hd * std::pow(10.0, ex);
      for (int c = 0; c < 3; ++c) tx[3 * j + c] = x[c] + d * n[c] / nn;
      tri[j] = k;
      dist[j] = d;
    }
    return 0;
  } catch (const std::exception& e) {
    qol::set_error(e.what());
    return 1;
  }
}

}  // extern "C"
