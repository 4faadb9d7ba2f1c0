// Host-side support for the MI355X RRQ close-evaluation library.
//
// Per-order tables (Gauss-Legendre, Vioreanu-Rokhlin nodes, the Koornwinder
// values<->coefficients maps and the four refinement levels of
// patch_refine_mod.f90:100-421) are built once per patch order on the host and
// uploaded; everything per panel and per target runs on the GPU.
#pragma once
#include <cstdint>
#include <memory>
#include <vector>

namespace qol {

typedef int64_t i64;

struct LevelTable {
  i64 n = 0, nsub = 0;
  std::vector<double> uv, wt, interpmat;  // uv(2,n), wt(n), interpmat(n,h) column-major
  std::vector<i64> suboff, subtyp;        // suboff(nsub+1) 0-based
};

struct OrderTables {
  i64 p = 0, h = 0, nq = 0, korder = 0, orderff = 0, ncoef = 0;
  std::vector<double> tgl, wgl, Dgl, wbc, Legmat;  // nq, nq, nq*nq, nq, nq*nq
  std::vector<double> uvn, wn;                    // VR nodes of order p-1
  std::vector<double> umatr, vmatr;               // h*h
  LevelTable lev[4];
  i64 ntot = 0, nsubtot = 0;                       // sums over levels
  bool square = true;                              // levels 2-4 regrouped into squares
};

// Thread-safe, cached per order and level form.  Orders 2..20.  square = false
// is the reference's PR_SQUARE = .false. (patch_refine_mod.f90:229-240): every
// sub-patch of levels 2-4 keeps its triangle nodes.
const OrderTables& order_tables(i64 p, bool square = true);

// primitives (also used by the workload generator)
void gl_nodes(i64 n, double* t, double* w, double* D);
void koornwinder(double u, double v, i64 nmax, double* f, double* fu, double* fv);
void vr_table(i64 order, double* uv, double* w);
void invert(i64 n, double* A);

// ---------------- workload (stellarator surfaces) ----------------
struct Surface {
  // nmode == 0: the reference's built-in 7-term surface; otherwise VMEC form
  // R = sum rc cos(m th - nfp n ph), Z = sum zs sin(m th - nfp n ph)
  i64 nfp = 1;
  std::vector<double> m, n, rc, zs;
};

struct Chart {
  double a1, b1, a2, b2, jj, ii;
};

std::vector<Chart> build_charts(const Surface& s, i64 mp, i64 np, i64 p, bool iterate_aspect);
// iterated aspect splits, then, for restol > 0, curvature refinement and 2:1
// balance (utils/stellarator_geo_mex.c:427-546); empty when the list would pass cap
// gpu: each curvature pass evaluates its charts' tails in one GPU launch (the
// tests of one pass depend only on the chart tested, so the result is the same)
std::vector<Chart> build_charts_refined(const Surface& s, i64 mp, i64 np, i64 p, double restol, i64 cap, bool* over,
                                        bool gpu = false);
// the relative weight of the last two degrees in the tensor-Legendre spectrum of
// the chart's mean curvature on its p x p grid (utils/stellarator_geo_mex.c:359-411)
double chart_curvature_tail(const Surface& s, const Chart& c, i64 mp, i64 np, i64 p);
// the tables it uses: Gauss-Legendre nodes x0 (p), differentiation matrix D (p x p)
// and the inverse of the Legendre-Vandermonde matrix (p x p), column-major
void curvature_tables(i64 p, std::vector<double>& x0, std::vector<double>& D, std::vector<double>& Vinv);
// the same tail for every chart of cl, one GPU wave per chart (device 0); equal
// to chart_curvature_tail up to the rounding of the device sin/cos
std::vector<double> curvature_tails_gpu(const Surface& s, const std::vector<Chart>& cl, i64 mp, i64 np, i64 p);

struct Geometry {
  i64 p = 0, h = 0, ntri = 0, nsrc = 0, mp = 0, np = 0;
  std::vector<double> sx, snx, sw, rts, rps;  // (3,nsrc) / (nsrc)
  std::vector<Chart> charts;
};

void build_geometry(const Surface& s, i64 mp, i64 np, i64 p, bool iterate_aspect, Geometry& g);
void build_geometry(const Surface& s, i64 mp, i64 np, i64 p, const std::vector<Chart>& charts, Geometry& g);

// What the per-chart work needs, built on the host: the chart list (aspect
// splits), each chart's diagonal (0: A), the Gauss-Legendre grid x0 and its
// differentiation matrix D (p x p), the tensor interpolation weights of the
// four (diagonal, side) cases, W[(2*diag + side)*p^2*h + m + i*p^2], and the
// VR weights wn.  build_geometry (host) and build_geometry_gpu use the same plan.
struct GeometryPlan {
  i64 p = 0, h = 0;
  double ts1 = 0.0, ts2 = 0.0;
  std::vector<Chart> charts;
  std::vector<int> diag;
  std::vector<double> x0, D, W, wn;
};
void plan_geometry(const Surface& s, i64 mp, i64 np, i64 p, bool iterate_aspect, GeometryPlan& g);
void plan_geometry(const Surface& s, i64 mp, i64 np, i64 p, const std::vector<Chart>& charts, GeometryPlan& g);
// the per-chart grids, spectral derivatives and node interpolation on the GPU
// (device 0, default stream); outputs to host arrays laid out as build_geometry's
void build_geometry_gpu(const Surface& s, const GeometryPlan& P, double* sx, double* snx, double* sw, double* rts,
                        double* rps);
// exact chart points of triangle itri (0-based) at reference-triangle (u,v)
void chart_uv2x(const Surface& s, const Geometry& g, i64 itri, i64 nuv, const double* uv, double* x);

}  // namespace qol
