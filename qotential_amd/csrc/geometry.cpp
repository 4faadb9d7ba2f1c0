// Workload generator: stellarator surfaces split into quad charts, each chart
// into two Vioreanu-Rokhlin triangles with p(p+1)/2 nodes.
//
// Restates the geometry the reference's drivers feed to rrq_r64:
//   surface evaluation          utils/stellarator_geo_mex.c:60-102 (built-in 7-term
//                               surface and the VMEC Fourier form)
//   chart list, aspect splits   :124-159 (single pass) and :427-479 (iterated)
//   curvature refinement, 2:1   :282-301, 359-411, 480-546 (restol > 0)
//   balance
//   chart -> two triangles      :636-749 (tensor Gauss-Legendre grid, spectral
//                               derivatives, diagonal choice, sw = |t1 x t2| w 4)
//   exact chart boundary        :555-582 (uv2x_core), used for ichart = 1
// The 1-D tensor interpolation uses the barycentric form; the reference keeps a
// monomial solve (BARYCENTRIC 0, :147), which agrees to rounding for p <= 16.
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "qol_host.h"

namespace qol {

namespace {

const double kPi = 3.14159265358979323846;
const double kAspect = 1.65;

void builtin_surface(double t, double p, double* x) {
  const double ct = std::cos(t), st = std::sin(t);
  const double c2pt = std::cos(2.0 * p - t), s2pt = std::sin(2.0 * p - t);
  const double c2p = std::cos(2.0 * p), s2p = std::sin(2.0 * p);
  const double cp = std::cos(p), sp = std::sin(p);
  const double cpt = std::cos(p + t), spt = std::sin(p + t);
  const double cmpt = std::cos(-p + t), smpt = std::sin(-p + t);
  x[0] = 0.17 * ct * c2pt + 0.11 * ct * c2p + 1.0 * ct * cp + 4.5 * ct - 0.25 * ct * cp + 0.07 * ct * cpt - 0.45 * ct * cmpt;
  x[1] = 0.17 * st * c2pt + 0.11 * st * c2p + 1.0 * st * cp + 4.5 * st - 0.25 * st * cp + 0.07 * st * cpt - 0.45 * st * cmpt;
  x[2] = 0.17 * s2pt + 0.11 * s2p + 1.0 * sp + 0.0 + 0.25 * sp + 0.07 * spt - 0.45 * smpt;
}

void surface_eval(const Surface& s, double t, double p, double* x) {
  if (s.rc.empty()) { builtin_surface(t, p, x); return; }
  double R = 0.0, Z = 0.0;
  for (size_t k = 0; k < s.rc.size(); ++k) {
    const double ang = s.m[k] * p - (double)s.nfp * s.n[k] * t;
    R += s.rc[k] * std::cos(ang);
    Z += s.zs[k] * std::sin(ang);
  }
  x[0] = R * std::cos(t);
  x[1] = R * std::sin(t);
  x[2] = Z;
}

void chart_eval(const Surface& s, const Chart& c, double ts1, double ts2, double t1, double t2, double* x) {
  const double u = ((c.a1 * t1 + c.b1) + 2.0 * c.jj) * ts1;
  const double v = ((c.a2 * t2 + c.b2) + 2.0 * c.ii) * ts2;
  surface_eval(s, u, v, x);
}

double dist3(const double* a, const double* b) {
  return std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}

// barycentric Lagrange interpolation matrix L(i,j) = l_j(t_i) on nodes s (p)
void lagrange_matrix(const double* t, i64 q, const double* s, i64 p, double* L) {
  std::vector<double> bw(p);
  for (i64 j = 0; j < p; ++j) {
    double pr = 1.0;
    for (i64 k = 0; k < p; ++k) if (k != j) pr *= s[j] - s[k];
    bw[j] = 1.0 / pr;
  }
  for (i64 i = 0; i < q; ++i) {
    i64 hit = -1;
    for (i64 j = 0; j < p; ++j) if (t[i] == s[j]) hit = j;
    if (hit >= 0) {
      for (i64 j = 0; j < p; ++j) L[i + q * j] = (j == hit) ? 1.0 : 0.0;
      continue;
    }
    double den = 0.0;
    for (i64 j = 0; j < p; ++j) { L[i + q * j] = bw[j] / (t[i] - s[j]); den += L[i + q * j]; }
    for (i64 j = 0; j < p; ++j) L[i + q * j] /= den;
  }
}

// tensor interpolation weights Mt(m,i), grid node m = a + b p at (t1,t2) = (x0[b], x0[a])
void tensor_weights(const std::vector<double>& t1, const std::vector<double>& t2, const double* x0, i64 p,
                    std::vector<double>& Mt) {
  const i64 q = (i64)t1.size();
  std::vector<double> L1(q * p), L2(q * p);
  lagrange_matrix(t1.data(), q, x0, p, L1.data());
  lagrange_matrix(t2.data(), q, x0, p, L2.data());
  Mt.assign(p * p * q, 0.0);
  for (i64 i = 0; i < q; ++i)
    for (i64 b = 0; b < p; ++b)
      for (i64 a = 0; a < p; ++a) Mt[(a + b * p) + i * p * p] = L2[i + q * a] * L1[i + q * b];
}

// the triangle's corner map: side 0/1, diagonal A (d1 >= d2) or B
void corner_map(bool diagA, i64 side, double u, double v, double* t1, double* t2) {
  const double up = 2.0 * u - 1.0, vp = 2.0 * v - 1.0;
  double a = diagA ? up : vp, b = diagA ? vp : -up;
  if (side == 1) { a = -a; b = -b; }
  *t1 = a;
  *t2 = b;
}

bool chart_diagonal_A(const Surface& s, const Chart& c, double ts1, double ts2, const double* x0, i64 p) {
  double ll[3], ur[3], ul[3], lr[3];
  chart_eval(s, c, ts1, ts2, x0[0], x0[0], ll);
  chart_eval(s, c, ts1, ts2, x0[p - 1], x0[p - 1], ur);
  chart_eval(s, c, ts1, ts2, x0[0], x0[p - 1], ul);
  chart_eval(s, c, ts1, ts2, x0[p - 1], x0[0], lr);
  double d1 = 0.0, d2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    d1 += (ll[d] - ur[d]) * (ll[d] - ur[d]);
    d2 += (ul[d] - lr[d]) * (ul[d] - lr[d]);
  }
  return d1 + 1e-13 > d2;
}

// V(i, k) = P_k(x0_i), the Legendre polynomials at the p grid nodes (column-major)
void legendre_vandermonde(const double* x0, i64 p, double* V) {
  for (i64 i = 0; i < p; ++i) {
    V[i] = 1.0;
    if (p > 1) V[i + p] = x0[i];
    for (i64 k = 1; k + 1 < p; ++k)
      V[i + (k + 1) * p] = ((2.0 * k + 1.0) * x0[i] * V[i + k * p] - k * V[i + (k - 1) * p]) / (k + 1.0);
  }
}

// inverse of a p x p column-major matrix: LU with partial pivoting, then p solves
std::vector<double> lu_inverse(std::vector<double> A, i64 p) {
  std::vector<double> X(p * p, 0.0);
  for (i64 i = 0; i < p; ++i) X[i + i * p] = 1.0;
  for (i64 k = 0; k < p; ++k) {
    i64 piv = k;
    for (i64 i = k + 1; i < p; ++i)
      if (std::fabs(A[i + k * p]) > std::fabs(A[piv + k * p])) piv = i;
    if (piv != k)
      for (i64 j = 0; j < p; ++j) {
        std::swap(A[k + j * p], A[piv + j * p]);
        std::swap(X[k + j * p], X[piv + j * p]);
      }
    for (i64 i = k + 1; i < p; ++i) {
      const double f = A[i + k * p] / A[k + k * p];
      for (i64 j = k + 1; j < p; ++j) A[i + j * p] -= f * A[k + j * p];
      for (i64 j = 0; j < p; ++j) X[i + j * p] -= f * X[k + j * p];
    }
  }
  for (i64 j = 0; j < p; ++j)
    for (i64 i = p - 1; i >= 0; --i) {
      double v = X[i + j * p];
      for (i64 k = i + 1; k < p; ++k) v -= A[i + k * p] * X[k + j * p];
      X[i + j * p] = v / A[i + i * p];
    }
  return X;
}

// the four quarters of a chart: quarter q takes the upper half in t1 when q & 1
// and in t2 when q & 2
Chart quarter(const Chart& c, int q) {
  Chart k = c;
  k.a1 = c.a1 / 2.0;
  k.a2 = c.a2 / 2.0;
  k.b1 = c.b1 + ((q & 1) ? c.a1 / 2.0 : -c.a1 / 2.0);
  k.b2 = c.b2 + ((q & 2) ? c.a2 / 2.0 : -c.a2 / 2.0);
  return k;
}

// parameter box of a chart: [t_lo, t_hi, p_lo, p_hi] in surface angles
void chart_box(const Chart& c, double ts1, double ts2, double* bx) {
  bx[0] = ((c.b1 - c.a1) + 2.0 * c.jj) * ts1;
  bx[1] = ((c.b1 + c.a1) + 2.0 * c.jj) * ts1;
  bx[2] = ((c.b2 - c.a2) + 2.0 * c.ii) * ts2;
  bx[3] = ((c.b2 + c.a2) + 2.0 * c.ii) * ts2;
}

struct CurvaturePlan {
  i64 p;
  double ts1, ts2;
  std::vector<double> x0, D, Vinv;
  CurvaturePlan(i64 mp, i64 np, i64 p_) : p(p_), ts1(kPi / (double)np), ts2(kPi / (double)mp) {
    x0.assign(p, 0.0);
    D.assign(p * p, 0.0);
    std::vector<double> w0(p), V(p * p);
    gl_nodes(p, x0.data(), w0.data(), D.data());
    legendre_vandermonde(x0.data(), p, V.data());
    Vinv = lu_inverse(V, p);
  }
};

// Mean curvature H = (E N - 2 F M + G L) / (2 (E G - F^2)) at the p x p grid
// nodes from spectral first and second derivatives of the chart map; its tensor
// Legendre coefficients C = V^-1 H V^-T; the result is the root of the share of
// sum C^2 held by degrees p-2 and p-1 in either variable.
double curvature_tail(const Surface& s, const Chart& c, const CurvaturePlan& P) {
  const i64 p = P.p, n2 = p * p;
  const double* D = P.D.data();
  std::vector<double> X(3 * n2), X1(3 * n2), X2(3 * n2), X11(3 * n2), X12(3 * n2), X22(3 * n2);
  // grid node m = a + b p sits at (t1, t2) = (x0[b], x0[a])
  for (i64 b = 0; b < p; ++b)
    for (i64 a = 0; a < p; ++a) chart_eval(s, c, P.ts1, P.ts2, P.x0[b], P.x0[a], &X[3 * (a + b * p)]);
  // d/dt1 runs over b, d/dt2 over a
  auto d_t1 = [&](const std::vector<double>& f, std::vector<double>& g) {
    for (i64 b = 0; b < p; ++b)
      for (i64 a = 0; a < p; ++a)
        for (int d = 0; d < 3; ++d) {
          double v = 0.0;
          for (i64 k = 0; k < p; ++k) v += D[b + k * p] * f[3 * (a + k * p) + d];
          g[3 * (a + b * p) + d] = v;
        }
  };
  auto d_t2 = [&](const std::vector<double>& f, std::vector<double>& g) {
    for (i64 b = 0; b < p; ++b)
      for (i64 a = 0; a < p; ++a)
        for (int d = 0; d < 3; ++d) {
          double v = 0.0;
          for (i64 k = 0; k < p; ++k) v += D[a + k * p] * f[3 * (k + b * p) + d];
          g[3 * (a + b * p) + d] = v;
        }
  };
  d_t1(X, X1);
  d_t2(X, X2);
  d_t1(X1, X11);
  d_t1(X2, X12);
  d_t2(X2, X22);
  std::vector<double> H(n2), T(n2);
  for (i64 m = 0; m < n2; ++m) {
    const double* u = &X1[3 * m];
    const double* v = &X2[3 * m];
    double nv[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    double nn = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    if (nn <= 0.0) nn = 1.0;
    for (double& x : nv) x /= nn;
    auto dot = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    const double E = dot(u, u), F = dot(u, v), G = dot(v, v);
    const double L = dot(nv, &X11[3 * m]), M = dot(nv, &X12[3 * m]), N = dot(nv, &X22[3 * m]);
    double den = 2.0 * (E * G - F * F);
    if (den == 0.0) den = 1.0;
    H[m] = (E * N - 2.0 * F * M + G * L) / den;
  }
  // C = Vinv H Vinv^T, H(a, b) = H[a + b p]
  const double* W = P.Vinv.data();
  for (i64 b = 0; b < p; ++b)
    for (i64 a = 0; a < p; ++a) {
      double v = 0.0;
      for (i64 k = 0; k < p; ++k) v += W[a + k * p] * H[k + b * p];
      T[a + b * p] = v;
    }
  double tail = 0.0, tot = 0.0;
  for (i64 b = 0; b < p; ++b)
    for (i64 a = 0; a < p; ++a) {
      double v = 0.0;
      for (i64 k = 0; k < p; ++k) v += T[a + k * p] * W[b + k * p];
      tot += v * v;
      if (a >= p - 2 || b >= p - 2) tail += v * v;
    }
  return std::sqrt(tail / (tot > 0.0 ? tot : 1.0));
}

}  // namespace

double chart_curvature_tail(const Surface& s, const Chart& c, i64 mp, i64 np, i64 p) {
  return curvature_tail(s, c, CurvaturePlan(mp, np, p));
}

void curvature_tables(i64 p, std::vector<double>& x0, std::vector<double>& D, std::vector<double>& Vinv) {
  const CurvaturePlan P(1, 1, p);
  x0 = P.x0;
  D = P.D;
  Vinv = P.Vinv;
}

std::vector<Chart> build_charts_refined(const Surface& s, i64 mp, i64 np, i64 p, double restol, i64 cap,
                                        bool* over, bool gpu) {
  if (cap <= 0) cap = 200000;
  *over = false;
  std::vector<Chart> L = build_charts(s, mp, np, p, true);
  if ((i64)L.size() > cap) { *over = true; return {}; }
  if (!(restol > 0.0)) return L;
  const CurvaturePlan P(mp, np, p);
  // curvature: split in four until the tail passes; a chart that passed is not
  // tested again (at most 12 levels)
  std::vector<char> ok(L.size(), 0);
  for (int level = 0; level < 12; ++level) {
    const i64 n0 = (i64)L.size();
    i64 nsplit = 0;
    // this pass's tails: a chart is tested before anything of this pass
    // replaces it, so all of them can be evaluated up front
    std::vector<double> tail(n0, 0.0);
    if (gpu) {
      std::vector<Chart> todo;
      std::vector<i64> at;
      for (i64 k = 0; k < n0; ++k)
        if (!ok[k]) { todo.push_back(L[k]); at.push_back(k); }
      const std::vector<double> t = curvature_tails_gpu(s, todo, mp, np, p);
      for (size_t j = 0; j < at.size(); ++j) tail[at[j]] = t[j];
    } else {
      for (i64 k = 0; k < n0; ++k)
        if (!ok[k]) tail[k] = curvature_tail(s, L[k], P);
    }
    for (i64 k = 0; k < n0; ++k) {
      if (ok[k]) continue;
      if (tail[k] <= restol) { ok[k] = 1; continue; }
      if ((i64)L.size() + 3 > cap) { *over = true; return {}; }
      const Chart c = L[k];
      L[k] = quarter(c, 0);
      for (int q = 1; q < 4; ++q) { L.push_back(quarter(c, q)); ok.push_back(0); }
      ++nsplit;
    }
    if (!nsplit) break;
  }
  // 2:1 balance: split a chart whose extent along a face it shares with a
  // neighbour (one base cell apart at most, periodic in both angles) exceeds
  // twice the neighbour's; sweeps until none is split (at most 20)
  // Candidates are the charts of the base cells within one cell of k's, the
  // neighbourhood of the reference's index filter (|djj| <= 1 or wrapped, the
  // same in ii); bucketing the charts by base cell finds them without the
  // all-pairs scan, and a chart is split when any candidate asks for it, so the
  // visiting order does not change the result.
  const double et1 = 1e-9 * P.ts1, et2 = 1e-9 * P.ts2;
  for (int sweep = 0; sweep < 20; ++sweep) {
    const i64 n0 = (i64)L.size();
    std::vector<double> box(4 * n0);
    for (i64 k = 0; k < n0; ++k) chart_box(L[k], P.ts1, P.ts2, &box[4 * k]);
    // charts by base cell (jj - 1) + np (ii - 1), as a counting sort
    std::vector<i64> cstart(mp * np + 1, 0), cidx(n0);
    auto cell = [&](const Chart& c) { return ((i64)c.jj - 1) + np * ((i64)c.ii - 1); };
    for (i64 k = 0; k < n0; ++k) ++cstart[cell(L[k]) + 1];
    for (i64 c = 0; c < mp * np; ++c) cstart[c + 1] += cstart[c];
    {
      std::vector<i64> fill(cstart.begin(), cstart.end() - 1);
      for (i64 k = 0; k < n0; ++k) cidx[fill[cell(L[k])]++] = k;
    }
    std::vector<char> split(n0, 0);
    for (i64 k = 0; k < n0; ++k) {
      const double* bk = &box[4 * k];
      i64 cj[3], ci[3];
      int nj = 0, ni = 0;
      for (i64 d = -1; d <= 1; ++d) {
        const i64 j = (((i64)L[k].jj - 1 + d) % np + np) % np, i = (((i64)L[k].ii - 1 + d) % mp + mp) % mp;
        if (std::find(cj, cj + nj, j) == cj + nj) cj[nj++] = j;
        if (std::find(ci, ci + ni, i) == ci + ni) ci[ni++] = i;
      }
      for (int a = 0; a < nj && !split[k]; ++a)
        for (int b = 0; b < ni && !split[k]; ++b)
          for (i64 e = cstart[cj[a] + np * ci[b]]; e < cstart[cj[a] + np * ci[b] + 1] && !split[k]; ++e) {
        const i64 q = cidx[e];
        if (q == k) continue;
        const double* bq = &box[4 * q];
        if ((std::fabs(bk[1] - bq[0]) < et1 || std::fabs(bk[0] - bq[1]) < et1) &&
            std::fmin(bk[3], bq[3]) - std::fmax(bk[2], bq[2]) > et2) {
          // a face at constant t1: compare the extents in t2
          if (bk[3] - bk[2] > 2.0 * (bq[3] - bq[2]) + et2) split[k] = 1;
        } else if ((std::fabs(bk[3] - bq[2]) < et2 || std::fabs(bk[2] - bq[3]) < et2) &&
                   std::fmin(bk[1], bq[1]) - std::fmax(bk[0], bq[0]) > et1) {
          if (bk[1] - bk[0] > 2.0 * (bq[1] - bq[0]) + et1) split[k] = 1;
        }
          }
    }
    i64 nsplit = 0;
    for (i64 k = 0; k < n0; ++k) {
      if (!split[k]) continue;
      if ((i64)L.size() + 3 > cap) { *over = true; return {}; }
      const Chart c = L[k];
      L[k] = quarter(c, 0);
      for (int q = 1; q < 4; ++q) L.push_back(quarter(c, q));
      ++nsplit;
    }
    if (!nsplit) break;
  }
  return L;
}

std::vector<Chart> build_charts(const Surface& s, i64 mp, i64 np, i64 p, bool iterate_aspect) {
  std::vector<double> x0(p), w0(p);
  gl_nodes(p, x0.data(), w0.data(), nullptr);
  const double ts1 = kPi / (double)np, ts2 = kPi / (double)mp;
  std::vector<Chart> L;
  for (i64 j = 1; j <= np; ++j)
    for (i64 i = 1; i <= mp; ++i) L.push_back(Chart{1.0, 0.0, 1.0, 0.0, (double)j, (double)i});
  auto aspect = [&](const Chart& c, double& dt, double& dp) {
    double ll[3], ul[3], lr[3];
    chart_eval(s, c, ts1, ts2, x0[0], x0[0], ll);
    chart_eval(s, c, ts1, ts2, x0[0], x0[p - 1], ul);
    chart_eval(s, c, ts1, ts2, x0[p - 1], x0[0], lr);
    dt = dist3(ll, lr);
    dp = dist3(ll, ul);
  };
  if (!iterate_aspect) {
    // one pass; halves are emitted in place (utils/stellarator_geo_mex.c:124-159)
    std::vector<Chart> out;
    for (const Chart& c : L) {
      double dt, dp;
      aspect(c, dt, dp);
      if (dt / dp > kAspect) {
        Chart a = c, b = c;
        a.a1 = 0.5; a.b1 = -0.5; b.a1 = 0.5; b.b1 = 0.5;
        out.push_back(a); out.push_back(b);
      } else if (dp / dt > kAspect) {
        Chart a = c, b = c;
        a.a2 = 0.5; a.b2 = -0.5; b.a2 = 0.5; b.b2 = 0.5;
        out.push_back(a); out.push_back(b);
      } else out.push_back(c);
    }
    return out;
  }
  // iterated passes; the second half is appended (:451-479)
  for (int pass = 0; pass < 8; ++pass) {
    i64 nsplit = 0;
    const i64 n0 = (i64)L.size();
    for (i64 k = 0; k < n0; ++k) {
      const Chart c = L[k];
      double dt, dp;
      aspect(c, dt, dp);
      if (dt / dp > kAspect) {
        L[k].a1 = c.a1 / 2.0; L[k].b1 = c.b1 - c.a1 / 2.0;
        Chart b = c; b.a1 = c.a1 / 2.0; b.b1 = c.b1 + c.a1 / 2.0;
        L.push_back(b);
        ++nsplit;
      } else if (dp / dt > kAspect) {
        L[k].a2 = c.a2 / 2.0; L[k].b2 = c.b2 - c.a2 / 2.0;
        Chart b = c; b.a2 = c.a2 / 2.0; b.b2 = c.b2 + c.a2 / 2.0;
        L.push_back(b);
        ++nsplit;
      }
    }
    if (!nsplit) break;
  }
  return L;
}

void plan_geometry(const Surface& s, i64 mp, i64 np, i64 p, bool iterate_aspect, GeometryPlan& g) {
  plan_geometry(s, mp, np, p, build_charts(s, mp, np, p, iterate_aspect), g);
}

void plan_geometry(const Surface& s, i64 mp, i64 np, i64 p, const std::vector<Chart>& charts, GeometryPlan& g) {
  const OrderTables& T = order_tables(p);
  const i64 h = T.h;
  g.p = p; g.h = h;
  g.ts1 = kPi / (double)np;
  g.ts2 = kPi / (double)mp;
  g.charts = charts;
  g.x0.assign(p, 0.0);
  g.D.assign(p * p, 0.0);
  std::vector<double> w0(p);
  gl_nodes(p, g.x0.data(), w0.data(), g.D.data());
  g.diag.resize(g.charts.size());
  for (size_t k = 0; k < g.charts.size(); ++k)
    g.diag[k] = chart_diagonal_A(s, g.charts[k], g.ts1, g.ts2, g.x0.data(), p) ? 0 : 1;
  // interpolation weights for the four (diagonal, side) combinations
  g.W.clear();
  for (int dg = 0; dg < 2; ++dg)
    for (int sd = 0; sd < 2; ++sd) {
      std::vector<double> t1(h), t2(h), Mt;
      for (i64 i = 0; i < h; ++i) corner_map(dg == 0, sd, T.uvn[2 * i], T.uvn[2 * i + 1], &t1[i], &t2[i]);
      tensor_weights(t1, t2, g.x0.data(), p, Mt);
      g.W.insert(g.W.end(), Mt.begin(), Mt.end());
    }
  g.wn = T.wn;
}

void build_geometry(const Surface& s, i64 mp, i64 np, i64 p, bool iterate_aspect, Geometry& g) {
  build_geometry(s, mp, np, p, build_charts(s, mp, np, p, iterate_aspect), g);
}

void build_geometry(const Surface& s, i64 mp, i64 np, i64 p, const std::vector<Chart>& charts, Geometry& g) {
  GeometryPlan P;
  plan_geometry(s, mp, np, p, charts, P);
  const i64 h = P.h, p2 = p * p;
  g.p = p; g.h = h; g.mp = mp; g.np = np;
  g.charts = P.charts;
  const i64 nq = (i64)g.charts.size();
  g.ntri = 2 * nq;
  g.nsrc = g.ntri * h;
  g.sx.assign(3 * g.nsrc, 0.0); g.snx.assign(3 * g.nsrc, 0.0); g.sw.assign(g.nsrc, 0.0);
  g.rts.assign(3 * g.nsrc, 0.0); g.rps.assign(3 * g.nsrc, 0.0);
  const double* x0 = P.x0.data();
  const double* D = P.D.data();
  std::vector<double> grid(3 * p2), d2(3 * p2), d1(3 * p2);
  for (i64 k = 0; k < nq; ++k) {
    const Chart& c = g.charts[k];
    for (i64 b = 0; b < p; ++b)
      for (i64 a = 0; a < p; ++a) chart_eval(s, c, P.ts1, P.ts2, x0[b], x0[a], &grid[3 * (a + b * p)]);
    // d/dt2 (first grid index) and d/dt1 (second grid index), spectral
    for (i64 b = 0; b < p; ++b)
      for (i64 a = 0; a < p; ++a)
        for (int d = 0; d < 3; ++d) {
          double s2 = 0.0, s1 = 0.0;
          for (i64 q = 0; q < p; ++q) s2 += D[a + q * p] * grid[3 * (q + b * p) + d];
          for (i64 q = 0; q < p; ++q) s1 += D[b + q * p] * grid[3 * (a + q * p) + d];
          d2[3 * (a + b * p) + d] = s2;
          d1[3 * (a + b * p) + d] = s1;
        }
    const int dg = P.diag[k];
    for (int sd = 0; sd < 2; ++sd) {
      const double* W = P.W.data() + (size_t)(2 * dg + sd) * p2 * h;
      const i64 off = (2 * k + sd) * h;
      for (i64 i = 0; i < h; ++i) {
        double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (i64 m = 0; m < p2; ++m) {
          const double w = W[m + i * p2];
          for (int d = 0; d < 3; ++d) {
            acc[d] += grid[3 * m + d] * w;
            acc[3 + d] += d2[3 * m + d] * w;
            acc[6 + d] += d1[3 * m + d] * w;
          }
        }
        // "rts" carries d/dt2, "rps" d/dt1, each times 2; normal = rps x rts
        const double* xt = acc + 3;
        const double* xp = acc + 6;
        double nx[3] = {xp[1] * xt[2] - xp[2] * xt[1], xp[2] * xt[0] - xp[0] * xt[2], xp[0] * xt[1] - xp[1] * xt[0]};
        const double sp = std::sqrt(nx[0] * nx[0] + nx[1] * nx[1] + nx[2] * nx[2]);
        for (int d = 0; d < 3; ++d) {
          g.sx[3 * (off + i) + d] = acc[d];
          g.snx[3 * (off + i) + d] = nx[d] / sp;
          g.rts[3 * (off + i) + d] = 2.0 * xt[d];
          g.rps[3 * (off + i) + d] = 2.0 * xp[d];
        }
        g.sw[off + i] = sp * P.wn[i] * 4.0;
      }
    }
  }
}

void chart_uv2x(const Surface& s, const Geometry& g, i64 itri, i64 nuv, const double* uv, double* x) {
  if (itri < 0 || itri >= g.ntri) throw std::runtime_error("chart_uv2x: triangle index out of range");
  const i64 p = g.p;
  std::vector<double> x0(p), w0(p);
  gl_nodes(p, x0.data(), w0.data(), nullptr);
  const double ts1 = kPi / (double)g.np, ts2 = kPi / (double)g.mp;
  const Chart& c = g.charts[itri / 2];
  const bool dA = chart_diagonal_A(s, c, ts1, ts2, x0.data(), p);
  for (i64 i = 0; i < nuv; ++i) {
    double t1, t2;
    corner_map(dA, itri % 2, uv[2 * i], uv[2 * i + 1], &t1, &t2);
    chart_eval(s, c, ts1, ts2, t1, t2, &x[3 * i]);
  }
}

}  // namespace qol
