/* Oracle: recursive-reduction quadrature (RRQ) for 3D Laplace S and D on one
 * curved triangular patch.  TEST INFRASTRUCTURE ONLY (see qol_oracle.h).
 *
 * Restates rrq_r64 (lap3d_close_mod.f90:622-881) step for step:
 *   levels (build_patch_levels / lap3dsdlpmat_levels)        :712-773
 *   frame + boundary + harmonic basis + quaternion matrix      :716-768, :472-620
 *   singularity-swap weights for the 1/r kernel                 :797-799
 *   solid angle of the patch (Dirac-string line integral)       :808-817
 *   chi / alpha boundary moments                                :832-834, :336-470
 *   moment recursion ("omega")                                  :838-839
 *   moments -> weights: solve with M^T (4h) then Msl^T (h)      :842-870
 * The helpers the reference takes from the absent QuatApproximation /
 * LineQuaaadrature submodules are re-derived here from the mathematics:
 *  - basis: regular solid harmonics H_n^m = coefficients of e^{-i m a} in
 *    (Z + i/2 (W e^{-ia} + conj(W) e^{ia}))^n, polar axis = frame x,
 *    W = frame y + i frame z (the reference's (y,z,x) permutation, :570);
 *  - for F harmonic and Z = (y-x).grad the Euler field about the target x,
 *    Cartan's formula gives  W(Z F) = int_{dS} grad F.(z x dy)/|z|  and
 *    Q[Z grad F] = int_{dS} E(z)(z x dy) grad F, with W(F)=int G dF/dn - F dG/dn,
 *    Q[Phi] = int E n Phi (quaternion), E(z) = z/|z|^3.  Z H_n^m = n H_n^m -
 *    (x.grad) H_n^m and (x.grad) lowers the degree, so the surface moments follow
 *    by an upward recursion in degree from the line moments, seeded by the solid
 *    angle (degree 0) and int E n dS (degree 1);
 *  - the 1/r^3 line moment is integrated by parts around the closed boundary into
 *    a 1/r moment of z (x) (grad grad H . dy), so one singularity-swap rule (1/r)
 *    serves every line integral.
 */
#include "qol_oracle.h"
#include <algorithm>
#include <mutex>
#include <chrono>
#include <cmath>
#include <cstdio>

namespace orc {

static const double PI = 3.14159265358979323846;
static const cplx IU(0.0, 1.0);

/* side convention for on-surface targets: exterior => limit from the side the
   panel normals snx point to (the stellarator charts have outward normals;
   the order-4 fixture GRF residual pins this choice). */
static const double EXTERIOR_SNX_SIDE = 1.0;

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

/* ---------- harmonics: H'_n^m = n! H_n^m, m = 0..n, stored at idx(n,m) ---------- */
static inline i64 hidx(i64 n, i64 m) { return n * (n + 1) / 2 + m; }

/* values of H'_n^m for m in -1..n+1 handled through conj(H'_n^m)=(-1)^m H'_n^{-m} */
static inline cplx hget(const std::vector<cplx>& H, i64 n, i64 m) {
  if (n < 0) return 0.0;
  if (m > n || m < -n) return 0.0;
  if (m >= 0) return H[hidx(n, m)];
  cplx v = std::conj(H[hidx(n, -m)]);
  return ((-m) % 2) ? -v : v;
}

/* X = (x,y,z) frame point.  Z = x, W = y + i z. */
static void harm_vals(const double* X, i64 p, std::vector<cplx>& H) {
  H.assign((p + 1) * (p + 2) / 2, 0.0);
  cplx W(X[1], X[2]);
  double Z = X[0];
  H[0] = 1.0;
  for (i64 n = 1; n <= p; ++n)
    for (i64 m = 0; m <= n; ++m) {
      cplx v = Z * hget(H, n - 1, m) + 0.5 * IU * (W * hget(H, n - 1, m - 1) + std::conj(W) * hget(H, n - 1, m + 1));
      H[hidx(n, m)] = v;
    }
}

/* gradient (d/dx, d/dy, d/dz in the frame) of H'_n^m from degree n-1 values */
static void harm_grad(const std::vector<cplx>& H, i64 n, i64 m, cplx* g) {
  double dn = (double)n;
  g[0] = dn * hget(H, n - 1, m);
  g[1] = dn * 0.5 * IU * (hget(H, n - 1, m - 1) + hget(H, n - 1, m + 1));
  g[2] = dn * 0.5 * (hget(H, n - 1, m + 1) - hget(H, n - 1, m - 1));
}

/* ---------- Legendre expansions ---------- */
template <class T>
static void leg_eval(const double* a, i64 n, T s, T& val, T& der) {
  /* a: n coefficients; returns sum a_k P_k(s) and its derivative */
  T p0 = 1.0, p1 = s, d0 = 0.0, d1 = 1.0;
  val = a[0] * p0;
  der = 0.0;
  if (n > 1) { val += a[1] * p1; der += a[1] * d1; }
  for (i64 k = 1; k + 1 < n; ++k) {
    double dk = (double)k;
    T p2 = ((2.0 * dk + 1.0) * s * p1 - dk * p0) / (dk + 1.0);
    T d2 = ((2.0 * dk + 1.0) * (p1 + s * d1) - dk * d0) / (dk + 1.0);
    val += a[k + 1] * p2;
    der += a[k + 1] * d2;
    p0 = p1; p1 = p2; d0 = d1; d1 = d2;
  }
}

static double bernstein_rho(cplx s) {
  cplx r = s + std::sqrt(s - 1.0) * std::sqrt(s + 1.0);
  double a = std::abs(r);
  return a >= 1.0 ? a : 1.0 / a;
}

/* ---------- one panel ---------- */
struct Edge {
  std::vector<double> coef; /* 3*nq Legendre coefficients, coef[c*nq+k] */
  std::vector<double> X, dX; /* frame positions/derivatives at GL nodes, 3*nq */
};

struct Panel {
  i64 p = 0, h = 0, nq = 0, ncoef = 0;
  double alpha = 0.0, sigma = 1.0, L = 0.0;
  double R[9], c[3];
  std::vector<double> cx;    /* Koornwinder coefficients of the geometry, h x 3 */
  std::vector<double> tgl, wgl, Dgl, Legmat;
  Edge e[3];
  /* basis data at boundary nodes, per node i (3*nq), per coefficient k (ncoef):
     gradH[(i*ncoef+k)*3+c], V[(i*ncoef+k)*3+c] */
  std::vector<cplx> gradH, Vd;
  /* quaternion system */
  std::vector<double> MDT, Mslt, Fnode; /* LU-factored M^T (4h), Msl^T (h); F(i,k) */
  std::vector<double> MDT0, Mslt0;      /* the same before factorisation */
  std::vector<i64> piv4, piv1;
  /* monomial Vandermonde at GL nodes (for SSQ), LU of V^T */
  std::vector<double> VT;
  std::vector<i64> pivV;
};

static void lu_factor(i64 n, double* A, i64* piv) {
  for (i64 k = 0; k < n; ++k) {
    i64 ip = k;
    double pm = std::fabs(A[k + n * k]);
    for (i64 i = k + 1; i < n; ++i)
      if (std::fabs(A[i + n * k]) > pm) { pm = std::fabs(A[i + n * k]); ip = i; }
    piv[k] = ip;
    if (ip != k)
      for (i64 j = 0; j < n; ++j) std::swap(A[k + n * j], A[ip + n * j]);
    for (i64 i = k + 1; i < n; ++i) {
      double f = A[i + n * k] / A[k + n * k];
      A[i + n * k] = f;
      for (i64 j = k + 1; j < n; ++j) A[i + n * j] -= f * A[k + n * j];
    }
  }
}

static void lu_solve(i64 n, const double* A, const i64* piv, double* b) {
  /* rows were swapped whole during factorisation, so apply every swap first */
  for (i64 k = 0; k < n; ++k)
    if (piv[k] != k) std::swap(b[k], b[piv[k]]);
  for (i64 k = 0; k < n; ++k)
    for (i64 i = k + 1; i < n; ++i) b[i] -= A[i + n * k] * b[k];
  for (i64 i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (i64 k = i + 1; k < n; ++k) s -= A[i + n * k] * b[k];
    b[i] = s / A[i + n * i];
  }
}

static void geom_eval(const Panel& P, double u, double v, double* x, double* xu, double* xv) {
  std::vector<double> pol(P.h), pu(P.h), pv(P.h);
  koorn_pols(u, v, P.p - 1, pol.data(), pu.data(), pv.data());
  for (int c = 0; c < 3; ++c) {
    double s = 0.0, su = 0.0, sv = 0.0;
    for (i64 k = 0; k < P.h; ++k) {
      s += pol[k] * P.cx[k + P.h * c];
      su += pu[k] * P.cx[k + P.h * c];
      sv += pv[k] * P.cx[k + P.h * c];
    }
    x[c] = s;
    if (xu) xu[c] = su;
    if (xv) xv[c] = sv;
  }
}

static void to_frame(const Panel& P, const double* x, double* X) {
  double d[3] = {x[0] - P.c[0], x[1] - P.c[1], x[2] - P.c[2]};
  for (int r = 0; r < 3; ++r) X[r] = P.alpha * (P.R[3 * r] * d[0] + P.R[3 * r + 1] * d[1] + P.R[3 * r + 2] * d[2]);
}
static void rot(const Panel& P, const double* v, double* V) {
  for (int r = 0; r < 3; ++r) V[r] = P.R[3 * r] * v[0] + P.R[3 * r + 1] * v[1] + P.R[3 * r + 2] * v[2];
}

static inline void cross(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

static void setup_panel(Panel& P, i64 n, const double* sx, const double* snx, const double* sw,
                        i64 order, i64 nquad, i64 ichart, const double* sxbd_chart,
                        const double* rv_chart, const double* tgl, const double* wgl,
                        const double* Dgl, const double* Legmat, const double* umatr) {
  P.p = order; P.h = n; P.nq = nquad; P.ncoef = (order + 1) * (order + 2) / 2;
  P.tgl.assign(tgl, tgl + nquad); P.wgl.assign(wgl, wgl + nquad);
  P.Dgl.assign(Dgl, Dgl + nquad * nquad); P.Legmat.assign(Legmat, Legmat + nquad * nquad);
  double swsum = 0.0;
  for (i64 i = 0; i < n; ++i) swsum += sw[i];
  P.alpha = 1.25 / std::sqrt(2.0 * swsum);
  /* geometry polynomial */
  P.cx.assign(n * 3, 0.0);
  for (int c = 0; c < 3; ++c)
    for (i64 k = 0; k < n; ++k) {
      double s = 0.0;
      for (i64 i = 0; i < n; ++i) s += umatr[k + n * i] * sx[3 * i + c];
      P.cx[k + n * c] = s;
    }
  /* vertices */
  double rv[9];
  if (ichart == 1) for (int i = 0; i < 9; ++i) rv[i] = rv_chart[i];
  else {
    const double uvv[6] = {0, 0, 1, 0, 0, 1};
    for (int j = 0; j < 3; ++j) geom_eval(P, uvv[2 * j], uvv[2 * j + 1], rv + 3 * j, nullptr, nullptr);
  }
  /* circumcircle frame */
  double a[3], b[3], nrm[3];
  for (int c = 0; c < 3; ++c) { a[c] = rv[3 + c] - rv[c]; b[c] = rv[6 + c] - rv[c]; }
  cross(a, b, nrm);
  double nn = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
  double e3[3] = {nrm[0] / nn, nrm[1] / nn, nrm[2] / nn};
  double mn[3] = {0, 0, 0};
  for (i64 i = 0; i < n; ++i) for (int c = 0; c < 3; ++c) mn[c] += snx[3 * i + c];
  if (mn[0] * e3[0] + mn[1] * e3[1] + mn[2] * e3[2] < 0.0) for (int c = 0; c < 3; ++c) e3[c] = -e3[c];
  double la = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  double e1[3] = {a[0] / la, a[1] / la, a[2] / la}, e2[3];
  cross(e3, e1, e2);
  /* circumcenter: c0 = rv0 + (|a|^2 (b x n) x ... ) standard formula */
  double axb[3];
  cross(a, b, axb);
  double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2], bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
  double t1[3], t2[3], num[3];
  cross(axb, a, t1);
  cross(b, axb, t2);
  double den = 2.0 * (axb[0] * axb[0] + axb[1] * axb[1] + axb[2] * axb[2]);
  for (int c = 0; c < 3; ++c) num[c] = bb * t1[c] + aa * t2[c];
  double c0[3];
  for (int c = 0; c < 3; ++c) c0[c] = rv[c] + num[c] / den;
  for (int c = 0; c < 3; ++c) { P.R[c] = e1[c]; P.R[3 + c] = e2[c]; P.R[6 + c] = e3[c]; }
  for (int c = 0; c < 3; ++c) P.c[c] = c0[c];
  double zmin = 1e300;
  for (i64 i = 0; i < n; ++i) { double X[3]; to_frame(P, sx + 3 * i, X); zmin = std::min(zmin, X[2]); }
  double sx3min = std::max(0.0, -zmin);
  for (int c = 0; c < 3; ++c) P.c[c] = c0[c] - (1.25 * sx3min / P.alpha) * e3[c];
  /* patch size */
  P.L = 0.0;
  for (int i = 0; i < 3; ++i) for (int j = i + 1; j < 3; ++j) {
    double d = 0; for (int c = 0; c < 3; ++c) d += (rv[3 * i + c] - rv[3 * j + c]) * (rv[3 * i + c] - rv[3 * j + c]);
    P.L = std::max(P.L, std::sqrt(d));
  }
  /* orientation: sigma = sign(n_uv . snx) */
  {
    std::vector<double> uvn(2 * n);
    vr_nodes(order - 1, uvn.data(), nullptr);
    double acc = 0.0;
    for (i64 i = 0; i < n; ++i) {
      double x[3], xu[3], xv[3], nuv[3];
      geom_eval(P, uvn[2 * i], uvn[2 * i + 1], x, xu, xv);
      cross(xu, xv, nuv);
      acc += nuv[0] * snx[3 * i] + nuv[1] * snx[3 * i + 1] + nuv[2] * snx[3 * i + 2];
    }
    P.sigma = acc >= 0.0 ? 1.0 : -1.0;
  }
  /* boundary: ccw in uv, (0,1)->(0,0)->(1,0)->(0,1) */
  const double EA[3][2] = {{0, 1}, {0, 0}, {1, 0}}, EB[3][2] = {{0, 0}, {1, 0}, {0, 1}};
  for (int e = 0; e < 3; ++e) {
    Edge& E = P.e[e];
    E.X.assign(3 * nquad, 0.0); E.dX.assign(3 * nquad, 0.0); E.coef.assign(3 * nquad, 0.0);
    for (i64 i = 0; i < nquad; ++i) {
      double x[3];
      if (ichart == 1) for (int c = 0; c < 3; ++c) x[c] = sxbd_chart[3 * (e * nquad + i) + c];
      else {
        double s01 = 0.5 * (tgl[i] + 1.0);
        double u = EA[e][0] + (EB[e][0] - EA[e][0]) * s01, v = EA[e][1] + (EB[e][1] - EA[e][1]) * s01;
        geom_eval(P, u, v, x, nullptr, nullptr);
      }
      to_frame(P, x, &E.X[3 * i]);
    }
    for (int c = 0; c < 3; ++c)
      for (i64 k = 0; k < nquad; ++k) {
        double s = 0.0;
        for (i64 i = 0; i < nquad; ++i) s += Legmat[k + nquad * i] * E.X[3 * i + c];
        E.coef[c * nquad + k] = s;
      }
    for (i64 i = 0; i < nquad; ++i)
      for (int c = 0; c < 3; ++c) {
        double v, d;
        leg_eval<double>(&E.coef[c * nquad], nquad, tgl[i], v, d);
        E.dX[3 * i + c] = d;
      }
  }
  /* basis data at boundary nodes */
  i64 nb = 3 * nquad, nc = P.ncoef, p = order;
  P.gradH.assign(nb * nc * 3, 0.0);
  P.Vd.assign(nb * nc * 3, 0.0);
  std::vector<cplx> H;
  for (int e = 0; e < 3; ++e)
    for (i64 i = 0; i < nquad; ++i) {
      i64 ib = e * nquad + i;
      const double* X = &P.e[e].X[3 * i];
      const double* dX = &P.e[e].dX[3 * i];
      harm_vals(X, p, H);
      /* gradients of all degrees 0..p (need degree n-1 gradients for V) */
      std::vector<cplx> G((p + 1) * (p + 2) / 2 * 3 + 3, 0.0);
      for (i64 nn2 = 0; nn2 <= p; ++nn2)
        for (i64 m = 0; m <= nn2; ++m) harm_grad(H, nn2, m, &G[3 * hidx(nn2, m)]);
      auto gget = [&](i64 n2, i64 m, int c) -> cplx {
        if (n2 < 0 || m > n2 || m < -n2) return 0.0;
        if (m >= 0) return G[3 * hidx(n2, m) + c];
        cplx v = std::conj(G[3 * hidx(n2, -m) + c]);
        return ((-m) % 2) ? -v : v;
      };
      cplx dW(dX[1], dX[2]);
      for (i64 nn2 = 0; nn2 <= p; ++nn2)
        for (i64 m = 0; m <= nn2; ++m) {
          i64 k = hidx(nn2, m);
          for (int c = 0; c < 3; ++c) {
            P.gradH[(ib * nc + k) * 3 + c] = G[3 * k + c];
            cplx vv = (double)nn2 * (dX[0] * gget(nn2 - 1, m, c) +
                                     0.5 * IU * (dW * gget(nn2 - 1, m - 1, c) + std::conj(dW) * gget(nn2 - 1, m + 1, c)));
            P.Vd[(ib * nc + k) * 3 + c] = vv;
          }
        }
    }
  /* quaternion system at the nodes: F_k = Im((-i)^m H'_n^m), n=1..p, m=1..n */
  i64 h = n;
  std::vector<double> F(h * h), F1(h * h), F2(h * h), F3(h * h), Msl(h * h);
  for (i64 i = 0; i < h; ++i) {
    double X[3], nX[3];
    to_frame(P, sx + 3 * i, X);
    rot(P, snx + 3 * i, nX);
    harm_vals(X, p, H);
    i64 k = 0;
    for (i64 nn2 = 1; nn2 <= p; ++nn2)
      for (i64 m = 1; m <= nn2; ++m, ++k) {
        cplx rotm = std::pow(-IU, (double)m);
        cplx g[3];
        harm_grad(H, nn2, m, g);
        F[i + h * k] = std::imag(rotm * H[hidx(nn2, m)]);
        F1[i + h * k] = std::imag(rotm * g[0]);
        F2[i + h * k] = std::imag(rotm * g[1]);
        F3[i + h * k] = std::imag(rotm * g[2]);
        Msl[i + h * k] = nX[0] * F1[i + h * k] + nX[1] * F2[i + h * k] + nX[2] * F3[i + h * k];
      }
  }
  P.Fnode = F;
  i64 H4 = 4 * h;
  std::vector<double> M(H4 * H4, 0.0);
  auto put = [&](i64 rb, i64 cb, const std::vector<double>& B, double sg) {
    for (i64 j = 0; j < h; ++j)
      for (i64 i = 0; i < h; ++i) M[(rb * h + i) + H4 * (cb * h + j)] = sg * B[i + h * j];
  };
  put(0, 1, F1, -1); put(0, 2, F2, -1); put(0, 3, F3, -1);
  put(1, 0, F1, 1);  put(1, 2, F3, -1); put(1, 3, F2, 1);
  put(2, 0, F2, 1);  put(2, 1, F3, 1);  put(2, 3, F1, -1);
  put(3, 0, F3, 1);  put(3, 1, F2, -1); put(3, 2, F1, 1);
  P.MDT.assign(H4 * H4, 0.0);
  for (i64 j = 0; j < H4; ++j) for (i64 i = 0; i < H4; ++i) P.MDT[i + H4 * j] = M[j + H4 * i];
  P.MDT0 = P.MDT;
  P.piv4.assign(H4, 0);
  lu_factor(H4, P.MDT.data(), P.piv4.data());
  P.Mslt.assign(h * h, 0.0);
  for (i64 j = 0; j < h; ++j) for (i64 i = 0; i < h; ++i) P.Mslt[i + h * j] = Msl[j + h * i];
  P.Mslt0 = P.Mslt;
  P.piv1.assign(h, 0);
  lu_factor(h, P.Mslt.data(), P.piv1.data());
}

/* ---------- per-target pieces ---------- */

/* complex root of |X_e(s) - x|^2 near [-1,1]; returns false if Newton fails.
   The start is the root of the chord's |C + s D|^2: s = (-C.D + i |C x D|) / D.D.
   Its imaginary part is taken from the cross product: the equal form
   sqrt(|C|^2 |D|^2 - (C.D)^2) cancels for a target near the chord's line (near a
   vertex), can round to 0 there, and Newton from a real start stays real. */
static bool edge_root(const Edge& E, i64 nq, const double* x, cplx& s0) {
  double A[3], B[3], D[3], C[3];
  for (int c = 0; c < 3; ++c) {
    double v, d;
    leg_eval<double>(&E.coef[c * nq], nq, -1.0, v, d); A[c] = v;
    leg_eval<double>(&E.coef[c * nq], nq, 1.0, v, d); B[c] = v;
    D[c] = 0.5 * (B[c] - A[c]);
    C[c] = 0.5 * (A[c] + B[c]) - x[c];
  }
  double dd = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
  double cd = C[0] * D[0] + C[1] * D[1] + C[2] * D[2];
  double cxd[3];
  cross(C, D, cxd);
  double im = std::sqrt(cxd[0] * cxd[0] + cxd[1] * cxd[1] + cxd[2] * cxd[2]);
  cplx s(-cd / dd, im / dd);
  for (int it = 0; it < 30; ++it) {
    cplx f = 0.0, fp = 0.0;
    for (int c = 0; c < 3; ++c) {
      cplx v, d;
      leg_eval<cplx>(&E.coef[c * nq], nq, s, v, d);
      v -= x[c];
      f += v * v;
      fp += 2.0 * v * d;
    }
    cplx ds = f / fp;
    s -= ds;
    if (std::abs(ds) < 1e-15 * (1.0 + std::abs(s))) { s0 = s; return std::isfinite(s.real()) && std::isfinite(s.imag()); }
  }
  s0 = s;
  return false;
}

/* singularity-swap weights for int f(s)/|X(s)-x| ds on one edge:
   w_i applies to f_i/R_i.  Basis ((s-a)/sg)^k about Re(s0)=a. */
static void ssq_w1(const Panel& P, const Edge& E, const double* x, double* w, double rho_thr) {
  i64 n = P.nq;
  cplx s0;
  bool ok = edge_root(E, n, x, s0);
  if (!ok || bernstein_rho(s0) >= rho_thr) {
    for (i64 i = 0; i < n; ++i) w[i] = P.wgl[i];
    return;
  }
  double a = s0.real(), b = std::fabs(s0.imag());
  double sg = 1.0 + std::fabs(a);
  double x1 = 1.0 - a, x2 = -1.0 - a;
  double r1 = std::sqrt(x1 * x1 + b * b), r2 = std::sqrt(x2 * x2 + b * b);
  std::vector<double> I(n);
  /* I'_k = int_{x2}^{x1} t^k / sqrt(t^2+b^2) dt */
  double I0;
  if (x2 >= 0.0) I0 = std::log((x1 + r1) / (x2 + r2));
  else if (x1 <= 0.0) I0 = std::log((r2 - x2) / (r1 - x1));
  else I0 = std::log((x1 + r1) * (r2 - x2) / (b * b));
  I[0] = I0;
  if (n > 1) I[1] = r1 - r2;
  double p1 = x1, p2 = x2; /* t^{k-1} at ends */
  for (i64 k = 2; k < n; ++k) {
    double dk = (double)k;
    I[k] = (p1 * r1 - p2 * r2 - (dk - 1.0) * b * b * I[k - 2]) / dk;
    p1 *= x1; p2 *= x2;
  }
  /* scale to the ((s-a)/sg)^k basis */
  double sc = 1.0;
  for (i64 k = 0; k < n; ++k) { I[k] *= sc; sc /= sg; }
  /* solve V^T lam = I, V(i,k) = ((s_i-a)/sg)^k */
  std::vector<double> VT(n * n);
  for (i64 i = 0; i < n; ++i) {
    double t = (P.tgl[i] - a) / sg, pk = 1.0;
    for (i64 k = 0; k < n; ++k) { VT[k + n * i] = pk; pk *= t; }
  }
  std::vector<i64> piv(n);
  lu_factor(n, VT.data(), piv.data());
  lu_solve(n, VT.data(), piv.data(), I.data());
  for (i64 i = 0; i < n; ++i) {
    double ds = P.tgl[i] - a;
    w[i] = I[i] * std::sqrt(ds * ds + b * b);
  }
}

/* solid angle int_S z.n_uv/|z|^3 dS via the Dirac-string potential
   A(y) = (q x z)/(|z|(|z| + q.z)), z = y - x, whose string x - s q leaves the
   patch on the target's side.  Each edge is integrated left to right with
   Gauss-Legendre panels that shrink geometrically toward the edge's near point
   (accepted when the complex root lies outside the rho_thr Bernstein ellipse of
   the panel) and grow again past it. */
static double solid_angle(const Panel& P, const double* x, const double* q, double rho_thr) {
  i64 n = P.nq;
  double total = 0.0;
  for (int e = 0; e < 3; ++e) {
    const Edge& E = P.e[e];
    cplx s0;
    bool ok = edge_root(E, n, x, s0);
    double lo = -1.0, len = 2.0;
    int guard = 0;
    while (lo < 1.0 && guard < 512) {
      ++guard;
      double hi = std::min(1.0, lo + len);
      if (ok) {
        while (hi - lo > 1e-14) {
          double cc = 0.5 * (lo + hi), hh = 0.5 * (hi - lo);
          if (bernstein_rho((s0 - cc) / hh) >= rho_thr) break;
          hi = lo + 0.5 * (hi - lo);
        }
      }
      double cc = 0.5 * (lo + hi), hh = 0.5 * (hi - lo);
      for (i64 i = 0; i < n; ++i) {
        double s = cc + hh * P.tgl[i];
        double X[3], dX[3];
        for (int c = 0; c < 3; ++c) leg_eval<double>(&E.coef[c * n], n, s, X[c], dX[c]);
        double z[3] = {X[0] - x[0], X[1] - x[1], X[2] - x[2]};
        double zx[3];
        cross(z, dX, zx);
        double rz = std::sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
        double qz = q[0] * z[0] + q[1] * z[1] + q[2] * z[2];
        double qzx = q[0] * zx[0] + q[1] * zx[1] + q[2] * zx[2];
        total += hh * P.wgl[i] * qzx / (rz * (rz + qz));
      }
      len = 2.0 * (hi - lo);
      lo = hi;
    }
  }
  return total;
}

/* projections that had not met the stop test by their 40th step, and the
   largest such last step */
static std::mutex g_capped_mu;
static i64 g_unconverged = 0;
static double g_capped_step = 0.0;
i64 projection_unconverged(bool reset, double* max_step) {
  std::lock_guard<std::mutex> g(g_capped_mu);
  const i64 n = g_unconverged;
  if (max_step) *max_step = g_capped_step;
  if (reset) { g_unconverged = 0; g_capped_step = 0.0; }
  return n;
}

/* Gauss-Newton projection of x (original coordinates) onto the patch.  It stops
   when the step |du| + |dv| falls below 1e-15 or below 8 eps |x| sqrt((a11 +
   a22) / det), eight times the step that rounding of the residual (eps |x|) alone
   makes (1 / the smallest singular value of the Jacobian is at most
   sqrt((a11 + a22) / det)); a fixed 1e-15 is below that floor on small panels,
   where the iteration would cycle to the step cap.  At the cap (40 steps) the
   last iterate is kept, and counted. */
static void project(const Panel& P, const double* x, const double* uvn, double* ystar,
                    double* nstar, double* dist) {
  i64 h = P.h;
  /* start at the nearest node */
  double best = 1e300, u = 1.0 / 3.0, v = 1.0 / 3.0;
  for (i64 i = 0; i < h; ++i) {
    double y[3];
    geom_eval(P, uvn[2 * i], uvn[2 * i + 1], y, nullptr, nullptr);
    double d = (y[0] - x[0]) * (y[0] - x[0]) + (y[1] - x[1]) * (y[1] - x[1]) + (y[2] - x[2]) * (y[2] - x[2]);
    if (d < best) { best = d; u = uvn[2 * i]; v = uvn[2 * i + 1]; }
  }
  double y[3], yu[3], yv[3], step = 0.0;
  const double xn = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  bool done = false;
  for (int it = 0; it < 40 && !done; ++it) {
    geom_eval(P, u, v, y, yu, yv);
    double r[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
    double a11 = yu[0] * yu[0] + yu[1] * yu[1] + yu[2] * yu[2];
    double a12 = yu[0] * yv[0] + yu[1] * yv[1] + yu[2] * yv[2];
    double a22 = yv[0] * yv[0] + yv[1] * yv[1] + yv[2] * yv[2];
    double b1 = -(yu[0] * r[0] + yu[1] * r[1] + yu[2] * r[2]);
    double b2 = -(yv[0] * r[0] + yv[1] * r[1] + yv[2] * r[2]);
    double det = a11 * a22 - a12 * a12;
    double du = (a22 * b1 - a12 * b2) / det, dv = (a11 * b2 - a12 * b1) / det;
    double un = u + du, vn = v + dv;
    if (un < 0.0) un = 0.0;
    if (vn < 0.0) vn = 0.0;
    if (un + vn > 1.0) { double ex = 0.5 * (un + vn - 1.0); un -= ex; vn -= ex; if (un < 0) { vn += un; un = 0; } if (vn < 0) { un += vn; vn = 0; } }
    step = std::fabs(un - u) + std::fabs(vn - v);
    u = un; v = vn;
    const double floor8 = 8.0 * 2.220446049250313e-16 * xn * std::sqrt((a11 + a22) / det);
    done = step < std::max(1e-15, floor8);
  }
  if (!done) {
    std::lock_guard<std::mutex> g(g_capped_mu);
    ++g_unconverged;
    g_capped_step = std::max(g_capped_step, step);
  }
  geom_eval(P, u, v, y, yu, yv);
  double nn[3];
  cross(yu, yv, nn);
  double l = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
  for (int c = 0; c < 3; ++c) { nstar[c] = nn[c] / l; ystar[c] = y[c]; }
  *dist = (x[0] - y[0]) * nstar[0] + (x[1] - y[1]) * nstar[1] + (x[2] - y[2]) * nstar[2];
}

struct Moments {
  std::vector<double> W;  /* h  : W~(F_k) = int (dF/dn - F d/dn)(1/|z|) dS, snx orientation */
  std::vector<double> Q;  /* 4h : Q[grad F_k] components, Q[4*k + c] */
  double omega = 0.0;     /* int z.snx/|z|^3 dS */
};

/* all moments for one frame target X (x original for the projection) */
static void target_moments(const Panel& P, const double* X, const double* x, const double* uvn,
                           bool exterior, double side_tol, Moments& Mo, double* tinfo) {
  i64 p = P.p, h = P.h, nq = P.nq, nc = P.ncoef, nb = 3 * nq;
  double rho_thr = std::pow(8.0, 8.0 / (double)nq);
  double t0 = now();
  std::vector<double> w1(nb);
  for (int e = 0; e < 3; ++e) ssq_w1(P, P.e[e], X, &w1[e * nq], rho_thr);
  double t1 = now();
  tinfo[3] += t1 - t0;
  /* side and string direction */
  double ys[3], ns[3], d;
  project(P, x, uvn, ys, ns, &d);
  double side;
  if (std::fabs(d) > side_tol * P.L) side = d > 0.0 ? 1.0 : -1.0;
  else side = (exterior ? EXTERIOR_SNX_SIDE : -EXTERIOR_SNX_SIDE) * P.sigma;
  double qo[3] = {-side * ns[0], -side * ns[1], -side * ns[2]}, q[3];
  rot(P, qo, q);
  double om_uv = solid_angle(P, X, q, rho_thr);
  double t2 = now();
  tinfo[4] += t2 - t1;
  /* line moments chi (1 complex) and alpha (4 complex) per coefficient, Lambda */
  std::vector<cplx> chi(nc, 0.0), al(4 * nc, 0.0);
  double lam[3] = {0, 0, 0};
  for (int e = 0; e < 3; ++e)
    for (i64 i = 0; i < nq; ++i) {
      i64 ib = e * nq + i;
      const double* Y = &P.e[e].X[3 * i];
      const double* dY = &P.e[e].dX[3 * i];
      double z[3] = {Y[0] - X[0], Y[1] - X[1], Y[2] - X[2]};
      double R = std::sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
      double wr = w1[ib] / R;
      double zx[3];
      cross(z, dY, zx);
      for (int c = 0; c < 3; ++c) lam[c] += wr * dY[c];
      for (i64 k = 0; k < nc; ++k) {
        const cplx* g = &P.gradH[(ib * nc + k) * 3];
        const cplx* V = &P.Vd[(ib * nc + k) * 3];
        chi[k] += wr * (zx[0] * g[0] + zx[1] * g[1] + zx[2] * g[2]);
        /* z (x) V = (-z.V, z x V) */
        al[4 * k + 0] += wr * (-(z[0] * V[0] + z[1] * V[1] + z[2] * V[2]));
        al[4 * k + 1] += wr * (z[1] * V[2] - z[2] * V[1]);
        al[4 * k + 2] += wr * (z[2] * V[0] - z[0] * V[2]);
        al[4 * k + 3] += wr * (z[0] * V[1] - z[1] * V[0]);
      }
    }
  double t3 = now();
  tinfo[5] += t3 - t2;
  /* recursion */
  std::vector<cplx> Wt(nc, 0.0), Qt(4 * nc, 0.0);
  auto wget = [&](i64 n2, i64 m) -> cplx {
    if (n2 < 0 || m > n2 || m < -n2) return 0.0;
    if (m >= 0) return Wt[hidx(n2, m)];
    cplx v = std::conj(Wt[hidx(n2, -m)]);
    return ((-m) % 2) ? -v : v;
  };
  auto qget = [&](i64 n2, i64 m, int c) -> cplx {
    if (n2 < 0 || m > n2 || m < -n2) return 0.0;
    if (m >= 0) return Qt[4 * hidx(n2, m) + c];
    cplx v = std::conj(Qt[4 * hidx(n2, -m) + c]);
    return ((-m) % 2) ? -v : v;
  };
  double xZ = X[0];
  cplx w(X[1], X[2]);
  Wt[0] = om_uv;
  /* Lambda = -Omega + (int dy/|z|) ; Q_{1,m} = Lambda (x) grad H'_1^m */
  for (i64 m = 0; m <= 1; ++m) {
    cplx g[3];
    if (m == 0) { g[0] = 1.0; g[1] = 0.0; g[2] = 0.0; }
    else { g[0] = 0.0; g[1] = 0.5 * IU; g[2] = -0.5; }
    double L0 = -om_uv;
    /* (L0 + l)(g) = -l.g + L0 g + l x g */
    cplx* Qm = &Qt[4 * hidx(1, m)];
    Qm[0] = -(lam[0] * g[0] + lam[1] * g[1] + lam[2] * g[2]);
    Qm[1] = L0 * g[0] + (lam[1] * g[2] - lam[2] * g[1]);
    Qm[2] = L0 * g[1] + (lam[2] * g[0] - lam[0] * g[2]);
    Qm[3] = L0 * g[2] + (lam[0] * g[1] - lam[1] * g[0]);
  }
  for (i64 n2 = 1; n2 <= p; ++n2)
    for (i64 m = 0; m <= n2; ++m) {
      double dn = (double)n2;
      Wt[hidx(n2, m)] = chi[hidx(n2, m)] / dn + xZ * wget(n2 - 1, m) +
                        0.5 * IU * (w * wget(n2 - 1, m - 1) + std::conj(w) * wget(n2 - 1, m + 1));
      if (n2 >= 2)
        for (int c = 0; c < 4; ++c)
          Qt[4 * hidx(n2, m) + c] =
              (al[4 * hidx(n2, m) + c] +
               dn * (xZ * qget(n2 - 1, m, c) + 0.5 * IU * (w * qget(n2 - 1, m - 1, c) + std::conj(w) * qget(n2 - 1, m + 1, c)))) /
              (dn - 1.0);
    }
  Mo.W.assign(h, 0.0);
  Mo.Q.assign(4 * h, 0.0);
  i64 k = 0;
  for (i64 n2 = 1; n2 <= p; ++n2)
    for (i64 m = 1; m <= n2; ++m, ++k) {
      cplx rotm = std::pow(-IU, (double)m);
      Mo.W[k] = P.sigma * std::imag(rotm * Wt[hidx(n2, m)]);
      for (int c = 0; c < 4; ++c) Mo.Q[4 * k + c] = P.sigma * std::imag(rotm * Qt[4 * hidx(n2, m) + c]);
    }
  Mo.omega = P.sigma * om_uv;
  tinfo[6] += now() - t3;
}

static double g_side_tol = 1e300; /* reference semantics: every target takes the flag side */

void rrq(i64 m, const double* tx, i64 n, const double* sx, const double* snx,
         const double* sw, const double* rts, const double* rps, i64 order, i64 nquad,
         i64 orderff, double distff, bool exterior, i64 ichart, const double* sxbd_chart,
         const double* rv_chart, const double* tgl, const double* wgl, const double* Dgl,
         const double* w_bclag, const double* Legmat, const double* umatr,
         const double* vmatr, double* As, double* Ad, double* Omega, double* Ias,
         double* timeinfo, i64* label) {
  (void)w_bclag; (void)vmatr;
  i64 h = n, korder = order - 1;
  for (i64 i = 0; i < m * n; ++i) { As[i] = 0.0; Ad[i] = 0.0; }
  for (i64 i = 0; i < 4 * h * m; ++i) Omega[i] = 0.0;
  for (i64 i = 0; i < m; ++i) Ias[i] = 0.0;
  double ta = now();
  PanelLevels lv;
  build_patch_levels(korder, n, sx, rts, rps, orderff, distff, 4, lv);
  const PatchRef& ref = patch_ref(korder, n, orderff, 4);
  double tb = now();
  timeinfo[0] += tb - ta;
  Panel P;
  setup_panel(P, n, sx, snx, sw, order, nquad, ichart, sxbd_chart, rv_chart, tgl, wgl, Dgl, Legmat, umatr);
  double tc = now();
  timeinfo[1] += tc - tb;
  std::vector<i64> idxs(m);
  i64 ms = 0;
  lap3dsdlpmat_levels(m, tx, n, lv, ref, As, Ad, idxs.data(), &ms, label);
  timeinfo[2] += now() - tc;
  if (ms == 0) return;
  std::vector<double> uvn(2 * h);
  vr_nodes(order - 1, uvn.data(), nullptr);
  double fac = 1.0 / (4.0 * PI);
  i64 H4 = 4 * h;
  std::vector<double> rhs(H4), rs(h);
  Moments Mo;
  for (i64 j = 0; j < ms; ++j) {
    i64 t = idxs[j] - 1;
    double X[3];
    to_frame(P, tx + 3 * t, X);
    target_moments(P, X, tx + 3 * t, uvn.data(), exterior, g_side_tol, Mo, timeinfo);
    double t5 = now();
    for (i64 k = 0; k < h; ++k) {
      rhs[k] = Mo.Q[4 * k + 0];
      rhs[h + k] = -Mo.Q[4 * k + 1];
      rhs[2 * h + k] = -Mo.Q[4 * k + 2];
      rhs[3 * h + k] = -Mo.Q[4 * k + 3];
    }
    for (i64 k = 0; k < H4; ++k) Omega[k + H4 * t] = rhs[k] * fac / P.alpha;
    lu_solve(H4, P.MDT.data(), P.piv4.data(), rhs.data());
    for (i64 i = 0; i < h; ++i) Ad[t + m * i] = rhs[i] * fac;
    for (i64 k = 0; k < h; ++k) {
      double s = Mo.W[k] * fac;
      for (i64 i = 0; i < h; ++i) s += Ad[t + m * i] * P.Fnode[i + h * k];
      rs[k] = s;
    }
    lu_solve(h, P.Mslt.data(), P.piv1.data(), rs.data());
    for (i64 i = 0; i < h; ++i) As[t + m * i] = rs[i] / P.alpha;
    Ias[t] = Mo.omega;
    timeinfo[7] += now() - t5;
  }
}

void set_side_tol(double tol) { g_side_tol = tol; }

}  // namespace orc

namespace orc {
/* Debug / test hook: the moments W~(F_k), Q[grad F_k] (snx orientation) of one
   target from the line-integral recursion, and the same moments by brute-force
   surface quadrature (nsub^2 sub-triangles x VR order qord) for comparison. */
void debug_moments(const double* x, i64 n, const double* sx, const double* snx, const double* sw, i64 order,
                   i64 nquad, bool exterior, i64 nsub, i64 qord, double* W, double* Q, double* Wb, double* Qb,
                   double* om) {
  std::vector<double> tgl(nquad), wgl(nquad), Dgl(nquad * nquad), wb(nquad), Leg(nquad * nquad), um(n * n), vm(n * n);
  simplex_precomp(nquad, order - 1, n, tgl.data(), wgl.data(), Dgl.data(), wb.data(), Leg.data(), um.data(), vm.data());
  Panel P;
  std::vector<double> sxbd(9 * nquad, 0.0), rv(9, 0.0);
  setup_panel(P, n, sx, snx, sw, order, nquad, 0, sxbd.data(), rv.data(), tgl.data(), wgl.data(), Dgl.data(), Leg.data(), um.data());
  std::vector<double> uvn(2 * n);
  vr_nodes(order - 1, uvn.data(), nullptr);
  double X[3], ti[20] = {0};
  to_frame(P, x, X);
  Moments Mo;
  target_moments(P, X, x, uvn.data(), exterior, 1e-10, Mo, ti);
  for (i64 k = 0; k < n; ++k) { W[k] = Mo.W[k]; for (int c = 0; c < 4; ++c) Q[4 * k + c] = Mo.Q[4 * k + c]; }
  om[0] = Mo.omega;
  /* brute force in the frame */
  i64 p = order, h = n;
  i64 nqn = vr_count(qord);
  std::vector<double> quv(2 * nqn), qw(nqn);
  vr_nodes(qord, quv.data(), qw.data());
  for (i64 k = 0; k < h; ++k) { Wb[k] = 0; for (int c = 0; c < 4; ++c) Qb[4 * k + c] = 0; }
  om[1] = 0.0;
  double hs = 1.0 / (double)nsub;
  std::vector<cplx> H;
  for (i64 a = 0; a < nsub; ++a)
    for (i64 b = 0; b < nsub - a; ++b)
      for (int up = 0; up < 2; ++up) {
        if (up == 1 && a + b >= nsub - 1) continue;
        for (i64 q = 0; q < nqn; ++q) {
          double s = quv[2 * q], r = quv[2 * q + 1], u, v;
          if (up == 0) { u = (a + s) * hs; v = (b + r) * hs; }
          else { u = (a + 1 - s) * hs; v = (b + 1 - r) * hs; }
          double y[3], yu[3], yv[3], nn[3];
          geom_eval(P, u, v, y, yu, yv);
          cross(yu, yv, nn);
          double J = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
          double nuv[3] = {nn[0] / J, nn[1] / J, nn[2] / J};
          double Y[3], N[3];
          to_frame(P, y, Y);
          rot(P, nuv, N);
          for (int c = 0; c < 3; ++c) N[c] *= P.sigma; /* snx orientation */
          double w = qw[q] * hs * hs * J * P.alpha * P.alpha;
          double z[3] = {Y[0] - X[0], Y[1] - X[1], Y[2] - X[2]};
          double rr = z[0] * z[0] + z[1] * z[1] + z[2] * z[2], r1 = std::sqrt(rr), r3 = r1 * rr;
          double zn = z[0] * N[0] + z[1] * N[1] + z[2] * N[2];
          om[1] += w * zn / r3;
          harm_vals(Y, p, H);
          i64 k = 0;
          for (i64 n2 = 1; n2 <= p; ++n2)
            for (i64 m = 1; m <= n2; ++m, ++k) {
              cplx rotm = std::pow(-IU, (double)m);
              cplx g[3];
              harm_grad(H, n2, m, g);
              double F = std::imag(rotm * H[hidx(n2, m)]);
              double G[3] = {std::imag(rotm * g[0]), std::imag(rotm * g[1]), std::imag(rotm * g[2])};
              double dFn = G[0] * N[0] + G[1] * N[1] + G[2] * N[2];
              /* W = int dF/dn /|z| - F d/dn(1/|z|), d/dn_y (1/|z|) = -z.n/|z|^3 */
              Wb[k] += w * (dFn / r1 + F * zn / r3);
              /* Q = int E n gradF, E = z/|z|^3: (z (x) N) (x) G */
              double e0 = -zn / r3, e[3] = {(z[1] * N[2] - z[2] * N[1]) / r3, (z[2] * N[0] - z[0] * N[2]) / r3,
                                            (z[0] * N[1] - z[1] * N[0]) / r3};
              /* (e0 + e)(G) = -e.G + e0 G + e x G */
              Qb[4 * k + 0] += w * (-(e[0] * G[0] + e[1] * G[1] + e[2] * G[2]));
              Qb[4 * k + 1] += w * (e0 * G[0] + e[1] * G[2] - e[2] * G[1]);
              Qb[4 * k + 2] += w * (e0 * G[1] + e[2] * G[0] - e[0] * G[2]);
              Qb[4 * k + 3] += w * (e0 * G[2] + e[0] * G[1] - e[1] * G[0]);
            }
        }
      }
}
}  // namespace orc

extern "C" void orc_debug_moments(const double* x, i64 n, const double* sx, const double* snx, const double* sw,
                                  i64 order, i64 nquad, i64 exterior, i64 nsub, i64 qord, double* W, double* Q,
                                  double* Wb, double* Qb, double* om) {
  orc::debug_moments(x, n, sx, snx, sw, order, nquad, exterior != 0, nsub, qord, W, Q, Wb, Qb, om);
}

extern "C" void orc_debug_panel(i64 n, const double* sx, const double* snx, const double* sw, i64 order, i64 nquad,
                                double* Xb, double* dXb, double* Rc) {
  using namespace orc;
  std::vector<double> tgl(nquad), wgl(nquad), Dgl(nquad * nquad), wb(nquad), Leg(nquad * nquad), um(n * n), vm(n * n);
  simplex_precomp(nquad, order - 1, n, tgl.data(), wgl.data(), Dgl.data(), wb.data(), Leg.data(), um.data(), vm.data());
  Panel P;
  std::vector<double> sxbd(9 * nquad, 0.0), rv(9, 0.0);
  setup_panel(P, n, sx, snx, sw, order, nquad, 0, sxbd.data(), rv.data(), tgl.data(), wgl.data(), Dgl.data(), Leg.data(), um.data());
  for (int e = 0; e < 3; ++e)
    for (i64 i = 0; i < 3 * nquad; ++i) { Xb[e * 3 * nquad + i] = P.e[e].X[i]; dXb[e * 3 * nquad + i] = P.e[e].dX[i]; }
  for (int i = 0; i < 9; ++i) Rc[i] = P.R[i];
  for (int i = 0; i < 3; ++i) Rc[9 + i] = P.c[i];
  Rc[12] = P.alpha; Rc[13] = P.sigma;
}

extern "C" void orc_debug_ssq(const double* coef, i64 nq, const double* x, double* w, double* s0out) {
  using namespace orc;
  Panel P;
  P.nq = nq;
  P.tgl.resize(nq); P.wgl.resize(nq);
  gauss(nq, P.tgl.data(), P.wgl.data(), nullptr);
  Edge E;
  E.coef.assign(coef, coef + 3 * nq);
  cplx s0;
  bool ok = edge_root(E, nq, x, s0);
  s0out[0] = s0.real(); s0out[1] = s0.imag(); s0out[2] = ok; s0out[3] = bernstein_rho(s0);
  ssq_w1(P, E, x, w, 1e300);
}

extern "C" void orc_debug_system(i64 n, const double* sx, const double* snx, const double* sw, i64 order, i64 nquad,
                                 double* MT, double* MslT) {
  using namespace orc;
  std::vector<double> tgl(nquad), wgl(nquad), Dgl(nquad * nquad), wb(nquad), Leg(nquad * nquad), um(n * n), vm(n * n);
  simplex_precomp(nquad, order - 1, n, tgl.data(), wgl.data(), Dgl.data(), wb.data(), Leg.data(), um.data(), vm.data());
  Panel P;
  std::vector<double> sxbd(9 * nquad, 0.0), rv(9, 0.0);
  setup_panel(P, n, sx, snx, sw, order, nquad, 0, sxbd.data(), rv.data(), tgl.data(), wgl.data(), Dgl.data(), Leg.data(), um.data());
  for (size_t i = 0; i < P.MDT0.size(); ++i) MT[i] = P.MDT0[i];
  for (size_t i = 0; i < P.Mslt0.size(); ++i) MslT[i] = P.Mslt0[i];
}
