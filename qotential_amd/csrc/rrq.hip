// Recursive-reduction quadrature (RRQ) for the targets no refinement level can
// serve: per-panel set-up, the quaternion system and its weights operator, and
// the per-target boundary moments, recursion and weights.
//
// Reference: rrq_r64                       lap3d_close_mod.f90:622-881
//            build_closepanel_precomp_r64  lap3d_close_mod.f90:472-620
//            build_tc_chialpha_r64         lap3d_close_mod.f90:336-470
//            (SSQ weights, solid angle, omega moments: absent LQ/QA modules,
//             re-derived -- see the oracle header for the identities used)
// FP64 throughout.
#include <cstdlib>
#include <stdexcept>

#include "qol_kernels.h"

namespace qol {

// ---------------------------------------------------------------------------
// geometry polynomial x(u,v) = sum_k K_k(u,v) cx[c][k] and its uv derivatives,
// accumulated while the Koornwinder basis is generated (no h-sized arrays).
// The recurrence constants come from the per-order table kt (koorn_eval_table:
// no divisions or square roots here): kt[2am], kt[2am+1] scale the Legendre
// step to am, and term t (loop order am, bd) has its normalisation and Jacobi
// coefficients at kt[2N + 4t .. 2N + 4t + 3].
__device__ void geom_eval(const double* cx, const double* __restrict__ kt, int h, int korder, double u, double v,
                          double* x, double* xu, double* xv) {
  const int N = korder + 1;
  const double a = 2.0 * u + v - 1.0, b = 1.0 - v, bb = b * b;
  for (int c = 0; c < 3; ++c) { x[c] = 0.0; if (xu) { xu[c] = 0.0; xv[c] = 0.0; } }
  const double xx = 2.0 * v - 1.0;
  // Q_am(a, b) (scaled Legendre in a/b) and its u, v derivatives advance with am;
  // for each am, Jacobi P_bd^(2am+1,0)(xx) by its three-term recurrence in bd.
  // Basis index of (degree d = am + bd, bd) is d(d+1)/2 + bd.
  double Q = 1.0, Qu = 0.0, Qv = 0.0, Qm = 0.0, Qum = 0.0, Qvm = 0.0;
  const double* tt = kt + 2 * N;
  for (int am = 0; am < N; ++am) {
    if (am == 1) {
      Qm = Q; Qum = Qu; Qvm = Qv;
      Q = a; Qu = 2.0; Qv = 1.0;
    } else if (am >= 2) {
      const double e1 = kt[2 * am], e2 = kt[2 * am + 1];
      const double Qn = e1 * a * Q - e2 * bb * Qm;
      const double Qun = e1 * (2.0 * Q + a * Qu) - e2 * bb * Qum;
      const double Qvn = e1 * (Q + a * Qv) - e2 * (bb * Qvm - 2.0 * b * Qm);
      Qm = Q; Qum = Qu; Qvm = Qv;
      Q = Qn; Qu = Qun; Qv = Qvn;
    }
    const double al = 2.0 * am + 1.0;
    double p = 1.0, dp = 0.0, pm = 1.0, dpm = 0.0;
    for (int bd = 0; am + bd < N; ++bd, tt += 4) {
      if (bd == 1) {
        p = 0.5 * ((al + 2.0) * xx + al);
        dp = al + 2.0;
      } else if (bd >= 2) {
        const double w = tt[1] + tt[2] * xx;
        const double pn = w * p - tt[3] * pm;
        const double dpn = w * dp + 2.0 * tt[2] * p - tt[3] * dpm;
        pm = p; p = pn; dpm = dp; dp = dpn;
      }
      const int d = am + bd, idx = d * (d + 1) / 2 + bd;
      const double sc = tt[0];
      const double f = sc * Q * p;
      for (int c = 0; c < 3; ++c) x[c] += f * cx[c * h + idx];
      if (xu) {
        const double fu = sc * Qu * p, fv = sc * (Qv * p + Q * dp);
        for (int c = 0; c < 3; ++c) { xu[c] += fu * cx[c * h + idx]; xv[c] += fv * cx[c * h + idx]; }
      }
    }
  }
}

std::vector<double> koorn_eval_table(int N) {
  std::vector<double> t(2 * (size_t)N, 0.0);
  for (int am = 2; am < N; ++am) {
    const int m = am - 1;
    const double inv = 1.0 / (double)(m + 1);
    t[2 * am] = (2.0 * m + 1.0) * inv;
    t[2 * am + 1] = (double)m * inv;
  }
  for (int am = 0; am < N; ++am) {
    const double al = 2.0 * am + 1.0;
    for (int bd = 0; am + bd < N; ++bd) {
      const int d = am + bd;
      double ca = 0.0, cb = 0.0, cc = 0.0;
      if (bd >= 2) {
        const double dk = (double)(bd - 1);
        const double a1 = 2.0 * (dk + 1.0) * (dk + al + 1.0) * (2.0 * dk + al);
        const double a2 = (2.0 * dk + al + 1.0) * al * al;
        const double a3 = (2.0 * dk + al) * (2.0 * dk + al + 1.0) * (2.0 * dk + al + 2.0);
        const double a4 = 2.0 * (dk + al) * dk * (2.0 * dk + al + 2.0);
        ca = a2 / a1; cb = a3 / a1; cc = a4 / a1;
      }
      t.insert(t.end(), {std::sqrt((2.0 * am + 1.0) * 2.0 * (double)(d + 1)), ca, cb, cc});
    }
  }
  return t;
}

__device__ __forceinline__ void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// frame[] = R (rows e1,e2,e3), c, alpha, sigma, L
__device__ __forceinline__ void to_frame(const double* fr, const double* x, double* X) {
  const double d0 = x[0] - fr[9], d1 = x[1] - fr[10], d2 = x[2] - fr[11];
  for (int r = 0; r < 3; ++r) X[r] = fr[12] * (fr[3 * r] * d0 + fr[3 * r + 1] * d1 + fr[3 * r + 2] * d2);
}
__device__ __forceinline__ void rotate(const double* fr, const double* v, double* V) {
  for (int r = 0; r < 3; ++r) V[r] = fr[3 * r] * v[0] + fr[3 * r + 1] * v[1] + fr[3 * r + 2] * v[2];
}

// harmonic values H'_n^m (m=0..n, n=0..p) at frame point X into H (ncoef complex)
__device__ void harm_vals(const double* X, int p, cd* H) {
  const cd W = cmk(X[1], X[2]);
  const double Z = X[0];
  H[0] = cmk(1.0, 0.0);
  for (int n = 1; n <= p; ++n)
    for (int m = 0; m <= n; ++m) {
      const cd a = cscale(hget(H, n - 1, m), Z);
      const cd b = cadd(cmul(W, hget(H, n - 1, m - 1)), cmul(cconj(W), hget(H, n - 1, m + 1)));
      H[hidx(n, m)] = cadd(a, cscale(cmuli(b), 0.5));
    }
}

// gradient of H'_n^m (frame x,y,z) from degree n-1 values
__device__ __forceinline__ void harm_grad(const cd* H, int n, int m, cd* g) {
  const double dn = (double)n;
  const cd lo = hget(H, n - 1, m - 1), hi = hget(H, n - 1, m + 1);
  g[0] = cscale(hget(H, n - 1, m), dn);
  g[1] = cscale(cmuli(cadd(lo, hi)), 0.5 * dn);
  g[2] = cscale(csub(hi, lo), 0.5 * dn);
}

// Im((-i)^m z)
__device__ __forceinline__ double imrot(int m, cd z) {
  switch (m & 3) {
    case 0: return z.im;
    case 1: return -z.re;
    case 2: return -z.im;
    default: return z.re;
  }
}

// ---------------------------------------------------------------------------
size_t rrq_frame_smem(int h, int nq) {
  const int nb = 3 * nq;
  return (size_t)(7 * h + 3 * h + 16 + 9 * nq + 6 * nb + h) * sizeof(double);
}
size_t rrq_system_smem(int h, int nq, int ncoef) {
  const int nb = 3 * nq;
  size_t a = (size_t)(6 * h + 16 + 6 * nb) * sizeof(double);
  size_t b = (size_t)nb * ncoef * sizeof(cd);
  size_t c = (size_t)64 * ncoef * sizeof(cd);
  return a + (b > c ? b : c);
}

// One workgroup per panel (with RRQ pairs, or every panel when nrrq is null):
// Koornwinder coefficients, orientation, frame, boundary nodes and the edges'
// Legendre coefficients -- what the per-pair projection and edge kernels need.
__global__ void __launch_bounds__(256) rrq_panel_frame_kernel(RRQTables T, RRQPanels P, int npan,
                                                              const int* __restrict__ nrrq,
                                                              const double* __restrict__ sx,
                                                              const double* __restrict__ snx,
                                                              const double* __restrict__ sw, int ichart,
                                                              const double* __restrict__ sxbd_chart,
                                                              const double* __restrict__ rv_chart) {
  const int k = blockIdx.x;
  if (nrrq && nrrq[k] == 0) return;  // null: every panel
  extern __shared__ double sm[];
  const int h = T.h, nq = T.nq, p = T.p, nb = 3 * nq, ncoef = T.ncoef, H4 = 4 * h;
  double* ssx = sm;             // [3][h]
  double* ssn = ssx + 3 * h;    // [3][h]
  double* ssw = ssn + 3 * h;    // [h]
  double* scx = ssw + h;        // [3][h]
  double* fr = scx + 3 * h;     // 16
  double* ecf = fr + 16;        // [3 edges][3][nq]
  double* bx = ecf + 9 * nq;    // [nb][6]
  double* sg = bx + 6 * nb;     // [h] sigma contributions
  cd* Hs = (cd*)(sg + h);       // [nb or 64][ncoef]
  const size_t gk = (size_t)k * h;
  for (int i = threadIdx.x; i < h; i += blockDim.x) {
    for (int c = 0; c < 3; ++c) {
      ssx[c * h + i] = sx[3 * (gk + i) + c];
      ssn[c * h + i] = snx[3 * (gk + i) + c];
    }
    ssw[i] = sw[gk + i];
  }
  __syncthreads();
  // Koornwinder coefficients cx = umatr * sx
  for (int i = threadIdx.x; i < 3 * h; i += blockDim.x) {
    const int c = i / h, a = i % h;
    double s = 0.0;
    for (int q = 0; q < h; ++q) s += T.umatr[a + (size_t)h * q] * ssx[c * h + q];
    scx[c * h + a] = s;
    P.cx[(size_t)k * 3 * h + c * h + a] = s;
  }
  __syncthreads();
  // orientation: sigma = sign(sum (x_u x x_v) . snx) over the nodes
  for (int i = threadIdx.x; i < h; i += blockDim.x) {
    double x[3], xu[3], xv[3], n[3];
    geom_eval(scx, T.ktab, h, p - 1, T.uvn[2 * i], T.uvn[2 * i + 1], x, xu, xv);
    cross3(xu, xv, n);
    sg[i] = n[0] * ssn[i] + n[1] * ssn[h + i] + n[2] * ssn[2 * h + i];
  }
  // triangle vertices, one thread each (the patch at (0,0), (1,0), (0,1), or the chart's)
  __shared__ double srv[9];
  if (threadIdx.x < 3) {
    const int j = threadIdx.x;
    if (ichart == 1) {
      for (int c = 0; c < 3; ++c) srv[3 * j + c] = rv_chart[9 * (size_t)k + 3 * j + c];
    } else {
      double xv[3];
      geom_eval(scx, T.ktab, h, p - 1, j == 1 ? 1.0 : 0.0, j == 2 ? 1.0 : 0.0, xv, nullptr, nullptr);
      for (int c = 0; c < 3; ++c) srv[3 * j + c] = xv[c];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double rv[9];
    for (int i = 0; i < 9; ++i) rv[i] = srv[i];
    double swsum = 0.0, acc = 0.0, mn[3] = {0, 0, 0};
    for (int i = 0; i < h; ++i) {
      swsum += ssw[i];
      acc += sg[i];
      for (int c = 0; c < 3; ++c) mn[c] += ssn[c * h + i];
    }
    const double alpha = 1.25 / sqrt(2.0 * swsum);
    double a[3], b[3], nrm[3], axb[3], t1[3], t2[3];
    for (int c = 0; c < 3; ++c) { a[c] = rv[3 + c] - rv[c]; b[c] = rv[6 + c] - rv[c]; }
    cross3(a, b, nrm);
    const double nn = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    double e3[3] = {nrm[0] / nn, nrm[1] / nn, nrm[2] / nn};
    if (mn[0] * e3[0] + mn[1] * e3[1] + mn[2] * e3[2] < 0.0) for (int c = 0; c < 3; ++c) e3[c] = -e3[c];
    const double la = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    double e1[3] = {a[0] / la, a[1] / la, a[2] / la}, e2[3];
    cross3(e3, e1, e2);
    cross3(a, b, axb);
    const double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2], bbv = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    cross3(axb, a, t1);
    cross3(b, axb, t2);
    const double den = 2.0 * (axb[0] * axb[0] + axb[1] * axb[1] + axb[2] * axb[2]);
    double c0[3];
    for (int c = 0; c < 3; ++c) c0[c] = rv[c] + (bbv * t1[c] + aa * t2[c]) / den;
    for (int c = 0; c < 3; ++c) { fr[c] = e1[c]; fr[3 + c] = e2[c]; fr[6 + c] = e3[c]; fr[9 + c] = c0[c]; }
    fr[12] = alpha;
    double zmin = 1e300;
    for (int i = 0; i < h; ++i) {
      double x[3] = {ssx[i], ssx[h + i], ssx[2 * h + i]}, X[3];
      to_frame(fr, x, X);
      zmin = fmin(zmin, X[2]);
    }
    const double s3 = fmax(0.0, -zmin);
    for (int c = 0; c < 3; ++c) fr[9 + c] = c0[c] - (1.25 * s3 / alpha) * e3[c];
    fr[13] = acc >= 0.0 ? 1.0 : -1.0;
    double L = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = i + 1; j < 3; ++j) {
        double d = 0.0;
        for (int c = 0; c < 3; ++c) d += (rv[3 * i + c] - rv[3 * j + c]) * (rv[3 * i + c] - rv[3 * j + c]);
        L = fmax(L, sqrt(d));
      }
    fr[14] = L;
    fr[15] = 0.0;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 16; i += blockDim.x) P.frame[(size_t)k * 16 + i] = fr[i];
  // boundary positions: edges (0,1)->(0,0), (0,0)->(1,0), (1,0)->(0,1)
  for (int ib = threadIdx.x; ib < nb; ib += blockDim.x) {
    const int e = ib / nq, i = ib % nq;
    double x[3];
    if (ichart == 1) {
      for (int c = 0; c < 3; ++c) x[c] = sxbd_chart[(size_t)k * 9 * nq + 3 * ib + c];
    } else {
      const double EA[3][2] = {{0, 1}, {0, 0}, {1, 0}}, EB[3][2] = {{0, 0}, {1, 0}, {0, 1}};
      const double s01 = 0.5 * (T.tgl[i] + 1.0);
      const double u = EA[e][0] + (EB[e][0] - EA[e][0]) * s01, v = EA[e][1] + (EB[e][1] - EA[e][1]) * s01;
      geom_eval(scx, T.ktab, h, p - 1, u, v, x, nullptr, nullptr);
    }
    to_frame(fr, x, bx + 6 * ib);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 9 * nq; i += blockDim.x) {
    const int e = i / (3 * nq), c = (i / nq) % 3, kk = i % nq;
    double s = 0.0;
    for (int q = 0; q < nq; ++q) s += T.Legmat[kk + (size_t)nq * q] * bx[6 * (e * nq + q) + c];
    ecf[i] = s;
  }
  __syncthreads();
  // the edges in the power basis for the per-pair kernels (Horner's rule costs
  // about half the Legendre recurrence per point): coefficient j of edge
  // component ec is sum_k ecf_k [s^j] P_k, over k = j, j+2, ...
  for (int i = threadIdx.x; i < 9 * nq; i += blockDim.x) {
    const int ec = i / nq, j = i % nq;
    double m = 0.0;
    for (int q = j; q < nq; q += 2) m += ecf[ec * nq + q] * T.leg2mon[(size_t)q * nq + j];
    P.edge[(size_t)k * 9 * nq + i] = m;
  }
  for (int ib = threadIdx.x; ib < nb; ib += blockDim.x) {
    const int e = ib / nq, i = ib % nq;
    for (int c = 0; c < 3; ++c) {
      double v, d;
      leg_eval_r(ecf + (e * 3 + c) * nq, nq, T.tgl[i], v, d);
      bx[6 * ib + 3 + c] = d;
    }
    for (int c = 0; c < 6; ++c) P.bX[((size_t)k * nb + ib) * 6 + c] = bx[6 * ib + c];
  }
  __syncthreads();
}

// One workgroup per panel: harmonic values at the boundary nodes, and the quaternion
// matrix, Msl^T and Fn at the panel nodes, from the frame and boundary nodes
// written by rrq_panel_frame_kernel.
__global__ void __launch_bounds__(256) rrq_panel_system_kernel(RRQTables T, RRQPanels P, int npan,
                                                               const int* __restrict__ nrrq,
                                                               const double* __restrict__ sx,
                                                               const double* __restrict__ snx) {
  const int k = blockIdx.x;
  if (nrrq && nrrq[k] == 0) return;  // null: every panel
  extern __shared__ double sm[];
  const int h = T.h, nq = T.nq, p = T.p, nb = 3 * nq, ncoef = T.ncoef;
  double* ssx = sm;             // [3][h]
  double* ssn = ssx + 3 * h;    // [3][h]
  double* fr = ssn + 3 * h;     // 16
  double* bx = fr + 16;         // [nb][6]
  cd* Hs = (cd*)(bx + 6 * nb);  // [nb or 64][ncoef]
  const size_t gk = (size_t)k * h;
  for (int i = threadIdx.x; i < h; i += blockDim.x)
    for (int c = 0; c < 3; ++c) {
      ssx[c * h + i] = sx[3 * (gk + i) + c];
      ssn[c * h + i] = snx[3 * (gk + i) + c];
    }
  for (int i = threadIdx.x; i < 16; i += blockDim.x) fr[i] = P.frame[(size_t)k * 16 + i];
  for (int i = threadIdx.x; i < 6 * nb; i += blockDim.x) bx[i] = P.bX[(size_t)k * nb * 6 + i];
  __syncthreads();
  // harmonic values H'_n^m of degree n < p at the boundary nodes: the moments
  // of every pair are contracted against these (their gradients and V are
  // fixed combinations of them, applied after the contraction)
  for (int ib = threadIdx.x; ib < nb; ib += blockDim.x) harm_vals(bx + 6 * ib, p, Hs + (size_t)ib * ncoef);
  __syncthreads();
  const int nc1 = p * (p + 1) / 2;
  cd* Ho = (cd*)P.basis + (size_t)k * nb * nc1;
  for (int it = threadIdx.x; it < nb * nc1; it += blockDim.x) Ho[it] = Hs[(size_t)(it / nc1) * ncoef + it % nc1];
  __syncthreads();
  // quaternion system at the nodes.  F_k = Im((-i)^m H'_n^m), n=1..p, m=1..n.
  // Fq(i,k) = grad F_k(y_i) as a pure quaternion; the real 4h x 4h matrix of the
  // reference (Mmatrix with its first two diagonal blocks zeroed, rrq_r64:851-858)
  // is left multiplication by this h x h quaternion matrix.
  double* FQ = P.Fq + (size_t)k * 3 * h * h;
  double* MS = P.Msl + (size_t)k * h * h;
  double* FN = P.Fn + (size_t)k * h * h;
  // node positions and normals in the frame (in place), then per chunk of nodes:
  // harmonics at the nodes, and one thread per (node, basis function)
  for (int i = threadIdx.x; i < h; i += blockDim.x) {
    const double x[3] = {ssx[i], ssx[h + i], ssx[2 * h + i]}, nn0[3] = {ssn[i], ssn[h + i], ssn[2 * h + i]};
    double X[3], N[3];
    to_frame(fr, x, X);
    rotate(fr, nn0, N);
    for (int c = 0; c < 3; ++c) { ssx[c * h + i] = X[c]; ssn[c * h + i] = N[c]; }
  }
  __syncthreads();
  const int chunk = nb > 64 ? nb : 64;
  for (int i0 = 0; i0 < h; i0 += chunk) {
    const int ni = min(chunk, h - i0);
    for (int t = threadIdx.x; t < ni; t += blockDim.x) {
      const double X[3] = {ssx[i0 + t], ssx[h + i0 + t], ssx[2 * h + i0 + t]};
      harm_vals(X, p, Hs + (size_t)t * ncoef);
    }
    __syncthreads();
    for (int it = threadIdx.x; it < ni * h; it += blockDim.x) {
      const int t = it / h, kk = it % h, i = i0 + t;
      int n = 1;
      while ((n + 1) * n / 2 <= kk) ++n;
      const int m = kk - (n - 1) * n / 2 + 1;
      const cd* H = Hs + (size_t)t * ncoef;
      cd g[3];
      harm_grad(H, n, m, g);
      const double F = imrot(m, H[hidx(n, m)]);
      const double F1 = imrot(m, g[0]), F2 = imrot(m, g[1]), F3 = imrot(m, g[2]);
      FN[i + (size_t)h * kk] = F;
      MS[kk + (size_t)h * i] = ssn[i] * F1 + ssn[h + i] * F2 + ssn[2 * h + i] * F3;
      double* q = FQ + 3 * ((size_t)i * h + kk);
      q[0] = F1; q[1] = F2; q[2] = F3;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Weights operator of one panel by two Gauss-Jordan eliminations run in one loop:
//   [Fq | I] -> [I | X],  X = Fq^{-1}   (h x h pure-quaternion matrix, elements
//                                        combined by left multiplication), and
//   [Msl^T | I] -> [I | B]              (real).
// The matrices live in registers on a 16 x 16 thread grid, thread (ty, tx)
// holding rows ty+16a and columns tx+16b.  Partial pivoting without moving rows:
// step k picks the unused row with the largest |element| in column k (lowest
// index on ties), eliminates column k from every other row and normalises the
// pivot row, so row piv[k] ends as row k of the inverse.  Two barriers per step;
// LDS buffers alternate between steps.
//
// With G1(i, (b,k)) = (X_ki)_b, Ad = G1 rhs / 4pi solves M^T w = rhs (first h
// components), and As = B (W/4pi + Fn^T Ad) / alpha = (C rhs + B W) / (4pi alpha),
// C = B Fn^T G1.  Rows j < h of G: As, rows h..2h-1: Ad.
__device__ __forceinline__ void qmul_fma(const double* a, const double* b, double* c) {
  c[0] = fma(a[0], b[0], fma(-a[1], b[1], fma(-a[2], b[2], -a[3] * b[3])));
  c[1] = fma(a[0], b[1], fma(a[1], b[0], fma(a[2], b[3], -a[3] * b[2])));
  c[2] = fma(a[0], b[2], fma(-a[1], b[3], fma(a[2], b[0], a[3] * b[1])));
  c[3] = fma(a[0], b[3], fma(a[1], b[2], fma(-a[2], b[1], a[3] * b[0])));
}
// e -= a * b (quaternions)
__device__ __forceinline__ void qsub_mul(double* e, const double* a, const double* b) {
  e[0] = fma(-a[0], b[0], fma(a[1], b[1], fma(a[2], b[2], fma(a[3], b[3], e[0]))));
  e[1] = fma(-a[0], b[1], fma(-a[1], b[0], fma(-a[2], b[3], fma(a[3], b[2], e[1]))));
  e[2] = fma(-a[0], b[2], fma(a[1], b[3], fma(-a[2], b[0], fma(-a[3], b[1], e[2]))));
  e[3] = fma(-a[0], b[3], fma(-a[1], b[2], fma(a[2], b[1], fma(-a[3], b[0], e[3]))));
}

// lowest row not yet pivoted (bit clear in u[0..1]); the fallback pivot when
// NaN entries leave the search without a candidate
__device__ __forceinline__ int first_unused(const unsigned long long* u) {
  return u[0] != ~0ull ? __builtin_ctzll(~u[0]) : 64 + __builtin_ctzll(~u[1]);
}

bool rrq_solve_g1_in_lds(int h) { return (size_t)6 * h * h * sizeof(double) <= 120 * 1024; }
size_t rrq_solve_smem(int h) {
  return sizeof(double) * (size_t)((rrq_solve_g1_in_lds(h) ? 4 * h * h : 0) + 2 * h * h);
}

// Shared tail of the solve kernels: C = B Fn^T, then G (stored transposed,
// G(row, c) at c*2h + row): rows j<h  As = (C rhs + B W) * fa, rows h..2h-1
// Ad = G1 rhs * fac.  Bt = B^T (Bt[r*h + j] = B(j, r)) and Ct = C^T in LDS, G1 in
// LDS or scratch.  Threads run over consecutive j, so the G stores coalesce and
// the LDS reads of Bt / Ct are conflict-free (G1 and Fn values are broadcasts).
__device__ void solve_tail(int h, const double* Bt, double* Ct, const double* G1, const double* __restrict__ FN,
                           double* __restrict__ G, double fa, double fac) {
  const int H4 = 4 * h, H5 = 5 * h, m = 2 * h;
  // Ct[i*h + j] = C(j, i) = sum_r B(j, r) Fn(i, r)
  for (int e = threadIdx.x; e < h * h; e += blockDim.x) {
    const int j = e % h, i = e / h;
    double s = 0.0;
    for (int r = 0; r < h; ++r) s = fma(Bt[(size_t)r * h + j], FN[i + (size_t)h * r], s);
    Ct[(size_t)i * h + j] = s;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < h * H5; e += blockDim.x) {
    const int j = e % h, c = e / h;
    double* Gc = G + (size_t)c * m;
    Gc[h + j] = c < H4 ? G1[(size_t)j * H4 + c] * fac : 0.0;
    double s = 0.0;
    if (c < H4) {
      for (int i = 0; i < h; ++i) s = fma(Ct[(size_t)i * h + j], G1[(size_t)i * H4 + c], s);
    } else {
      s = Bt[(size_t)(c - H4) * h + j];
    }
    Gc[j] = s * fa;
  }
}

template <int RPT, int CPT>
__global__ void __launch_bounds__(256) rrq_panel_solve_kernel(RRQTables T, RRQPanels P, int npan,
                                                              const int* __restrict__ nrrq) {
  const int k = blockIdx.x;
  if (nrrq && nrrq[k] == 0) return;  // null: every panel
  const int h = T.h, H4 = 4 * h, H5 = 5 * h, m = 2 * h;
  constexpr int NR = 16 * RPT, NC = 16 * CPT;
  __shared__ double colq[2][4 * NR], colr[2][NR], prq[2][4 * NC], prr[2][NC];
  __shared__ double cvq[2][4], cvr[2][4];
  __shared__ int ciq[2][4], cir[2][4];
  __shared__ int pivq[NR], pivr[NR];
  extern __shared__ double dyn[];
  double* Bm = dyn;                      // [h][h] B(r,c) at r*h+c
  double* Cm = Bm + (size_t)h * h;       // [h][h] (B Fn^T)(j,i) at j*h+i
  double* G1 = P.scr ? P.scr + (size_t)k * 4 * h * h : Cm + (size_t)h * h;  // [h][4h]
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const double* FQ = P.Fq + (size_t)k * 3 * h * h;
  const double* MS = P.Msl + (size_t)k * h * h;
  const double* FN = P.Fn + (size_t)k * h * h;
  double* G = P.G + (size_t)k * 2 * h * H5;
  const double alpha = P.frame[(size_t)k * 16 + 12];
  const double fac = 1.0 / (16.0 * atan(1.0));
  double Aq[RPT][CPT][4], Ar[RPT][CPT];
#pragma unroll
  for (int a = 0; a < RPT; ++a)
#pragma unroll
    for (int b = 0; b < CPT; ++b) {
      const int r = ty + 16 * a, c = tx + 16 * b;
      double* e = Aq[a][b];
      e[0] = e[1] = e[2] = e[3] = 0.0;
      Ar[a][b] = 0.0;
      if (r < h && c < h) {
        const double* q = FQ + 3 * ((size_t)r * h + c);
        e[1] = q[0]; e[2] = q[1]; e[3] = q[2];
        Ar[a][b] = MS[r + (size_t)h * c];
      }
    }
  // rows already used as pivots (same in every thread)
  unsigned long long uq[2] = {0ull, 0ull}, ur[2] = {0ull, 0ull};
  const int wv = threadIdx.x >> 6;
  for (int kk = 0; kk < h; ++kk) {
    const int buf = kk & 1, kt = kk & 15, kb = kk >> 4;
    // column kk: its owners (tx == kt, four per wave) publish the multipliers and
    // their best unused row; the wave reduces the four candidates by shuffles
    double bq = -1.0, br = -1.0;
    int iq = NR, ir = NR;
    const bool own = tx == kt;
#pragma unroll
    for (int b = 0; b < CPT; ++b) {
      if (b != kb) continue;  // uniform
#pragma unroll
      for (int a = 0; a < RPT; ++a) {
        const int r = ty + 16 * a;
        if (own & (r < h)) {
          const double* e = Aq[a][b];
          for (int q = 0; q < 4; ++q) colq[buf][4 * r + q] = e[q];
          colr[buf][r] = Ar[a][b];
        }
        const bool ok = own & (r < h);
        const bool fq = (uq[r >> 6] >> (r & 63)) & 1ull, fr = (ur[r >> 6] >> (r & 63)) & 1ull;
        const double* e = Aq[a][b];
        const double vq = e[0] * e[0] + e[1] * e[1] + e[2] * e[2] + e[3] * e[3], vr = fabs(Ar[a][b]);
        const bool tq = ok & !fq & (vq > bq), tr = ok & !fr & (vr > br);
        bq = tq ? vq : bq; iq = tq ? r : iq;
        br = tr ? vr : br; ir = tr ? r : ir;
      }
    }
#pragma unroll
    for (int off = 16; off <= 32; off <<= 1) {
      const double oq = __shfl_xor(bq, off, 64), orr = __shfl_xor(br, off, 64);
      const int oiq = __shfl_xor(iq, off, 64), oir = __shfl_xor(ir, off, 64);
      const bool tq = (oq > bq) | ((oq == bq) & (oiq < iq)), tr = (orr > br) | ((orr == br) & (oir < ir));
      bq = tq ? oq : bq; iq = tq ? oiq : iq;
      br = tr ? orr : br; ir = tr ? oir : ir;
    }
    if ((threadIdx.x & 63) == kt) {
      cvq[buf][wv] = bq; ciq[buf][wv] = iq;
      cvr[buf][wv] = br; cir[buf][wv] = ir;
    }
    __syncthreads();
    int pq = NR, pr = NR;
    {
      double vbq = -1.0, vbr = -1.0;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const double vq = cvq[buf][t], vr = cvr[buf][t];
        const int i1 = ciq[buf][t], i2 = cir[buf][t];
        const bool tq = (vq > vbq) | ((vq == vbq) & (i1 < pq)), tr = (vr > vbr) | ((vr == vbr) & (i2 < pr));
        vbq = tq ? vq : vbq; pq = tq ? i1 : pq;
        vbr = tr ? vr : vbr; pr = tr ? i2 : pr;
      }
    }
    // identical in every lane: keep them in scalar registers
    pq = __builtin_amdgcn_readfirstlane(pq);
    pr = __builtin_amdgcn_readfirstlane(pr);
    if (pq >= h) pq = first_unused(uq);
    if (pr >= h) pr = first_unused(ur);
    uq[pq >> 6] |= 1ull << (pq & 63);
    ur[pr >> 6] |= 1ull << (pr & 63);
    if (threadIdx.x == 0) { pivq[kk] = pq; pivr[kk] = pr; }
    // pivot rows to LDS (row pq lives in slot a = pq/16 of the threads with ty =
    // pq%16); all 16*CPT columns are written (the buffers have room for them)
    // The right block is kept in pivot order: its column h+kk is the identity
    // column of the row pivoted at step kk, materialised now (1 in the pivot row,
    // 0 elsewhere, which is what every other row still holds there).
#pragma unroll
    for (int a = 0; a < RPT; ++a) {
      if (a == (pq >> 4) && ty == (pq & 15)) {
#pragma unroll
        for (int b = 0; b < CPT; ++b) {
          const bool unit = tx + 16 * b == h + kk;
#pragma unroll
          for (int q = 0; q < 4; ++q) prq[buf][4 * (tx + 16 * b) + q] = unit ? (q == 0 ? 1.0 : 0.0) : Aq[a][b][q];
        }
      }
      if (a == (pr >> 4) && ty == (pr & 15)) {
#pragma unroll
        for (int b = 0; b < CPT; ++b) prr[buf][tx + 16 * b] = tx + 16 * b == h + kk ? 1.0 : Ar[a][b];
      }
    }
    __syncthreads();
    // inverse of the pivots
    double iq4[4];
    {
      const double* e = prq[buf] + 4 * kk;
      const double n2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2] + e[3] * e[3];
      const double in2 = 1.0 / n2;
      iq4[0] = e[0] * in2; iq4[1] = -e[1] * in2; iq4[2] = -e[2] * in2; iq4[3] = -e[3] * in2;
    }
    const double irr = 1.0 / prr[buf][kk];
    // every row: A_r -= A_rk * (pivot^-1 * pivot row); the pivot row itself is
    // overwritten below.  Live columns are the window (kk, h+kk]: finished left
    // columns and right columns of rows not yet pivoted (still unit columns,
    // untouched by this step) get a zero normalised row, which leaves them
    // unchanged; chunks of 16 columns outside the window are skipped outright.
#pragma unroll
    for (int b = 0; b < CPT; ++b) {
      if (16 * b + 15 <= kk || 16 * b > h + kk) continue;  // uniform
      const int c = tx + 16 * b;
      const bool colok = (c > kk) & (c <= h + kk) & (c < m);
      const int cc = c < m ? c : m - 1;
      double nq[4];
      qmul_fma(iq4, prq[buf] + 4 * cc, nq);
#pragma unroll
      for (int q = 0; q < 4; ++q) nq[q] = colok ? nq[q] : 0.0;
      const double nr = colok ? irr * prr[buf][cc] : 0.0;
#pragma unroll
      for (int a = 0; a < RPT; ++a) {
        const int r = ty + 16 * a;
        const int rc = r < h ? r : h - 1;
        qsub_mul(Aq[a][b], colq[buf] + 4 * rc, nq);
        Ar[a][b] = fma(-colr[buf][rc], nr, Ar[a][b]);
      }
    }
#pragma unroll
    for (int a = 0; a < RPT; ++a) {
      if (a == (pq >> 4) && ty == (pq & 15)) {
#pragma unroll
        for (int b = 0; b < CPT; ++b) {
          const int c = tx + 16 * b, cc = c < m ? c : m - 1;
          double nq[4];
          qmul_fma(iq4, prq[buf] + 4 * cc, nq);
          // columns right of the window stay zero (unit columns of later pivots)
          const bool keep = c <= h + kk;
#pragma unroll
          for (int q = 0; q < 4; ++q) Aq[a][b][q] = keep ? nq[q] : 0.0;
        }
      }
      if (a == (pr >> 4) && ty == (pr & 15)) {
#pragma unroll
        for (int b = 0; b < CPT; ++b) {
          const int c = tx + 16 * b, cc = c < m ? c : m - 1;
          Ar[a][b] = c <= h + kk ? irr * prr[buf][cc] : 0.0;
        }
      }
    }
  }
  __syncthreads();
  // row piv[kk] holds row kk of the inverse: scatter by the inverse permutation
  __shared__ int posq[NR], posr[NR];
  for (int kk = threadIdx.x; kk < h; kk += blockDim.x) { posq[pivq[kk]] = kk; posr[pivr[kk]] = kk; }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < RPT; ++a)
#pragma unroll
    for (int b = 0; b < CPT; ++b) {
      const int r = ty + 16 * a, c = tx + 16 * b;
      if (r < h && c >= h && c < m) {
        // row piv[k] holds row k of the inverse, column h+j' holds its column piv[j']
        const int jq = pivq[c - h], jr = pivr[c - h];
        const int kq = posq[r], kr = posr[r];
        for (int q = 0; q < 4; ++q) G1[(size_t)jq * H4 + q * h + kq] = Aq[a][b][q];
        Bm[(size_t)jr * h + kr] = Ar[a][b];  // transposed: Bt[r*h + j] = B(j, r)
      }
    }
  __syncthreads();
  solve_tail(h, Bm, Cm, G1, FN, G, fac / alpha, fac);
}

// Entry (a, b) of the left-multiplication matrix of quaternion q (q*x = L(q) x):
// L[a][b] = +/- q[a ^ b], negative at (0,1) (0,2) (0,3) (1,2) (2,3) (3,1).
__device__ __forceinline__ double lmat(const double* q, int a, int b) {
  const unsigned neg = 0x0Eu | (0x4u << 4) | (0x8u << 8) | (0x2u << 12);
  const double v = q[a ^ b];
  return ((neg >> (a * 4 + b)) & 1u) ? -v : v;
}

__device__ __forceinline__ double qd4_get(const qd4& d, int i) {
  return i == 0 ? d[0] : (i == 1 ? d[1] : (i == 2 ? d[2] : d[3]));
}

// The same weights operator with the quaternion elimination on the FP64 matrix
// cores.  In real form the quaternion rank-1 update A_rc -= m_r n_c is
// A(4h x 2h) -= L(m)(4h x 4) N(4 x 2h): one v_mfma_f64_16x16x4 per 16x16 tile.
// A is held as accumulator tiles (row 4r + component, column c): wave w owns the
// row tiles w, w+4, ...; lane l holds rows 4i + l/16 (i = 0..3, i.e. quaternion
// 4*tile + i, component l/16) of column l%16.  The real Msl^T elimination stays
// on the 16 x 16 thread grid as in rrq_panel_solve_kernel; both share the step
// loop (implicit partial pivoting, pivot-ordered right block, column window).
template <int RT, int CT>
__global__ void __launch_bounds__(256) rrq_panel_solve_mfma_kernel(RRQTables T, RRQPanels P, int npan,
                                                                   const int* __restrict__ nrrq) {
  const int k = blockIdx.x;
  if (nrrq && nrrq[k] == 0) return;  // null: every panel
  const int h = T.h, H4 = 4 * h, m = 2 * h;
  constexpr int RTW = (RT + 3) / 4, NQR = 4 * RT, NC = 16 * CT;
  __shared__ double colq[2][4 * NQR], prq[2][4 * NC];
  __shared__ double cvq[2][4];
  __shared__ int ciq[2][4];
  __shared__ int pivq[NQR], posq[NQR];
  extern __shared__ double dyn[];
  double* Bm = dyn;
  double* Cm = Bm + (size_t)h * h;
  double* G1 = P.scr ? P.scr + (size_t)k * 4 * h * h : Cm + (size_t)h * h;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
  const double* FQ = P.Fq + (size_t)k * 3 * h * h;
  const double* FN = P.Fn + (size_t)k * h * h;
  double* G = P.G + (size_t)k * 2 * h * 5 * h;
  const double alpha = P.frame[(size_t)k * 16 + 12];
  const double fac = 1.0 / (16.0 * atan(1.0));
  qd4 D[RTW][CT];
#pragma unroll
  for (int t = 0; t < RTW; ++t)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int rt = wv + 4 * t, r = 4 * rt + i, c = 16 * ct + li;
        D[t][ct][i] = (rt < RT && r < h && c < h && lk > 0) ? FQ[3 * ((size_t)r * h + c) + lk - 1] : 0.0;
      }
  // B = (Msl^T)^{-1} from its own kernel, to LDS transposed (Bt[r*h + j] = B(j, r))
  for (int e = threadIdx.x; e < h * h; e += blockDim.x) Bm[(size_t)(e % h) * h + e / h] = P.Binv[(size_t)k * h * h + e];
  unsigned long long uq[2] = {0ull, 0ull};
  for (int kk = 0; kk < h; ++kk) {
    const int buf = kk & 1, kt = kk & 15, kb = kk >> 4;
    // ---- column kk: multipliers and pivot candidates
    {
      double bq = -1.0;
      int iq = NQR;
#pragma unroll
      for (int t = 0; t < RTW; ++t) {
        const int rt = wv + 4 * t;
        if (rt >= RT) continue;  // uniform
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          if (ct != kb) continue;  // uniform
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const double v = D[t][ct][i];
            double s2 = v * v;
            s2 += __shfl_xor(s2, 16, 64);
            s2 += __shfl_xor(s2, 32, 64);
            const int r = 4 * rt + i;
            const bool here = (li == kt) & (r < h);
            if (here) colq[buf][4 * r + lk] = v;
            const bool used = (uq[r >> 6] >> (r & 63)) & 1ull;
            const bool tq = here & !used & (s2 > bq);
            bq = tq ? s2 : bq;
            iq = tq ? r : iq;
          }
        }
      }
      if (lane == kt) { cvq[buf][wv] = bq; ciq[buf][wv] = iq; }
    }
    __syncthreads();
    int pq = NQR;
    {
      double vbq = -1.0;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const double vq = cvq[buf][t];
        const int i1 = ciq[buf][t];
        const bool tq = (vq > vbq) | ((vq == vbq) & (i1 < pq));
        vbq = tq ? vq : vbq; pq = tq ? i1 : pq;
      }
    }
    pq = __builtin_amdgcn_readfirstlane(pq);
    if (pq >= h) pq = first_unused(uq);
    uq[pq >> 6] |= 1ull << (pq & 63);
    if (threadIdx.x == 0) pivq[kk] = pq;
    const int rtp = pq >> 2, ip = pq & 3, wp = rtp & 3, tp = rtp >> 2;
    // pivot row to LDS; the right block is in pivot order, its column h+kk is the
    // unit column of the row pivoted now (materialised here)
    if (wv == wp) {
#pragma unroll
      for (int t = 0; t < RTW; ++t) {
        if (t != tp) continue;  // uniform
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          const int c = 16 * ct + li;
          const double v = qd4_get(D[t][ct], ip);
          prq[buf][4 * c + lk] = c == h + kk ? (lk == 0 ? 1.0 : 0.0) : v;
        }
      }
    }
    __syncthreads();
    // ---- update
    double iq4[4];
    {
      const double* e = prq[buf] + 4 * kk;
      const double n2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2] + e[3] * e[3];
      const double in2 = 1.0 / n2;
      iq4[0] = e[0] * in2; iq4[1] = -e[1] * in2; iq4[2] = -e[2] * in2; iq4[3] = -e[3] * in2;
    }
    // B operand: component lk of the normalised pivot row at column 16ct + li
    // (zero outside the live window (kk, h+kk])
    double Bop[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
      const int c = 16 * ct + li;
      const int cc = c < m ? c : m - 1;
      const double* x = prq[buf] + 4 * cc;
      double v = 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j) v = fma(lmat(iq4, lk, j), x[j], v);
      const bool colok = (c > kk) & (c <= h + kk) & (c < m);
      Bop[ct] = colok ? v : 0.0;
    }
    // A operand: -L(m_r)[comp][lk] for row 4r + comp = 16 rt + li; the pivot row
    // gets zero (it is overwritten below)
#pragma unroll
    for (int t = 0; t < RTW; ++t) {
      const int rt = wv + 4 * t;
      if (rt >= RT) continue;  // uniform
      const int r = 4 * rt + (li >> 2), comp = li & 3;
      const double mv = lmat(colq[buf] + 4 * r, comp, lk);
      const double aop = ((r < h) & (r != pq)) ? -mv : 0.0;
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        if (16 * ct + 15 <= kk || 16 * ct > h + kk) continue;  // uniform: tile outside the window
        D[t][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(aop, Bop[ct], D[t][ct], 0, 0, 0);
      }
    }
    if (wv == wp) {
#pragma unroll
      for (int t = 0; t < RTW; ++t) {
        if (t != tp) continue;  // uniform
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
          for (int i = 0; i < 4; ++i) D[t][ct][i] = i == ip ? Bop[ct] : D[t][ct][i];
      }
    }
  }
  __syncthreads();
  for (int kk = threadIdx.x; kk < h; kk += blockDim.x) posq[pivq[kk]] = kk;
  __syncthreads();
  // row piv[k] holds row k of the inverse; column h+j' holds its column piv[j']
#pragma unroll
  for (int t = 0; t < RTW; ++t) {
    const int rt = wv + 4 * t;
    if (rt >= RT) continue;
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = 4 * rt + i, c = 16 * ct + li;
        if (r < h && c >= h && c < m) G1[(size_t)pivq[c - h] * H4 + lk * h + posq[r]] = D[t][ct][i];
      }
  }
  __syncthreads();
  solve_tail(h, Bm, Cm, G1, FN, G, fac / alpha, fac);
}

// B = (Msl^T)^{-1} (row-major, B(r,c) at r*h + c) by the same windowed
// Gauss-Jordan on the 16 x 16 thread grid; small enough to run many panels per CU.
template <int RPT, int CPT>
__global__ void __launch_bounds__(256) rrq_msl_inverse_kernel(RRQTables T, RRQPanels P, int npan,
                                                              const int* __restrict__ nrrq) {
  const int k = blockIdx.x;
  if (nrrq && nrrq[k] == 0) return;  // null: every panel
  const int h = T.h, m = 2 * h;
  constexpr int NR = 16 * RPT, NC = 16 * CPT;
  __shared__ double colr[2][NR], prr[2][NC], cvr[2][4];
  __shared__ int cir[2][4], pivr[NR], posr[NR];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const double* MS = P.Msl + (size_t)k * h * h;
  double Ar[RPT][CPT];
#pragma unroll
  for (int a = 0; a < RPT; ++a)
#pragma unroll
    for (int b = 0; b < CPT; ++b) {
      const int r = ty + 16 * a, c = tx + 16 * b;
      Ar[a][b] = (r < h && c < h) ? MS[r + (size_t)h * c] : 0.0;
    }
  unsigned long long ur[2] = {0ull, 0ull};
  for (int kk = 0; kk < h; ++kk) {
    const int buf = kk & 1, kt = kk & 15, kb = kk >> 4;
    double br = -1.0;
    int ir = NR;
    const bool own = tx == kt;
#pragma unroll
    for (int b = 0; b < CPT; ++b) {
      if (b != kb) continue;  // uniform
#pragma unroll
      for (int a = 0; a < RPT; ++a) {
        const int r = ty + 16 * a;
        if (own & (r < h)) colr[buf][r] = Ar[a][b];
        const bool fr = (ur[r >> 6] >> (r & 63)) & 1ull;
        const double vr = fabs(Ar[a][b]);
        const bool tr = own & (r < h) & !fr & (vr > br);
        br = tr ? vr : br;
        ir = tr ? r : ir;
      }
    }
#pragma unroll
    for (int off = 16; off <= 32; off <<= 1) {
      const double orr = __shfl_xor(br, off, 64);
      const int oir = __shfl_xor(ir, off, 64);
      const bool tr = (orr > br) | ((orr == br) & (oir < ir));
      br = tr ? orr : br;
      ir = tr ? oir : ir;
    }
    if (lane == kt) { cvr[buf][wv] = br; cir[buf][wv] = ir; }
    __syncthreads();
    int pr = NR;
    {
      double vbr = -1.0;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const double vr = cvr[buf][t];
        const int i2 = cir[buf][t];
        const bool tr = (vr > vbr) | ((vr == vbr) & (i2 < pr));
        vbr = tr ? vr : vbr; pr = tr ? i2 : pr;
      }
    }
    pr = __builtin_amdgcn_readfirstlane(pr);
    if (pr >= h) pr = first_unused(ur);
    ur[pr >> 6] |= 1ull << (pr & 63);
    if (threadIdx.x == 0) pivr[kk] = pr;
#pragma unroll
    for (int a = 0; a < RPT; ++a)
      if (a == (pr >> 4) && ty == (pr & 15)) {
#pragma unroll
        for (int b = 0; b < CPT; ++b) prr[buf][tx + 16 * b] = tx + 16 * b == h + kk ? 1.0 : Ar[a][b];
      }
    __syncthreads();
    const double irr = 1.0 / prr[buf][kk];
#pragma unroll
    for (int b = 0; b < CPT; ++b) {
      if (16 * b + 15 <= kk || 16 * b > h + kk) continue;  // uniform
      const int c = tx + 16 * b;
      const bool colok = (c > kk) & (c <= h + kk) & (c < m);
      const int cc = c < m ? c : m - 1;
      const double nr = colok ? irr * prr[buf][cc] : 0.0;
#pragma unroll
      for (int a = 0; a < RPT; ++a) {
        const int r = ty + 16 * a;
        const int rc = r < h ? r : h - 1;
        Ar[a][b] = fma(-colr[buf][rc], nr, Ar[a][b]);
      }
    }
#pragma unroll
    for (int a = 0; a < RPT; ++a)
      if (a == (pr >> 4) && ty == (pr & 15)) {
#pragma unroll
        for (int b = 0; b < CPT; ++b) {
          const int c = tx + 16 * b, cc = c < m ? c : m - 1;
          Ar[a][b] = c <= h + kk ? irr * prr[buf][cc] : 0.0;
        }
      }
  }
  __syncthreads();
  for (int kk = threadIdx.x; kk < h; kk += blockDim.x) posr[pivr[kk]] = kk;
  __syncthreads();
  double* Bk = P.Binv + (size_t)k * h * h;
#pragma unroll
  for (int a = 0; a < RPT; ++a)
#pragma unroll
    for (int b = 0; b < CPT; ++b) {
      const int r = ty + 16 * a, c = tx + 16 * b;
      if (r < h && c >= h && c < m) Bk[(size_t)posr[r] * h + pivr[c - h]] = Ar[a][b];
    }
}

// rows already pivoted by a register-resident Gauss-Jordan (h <= 128)
struct RowMask {
  unsigned long long lo = 0ull, hi = 0ull;
  __device__ __forceinline__ bool has(int r) const { return ((r < 64 ? lo : hi) >> (r & 63)) & 1ull; }
  __device__ __forceinline__ void add(int r) {
    if (r < 64) lo |= 1ull << r;
    else hi |= 1ull << (r & 63);
  }
  // lowest row not yet pivoted
  __device__ __forceinline__ int first_free() const { return ~lo ? __builtin_ctzll(~lo) : 64 + __builtin_ctzll(~hi); }
};

// The epilogue of the quaternion solve, shared by rrq_quat_inverse_kernel with
// EPI (X in registers) and rrq_quat_epilogue_kernel (X from the Ad rows of G):
// with
// B = (Msl^T)^{-1} from rrq_real_inverse_kernel and X = Fq^{-1},
//   Ad rows  G(h+j, q*h+k) = (X_kj)_q fac
//   As rows  G(j, q*h+k)   = (fac/alpha) sum_i C(j,i) (X_ki)_q,  C = B Fn^T
//   W block  G(j, 4h+r)    = (fac/alpha) B(j,r),  G(h+j, 4h+r) = 0
// G is stored transposed, G(row, c) at c*2h + row.  LDS (3h^2 doubles): C, then
// Bt and Fn (for C) or one component of X at a time (Xq aliases Bt).
struct SolveLds {
  double* Ct;  // Ct[i*h + j] = C(j, i)
  double* Bt;  // Bt[r*h + j] = B(j, r)
  double* Fs;  // Fs[i + h*r] = Fn(i, r)
  double* Xq;  // Xq[i*h + k] = (X_ki)_q
  __device__ SolveLds(double* dyn, int h)
      : Ct(dyn), Bt(dyn + (size_t)h * h), Fs(dyn + (size_t)2 * h * h), Xq(dyn + (size_t)h * h) {}
};

// C = B Fn^T into LDS and the W block of G; ends with C visible to every thread
// once the caller's next barrier has passed
__device__ __forceinline__ void solve_epilogue_c(const RRQPanels& P, int k, int h, const SolveLds& L, double fa,
                                                 double* G, int tid, int nth) {
  const int m = 2 * h;
  {
    const double* Bk = P.Binv + (size_t)k * h * h;
    const double* FN = P.Fn + (size_t)k * h * h;
    for (int e = tid; e < h * h; e += nth) {
      L.Bt[(size_t)(e % h) * h + e / h] = Bk[e];
      L.Fs[e] = FN[e];
    }
  }
  __syncthreads();
  const int nig = (h + 3) / 4;
  for (int it = tid; it < h * nig; it += nth) {
    const int j = it % h, i0 = 4 * (it / h);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int r = 0; r < h; ++r) {
      const double bv = L.Bt[(size_t)r * h + j];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] = fma(bv, L.Fs[(i0 + u < h ? i0 + u : h - 1) + (size_t)h * r], acc[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (i0 + u < h) L.Ct[(size_t)(i0 + u) * h + j] = acc[u];
  }
  for (int e = tid; e < h * h; e += nth) {
    const int j = e % h, r = e / h;
    double* w = G + (size_t)(4 * h + r) * m;
    w[j] = L.Bt[(size_t)r * h + j] * fa;
    w[h + j] = 0.0;
  }
}

// As rows of component q from C and Xq in LDS: tiles of 3 rows j x 4 columns k
__device__ __forceinline__ void solve_epilogue_as(int h, int q, const SolveLds& L, double fa, double* G, int tid,
                                                  int nth) {
  const int m = 2 * h;
  const int njg = (h + 2) / 3, nkg = (h + 3) / 4;
  for (int it = tid; it < njg * nkg; it += nth) {
    const int j0 = 3 * (it % njg), k0 = 4 * (it / njg);
    double acc[3][4];
#pragma unroll
    for (int v = 0; v < 3; ++v)
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[v][u] = 0.0;
    for (int i = 0; i < h; ++i) {
      double cv[3], xv[4];
#pragma unroll
      for (int v = 0; v < 3; ++v) cv[v] = L.Ct[(size_t)i * h + (j0 + v < h ? j0 + v : h - 1)];
#pragma unroll
      for (int u = 0; u < 4; ++u) xv[u] = L.Xq[(size_t)i * h + (k0 + u < h ? k0 + u : h - 1)];
#pragma unroll
      for (int v = 0; v < 3; ++v)
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[v][u] = fma(cv[v], xv[u], acc[v][u]);
    }
#pragma unroll
    for (int v = 0; v < 3; ++v)
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if ((j0 + v < h) & (k0 + u < h)) G[(size_t)(q * h + k0 + u) * m + j0 + v] = acc[v][u] * fa;
  }
}

// X = Fq^{-1} for the h x h pure-quaternion matrix Fq by Gauss-Jordan, the
// matrix in registers of NW waves (one panel per workgroup).  Lane (ly, lx) of
// an RL x CL lane grid (RL*CL <= 64; lanes beyond it idle) owns rows ly + RL*a
// (a < RA) and column slots lx + CL*b for the wave's CB/NW slot groups b;
// RL*RA >= h and CL*CB >= h, the grid chosen per order to waste few registers.
// Slot s holds left column s until step s eliminates it and right column h+s
// afterwards (the unit column of the row pivoted at step s, zero in every other
// row until then), so h slots carry the live window (kk, h+kk] at every step.
// Pivoting is implicit (no row moves): the unused row with the largest |A_r,kk|
// wins, ties to the lower row.  Column kk lies in one wave: it finds the pivot
// and publishes the column with a zero in the pivot row (so the rank-1 update,
// the same in every lane, leaves that row alone; it is scaled by the inverse
// pivot afterwards).  Each wave already holds its part of the pivot row, so one
// workgroup barrier per step suffices (LDS double-buffered by step parity).
// Afterwards row piv[k] holds row k of X with slot s holding column piv[s].
// EPI: the kernel also forms G (solve_epilogue_*, 3h^2 doubles of dynamic LDS);
// otherwise it writes X itself into the Ad rows of G, G(h+j, q*h+k) = (X_kj)_q,
// and rrq_quat_epilogue_kernel finishes G, so that the elimination needs no
// dynamic LDS and more panels share a CU.
template <int RL, int RA, int CL, int CB, int NW, bool EPI>
__global__ void __launch_bounds__(64 * NW) rrq_quat_inverse_kernel(RRQTables T, RRQPanels P, int npan,
                                                                   const int* __restrict__ nrrq) {
  static_assert(RL * CL <= 64 && RL * RA <= 128 && CB % NW == 0, "lane grid");
  constexpr int CW = CB / NW;  // slot groups per wave
  const int k = blockIdx.x;
  if (nrrq && nrrq[k] == 0) return;  // null: every panel
  const int h = T.h, m = 2 * h;
  constexpr int NRW = RL * RA, NSL = CL * CB;
  __shared__ double colq[2][4 * NRW], rowq[2][4 * NSL], pivel[2][4];
  __shared__ int pivs[2], piv[NRW], pos[NRW];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const bool act = lane < RL * CL;
  const int ly = lane / CL, lx = lane % CL;
  const double* FQ = P.Fq + (size_t)k * 3 * h * h;
  double A[RA][CW][4];
#pragma unroll
  for (int a = 0; a < RA; ++a)
#pragma unroll
    for (int b = 0; b < CW; ++b) {
      const int r = ly + RL * a, c = lx + CL * (wv * CW + b);
      double* e = A[a][b];
      e[0] = e[1] = e[2] = e[3] = 0.0;
      if (act & (r < h) & (c < h)) {
        const double* q = FQ + 3 * ((size_t)r * h + c);
        e[1] = q[0]; e[2] = q[1]; e[3] = q[2];
      }
    }
  RowMask used;
  for (int kk = 0; kk < h; ++kk) {
    const int buf = kk & 1, kt = kk % CL, kbg = kk / CL;
    const int ow = kbg / CW, kb = kbg % CW;  // wave and slot group holding column kk
    double* cq = colq[buf];
    double* rq = rowq[buf];
    if (wv == ow) {  // uniform
      const bool own = act & (lx == kt);
      double bv = -1.0;
      int bi = NRW;
#pragma unroll
      for (int b = 0; b < CW; ++b) {
        if (b != kb) continue;  // uniform
#pragma unroll
        for (int a = 0; a < RA; ++a) {
          const int r = ly + RL * a;
          const double* e = A[a][b];
          const double v = e[0] * e[0] + e[1] * e[1] + e[2] * e[2] + e[3] * e[3];
          const bool t = own & (r < h) & !used.has(r) & (v > bv);
          bv = t ? v : bv;
          bi = t ? r : bi;
        }
      }
      // best of the RL owner lanes (lane ids kt + CL*t)
      int p = NRW;
      {
        double pv = -1.0;
#pragma unroll
        for (int t = 0; t < RL; ++t) {
          const double v = __shfl(bv, kt + CL * t, 64);
          const int i = __shfl(bi, kt + CL * t, 64);
          const bool tk = (v > pv) | ((v == pv) & (i < p));
          pv = tk ? v : pv;
          p = tk ? i : p;
        }
      }
      p = __builtin_amdgcn_readfirstlane(p);
      // no candidate only with NaN entries: take the lowest unused row, so every
      // row is pivoted once and every index stays in range
      if (p >= h) p = used.first_free();
      const int pa = p / RL;
      const bool prow = act & (ly == p % RL);
      if (lane == 0) pivs[buf] = p;
      // column kk by row (zero in the pivot row), the pivot itself, and slot kk
      // turned into right column h+kk: zero, 1 in the pivot row
#pragma unroll
      for (int b = 0; b < CW; ++b) {
        if (b != kb) continue;  // uniform
#pragma unroll
        for (int a = 0; a < RA; ++a) {
          double* e = A[a][b];
          if (own) {
            const int r = ly + RL * a;
            const bool isp = (a == pa) & prow;
#pragma unroll
            for (int q = 0; q < 4; ++q) cq[4 * r + q] = isp ? 0.0 : e[q];
            if (isp) {
#pragma unroll
              for (int q = 0; q < 4; ++q) pivel[buf][q] = e[q];
            }
            e[0] = isp ? 1.0 : 0.0;
            e[1] = e[2] = e[3] = 0.0;
          }
        }
      }
    }
    __syncthreads();
    const int p = __builtin_amdgcn_readfirstlane(pivs[buf]);
    used.add(p);
    if (threadIdx.x == 0) piv[kk] = p;
    const int pa = p / RL;
    const bool prow = act & (ly == p % RL);
    // this wave's part of the pivot row, by slot (wave-local exchange)
#pragma unroll
    for (int a = 0; a < RA; ++a) {
      if (a != pa) continue;  // uniform
      if (prow) {
#pragma unroll
        for (int b = 0; b < CW; ++b) {
          const int c = lx + CL * (wv * CW + b);
#pragma unroll
          for (int q = 0; q < 4; ++q) rq[4 * c + q] = A[a][b][q];
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double pinv[4];
    {
      const double* e = pivel[buf];
      const double in2 = 1.0 / (e[0] * e[0] + e[1] * e[1] + e[2] * e[2] + e[3] * e[3]);
      pinv[0] = e[0] * in2; pinv[1] = -e[1] * in2; pinv[2] = -e[2] * in2; pinv[3] = -e[3] * in2;
    }
    // every other row: A_r -= A_r,kk * (pivot^-1 * pivot row); the normalised
    // row entries of this wave's slots first, then one column value per row
    double nq[CW][4];
#pragma unroll
    for (int b = 0; b < CW; ++b) qmul_fma(pinv, rq + 4 * (lx + CL * (wv * CW + b)), nq[b]);
#pragma unroll
    for (int a = 0; a < RA; ++a) {
      const double* c = cq + 4 * (act ? ly + RL * a : 0);  // lanes past the grid idle
      const double cv[4] = {c[0], c[1], c[2], c[3]};
#pragma unroll
      for (int b = 0; b < CW; ++b) qsub_mul(A[a][b], cv, nq[b]);
    }
    // the pivot row: pivot^-1 * row
#pragma unroll
    for (int a = 0; a < RA; ++a) {
      if (a != pa) continue;  // uniform
      if (prow) {
#pragma unroll
        for (int b = 0; b < CW; ++b) {
          double nq[4];
          qmul_fma(pinv, A[a][b], nq);
#pragma unroll
          for (int q = 0; q < 4; ++q) A[a][b][q] = nq[q];
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < h; i += 64 * NW) pos[piv[i]] = i;
  double* G = P.G + (size_t)k * m * 5 * h;
  if constexpr (!EPI) {
    __syncthreads();
#pragma unroll
    for (int a = 0; a < RA; ++a)
#pragma unroll
      for (int b = 0; b < CW; ++b) {
        const int r = ly + RL * a, c = lx + CL * (wv * CW + b);
        if (act & (r < h) & (c < h)) {
          const int kx = pos[r], j = piv[c];
#pragma unroll
          for (int q = 0; q < 4; ++q) G[(size_t)(q * h + kx) * m + h + j] = A[a][b][q];
        }
      }
  } else {
    extern __shared__ double dyn[];
    const SolveLds L(dyn, h);
    const double fac = 1.0 / (16.0 * atan(1.0));
    const double fa = fac / P.frame[(size_t)k * 16 + 12];
    solve_epilogue_c(P, k, h, L, fa, G, threadIdx.x, 64 * NW);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      __syncthreads();  // Bt/Fn (q = 0) or the previous component no longer read
#pragma unroll
      for (int a = 0; a < RA; ++a)
#pragma unroll
        for (int b = 0; b < CW; ++b) {
          const int r = ly + RL * a, c = lx + CL * (wv * CW + b);
          if (act & (r < h) & (c < h)) {
            const int kx = pos[r], j = piv[c];
            L.Xq[(size_t)j * h + kx] = A[a][b][q];
            G[(size_t)(q * h + kx) * m + h + j] = A[a][b][q] * fac;
          }
        }
      __syncthreads();
      solve_epilogue_as(h, q, L, fa, G, threadIdx.x, 64 * NW);
    }
  }
}

// The rest of G after rrq_quat_inverse_kernel<..., false> has left X in its Ad
// rows: the same epilogue as the fused form (so the same results), one
// workgroup of 16 waves per panel (its 3h^2 doubles of LDS leave one workgroup
// per CU at h = 66 and 78, so the waves hide the LDS latency: 256, 512 and 1024
// threads took 4.3, 2.8 and 2.0 ms at order 12), X read back one component at a
// time and the Ad rows scaled in place.
__global__ void __launch_bounds__(1024) rrq_quat_epilogue_kernel(RRQTables T, RRQPanels P, int npan,
                                                                const int* __restrict__ nrrq) {
  const int k = blockIdx.x;
  if (nrrq && nrrq[k] == 0) return;
  const int h = T.h, m = 2 * h;
  extern __shared__ double dyn[];
  const SolveLds L(dyn, h);
  const double fac = 1.0 / (16.0 * atan(1.0));
  const double fa = fac / P.frame[(size_t)k * 16 + 12];
  double* G = P.G + (size_t)k * m * 5 * h;
  solve_epilogue_c(P, k, h, L, fa, G, threadIdx.x, blockDim.x);
  for (int q = 0; q < 4; ++q) {
    __syncthreads();
    for (int e = threadIdx.x; e < h * h; e += blockDim.x) {
      const int j = e % h, kx = e / h;
      double* g = G + (size_t)(q * h + kx) * m + h + j;
      const double x = *g;
      L.Xq[(size_t)j * h + kx] = x;
      *g = x * fac;
    }
    __syncthreads();
    solve_epilogue_as(h, q, L, fa, G, threadIdx.x, blockDim.x);
  }
}

// dynamic LDS of rrq_quat_inverse_kernel: C plus Bt and Fn (or one component of X)
size_t rrq_quat_inverse_smem(int h) { return sizeof(double) * (size_t)3 * h * h; }

// B = (Msl^T)^{-1} by Gauss-Jordan in the registers of one wave per panel, the
// real counterpart of rrq_quat_inverse_kernel (same lane grid, slot reuse,
// implicit pivoting with ties to the lower row, zeroed pivot-row column entry).
// The system matrix is A(r, c) = Msl[r + h*c]; output B row-major, B(kx, j) at
// Binv[kx*h + j].
template <int RL, int RA, int CL, int CB>
__global__ void __launch_bounds__(64) rrq_real_inverse_kernel(RRQTables T, RRQPanels P, int npan,
                                                              const int* __restrict__ nrrq) {
  static_assert(RL * CL <= 64 && RL * RA <= 128, "lane grid");
  const int k = blockIdx.x;
  if (nrrq && nrrq[k] == 0) return;  // null: every panel
  const int h = T.h;
  constexpr int NRW = RL * RA, NSL = CL * CB;
  __shared__ double colr[2][NRW], rowr[2][NSL], pivel[2];
  __shared__ int piv[NRW], pos[NRW];
  const int lane = threadIdx.x;
  const bool act = lane < RL * CL;
  const int ly = lane / CL, lx = lane % CL;
  const double* MS = P.Msl + (size_t)k * h * h;
  double A[RA][CB];
#pragma unroll
  for (int a = 0; a < RA; ++a)
#pragma unroll
    for (int b = 0; b < CB; ++b) {
      const int r = ly + RL * a, c = lx + CL * b;
      A[a][b] = (act & (r < h) & (c < h)) ? MS[r + (size_t)h * c] : 0.0;
    }
  RowMask used;
  for (int kk = 0; kk < h; ++kk) {
    const int buf = kk & 1, kt = kk % CL, kb = kk / CL;
    const bool own = act & (lx == kt);
    double bv = -1.0;
    int bi = NRW;
#pragma unroll
    for (int b = 0; b < CB; ++b) {
      if (b != kb) continue;  // uniform
#pragma unroll
      for (int a = 0; a < RA; ++a) {
        const int r = ly + RL * a;
        const double v = fabs(A[a][b]);
        const bool t = own & (r < h) & !used.has(r) & (v > bv);
        bv = t ? v : bv;
        bi = t ? r : bi;
      }
    }
    int p = NRW;
    {
      double pv = -1.0;
#pragma unroll
      for (int t = 0; t < RL; ++t) {
        const double v = __shfl(bv, kt + CL * t, 64);
        const int i = __shfl(bi, kt + CL * t, 64);
        const bool tk = (v > pv) | ((v == pv) & (i < p));
        pv = tk ? v : pv;
        p = tk ? i : p;
      }
    }
    p = __builtin_amdgcn_readfirstlane(p);
    if (p >= h) p = used.first_free();  // NaN entries only
    used.add(p);
    if (lane == 0) piv[kk] = p;
    const int pa = p / RL;
    const bool prow = act & (ly == p % RL);
#pragma unroll
    for (int b = 0; b < CB; ++b) {
      if (b != kb) continue;  // uniform
#pragma unroll
      for (int a = 0; a < RA; ++a) {
        if (own) {
          const bool isp = (a == pa) & prow;
          colr[buf][ly + RL * a] = isp ? 0.0 : A[a][b];
          if (isp) pivel[buf] = A[a][b];
          A[a][b] = isp ? 1.0 : 0.0;
        }
      }
    }
#pragma unroll
    for (int a = 0; a < RA; ++a) {
      if (a != pa) continue;  // uniform
      if (prow) {
#pragma unroll
        for (int b = 0; b < CB; ++b) rowr[buf][lx + CL * b] = A[a][b];
      }
    }
    __syncthreads();  // one wave: orders the LDS stores before the loads
    const double ip = 1.0 / pivel[buf];
    double nr[CB];
#pragma unroll
    for (int b = 0; b < CB; ++b) nr[b] = ip * rowr[buf][lx + CL * b];
#pragma unroll
    for (int a = 0; a < RA; ++a) {
      const double c = colr[buf][act ? ly + RL * a : 0];  // lanes past the grid idle
#pragma unroll
      for (int b = 0; b < CB; ++b) A[a][b] = fma(-c, nr[b], A[a][b]);
    }
#pragma unroll
    for (int a = 0; a < RA; ++a) {
      if (a != pa) continue;  // uniform
      if (prow) {
#pragma unroll
        for (int b = 0; b < CB; ++b) A[a][b] *= ip;
      }
    }
  }
  __syncthreads();
  for (int i = lane; i < h; i += 64) pos[piv[i]] = i;
  __syncthreads();
  double* Bk = P.Binv + (size_t)k * h * h;
#pragma unroll
  for (int a = 0; a < RA; ++a)
#pragma unroll
    for (int b = 0; b < CB; ++b) {
      const int r = ly + RL * a, c = lx + CL * b;
      if (act & (r < h) & (c < h)) Bk[(size_t)pos[r] * h + piv[c]] = A[a][b];
    }
}

// whether one workgroup on the current device may take this much dynamic LDS
// (beside a few KB of static LDS)
static bool lds_fits(size_t bytes) {
  int dev = 0, mx = 0;
  if (hipGetDevice(&dev) != hipSuccess) return false;
  if (hipDeviceGetAttribute(&mx, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return false;
  return bytes + 16 * 1024 <= (size_t)mx;
}

void launch_rrq_panel_solve(RRQTables T, RRQPanels P, int npan, const int* nrrq, hipStream_t s, int variant) {
  const int h = T.h;
  const int rpt = (h + 15) / 16, cpt = (2 * h + 15) / 16;
  const size_t lds = rrq_solve_smem(h);
  if (!rrq_solve_g1_in_lds(h) && !P.scr) throw std::runtime_error("RRQ solve: scratch for G1 not allocated");
  // variant 2 (or QOL_SOLVE_MFMA=1) selects, for patch orders up to 8, the
  // matrix-core quaternion elimination (A as 4-wave accumulator tiles, B from a
  // separate real inverse), which is slower on the order-8 workload (5.3 vs 4.9
  // ms for 14,250 panels, against 2.4 ms for the register-resident inverses)
  const char* mf = std::getenv("QOL_SOLVE_MFMA");
  if (variant == 2 || (variant == 0 && mf && mf[0] == '1')) switch (T.p) {
#define QOL_SOLVE_MFMA(P_, RT_, CT_, RPT_)                                                                  \
  case P_:                                                                                                \
    hipLaunchKernelGGL((rrq_msl_inverse_kernel<RPT_, CT_>), dim3(npan), dim3(256), 0, s, T, P, npan, nrrq);  \
    hipLaunchKernelGGL((rrq_panel_solve_mfma_kernel<RT_, CT_>), dim3(npan), dim3(256), lds, s, T, P, npan,   \
                       nrrq);                                                                             \
    return;
    QOL_SOLVE_MFMA(2, 1, 1, 1)
    QOL_SOLVE_MFMA(3, 2, 1, 1)
    QOL_SOLVE_MFMA(4, 3, 2, 1)
    QOL_SOLVE_MFMA(5, 4, 2, 1)
    QOL_SOLVE_MFMA(6, 6, 3, 2)
    QOL_SOLVE_MFMA(7, 7, 4, 2)
    QOL_SOLVE_MFMA(8, 9, 5, 3)
    default: break;
  }
#undef QOL_SOLVE_MFMA
  // by default the real and the quaternion inverse with the matrix in registers;
  // variant 1 (or QOL_SOLVE_FUSED=1) selects the single fused kernel, which also
  // serves orders 9-12 where the epilogue's LDS does not fit
  const char* fu = std::getenv("QOL_SOLVE_FUSED");
  const bool fused = variant == 1 || (variant == 0 && fu && fu[0] == '1');
  if (h <= 36 && !fused) {
    // lane grid per order (rows RL x RA, slots CL x CB)
#define QOL_QINV(RL_, RA_, CL_, CB_, NW_)                                                                \
  hipLaunchKernelGGL((rrq_real_inverse_kernel<RL_, RA_, CL_, CB_>), dim3(npan), dim3(64), 0, s, T, P, npan, nrrq); \
  hipLaunchKernelGGL((rrq_quat_inverse_kernel<RL_, RA_, CL_, CB_, NW_, true>), dim3(npan), dim3(64 * NW_),       \
                     rrq_quat_inverse_smem(h), s, T, P, npan, nrrq)
    if (h <= 3) { QOL_QINV(3, 1, 3, 1, 1); }
    else if (h <= 6) { QOL_QINV(6, 1, 6, 1, 1); }
    else if (h <= 10) { QOL_QINV(5, 2, 10, 1, 1); }
    else if (h <= 15) { QOL_QINV(3, 5, 15, 1, 1); }
    else if (h <= 21) { QOL_QINV(3, 7, 21, 1, 1); }
    else if (h <= 28) { QOL_QINV(4, 7, 16, 2, 2); }
    else { QOL_QINV(6, 6, 10, 4, 2); }
#undef QOL_QINV
    return;
  }
  // patch orders 9 to 12: the same register-resident inverses over more waves
  // (about 48 quaternion entries per lane), the elimination without dynamic LDS
  // and G finished by rrq_quat_epilogue_kernel, whose 3h^2 doubles (146 KB at
  // h = 78) must fit one workgroup's LDS
  if (h <= 78 && !fused && lds_fits(rrq_quat_inverse_smem(h))) {
    const size_t el = rrq_quat_inverse_smem(h);
    if (el > 64 * 1024)
      (void)hipFuncSetAttribute((const void*)rrq_quat_epilogue_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)el);
#define QOL_QINV_SPLIT(RL_, RA_, CL_, CB_, NW_)                                                                   \
  hipLaunchKernelGGL((rrq_real_inverse_kernel<RL_, RA_, CL_, CB_>), dim3(npan), dim3(64), 0, s, T, P, npan, nrrq); \
  hipLaunchKernelGGL((rrq_quat_inverse_kernel<RL_, RA_, CL_, CB_, NW_, false>), dim3(npan), dim3(64 * NW_), 0, s, T, \
                     P, npan, nrrq);                                                                              \
  hipLaunchKernelGGL(rrq_quat_epilogue_kernel, dim3(npan), dim3(1024), el, s, T, P, npan, nrrq)
    if (h <= 45) { QOL_QINV_SPLIT(8, 6, 8, 6, 3); }
    else if (h <= 55) { QOL_QINV_SPLIT(9, 7, 7, 8, 4); }
    else if (h <= 66) { QOL_QINV_SPLIT(10, 7, 6, 12, 6); }
    else { QOL_QINV_SPLIT(12, 7, 5, 16, 8); }
#undef QOL_QINV_SPLIT
    return;
  }
  if (rpt <= 3 && cpt <= 5) {
    if (lds > 64 * 1024)
      (void)hipFuncSetAttribute((const void*)rrq_panel_solve_kernel<3, 5>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((rrq_panel_solve_kernel<3, 5>), dim3(npan), dim3(256), lds, s, T, P, npan, nrrq);
  } else if (rpt <= 4 && cpt <= 7) {
    (void)hipFuncSetAttribute((const void*)rrq_panel_solve_kernel<4, 7>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((rrq_panel_solve_kernel<4, 7>), dim3(npan), dim3(256), lds, s, T, P, npan, nrrq);
  } else if (rpt <= 5 && cpt <= 10) {
    (void)hipFuncSetAttribute((const void*)rrq_panel_solve_kernel<5, 10>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((rrq_panel_solve_kernel<5, 10>), dim3(npan), dim3(256), lds, s, T, P, npan, nrrq);
  } else {
    throw std::runtime_error("RRQ solve supports patch orders up to 12");
  }
}

// ---------------------------------------------------------------------------
// per-target pieces

// Edge helpers, specialised on the number of edge nodes NQ so that every
// nq-sized array is a register array and the Legendre recurrence constants
// (2k+1)/(k+1), k/(k+1) fold at compile time.  cf = 3 x NQ Legendre
// coefficients of one edge, cf[c*NQ + k].

// X(s), X'(s) of all three coordinates at real s
template <int NQ>
__device__ __forceinline__ void leg3_r(const double* cf, double s, double* val, double* der) {
  double p0 = 1.0, p1 = s, d0 = 0.0, d1 = 1.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    val[c] = cf[c * NQ] + cf[c * NQ + 1] * s;
    der[c] = cf[c * NQ + 1];
  }
#pragma unroll
  for (int k = 1; k + 1 < NQ; ++k) {
    const double c1 = (2.0 * k + 1.0) / (k + 1.0), c2 = (double)k / (k + 1.0);
    const double p2 = c1 * s * p1 - c2 * p0;
    const double d2 = c1 * (p1 + s * d1) - c2 * d0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      val[c] += cf[c * NQ + k + 1] * p2;
      der[c] += cf[c * NQ + k + 1] * d2;
    }
    p0 = p1; p1 = p2; d0 = d1; d1 = d2;
  }
}

// The edges' power-basis coefficients (cf[c * NQ + j] multiplies s^j): value and
// derivative of the three components by Horner's rule with fused multiply-adds
template <int NQ>
__device__ __forceinline__ void pow3_r(const double* cf, double s, double* val, double* der) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double v = cf[c * NQ + NQ - 1], d = 0.0;
#pragma unroll
    for (int k = NQ - 2; k >= 0; --k) {
      d = fma(d, s, v);
      v = fma(v, s, cf[c * NQ + k]);
    }
    val[c] = v;
    der[c] = d;
  }
}

// the same at complex s
template <int NQ>
__device__ __forceinline__ void pow3_c(const double* cf, cd s, cd* val, cd* der) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    cd v = cmk(cf[c * NQ + NQ - 1], 0.0), d = cmk(0.0, 0.0);
#pragma unroll
    for (int k = NQ - 2; k >= 0; --k) {
      // d = d s + v, v = v s + cf_k
      d = cmk(fma(d.re, s.re, fma(-d.im, s.im, v.re)), fma(d.re, s.im, fma(d.im, s.re, v.im)));
      v = cmk(fma(v.re, s.re, fma(-v.im, s.im, cf[c * NQ + k])), fma(v.re, s.im, v.im * s.re));
    }
    val[c] = v;
    der[c] = d;
  }
}

// the Legendre series at complex s
template <int NQ>
__device__ __forceinline__ void leg3_c(const double* cf, cd s, cd* val, cd* der) {
  cd p0 = cmk(1.0, 0.0), p1 = s, d0 = cmk(0.0, 0.0), d1 = cmk(1.0, 0.0);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    val[c] = cadd(cmk(cf[c * NQ], 0.0), cscale(s, cf[c * NQ + 1]));
    der[c] = cmk(cf[c * NQ + 1], 0.0);
  }
#pragma unroll
  for (int k = 1; k + 1 < NQ; ++k) {
    const double c1 = (2.0 * k + 1.0) / (k + 1.0), c2 = (double)k / (k + 1.0);
    const cd p2 = csub(cscale(cmul(s, p1), c1), cscale(p0, c2));
    const cd d2 = csub(cscale(cadd(p1, cmul(s, d1)), c1), cscale(d0, c2));
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      val[c] = cadd(val[c], cscale(p2, cf[c * NQ + k + 1]));
      der[c] = cadd(der[c], cscale(d2, cf[c * NQ + k + 1]));
    }
    p0 = p1; p1 = p2; d0 = d1; d1 = d2;
  }
}

// complex root of |X_e(s) - x|^2 near [-1,1] by Newton from the chord's root
// s = (-C.D + i |C x D|) / D.D.  The imaginary part comes from the cross product:
// the equal form sqrt(|C|^2 |D|^2 - (C.D)^2) cancels for a target near the
// chord's line (near a vertex), can round to 0, and Newton from a real start
// stays on the real axis.
template <int NQ>
__device__ bool edge_root(const double* cf, const double* x, cd& s0) {
  double A[3], B[3], dA[3], dB[3], D[3], C[3];
  pow3_r<NQ>(cf, -1.0, A, dA);
  pow3_r<NQ>(cf, 1.0, B, dB);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    D[c] = 0.5 * (B[c] - A[c]);
    C[c] = 0.5 * (A[c] + B[c]) - x[c];
  }
  const double dd = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
  const double cdv = C[0] * D[0] + C[1] * D[1] + C[2] * D[2];
  double cxd[3];
  cross3(C, D, cxd);
  cd s = cmk(-cdv / dd, sqrt(cxd[0] * cxd[0] + cxd[1] * cxd[1] + cxd[2] * cxd[2]) / dd);
  for (int it = 0; it < 30; ++it) {
    cd v[3], d[3];
    pow3_c<NQ>(cf, s, v, d);
    cd f = cmk(0, 0), fp = cmk(0, 0);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      v[c].re -= x[c];
      f = cadd(f, cmul(v[c], v[c]));
      fp = cadd(fp, cscale(cmul(v[c], d[c]), 2.0));
    }
    const cd ds = cdiv(f, fp);
    s = csub(s, ds);
    if (cabs_(ds) < 1e-15 * (1.0 + cabs_(s))) { s0 = s; return isfinite(s.re) && isfinite(s.im); }
  }
  s0 = s;
  return false;
}

// singularity-swap weights for the 1/|X(s)-x| kernel on one edge; w applies to f_i/R_i
template <int NQ>
__device__ void ssq_w1(const double* cf, const double* tgl, const double* wgl, const double* x, double rho_thr,
                       double* w, cd* root, bool* rok) {
  cd s0;
  const bool ok = edge_root<NQ>(cf, x, s0);
  *root = s0;
  *rok = ok;
  if (!ok || bernstein_rho(s0) >= rho_thr) {
#pragma unroll
    for (int i = 0; i < NQ; ++i) w[i] = wgl[i];
    return;
  }
  const double a = s0.re, b = fabs(s0.im), sg = 1.0 + fabs(a);
  const double x1 = 1.0 - a, x2 = -1.0 - a;
  const double r1 = sqrt(x1 * x1 + b * b), r2 = sqrt(x2 * x2 + b * b);
  double I[NQ];
  if (x2 >= 0.0) I[0] = log((x1 + r1) / (x2 + r2));
  else if (x1 <= 0.0) I[0] = log((r2 - x2) / (r1 - x1));
  else I[0] = log((x1 + r1) * (r2 - x2) / (b * b));
  if (NQ > 1) I[1] = r1 - r2;
  double p1 = x1, p2 = x2;
#pragma unroll
  for (int kk = 2; kk < NQ; ++kk) {
    const double dk = (double)kk;
    I[kk] = (p1 * r1 - p2 * r2 - (dk - 1.0) * b * b * I[kk - 2]) / dk;
    p1 *= x1;
    p2 *= x2;
  }
  double sc = 1.0;
#pragma unroll
  for (int kk = 0; kk < NQ; ++kk) { I[kk] *= sc; sc /= sg; }
  // dual Vandermonde system sum_i lam_i t_i^k = I_k (Bjorck-Pereyra), t_i = (s_i - a)/sg
  double t[NQ];
#pragma unroll
  for (int i = 0; i < NQ; ++i) t[i] = (tgl[i] - a) / sg;
  constexpr int N = NQ - 1;
#pragma unroll
  for (int kk = 0; kk < N; ++kk)
#pragma unroll
    for (int i = N; i > kk; --i) I[i] -= t[kk] * I[i - 1];
#pragma unroll
  for (int kk = N - 1; kk >= 0; --kk) {
#pragma unroll
    for (int i = kk + 1; i <= N; ++i) I[i] /= (t[i] - t[i - kk - 1]);
#pragma unroll
    for (int i = kk; i < N; ++i) I[i] -= I[i + 1];
  }
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    const double ds = tgl[i] - a;
    w[i] = I[i] * sqrt(ds * ds + b * b);
  }
}

// Dirac-string line integral of one edge, adaptive march toward the near point
template <int NQ>
__device__ double solid_angle_edge(const double* cf, const double* tgl, const double* wgl, const double* x,
                                   const double* q, cd s0, bool ok, double rho_thr) {
  double total = 0.0, lo = -1.0, len = 2.0;
  int guard = 0;
  while (lo < 1.0 && guard < 512) {
    ++guard;
    double hi = fmin(1.0, lo + len);
    if (ok) {
      while (hi - lo > 1e-14) {
        const double c = 0.5 * (lo + hi), hh = 0.5 * (hi - lo);
        const cd t0 = cdiv(csub(s0, cmk(c, 0.0)), cmk(hh, 0.0));
        if (bernstein_rho(t0) >= rho_thr) break;
        hi = lo + 0.5 * (hi - lo);
      }
    }
    const double c = 0.5 * (lo + hi), hh = 0.5 * (hi - lo);
#pragma unroll 2
    for (int i = 0; i < NQ; ++i) {
      const double s = c + hh * tgl[i];
      double X[3], dX[3];
      pow3_r<NQ>(cf, s, X, dX);
      const double z[3] = {X[0] - x[0], X[1] - x[1], X[2] - x[2]};
      double zx[3];
      cross3(z, dX, zx);
      const double rz = sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
      const double qz = q[0] * z[0] + q[1] * z[1] + q[2] * z[2];
      const double qzx = q[0] * zx[0] + q[1] * zx[1] + q[2] * zx[2];
      total += hh * wgl[i] * qzx * rcp_nr(rz * (rz + qz));
    }
    len = 2.0 * (hi - lo);
    lo = hi;
  }
  return total;
}

// Gauss-Newton projection of x (original coordinates) onto the patch, started
// at the nearest node (ssx: the panel's nodes, x(3,h)), in three pieces so the
// iterations can be split over two launches: the start, Gauss-Newton steps
// it0 <= it < it1 (true once a step is below 1e-15), and the end (the normal at
// the projection and the signed distance along it).
__device__ void project_start(const double* ssx, const double* uvn, int h, const double* x, double& u, double& v) {
  double best = 1e300;
  u = 1.0 / 3.0;
  v = 1.0 / 3.0;
  for (int i = 0; i < h; ++i) {
    const double d0 = ssx[3 * i] - x[0], d1 = ssx[3 * i + 1] - x[1], d2 = ssx[3 * i + 2] - x[2];
    const double d = d0 * d0 + d1 * d1 + d2 * d2;
    if (d < best) { best = d; u = uvn[2 * i]; v = uvn[2 * i + 1]; }
  }
}

// One Gauss-Newton update of (u, v) from the point y and its tangents, clamped
// to the triangle; true when the step |du| + |dv| (also returned in st) is below
// 1e-15 or below 8 eps |x| sqrt((a11 + a22) / det) (xn = |x|): eight times the
// step that rounding of the residual alone makes, since 1 / the Jacobian's
// smallest singular value is at most sqrt((a11 + a22) / det).  A fixed 1e-15 is
// below that floor on small panels, where the iteration would cycle to the cap.
__device__ __forceinline__ bool gn_step(const double* x, double xn, const double* y, const double* yu,
                                        const double* yv, double& u, double& v, double& st) {
  const double r[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
  const double a11 = yu[0] * yu[0] + yu[1] * yu[1] + yu[2] * yu[2];
  const double a12 = yu[0] * yv[0] + yu[1] * yv[1] + yu[2] * yv[2];
  const double a22 = yv[0] * yv[0] + yv[1] * yv[1] + yv[2] * yv[2];
  const double b1 = -(yu[0] * r[0] + yu[1] * r[1] + yu[2] * r[2]);
  const double b2 = -(yv[0] * r[0] + yv[1] * r[1] + yv[2] * r[2]);
  const double det = a11 * a22 - a12 * a12;
  double un = u + (a22 * b1 - a12 * b2) / det, vn = v + (a11 * b2 - a12 * b1) / det;
  if (un < 0.0) un = 0.0;
  if (vn < 0.0) vn = 0.0;
  if (un + vn > 1.0) {
    const double ex = 0.5 * (un + vn - 1.0);
    un -= ex; vn -= ex;
    if (un < 0.0) { vn += un; un = 0.0; }
    if (vn < 0.0) { un += vn; vn = 0.0; }
  }
  st = fabs(un - u) + fabs(vn - v);
  u = un; v = vn;
  return st < fmax(1e-15, 8.0 * 2.220446049250313e-16 * xn * sqrt((a11 + a22) / det));
}

__device__ bool project_steps(const double* cx, const double* __restrict__ kt, int h, int korder, const double* x,
                              double& u, double& v, int it0, int it1) {
  const double xn = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  double y[3], yu[3], yv[3], st;
  for (int it = it0; it < it1; ++it) {
    geom_eval(cx, kt, h, korder, u, v, y, yu, yv);
    if (gn_step(x, xn, y, yu, yv, u, v, st)) return true;
  }
  return false;
}

// A projection that ends at the step cap: stat[0] counts them, the 8 bytes at
// stat + 2 hold the largest last step (a non-negative double's bits order as its
// value, so an unsigned maximum of the bits is the maximum of the values).
__device__ __forceinline__ void note_capped(int* stat, double st) {
  atomicAdd(stat, 1);
  atomicMax(reinterpret_cast<unsigned long long*>(stat + 2), (unsigned long long)__double_as_longlong(st));
}

__device__ double project_end(const double* cx, const double* __restrict__ kt, int h, int korder, const double* x,
                              double u, double v, double* nstar) {
  double y[3], yu[3], yv[3];
  geom_eval(cx, kt, h, korder, u, v, y, yu, yv);
  double nn[3];
  cross3(yu, yv, nn);
  const double l = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
  for (int c = 0; c < 3; ++c) nstar[c] = nn[c] / l;
  return (x[0] - y[0]) * nstar[0] + (x[1] - y[1]) * nstar[1] + (x[2] - y[2]) * nstar[2];
}

// ---------------------------------------------------------------------------
// Per-target work in three launches over the RRQ pair list (label 5, grouped by
// panel; r = position in that list):
//   rrq_project_kernel  one thread per pair: projection onto the patch, side,
//                       Dirac-string direction q (frame)            -> pq[r][4]
//   rrq_edge_kernel     one thread per (pair, edge): complex root, SSQ weights
//                       for the 1/r line kernel and the edge's solid-angle
//                       line integral                              -> w1[r][3nq], oe[r][3]
//   rrq_target_kernel   one workgroup per panel, pairs in tiles: boundary
//                       moments, degree recursion, weights operator.

__device__ __forceinline__ int rrq_list_base(const int* ltot) { return ltot[0] + ltot[1] + ltot[2] + ltot[3]; }

__global__ void zero_ints_kernel(int* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Projection in two launches: every pair takes up to kPI Gauss-Newton steps;
// the few that have not converged (a list built with an atomic counter) continue
// in the second launch from the saved (u, v), so a wave no longer runs as long
// as its slowest pair.  The step sequence of every pair is the single loop's,
// so the results are the same.  A pair that has not met the stop test of
// gn_step after kPMax steps keeps its last iterate, as in the single loop, and is
// counted in `nfail` (with the largest such last step) so that the caller can
// see it.
constexpr int kPI = 8, kPMax = 40;

__device__ void project_write(RRQTables T, RRQPanels P, int k, const double* x, double u, double v, int exterior,
                              double side_tol, double* o) {
  const int h = T.h;
  const double* fr = P.frame + (size_t)k * 16;
  double ns[3];
  const double d = project_end(P.cx + (size_t)k * 3 * h, T.ktab, h, T.p - 1, x, u, v, ns);
  double side;
  if (fabs(d) > side_tol * fr[14]) side = d > 0.0 ? 1.0 : -1.0;
  else side = (exterior ? 1.0 : -1.0) * fr[13];
  const double qo[3] = {-side * ns[0], -side * ns[1], -side * ns[2]};
  double q[3];
  rotate(fr, qo, q);
  o[0] = q[0]; o[1] = q[1]; o[2] = q[2]; o[3] = side;
}

__global__ void __launch_bounds__(256) rrq_project_kernel(RRQTables T, RRQPanels P, const int* __restrict__ lpairs,
                                                          const int* __restrict__ ltot,
                                                          const int* __restrict__ ppanel,
                                                          const int* __restrict__ tidx,
                                                          const double* __restrict__ tx,
                                                          const double* __restrict__ sx, int exterior,
                                                          double side_tol, double* __restrict__ pq,
                                                          int* __restrict__ rest, int* __restrict__ nrest) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= ltot[4]) return;
  const int pair = lpairs[rrq_list_base(ltot) + r];
  const int k = ppanel[pair], h = T.h;
  const size_t g = (size_t)tidx[pair];
  const double x[3] = {tx[3 * g], tx[3 * g + 1], tx[3 * g + 2]};
  double u, v;
  project_start(sx + (size_t)k * 3 * h, T.uvn, h, x, u, v);
  double* o = pq + 4 * (size_t)r;
  if (project_steps(P.cx + (size_t)k * 3 * h, T.ktab, h, T.p - 1, x, u, v, 0, kPI)) {
    project_write(T, P, k, x, u, v, exterior, side_tol, o);
  } else {
    o[0] = u; o[1] = v;  // the state, for the second launch
    rest[atomicAdd(nrest, 1)] = r;
  }
}

__global__ void __launch_bounds__(256) rrq_project_rest_kernel(RRQTables T, RRQPanels P,
                                                               const int* __restrict__ lpairs,
                                                               const int* __restrict__ ltot,
                                                               const int* __restrict__ ppanel,
                                                               const int* __restrict__ tidx,
                                                               const double* __restrict__ tx, int exterior,
                                                               double side_tol, double* __restrict__ pq,
                                                               const int* __restrict__ rest,
                                                               const int* __restrict__ nrest,
                                                               int* __restrict__ nfail) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= *nrest) return;
  const int r = rest[gid];
  const int pair = lpairs[rrq_list_base(ltot) + r];
  const int k = ppanel[pair], h = T.h;
  const size_t g = (size_t)tidx[pair];
  const double x[3] = {tx[3 * g], tx[3 * g + 1], tx[3 * g + 2]};
  double* o = pq + 4 * (size_t)r;
  double u = o[0], v = o[1];
  const double* cx = P.cx + (size_t)k * 3 * h;
  const double xn = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  double y[3], yu[3], yv[3], st = 0.0;
  bool done = false;
  for (int it = kPI; it < kPMax && !done; ++it) {
    geom_eval(cx, T.ktab, h, T.p - 1, u, v, y, yu, yv);
    done = gn_step(x, xn, y, yu, yv, u, v, st);
  }
  if (!done) note_capped(nfail, st);
  project_write(T, P, k, x, u, v, exterior, side_tol, o);
}

// The second launch runs few pairs and so is bound by each step's latency: for
// the orders up to 8 it holds the panel's coefficients in registers and unrolls
// the evaluation (the same operations in the same order as geom_eval, so the
// same results).  In the first launch, which is bound by throughput, the same
// unrolled form measured slower (it needs more registers, so fewer waves).
template <int N>
__device__ __forceinline__ void geom_eval_reg(const double (&cr)[3][N * (N + 1) / 2], const double* __restrict__ kt,
                                              double u, double v, double* x, double* xu, double* xv) {
  const double a = 2.0 * u + v - 1.0, b = 1.0 - v, bb = b * b;
#pragma unroll
  for (int c = 0; c < 3; ++c) { x[c] = 0.0; xu[c] = 0.0; xv[c] = 0.0; }
  const double xx = 2.0 * v - 1.0;
  double Q = 1.0, Qu = 0.0, Qv = 0.0, Qm = 0.0, Qum = 0.0, Qvm = 0.0;
#pragma unroll
  for (int am = 0; am < N; ++am) {
    if (am == 1) {
      Qm = Q; Qum = Qu; Qvm = Qv;
      Q = a; Qu = 2.0; Qv = 1.0;
    } else if (am >= 2) {
      const double e1 = kt[2 * am], e2 = kt[2 * am + 1];
      const double Qn = e1 * a * Q - e2 * bb * Qm;
      const double Qun = e1 * (2.0 * Q + a * Qu) - e2 * bb * Qum;
      const double Qvn = e1 * (Q + a * Qv) - e2 * (bb * Qvm - 2.0 * b * Qm);
      Qm = Q; Qum = Qu; Qvm = Qv;
      Q = Qn; Qu = Qun; Qv = Qvn;
    }
    const double al = 2.0 * am + 1.0;
    double p = 1.0, dp = 0.0, pm = 1.0, dpm = 0.0;
#pragma unroll
    for (int bd = 0; am + bd < N; ++bd) {
      // table entry of (am, bd): terms are stored am-major after the 2N factors
      const double* tt = kt + 2 * N + 4 * (am * N - am * (am - 1) / 2 + bd);
      if (bd == 1) {
        p = 0.5 * ((al + 2.0) * xx + al);
        dp = al + 2.0;
      } else if (bd >= 2) {
        const double w = tt[1] + tt[2] * xx;
        const double pn = w * p - tt[3] * pm;
        const double dpn = w * dp + 2.0 * tt[2] * p - tt[3] * dpm;
        pm = p; p = pn; dpm = dp; dp = dpn;
      }
      const int d = am + bd, idx = d * (d + 1) / 2 + bd;
      const double sc = tt[0];
      const double f = sc * Q * p;
#pragma unroll
      for (int c = 0; c < 3; ++c) x[c] += f * cr[c][idx];
      const double fu = sc * Qu * p, fv = sc * (Qv * p + Q * dp);
#pragma unroll
      for (int c = 0; c < 3; ++c) { xu[c] += fu * cr[c][idx]; xv[c] += fv * cr[c][idx]; }
    }
  }
}

template <int N>
__global__ void __launch_bounds__(64) rrq_project_rest_reg_kernel(RRQTables T, RRQPanels P,
                                                                  const int* __restrict__ lpairs,
                                                                  const int* __restrict__ ltot,
                                                                  const int* __restrict__ ppanel,
                                                                  const int* __restrict__ tidx,
                                                                  const double* __restrict__ tx, int exterior,
                                                                  double side_tol, double* __restrict__ pq,
                                                                  const int* __restrict__ rest,
                                                                  const int* __restrict__ nrest,
                                                                  int* __restrict__ nfail) {
  constexpr int H = N * (N + 1) / 2;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= *nrest) return;
  const int r = rest[gid];
  const int pair = lpairs[rrq_list_base(ltot) + r];
  const int k = ppanel[pair];
  const size_t g = (size_t)tidx[pair];
  const double x[3] = {tx[3 * g], tx[3 * g + 1], tx[3 * g + 2]};
  const double* cx = P.cx + (size_t)k * 3 * H;
  double cr[3][H];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int i = 0; i < H; ++i) cr[c][i] = cx[c * H + i];
  double* o = pq + 4 * (size_t)r;
  double u = o[0], v = o[1];
  const double xn = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  double y[3], yu[3], yv[3], st = 0.0;
  bool done = false;
  for (int it = kPI; it < kPMax && !done; ++it) {
    geom_eval_reg<N>(cr, T.ktab, u, v, y, yu, yv);
    done = gn_step(x, xn, y, yu, yv, u, v, st);
  }
  if (!done) note_capped(nfail, st);
  project_write(T, P, k, x, u, v, exterior, side_tol, o);
}

void launch_rrq_project(RRQTables T, RRQPanels P, const int* lpairs, const int* ltot, const int* ppanel,
                        const int* tidx, const double* tx, const double* sx, int exterior, double side_tol,
                        double* pq, int* order, int* nfail, int64_t nnz, hipStream_t s) {
  int* nrest = order;      // [1]
  int* rest = order + 1;   // [nnz]
  const dim3 grid((unsigned)((nnz + 255) / 256)), blk(256);
  hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, s, nrest, 1);
  hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, s, nfail, 4);
  hipLaunchKernelGGL(rrq_project_kernel, grid, blk, 0, s, T, P, lpairs, ltot, ppanel, tidx, tx, sx, exterior,
                     side_tol, pq, rest, nrest);
  const dim3 rgrid((unsigned)((nnz + 63) / 64)), rblk(64);
#define QOL_PROJ(NN)                                                                                            \
  case NN:                                                                                                      \
    hipLaunchKernelGGL(rrq_project_rest_reg_kernel<NN>, rgrid, rblk, 0, s, T, P, lpairs, ltot, ppanel, tidx, tx, \
                       exterior, side_tol, pq, (const int*)rest, (const int*)nrest, nfail);                      \
    break;
  switch (T.p) {
    QOL_PROJ(2) QOL_PROJ(3) QOL_PROJ(4) QOL_PROJ(5) QOL_PROJ(6) QOL_PROJ(7) QOL_PROJ(8)
    default:
      hipLaunchKernelGGL(rrq_project_rest_kernel, grid, blk, 0, s, T, P, lpairs, ltot, ppanel, tidx, tx, exterior,
                         side_tol, pq, (const int*)rest, (const int*)nrest, nfail);
  }
#undef QOL_PROJ
}

// per (pair, edge): complex root and SSQ weights -> w1[r][3NQ], root[r][3][3] (re, im, ok)
template <int NQ>
__global__ void __launch_bounds__(256) rrq_root_kernel(RRQTables T, RRQPanels P, const int* __restrict__ lpairs,
                                                       const int* __restrict__ ltot, const int* __restrict__ ppanel,
                                                       const int* __restrict__ tidx, const double* __restrict__ tx,
                                                       double* __restrict__ w1, double* __restrict__ root) {
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= 3 * ltot[4]) return;
  const int r = it / 3, e = it % 3;
  const int pair = lpairs[rrq_list_base(ltot) + r];
  const int k = ppanel[pair];
  const double* fr = P.frame + (size_t)k * 16;
  const size_t g = (size_t)tidx[pair];
  const double x[3] = {tx[3 * g], tx[3 * g + 1], tx[3 * g + 2]};
  double X[3];
  to_frame(fr, x, X);
  const double* cf = P.edge + (size_t)k * 9 * NQ + e * 3 * NQ;
  const double rho_thr = pow(8.0, 8.0 / (double)NQ);
  cd s0;
  bool ok;
  ssq_w1<NQ>(cf, T.tgl, T.wgl, X, rho_thr, w1 + (size_t)r * 3 * NQ + e * NQ, &s0, &ok);
  double* o = root + 3 * (size_t)it;
  o[0] = s0.re; o[1] = s0.im; o[2] = ok ? 1.0 : 0.0;
}

// per (pair, edge): the edge's Dirac-string line integral of the solid angle -> oe[r][3]
template <int NQ>
__global__ void __launch_bounds__(256) rrq_solid_kernel(RRQTables T, RRQPanels P, const int* __restrict__ lpairs,
                                                        const int* __restrict__ ltot, const int* __restrict__ ppanel,
                                                        const int* __restrict__ tidx, const double* __restrict__ tx,
                                                        const double* __restrict__ pq, const double* __restrict__ root,
                                                        const int* __restrict__ perm, double* __restrict__ oe) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= 3 * ltot[4]) return;
  const int it = perm[gid];  // items in order of their cost bucket
  const int r = it / 3, e = it % 3;
  const int pair = lpairs[rrq_list_base(ltot) + r];
  const int k = ppanel[pair];
  const double* fr = P.frame + (size_t)k * 16;
  const size_t g = (size_t)tidx[pair];
  const double x[3] = {tx[3 * g], tx[3 * g + 1], tx[3 * g + 2]};
  double X[3];
  to_frame(fr, x, X);
  const double* cf = P.edge + (size_t)k * 9 * NQ + e * 3 * NQ;
  const double rho_thr = pow(8.0, 8.0 / (double)NQ);
  const double* rt = root + 3 * (size_t)it;
  const double q[3] = {pq[4 * (size_t)r], pq[4 * (size_t)r + 1], pq[4 * (size_t)r + 2]};
  oe[it] = solid_angle_edge<NQ>(cf, T.tgl, T.wgl, X, q, cmk(rt[0], rt[1]), rt[2] != 0.0, rho_thr);
}

// The solid-angle march of a (pair, edge) item takes about 2 log2(1/d) intervals,
// d the distance of the edge's complex root from [-1, 1] (one interval without a
// root).  Items are grouped by that cost -- a counting sort into kSB buckets,
// costliest first -- so the lanes of a wave run similar marches.  Each item
// writes only its own result, so the order changes nothing but the time.
constexpr int kSB = 64;

__device__ __forceinline__ int solid_bucket(const double* rt) {
  if (rt[2] == 0.0) return 0;
  const double dx = fmax(fabs(rt[0]) - 1.0, 0.0), im = fabs(rt[1]);
  const double d = sqrt(dx * dx + im * im);
  if (!(d > 0.0)) return kSB - 1;  // on the segment, or not a number
  int e;
  (void)frexp(d, &e);  // d in [2^(e-1), 2^e)
  return min(max(2 - e, 0), kSB - 1);
}

// counts per bucket: a histogram per workgroup in LDS, then one global add per
// non-empty bucket (kSI items per thread)
constexpr int kSI = 4;
__global__ void __launch_bounds__(256) solid_count_kernel(const int* __restrict__ ltot, const double* __restrict__ root,
                                                          int* __restrict__ cnt, int* __restrict__ key) {
  __shared__ int lh[kSB];
  for (int b = threadIdx.x; b < kSB; b += blockDim.x) lh[b] = 0;
  __syncthreads();
  const int n = 3 * ltot[4];
  const int base = blockIdx.x * blockDim.x * kSI + threadIdx.x;
#pragma unroll
  for (int u = 0; u < kSI; ++u) {
    const int it = base + u * blockDim.x;
    if (it < n) {
      const int b = solid_bucket(root + 3 * (size_t)it);
      key[it] = b;
      atomicAdd(lh + b, 1);
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kSB; b += blockDim.x)
    if (lh[b]) atomicAdd(cnt + b, lh[b]);
}

// bucket starts, costliest bucket first (one thread); cnt[kSB..2kSB) receive them
__global__ void solid_offsets_kernel(int* __restrict__ cnt) {
  if (threadIdx.x != 0) return;
  int s = 0;
  for (int b = kSB - 1; b >= 0; --b) {
    cnt[kSB + b] = s;
    s += cnt[b];
  }
}

// the permutation: ranks within the workgroup from LDS, one global range per
// non-empty bucket per workgroup
__global__ void __launch_bounds__(256) solid_scatter_kernel(const int* __restrict__ ltot, const int* __restrict__ key,
                                                            int* __restrict__ cnt, int* __restrict__ perm) {
  __shared__ int lh[kSB], lb[kSB];
  for (int b = threadIdx.x; b < kSB; b += blockDim.x) lh[b] = 0;
  __syncthreads();
  const int n = 3 * ltot[4];
  const int base = blockIdx.x * blockDim.x * kSI + threadIdx.x;
  int bk[kSI], rk[kSI];
#pragma unroll
  for (int u = 0; u < kSI; ++u) {
    const int it = base + u * blockDim.x;
    bk[u] = it < n ? key[it] : -1;
    rk[u] = bk[u] >= 0 ? atomicAdd(lh + bk[u], 1) : 0;
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kSB; b += blockDim.x) lb[b] = lh[b] ? atomicAdd(cnt + kSB + b, lh[b]) : 0;
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kSI; ++u)
    if (bk[u] >= 0) perm[lb[bk[u]] + rk[u]] = base + u * blockDim.x;
}

size_t rrq_edge_order_ints(int64_t nnz) { return (size_t)2 * kSB + (size_t)6 * nnz; }

void launch_rrq_edge(RRQTables T, RRQPanels P, const int* lpairs, const int* ltot, const int* ppanel,
                     const int* tidx, const double* tx, const double* pq, double* w1, double* root, double* oe,
                     int* order, int64_t nnz, hipStream_t s, hipEvent_t mid) {
  const dim3 grid((unsigned)((3 * nnz + 255) / 256)), blk(256);
  const dim3 gsort((unsigned)((3 * nnz + 256 * kSI - 1) / (256 * kSI)));
  int* cnt = order;                  // [2 kSB]: counts, then bucket starts
  int* key = order + 2 * kSB;        // [3 nnz]
  int* perm = key + 3 * nnz;         // [3 nnz]
#define QOL_EDGE(NQ_)                                                                                      \
  case NQ_:                                                                                                \
    hipLaunchKernelGGL(rrq_root_kernel<NQ_>, grid, blk, 0, s, T, P, lpairs, ltot, ppanel, tidx, tx, w1, root); \
    if (mid) (void)hipEventRecord(mid, s);                                                                 \
    hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(2 * kSB), 0, s, cnt, 2 * kSB);                      \
    hipLaunchKernelGGL(solid_count_kernel, gsort, blk, 0, s, ltot, (const double*)root, cnt, key);        \
    hipLaunchKernelGGL(solid_offsets_kernel, dim3(1), dim3(64), 0, s, cnt);                               \
    hipLaunchKernelGGL(solid_scatter_kernel, gsort, blk, 0, s, ltot, (const int*)key, cnt, perm);         \
    hipLaunchKernelGGL(rrq_solid_kernel<NQ_>, grid, blk, 0, s, T, P, lpairs, ltot, ppanel, tidx, tx, pq, root,  \
                       (const int*)perm, oe);                                                             \
    break;
  switch (T.nq) {
    QOL_EDGE(4) QOL_EDGE(5) QOL_EDGE(6) QOL_EDGE(7) QOL_EDGE(8) QOL_EDGE(9) QOL_EDGE(10) QOL_EDGE(11)
    QOL_EDGE(12) QOL_EDGE(13) QOL_EDGE(14)
    default: throw std::runtime_error("RRQ edge kernels: patch orders 2..12 only");
  }
#undef QOL_EDGE
}

// LDS plan of rrq_target_kernel: fixed panel data plus per-target slots.  Per
// target: the frame point (6), w1 (nb), the raw moments, later the moment vector
// (rstr doubles), solid angle and Lambda (4), and the node weights, later the
// boundary moments (mstr complex).
__host__ __device__ inline int rrq_raw_doubles(int p) { return 2 * (3 * (p * (p + 1) / 2) + 9 * ((p - 1) * p / 2)); }
__host__ __device__ inline int rrq_rstr(int p, int h) { const int r = rrq_raw_doubles(p); return r > 5 * h ? r : 5 * h; }
__host__ __device__ inline int rrq_mstr(int nb, int ncoef) { return 6 * nb > 5 * ncoef ? 6 * nb : 5 * ncoef; }
__host__ __device__ inline size_t rrq_target_fixed(int h, int nq) {
  return (size_t)(16 + 6 * 3 * nq) * 8;
}
__host__ __device__ inline size_t rrq_target_per(int p, int h, int nq, int ncoef) {
  const int nb = 3 * nq;
  return (size_t)(6 + nb + rrq_rstr(p, h) + 4) * 8 + (size_t)rrq_mstr(nb, ncoef) * 16;
}
__host__ __device__ int rrq_tile(int p, int h, int nq, int ncoef) {
  // about 48 KB per workgroup keeps three workgroups on a CU
  const size_t budget = 48 * 1024, fixed = rrq_target_fixed(h, nq), per = rrq_target_per(p, h, nq, ncoef);
  int t = budget > fixed + per ? (int)((budget - fixed) / per) : 1;
  return t > 16 ? 16 : (t < 1 ? 1 : t);
}
size_t rrq_target_smem(int p, int h, int nq, int ncoef) {
  return rrq_target_fixed(h, nq) + rrq_target_per(p, h, nq, ncoef) * rrq_tile(p, h, nq, ncoef);
}

// entry (n, m) of a raw-moment array of real weights, any m: conjugation and the
// sign (-1)^m for m < 0, zero outside |m| <= n (selects, no branches)
__device__ __forceinline__ cd rget(const cd* __restrict__ A, int n, int m) {
  const int am = m < 0 ? -m : m;
  const bool ok = n >= 0 && am <= n;
  const cd v = A[ok ? hidx(n, am) : 0];
  const double sr = (m < 0 && (am & 1)) ? -1.0 : 1.0, si = m < 0 ? -sr : sr;
  return ok ? cmk(sr * v.re, si * v.im) : cmk(0.0, 0.0);
}

// The boundary moments of one target, chi_n^m = sum_b a . grad H'_n^m and
// alpha_n^m = sum_b z (x) V_n^m with V_n^m = n [dX_Z grad H'_{n-1}^m + i/2 (dW
// grad H'_{n-1}^{m-1} + conj(dW) grad H'_{n-1}^{m+1})] (a = (w1/R)(z x dX), z
// scaled by w1/R, (x) the quaternion product (-z.V, z x V)), from raw moments
// of real node weights: R_c = sum_b a_c H'_{n'}^{m'} (n' < p) and Z_cj = sum_b
// z_c dX_j H'_{n'}^{m'} (n' < p - 1).  The gradient is linear with constant
// coefficients in the harmonics of one degree lower (harm_grad), and the
// weights are real, so the conjugation rule for negative orders carries over to
// the raw moments.
__device__ __forceinline__ void moments_from_raw(const cd* __restrict__ R, int p, int n, int m, cd* M) {
  const int nc1 = p * (p + 1) / 2, nc2 = (p - 1) * p / 2;
  for (int c = 0; c < 5; ++c) M[c] = cmk(0.0, 0.0);
  if (n < 1) return;
  // chi = sum_c (grad of R_c at degree n)_c
  {
    const cd a = rget(R, n - 1, m), lo = rget(R + nc1, n - 1, m - 1), hi = rget(R + nc1, n - 1, m + 1);
    const cd lo2 = rget(R + 2 * nc1, n - 1, m - 1), hi2 = rget(R + 2 * nc1, n - 1, m + 1);
    M[0] = cscale(cadd(cadd(a, cscale(cmuli(cadd(lo, hi)), 0.5)), cscale(csub(hi2, lo2), 0.5)), (double)n);
  }
  if (n < 2) return;
  // U_cd = sum_b z_c V_d; alpha = (-(U_00 + U_11 + U_22), U_12 - U_21, U_20 - U_02, U_01 - U_10)
  // added into M (LDS) term by term
  const int n2 = n - 2;
  const double g = 0.5 * (double)(n - 1);
#pragma unroll 1
  for (int c = 0; c < 3; ++c) {
    const cd* Z = R + 3 * nc1 + 3 * c * nc2;
    cd z0[3], z1[5], z2[5];
#pragma unroll
    for (int j = 0; j < 3; ++j) z0[j] = rget(Z, n2, m - 1 + j);
#pragma unroll
    for (int j = 0; j < 5; ++j) { z1[j] = rget(Z + nc2, n2, m - 2 + j); z2[j] = rget(Z + 2 * nc2, n2, m - 2 + j); }
    // with G the gradient at degree n-1 (harm_grad) and g = (n-1)/2:
    // U_cd = n (G(Z_c0)^m + i/2 [G(Z_c1)^{m-1} + G(Z_c1)^{m+1} + i (G(Z_c2)^{m-1} -
    // G(Z_c2)^{m+1})])_d = n g (t0 + i/2 a - b/2)_d, the components collected:
    //   t0 = (2 z0_m, i (z0_{m-1} + z0_{m+1}), z0_{m+1} - z0_{m-1})
    //   a  = (2 (z1_{m-1} + z1_{m+1}), i (z1_{m-2} + 2 z1_m + z1_{m+2}), z1_{m+2} - z1_{m-2})
    //   b  = (2 (z2_{m-1} - z2_{m+1}), i (z2_{m-2} - z2_{m+2}), 2 z2_m - z2_{m-2} - z2_{m+2})
    const double f = (double)n * g;
    cd u[3];
    {
      const cd t0 = cscale(z0[1], 2.0), a = cscale(cadd(z1[1], z1[3]), 2.0), b = cscale(csub(z2[1], z2[3]), 2.0);
      u[0] = cscale(cadd(t0, csub(cscale(cmuli(a), 0.5), cscale(b, 0.5))), f);
    }
    {
      const cd t0 = cmuli(cadd(z0[0], z0[2])), a = cmuli(cadd(cadd(z1[0], z1[4]), cscale(z1[2], 2.0)));
      const cd b = cmuli(csub(z2[0], z2[4]));
      u[1] = cscale(cadd(t0, csub(cscale(cmuli(a), 0.5), cscale(b, 0.5))), f);
    }
    {
      const cd t0 = csub(z0[2], z0[0]), a = csub(z1[4], z1[0]), b = csub(cscale(z2[2], 2.0), cadd(z2[0], z2[4]));
      u[2] = cscale(cadd(t0, csub(cscale(cmuli(a), 0.5), cscale(b, 0.5))), f);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d == c) M[1] = csub(M[1], u[d]);
      else if ((d - c + 3) % 3 == 1) M[1 + (c + 2) % 3 + 1] = cadd(M[1 + (c + 2) % 3 + 1], u[d]);  // U_01, U_12, U_20
      else M[1 + (c + 1) % 3 + 1] = csub(M[1 + (c + 1) % 3 + 1], u[d]);                            // U_02, U_10, U_21
    }
  }
}

__global__ void __launch_bounds__(256) rrq_target_kernel(RRQTables T, RRQPanels P, const int* __restrict__ lpairs,
                                                         const int* __restrict__ lstart, const int* __restrict__ ltot,
                                                         int npan, const int* __restrict__ tidx,
                                                         const double* __restrict__ tx, const int* __restrict__ nrrq,
                                                         const double* __restrict__ w1g, const double* __restrict__ oeg,
                                                         double* __restrict__ vt, double* __restrict__ omega_out,
                                                         double* __restrict__ ias_out) {
  const int k = blockIdx.x;
  const int nmine = nrrq[k];
  if (nmine == 0) return;
  const int h = T.h, nq = T.nq, p = T.p, nb = 3 * nq, ncoef = T.ncoef, H4 = 4 * h, H5 = 5 * h;
  const int nc1 = p * (p + 1) / 2, nc2 = (p - 1) * p / 2;
  const int TT = rrq_tile(p, h, nq, ncoef), rstr = rrq_rstr(p, h), mstr = rrq_mstr(nb, ncoef);
  extern __shared__ double sm[];
  double* fr = sm;                        // 16
  double* bx = fr + 16;                   // [nb][6]
  double* tX = bx + 6 * nb;               // [TT][6]   frame X, original x
  double* w1 = tX + 6 * TT;               // [TT][nb]
  double* rv = w1 + (size_t)nb * TT;      // [TT][rstr]: raw moments, later the moment vector v
  double* tq = rv + (size_t)rstr * TT;    // [TT][4]: solid angle, Lambda
  cd* mom = (cd*)(tq + 4 * TT);           // [TT][mstr]: node weights [nb][12], later moments [ncoef][5]
  __shared__ int plist[16];
  for (int i = threadIdx.x; i < 16; i += blockDim.x) fr[i] = P.frame[(size_t)k * 16 + i];
  for (int i = threadIdx.x; i < 6 * nb; i += blockDim.x) bx[i] = P.bX[(size_t)k * nb * 6 + i];
  const double sigma = P.frame[(size_t)k * 16 + 13], alpha = P.frame[(size_t)k * 16 + 12];
  const cd* Hb = (const cd*)P.basis + (size_t)k * nb * nc1;
  const int r0 = lstart[4 * (size_t)npan + k];
  const int* mine = lpairs + rrq_list_base(ltot) + r0;
  for (int tile0 = 0; tile0 < nmine; tile0 += TT) {
    const int nt = min(TT, nmine - tile0);
    __syncthreads();
    for (int t = threadIdx.x; t < nt; t += blockDim.x) plist[t] = mine[tile0 + t];
    for (int i = threadIdx.x; i < nt * nb; i += blockDim.x)
      w1[i] = w1g[(size_t)(r0 + tile0) * nb + i];
    for (int t = threadIdx.x; t < nt; t += blockDim.x) {
      const double* o = oeg + 3 * (size_t)(r0 + tile0 + t);
      tq[4 * t] = o[0] + o[1] + o[2];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nt; t += blockDim.x) {
      const size_t g = (size_t)tidx[plist[t]];
      const double x[3] = {tx[3 * g], tx[3 * g + 1], tx[3 * g + 2]};
      to_frame(fr, x, tX + 6 * t);
      tX[6 * t + 3] = x[0]; tX[6 * t + 4] = x[1]; tX[6 * t + 5] = x[2];
    }
    __syncthreads();
    // node weights: a = (w1/R)(z x dX) and (w1/R) z_c dX_j
    for (int it = threadIdx.x; it < nt * nb; it += blockDim.x) {
      const int t = it / nb, ib = it % nb;
      const double* Y = bx + 6 * ib;
      const double z[3] = {Y[0] - tX[6 * t], Y[1] - tX[6 * t + 1], Y[2] - tX[6 * t + 2]};
      const double wr = w1[(size_t)t * nb + ib] / sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
      w1[(size_t)t * nb + ib] = wr;  // w1/R from here on
      double zx[3];
      cross3(z, Y + 3, zx);
      double* o = (double*)(mom + (size_t)t * mstr) + 12 * ib;
      for (int c = 0; c < 3; ++c) o[c] = wr * zx[c];
      for (int c = 0; c < 3; ++c)
        for (int j = 0; j < 3; ++j) o[3 + 3 * c + j] = (wr * z[c]) * Y[3 + j];
    }
    __syncthreads();
    // Lambda = sum (w1/R) dX per (target, component), into the spare slots of tq
    for (int it = threadIdx.x; it < nt * 3; it += blockDim.x) {
      const int t = it / 3, c = it % 3;
      double l = 0.0;
      for (int ib = 0; ib < nb; ++ib) l += w1[(size_t)t * nb + ib] * bx[6 * ib + 3 + c];
      tq[4 * t + 1 + c] = l;
    }
    // raw moments, one thread per (target, harmonic of degree < p): the three
    // a-weights always, the nine z dX weights for degrees < p - 1; consecutive
    // threads read consecutive harmonics of a node
    for (int it = threadIdx.x; it < nt * nc1; it += blockDim.x) {
      const int t = it / nc1, kk = it % nc1;
      const bool zp = kk < nc2;
      const double* wt = (const double*)(mom + (size_t)t * mstr);
      double acc[24];
#pragma unroll
      for (int c = 0; c < 24; ++c) acc[c] = 0.0;
      // the harmonics of kHB nodes are loaded together (loads in flight)
      constexpr int kHB = 6;
      for (int ib0 = 0; ib0 < nb; ib0 += kHB) {
        cd hb[kHB];
#pragma unroll
        for (int u = 0; u < kHB; ++u) hb[u] = Hb[(size_t)min(ib0 + u, nb - 1) * nc1 + kk];
#pragma unroll
        for (int u = 0; u < kHB; ++u) {
          if (ib0 + u >= nb) break;
          const cd hv = hb[u];
          const double* w = wt + 12 * (ib0 + u);
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            acc[2 * q] = fma(w[q], hv.re, acc[2 * q]);
            acc[2 * q + 1] = fma(w[q], hv.im, acc[2 * q + 1]);
          }
          if (zp) {
#pragma unroll
            for (int q = 3; q < 12; ++q) {
              acc[2 * q] = fma(w[q], hv.re, acc[2 * q]);
              acc[2 * q + 1] = fma(w[q], hv.im, acc[2 * q + 1]);
            }
          }
        }
      }
      cd* R = (cd*)(rv + (size_t)t * rstr);
#pragma unroll
      for (int q = 0; q < 3; ++q) R[q * nc1 + kk] = cmk(acc[2 * q], acc[2 * q + 1]);
      if (zp) {
#pragma unroll
        for (int q = 3; q < 12; ++q) R[3 * nc1 + (q - 3) * nc2 + kk] = cmk(acc[2 * q], acc[2 * q + 1]);
      }
    }
    __syncthreads();
    // boundary moments chi (component 0) and alpha (1..4) per coefficient
    for (int it = threadIdx.x; it < nt * ncoef; it += blockDim.x) {
      const int t = it / ncoef, kk = it % ncoef;
      int n = 0;
      while (hidx(n + 1, 0) <= kk) ++n;
      moments_from_raw((const cd*)(rv + (size_t)t * rstr), p, n, kk - hidx(n, 0), mom + (size_t)t * mstr + 5 * kk);
    }
    __syncthreads();
    // seeds: W~_00 = solid angle, degree-1 Q = Lambda (x) grad H'_1^m with
    // Lambda = -omega + sum (w1/R) dX
    for (int t = threadIdx.x; t < nt; t += blockDim.x) {
      cd* M = mom + (size_t)t * mstr;
      const double om = tq[4 * t];
      const double lam[3] = {tq[4 * t + 1], tq[4 * t + 2], tq[4 * t + 3]};
      M[0] = cmk(om, 0.0);
      for (int c = 1; c < 5; ++c) M[c] = cmk(0.0, 0.0);
      for (int m = 0; m <= 1; ++m) {
        cd g[3];
        if (m == 0) { g[0] = cmk(1, 0); g[1] = cmk(0, 0); g[2] = cmk(0, 0); }
        else { g[0] = cmk(0, 0); g[1] = cmk(0, 0.5); g[2] = cmk(-0.5, 0); }
        cd* Q = M + (size_t)hidx(1, m) * 5 + 1;
        Q[0] = cscale(cadd(cadd(cscale(g[0], lam[0]), cscale(g[1], lam[1])), cscale(g[2], lam[2])), -1.0);
        Q[1] = cadd(cscale(g[0], -om), csub(cscale(g[2], lam[1]), cscale(g[1], lam[2])));
        Q[2] = cadd(cscale(g[1], -om), csub(cscale(g[0], lam[2]), cscale(g[2], lam[0])));
        Q[3] = cadd(cscale(g[2], -om), csub(cscale(g[1], lam[0]), cscale(g[0], lam[1])));
      }
    }
    // recursion in degree, all (target, m, component) of one degree in parallel
    for (int n = 1; n <= p; ++n) {
      __syncthreads();
      const double dn = (double)n;
      const int ncomp = n >= 2 ? 5 : 1;  // degree 1 of components 1..4 is the seed
      for (int it = threadIdx.x; it < nt * (n + 1) * ncomp; it += blockDim.x) {
        const int c = it % ncomp, tm = it / ncomp;
        const int t = tm / (n + 1), m = tm % (n + 1);
        cd* M = mom + (size_t)t * mstr;
        const double* X = tX + 6 * t;
        const double xZ = X[0];
        const cd w = cmk(X[1], X[2]), wc = cmk(X[1], -X[2]);
        // entries of degree n-1 at m-1, m, m+1 (negative m by conjugation)
        auto get = [&](int mm, int c) -> cd {
          const int n1 = n - 1;
          if (mm > n1 || mm < -n1) return cmk(0, 0);
          if (mm >= 0) return M[(size_t)hidx(n1, mm) * 5 + c];
          cd v = cconj(M[(size_t)hidx(n1, -mm) * 5 + c]);
          return ((-mm) & 1) ? cmk(-v.re, -v.im) : v;
        };
        cd* E = M + (size_t)hidx(n, m) * 5;
        const cd lo = get(m - 1, c), mid = get(m, c), hi = get(m + 1, c);
        const cd t2 = cscale(cmuli(cadd(cmul(w, lo), cmul(wc, hi))), 0.5);
        if (c == 0) E[0] = cadd(cadd(cscale(E[0], 1.0 / dn), cscale(mid, xZ)), t2);
        else E[c] = cscale(cadd(E[c], cscale(cadd(cscale(mid, xZ), t2), dn)), 1.0 / (dn - 1.0));
      }
    }
    __syncthreads();
    // moment vector v = [rhs (4h); W (h)] into the (now free) raw-moment slots
    for (int it = threadIdx.x; it < nt * h; it += blockDim.x) {
      const int t = it / h, kk = it % h;
      int n = 1;
      while ((n + 1) * n / 2 <= kk) ++n;
      const int m = kk - (n - 1) * n / 2 + 1;
      const cd* E = mom + (size_t)t * mstr + (size_t)hidx(n, m) * 5;
      double* v = rv + (size_t)t * rstr;
      v[kk] = sigma * imrot(m, E[1]);
      v[h + kk] = -sigma * imrot(m, E[2]);
      v[2 * h + kk] = -sigma * imrot(m, E[3]);
      v[3 * h + kk] = -sigma * imrot(m, E[4]);
      v[4 * h + kk] = sigma * imrot(m, E[0]);
    }
    __syncthreads();
    if (omega_out) {
      const double f = 1.0 / (16.0 * atan(1.0)) / alpha;
      for (int it = threadIdx.x; it < nt * H4; it += blockDim.x) {
        const int t = it / H4, i = it % H4;
        omega_out[(size_t)plist[t] * H4 + i] = rv[(size_t)t * rstr + i] * f;
      }
      for (int t = threadIdx.x; t < nt; t += blockDim.x) ias_out[plist[t]] = sigma * tq[4 * t];
    }
    // moment vectors to HBM, per panel transposed: vt[r0*5h + i*nrrq + t]
    for (int it = threadIdx.x; it < nt * H5; it += blockDim.x) {
      const int t = it % nt, i = it / nt;
      vt[(size_t)r0 * H5 + (size_t)i * nmine + tile0 + t] = rv[(size_t)t * rstr + i];
    }
  }
}

// RRQ weights of every RRQ pair of a panel: W (2h x nrrq) = G (2h x 5h) V (5h x nrrq)
// on the FP64 matrix cores.  One workgroup of kWM waves per panel; wave w owns
// output rows 16w..16w+15 and walks the panel's targets 16 at a time.  A operand:
// G (stored transposed, coalesced over rows); B operand: the panel's moment
// vectors (stored transposed, coalesced over targets).
constexpr int kWM = 5;  // 16*kWM >= 2h for p <= 8; larger orders loop over row tiles

__global__ void __launch_bounds__(64 * kWM) rrq_weights_kernel(RRQTables T, RRQPanels P,
                                                               const int* __restrict__ lpairs,
                                                               const int* __restrict__ lstart,
                                                               const int* __restrict__ ltot, int npan,
                                                               const int* __restrict__ nrrq,
                                                               const double* __restrict__ vt,
                                                               double* __restrict__ wout) {
  const int k = blockIdx.x;
  const int nm = nrrq[k];
  if (nm == 0) return;
  const int h = T.h, H5 = 5 * h, m2 = 2 * h;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int r0 = lstart[4 * (size_t)npan + k];
  const int* mine = lpairs + rrq_list_base(ltot) + r0;
  const double* G = P.G + (size_t)k * m2 * H5;  // G(j, r) at r*2h + j
  const double* V = vt + (size_t)r0 * H5;       // V(r, t) at r*nm + t
  for (int mt = w; 16 * mt < m2; mt += kWM) {
    const int j = 16 * mt + li;
    const bool jin = j < m2;
    const int jc = jin ? j : m2 - 1;
    // two 16-target tiles per pass: two independent accumulator chains share
    // every loaded G value
    for (int t0 = 0; t0 < nm; t0 += 32) {
      const int ta = t0 + li, tb = t0 + 16 + li;
      const bool ina = ta < nm, inb = tb < nm;
      const int tca = ina ? ta : nm - 1, tcb = inb ? tb : nm - 1;
      // K split into even and odd steps as well: four independent chains
      qd4 acc0 = qd4{0.0, 0.0, 0.0, 0.0}, acc1 = qd4{0.0, 0.0, 0.0, 0.0};
      qd4 acc2 = qd4{0.0, 0.0, 0.0, 0.0}, acc3 = qd4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 3
      for (int k0 = 0; k0 < H5; k0 += 8) {
        const int r = k0 + lk, s2 = k0 + 4 + lk;
        const bool rin = r < H5, sin_ = s2 < H5;
        const int rc = rin ? r : H5 - 1, sc = sin_ ? s2 : H5 - 1;
        const double a = (jin & rin) ? G[(size_t)rc * m2 + jc] : 0.0;
        const double b0 = (ina & rin) ? V[(size_t)rc * nm + tca] : 0.0;
        const double b1 = (inb & rin) ? V[(size_t)rc * nm + tcb] : 0.0;
        const double a2 = (jin & sin_) ? G[(size_t)sc * m2 + jc] : 0.0;
        const double b2 = (ina & sin_) ? V[(size_t)sc * nm + tca] : 0.0;
        const double b3 = (inb & sin_) ? V[(size_t)sc * nm + tcb] : 0.0;
        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, acc1, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, b2, acc2, 0, 0, 0);
        acc3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, b3, acc3, 0, 0, 0);
      }
      acc0 += acc2;
      acc1 += acc3;
      // D[4q + lk][li]: output row 16mt + 4q + lk, target t0 + li (and t0 + 16 + li)
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        const int t = half ? tb : ta;
        if (!(half ? inb : ina)) continue;
        const qd4& acc = half ? acc1 : acc0;
        double* o = wout + (size_t)mine[t] * m2;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int jj = 16 * mt + 4 * q + lk;
          if (jj < m2) o[jj] = acc[q];
        }
      }
    }
  }
}

void launch_rrq_weights(RRQTables T, RRQPanels P, const int* lpairs, const int* lstart, const int* ltot, int npan,
                        const int* nrrq, const double* vt, double* wout, hipStream_t s) {
  hipLaunchKernelGGL(rrq_weights_kernel, dim3(npan), dim3(64 * kWM), 0, s, T, P, lpairs, lstart, ltot, npan, nrrq,
                     vt, wout);
}

}  // namespace qol
