// Workload geometry on the GPU: the per-chart part of the stellarator surface
// generator (utils/stellarator_geo_mex.c:636-749, chart -> two triangles).  One
// wave per chart evaluates the surface on the p x p tensor Gauss-Legendre grid,
// takes the spectral derivatives along both grid directions, and interpolates
// positions and tangents to the p(p+1)/2 nodes of the chart's two triangles,
// from which it forms normals and weights.  The chart list, each chart's
// diagonal and the interpolation weights come from the host plan, so the
// triangulation is the host generator's; every sum runs in the host's order,
// and the nodes differ from the host's only through the device sin/cos.
#include <stdexcept>
#include <vector>

#include "qol_engine.h"
#include "qol_host.h"

namespace qol {

namespace {

struct GeoDev {
  int p, h, nmode, nfp;
  double ts1, ts2;
  const double* modes;  // nmode x (m, n, Rc, Zs)
  const double* x0;     // p
  const double* D;      // p x p
  const double* W;      // 4 x p^2 x h
  const double* wn;     // h
};

// the reference's built-in 7-term surface and the VMEC form (same expressions
// as the host generator)
__device__ void surface_dev(const GeoDev& G, double t, double p, double* x) {
  if (G.nmode == 0) {
    const double ct = cos(t), st = sin(t);
    const double c2pt = cos(2.0 * p - t), s2pt = sin(2.0 * p - t);
    const double c2p = cos(2.0 * p), s2p = sin(2.0 * p);
    const double cp = cos(p), sp = sin(p);
    const double cpt = cos(p + t), spt = sin(p + t);
    const double cmpt = cos(-p + t), smpt = sin(-p + t);
    x[0] = 0.17 * ct * c2pt + 0.11 * ct * c2p + 1.0 * ct * cp + 4.5 * ct - 0.25 * ct * cp + 0.07 * ct * cpt -
           0.45 * ct * cmpt;
    x[1] = 0.17 * st * c2pt + 0.11 * st * c2p + 1.0 * st * cp + 4.5 * st - 0.25 * st * cp + 0.07 * st * cpt -
           0.45 * st * cmpt;
    x[2] = 0.17 * s2pt + 0.11 * s2p + 1.0 * sp + 0.0 + 0.25 * sp + 0.07 * spt - 0.45 * smpt;
    return;
  }
  double R = 0.0, Z = 0.0;
  for (int k = 0; k < G.nmode; ++k) {
    const double* md = G.modes + 4 * k;
    const double ang = md[0] * p - (double)G.nfp * md[1] * t;
    R += md[2] * cos(ang);
    Z += md[3] * sin(ang);
  }
  x[0] = R * cos(t);
  x[1] = R * sin(t);
  x[2] = Z;
}

__global__ void __launch_bounds__(64) chart_geometry_kernel(GeoDev G, int nch, const double* __restrict__ charts,
                                                            const int* __restrict__ diag, double* __restrict__ sx,
                                                            double* __restrict__ snx, double* __restrict__ sw,
                                                            double* __restrict__ rts, double* __restrict__ rps) {
  extern __shared__ double sm[];
  const int k = blockIdx.x;
  if (k >= nch) return;
  const int p = G.p, p2 = p * p, h = G.h;
  double* grid = sm;          // [p2][3], node m = a + b p at (t1, t2) = (x0[b], x0[a])
  double* d2 = grid + 3 * p2;  // d/dt2
  double* d1 = d2 + 3 * p2;    // d/dt1
  const double* c = charts + 6 * (size_t)k;  // a1, b1, a2, b2, jj, ii
  for (int m = threadIdx.x; m < p2; m += blockDim.x) {
    const int a = m % p, b = m / p;
    const double u = ((c[0] * G.x0[b] + c[1]) + 2.0 * c[4]) * G.ts1;
    const double v = ((c[2] * G.x0[a] + c[3]) + 2.0 * c[5]) * G.ts2;
    surface_dev(G, u, v, grid + 3 * m);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 3 * p2; e += blockDim.x) {
    const int m = e / 3, d = e % 3, a = m % p, b = m / p;
    double s2 = 0.0, s1 = 0.0;
    for (int q = 0; q < p; ++q) s2 += G.D[a + q * p] * grid[3 * (q + b * p) + d];
    for (int q = 0; q < p; ++q) s1 += G.D[b + q * p] * grid[3 * (a + q * p) + d];
    d2[e] = s2;
    d1[e] = s1;
  }
  __syncthreads();
  const int dg = diag[k];
  for (int it = threadIdx.x; it < 2 * h; it += blockDim.x) {
    const int sd = it / h, i = it % h;
    const double* W = G.W + (size_t)(2 * dg + sd) * p2 * h + (size_t)i * p2;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int m = 0; m < p2; ++m) {
      const double w = W[m];
      for (int d = 0; d < 3; ++d) {
        acc[d] += grid[3 * m + d] * w;
        acc[3 + d] += d2[3 * m + d] * w;
        acc[6 + d] += d1[3 * m + d] * w;
      }
    }
    // "rts" carries d/dt2, "rps" d/dt1, each times 2; normal = rps x rts
    const double* xt = acc + 3;
    const double* xp = acc + 6;
    const double nx[3] = {xp[1] * xt[2] - xp[2] * xt[1], xp[2] * xt[0] - xp[0] * xt[2],
                          xp[0] * xt[1] - xp[1] * xt[0]};
    const double spn = sqrt(nx[0] * nx[0] + nx[1] * nx[1] + nx[2] * nx[2]);
    const size_t o = (size_t)(2 * k + sd) * h + i;
    for (int d = 0; d < 3; ++d) {
      sx[3 * o + d] = acc[d];
      snx[3 * o + d] = nx[d] / spn;
      rts[3 * o + d] = 2.0 * xt[d];
      rps[3 * o + d] = 2.0 * xp[d];
    }
    sw[o] = spn * G.wn[i] * 4.0;
  }
}

// The mean curvature's Legendre tail of one chart per workgroup (one wave), in
// the host's operations and order (chart_curvature_tail): the chart map on the
// p x p grid, first and second spectral derivatives, H at every node from the
// fundamental forms, C = Vinv H Vinv^T, and the share of sum C^2 in degrees p-2
// and p-1 of either variable, summed by one lane in the host's order.
__global__ void __launch_bounds__(64) chart_curvature_kernel(GeoDev G, int nch, const double* __restrict__ charts,
                                                             const double* __restrict__ Vinv,
                                                             double* __restrict__ tail_out) {
  extern __shared__ double sm[];
  const int k = blockIdx.x;
  if (k >= nch) return;
  const int p = G.p, p2 = p * p;
  double* X = sm;           // [p2][3], node m = a + b p at (t1, t2) = (x0[b], x0[a])
  double* X1 = X + 3 * p2;  // d/dt1 (over b)
  double* X2 = X1 + 3 * p2;  // d/dt2 (over a)
  double* X11 = X2 + 3 * p2;
  double* X12 = X11 + 3 * p2;
  double* X22 = X12 + 3 * p2;
  double* H = X22 + 3 * p2;  // [p2]
  double* T = H + p2;        // [p2]
  double* C = T + p2;        // [p2]
  const double* c = charts + 6 * (size_t)k;
  for (int m = threadIdx.x; m < p2; m += blockDim.x) {
    const int a = m % p, b = m / p;
    const double u = ((c[0] * G.x0[b] + c[1]) + 2.0 * c[4]) * G.ts1;
    const double v = ((c[2] * G.x0[a] + c[3]) + 2.0 * c[5]) * G.ts2;
    surface_dev(G, u, v, X + 3 * m);
  }
  __syncthreads();
  const double* D = G.D;
  for (int e = threadIdx.x; e < 3 * p2; e += blockDim.x) {
    const int m = e / 3, d = e % 3, a = m % p, b = m / p;
    double s1 = 0.0, s2 = 0.0;
    for (int q = 0; q < p; ++q) s1 += D[b + q * p] * X[3 * (a + q * p) + d];
    for (int q = 0; q < p; ++q) s2 += D[a + q * p] * X[3 * (q + b * p) + d];
    X1[e] = s1;
    X2[e] = s2;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 3 * p2; e += blockDim.x) {
    const int m = e / 3, d = e % 3, a = m % p, b = m / p;
    double s11 = 0.0, s12 = 0.0, s22 = 0.0;
    for (int q = 0; q < p; ++q) s11 += D[b + q * p] * X1[3 * (a + q * p) + d];
    for (int q = 0; q < p; ++q) s12 += D[b + q * p] * X2[3 * (a + q * p) + d];
    for (int q = 0; q < p; ++q) s22 += D[a + q * p] * X2[3 * (q + b * p) + d];
    X11[e] = s11;
    X12[e] = s12;
    X22[e] = s22;
  }
  __syncthreads();
  for (int m = threadIdx.x; m < p2; m += blockDim.x) {
    const double* u = X1 + 3 * m;
    const double* v = X2 + 3 * m;
    double nv[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    double nn = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    if (nn <= 0.0) nn = 1.0;
    nv[0] /= nn; nv[1] /= nn; nv[2] /= nn;
    const double E = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double F = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
    const double Gg = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const double* r11 = X11 + 3 * m;
    const double* r12 = X12 + 3 * m;
    const double* r22 = X22 + 3 * m;
    const double L = nv[0] * r11[0] + nv[1] * r11[1] + nv[2] * r11[2];
    const double M = nv[0] * r12[0] + nv[1] * r12[1] + nv[2] * r12[2];
    const double N = nv[0] * r22[0] + nv[1] * r22[1] + nv[2] * r22[2];
    double den = 2.0 * (E * Gg - F * F);
    if (den == 0.0) den = 1.0;
    H[m] = (E * N - 2.0 * F * M + Gg * L) / den;
  }
  __syncthreads();
  for (int m = threadIdx.x; m < p2; m += blockDim.x) {
    const int a = m % p, b = m / p;
    double v = 0.0;
    for (int q = 0; q < p; ++q) v += Vinv[a + q * p] * H[q + b * p];
    T[m] = v;
  }
  __syncthreads();
  for (int m = threadIdx.x; m < p2; m += blockDim.x) {
    const int a = m % p, b = m / p;
    double v = 0.0;
    for (int q = 0; q < p; ++q) v += T[a + q * p] * Vinv[b + q * p];
    C[m] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tail = 0.0, tot = 0.0;
    for (int b = 0; b < p; ++b)
      for (int a = 0; a < p; ++a) {
        const double v = C[a + b * p];
        tot += v * v;
        if (a >= p - 2 || b >= p - 2) tail += v * v;
      }
    tail_out[k] = sqrt(tail / (tot > 0.0 ? tot : 1.0));
  }
}

}  // namespace

std::vector<double> curvature_tails_gpu(const Surface& s, const std::vector<Chart>& cl, i64 mp, i64 np, i64 p) {
  const int nch = (int)cl.size();
  std::vector<double> out(nch, 0.0);
  if (nch == 0) return out;
  if (p < 2 || p > QOL_MAXP) throw std::runtime_error("curvature on the GPU: patch order out of range");
  std::vector<double> x0, D, Vinv;
  curvature_tables(p, x0, D, Vinv);
  std::vector<double> ch(6 * (size_t)nch), modes(4 * s.rc.size());
  for (int k = 0; k < nch; ++k) {
    const double v[6] = {cl[k].a1, cl[k].b1, cl[k].a2, cl[k].b2, cl[k].jj, cl[k].ii};
    for (int j = 0; j < 6; ++j) ch[6 * (size_t)k + j] = v[j];
  }
  for (size_t k = 0; k < s.rc.size(); ++k) {
    modes[4 * k] = s.m[k]; modes[4 * k + 1] = s.n[k]; modes[4 * k + 2] = s.rc[k]; modes[4 * k + 3] = s.zs[k];
  }
  DevBuf d_modes, d_x0, d_D, d_V, d_ch, d_out;
  auto up = [](DevBuf& b, const void* src, size_t bytes) {
    b.alloc(bytes > 0 ? bytes : 8);
    if (bytes) HIPCHK(hipMemcpy(b.ptr, src, bytes, hipMemcpyHostToDevice));
  };
  up(d_modes, modes.data(), sizeof(double) * modes.size());
  up(d_x0, x0.data(), sizeof(double) * x0.size());
  up(d_D, D.data(), sizeof(double) * D.size());
  up(d_V, Vinv.data(), sizeof(double) * Vinv.size());
  up(d_ch, ch.data(), sizeof(double) * ch.size());
  d_out.alloc(sizeof(double) * nch);
  const double pi = 3.14159265358979323846;
  GeoDev G{(int)p, 0, (int)s.rc.size(), (int)s.nfp, pi / (double)np, pi / (double)mp, (const double*)d_modes.ptr,
           (const double*)d_x0.ptr, (const double*)d_D.ptr, nullptr, nullptr};
  const size_t lds = sizeof(double) * 21 * (size_t)p * p;
  hipLaunchKernelGGL(chart_curvature_kernel, dim3(nch), dim3(64), lds, 0, G, nch, (const double*)d_ch.ptr,
                     (const double*)d_V.ptr, (double*)d_out.ptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(out.data(), d_out.ptr, sizeof(double) * nch, hipMemcpyDeviceToHost));
  return out;
}

void build_geometry_gpu(const Surface& s, const GeometryPlan& P, double* sx, double* snx, double* sw, double* rts,
                        double* rps) {
  const int nch = (int)P.charts.size(), p = (int)P.p, h = (int)P.h;
  if (nch == 0) return;
  const size_t lds = sizeof(double) * 9 * (size_t)p * p;
  if (lds > 64 * 1024) throw std::runtime_error("geometry on the GPU: patch order too high for one chart in LDS");
  std::vector<double> ch(6 * (size_t)nch), modes(4 * s.rc.size());
  for (int k = 0; k < nch; ++k) {
    const Chart& c = P.charts[k];
    const double v[6] = {c.a1, c.b1, c.a2, c.b2, c.jj, c.ii};
    for (int j = 0; j < 6; ++j) ch[6 * (size_t)k + j] = v[j];
  }
  for (size_t k = 0; k < s.rc.size(); ++k) {
    modes[4 * k] = s.m[k]; modes[4 * k + 1] = s.n[k]; modes[4 * k + 2] = s.rc[k]; modes[4 * k + 3] = s.zs[k];
  }
  const size_t nsrc = 2 * (size_t)nch * h;
  DevBuf d_modes, d_x0, d_D, d_W, d_wn, d_ch, d_diag, d_sx, d_snx, d_sw, d_rts, d_rps;
  auto up = [](DevBuf& b, const void* src, size_t bytes) {
    b.alloc(bytes > 0 ? bytes : 8);
    if (bytes) HIPCHK(hipMemcpy(b.ptr, src, bytes, hipMemcpyHostToDevice));
  };
  up(d_modes, modes.data(), sizeof(double) * modes.size());
  up(d_x0, P.x0.data(), sizeof(double) * P.x0.size());
  up(d_D, P.D.data(), sizeof(double) * P.D.size());
  up(d_W, P.W.data(), sizeof(double) * P.W.size());
  up(d_wn, P.wn.data(), sizeof(double) * P.wn.size());
  up(d_ch, ch.data(), sizeof(double) * ch.size());
  up(d_diag, P.diag.data(), sizeof(int) * P.diag.size());
  d_sx.alloc(sizeof(double) * 3 * nsrc);
  d_snx.alloc(sizeof(double) * 3 * nsrc);
  d_sw.alloc(sizeof(double) * nsrc);
  d_rts.alloc(sizeof(double) * 3 * nsrc);
  d_rps.alloc(sizeof(double) * 3 * nsrc);
  GeoDev G{p, h, (int)s.rc.size(), (int)s.nfp, P.ts1, P.ts2, (const double*)d_modes.ptr, (const double*)d_x0.ptr,
           (const double*)d_D.ptr, (const double*)d_W.ptr, (const double*)d_wn.ptr};
  hipLaunchKernelGGL(chart_geometry_kernel, dim3(nch), dim3(64), lds, 0, G, nch, (const double*)d_ch.ptr,
                     (const int*)d_diag.ptr, (double*)d_sx.ptr, (double*)d_snx.ptr, (double*)d_sw.ptr,
                     (double*)d_rts.ptr, (double*)d_rps.ptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(sx, d_sx.ptr, sizeof(double) * 3 * nsrc, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(snx, d_snx.ptr, sizeof(double) * 3 * nsrc, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(sw, d_sw.ptr, sizeof(double) * nsrc, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(rts, d_rts.ptr, sizeof(double) * 3 * nsrc, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(rps, d_rps.ptr, sizeof(double) * 3 * nsrc, hipMemcpyDeviceToHost));
}

}  // namespace qol
