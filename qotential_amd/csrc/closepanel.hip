// Close DLP matrix of a tensor-product patch (the reference's
// Lap3dDLP_closepanel_r64, src/lap3d_close_mod.f90:135-270, behind
// qol_lap3ddlp_closepanel_mex, src/lap3d_close_mex.f90:2-15).
//
// For targets x_k near one patch given by its nodes s_x on a p x p
// Gauss-Legendre tensor grid (node ip = b p + a, a the fast first-parameter
// index), Ac(k, ip) = 1/(4 pi) int (x_k - y).(y_u x y_v)/|x_k - y|^3 L_ip(u,v) du dv
// over [-1,1]^2, y the tensor interpolant of s_x and L_ip its Lagrange basis,
// so that sum_ip Ac(k, ip) sigma(y_ip) is the double-layer potential of the
// interpolated density.  The reference forms it from boundary moments in its
// absent submodules; here the integrals are computed directly by adaptive
// quadrature: a quadtree on the parameter square whose cells are split while
// the target is within ETA diameters of the cell's centre image (depth at most
// MAXDEPTH), each leaf integrated with a q x q Gauss rule, q = 2p + 2 ref.
//
// One workgroup per target.  Each level of the tree is tested in parallel (one
// thread per cell); the leaves then go through in batches of 256 quadrature
// nodes: every thread forms one node's geometry and kernel value and its
// Lagrange factors in LDS, then thread ip adds its basis function's products in
// node order (a fixed order: the result does not depend on timing).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "qol_engine.h"
#include "qol_host.h"
#include "qotential_hip.h"

namespace qol {
void set_status(int64_t s);
void set_error(const std::string& s);

namespace {

constexpr int kCpMaxP = 16;       // patch order limit (p^2 <= 256 basis functions)
constexpr int kCpCur = 256;       // cells per tree level
constexpr int kCpLeaves = 4096;   // leaves per target
constexpr int kCpMaxDepth = 48;
constexpr double kCpEta = 2.5;

struct CpArgs {
  int p, q, m;
  const double* t;    // p GL nodes
  const double* wb;   // p barycentric weights
  const double* gq;   // q Gauss nodes
  const double* gw;   // q Gauss weights
  const double* sx;   // 3 p^2 patch nodes
  const double* tx;   // 3 m targets
  double* leaves;     // m x kCpLeaves x 3 (u0, v0, side)
  double* Ac;         // m x p^2, column-major (m, p^2)
  int* flag;          // set when a target's tree passes the limits
};

// Lagrange basis on the GL nodes and its derivative at u (barycentric form;
// product form when u sits on a node)
__device__ void lagr(const CpArgs& A, double u, double* l, double* dl) {
  const int p = A.p;
  int hit = -1;
  for (int j = 0; j < p; ++j)
    if (fabs(u - A.t[j]) < 1e-14) hit = j;
  if (hit < 0) {
    double S = 0.0, S2 = 0.0;
    for (int j = 0; j < p; ++j) {
      const double r = 1.0 / (u - A.t[j]);
      const double a = A.wb[j] * r;
      l[j] = a;
      dl[j] = r;
      S += a;
      S2 += a * r;
    }
    const double iS = 1.0 / S, g = S2 * iS;
    for (int j = 0; j < p; ++j) {
      const double lj = l[j] * iS;
      dl[j] = lj * (g - dl[j]);
      l[j] = lj;
    }
  } else {
    for (int j = 0; j < p; ++j) {
      double v = 1.0, d = 0.0;
      for (int k = 0; k < p; ++k)
        if (k != j) v *= (u - A.t[k]) / (A.t[j] - A.t[k]);
      for (int r = 0; r < p; ++r) {
        if (r == j) continue;
        double w = 1.0 / (A.t[j] - A.t[r]);
        for (int k = 0; k < p; ++k)
          if (k != j && k != r) w *= (u - A.t[k]) / (A.t[j] - A.t[k]);
        d += w;
      }
      l[j] = v;
      dl[j] = d;
    }
  }
}

__device__ void cp_geom(const CpArgs& A, double u, double v, double* lu, double* lv, double* y, double* yu,
                        double* yv) {
  double dlu[kCpMaxP], dlv[kCpMaxP];
  lagr(A, u, lu, dlu);
  lagr(A, v, lv, dlv);
  for (int c = 0; c < 3; ++c) { y[c] = 0.0; yu[c] = 0.0; yv[c] = 0.0; }
  for (int b = 0; b < A.p; ++b)
    for (int a = 0; a < A.p; ++a) {
      const double* s = A.sx + 3 * (b * A.p + a);
      const double f = lu[a] * lv[b], fu = dlu[a] * lv[b], fv = lu[a] * dlv[b];
      for (int c = 0; c < 3; ++c) {
        y[c] += f * s[c];
        yu[c] += fu * s[c];
        yv[c] += fv * s[c];
      }
    }
}

__global__ void __launch_bounds__(256) closepanel_kernel(CpArgs A) {
  __shared__ double cur[kCpCur][3], nxt[kCpCur][3];
  __shared__ int split[kCpCur];
  __shared__ int ncur, nleaf, bad;
  extern __shared__ double sh[];  // 256 x (1 + 2p): kernel value, then lu, lv
  const int k = blockIdx.x, p = A.p, h = p * p, q = A.q;
  const double x0 = A.tx[3 * k], x1 = A.tx[3 * k + 1], x2 = A.tx[3 * k + 2];
  double* L = A.leaves + (size_t)k * kCpLeaves * 3;
  if (threadIdx.x == 0) {
    cur[0][0] = -1.0; cur[0][1] = -1.0; cur[0][2] = 2.0;
    ncur = 1; nleaf = 0; bad = 0;
  }
  __syncthreads();
  for (int depth = 0; ncur > 0; ++depth) {
    const int c = threadIdx.x;
    if (c < ncur) {
      const double u0 = cur[c][0], v0 = cur[c][1], s = cur[c][2];
      double lu[kCpMaxP], lv[kCpMaxP], yc[3], a1[3], a2[3], cy[4][3];
      cp_geom(A, u0 + 0.5 * s, v0 + 0.5 * s, lu, lv, yc, a1, a2);
      cp_geom(A, u0, v0, lu, lv, cy[0], a1, a2);
      cp_geom(A, u0 + s, v0, lu, lv, cy[1], a1, a2);
      cp_geom(A, u0, v0 + s, lu, lv, cy[2], a1, a2);
      cp_geom(A, u0 + s, v0 + s, lu, lv, cy[3], a1, a2);
      double diam2 = 0.0;
      for (int i = 0; i < 4; ++i)
        for (int j = i + 1; j < 4; ++j) {
          const double d0 = cy[i][0] - cy[j][0], d1 = cy[i][1] - cy[j][1], d2 = cy[i][2] - cy[j][2];
          diam2 = fmax(diam2, d0 * d0 + d1 * d1 + d2 * d2);
        }
      const double e0 = x0 - yc[0], e1 = x1 - yc[1], e2 = x2 - yc[2];
      split[c] = (depth < kCpMaxDepth && e0 * e0 + e1 * e1 + e2 * e2 < kCpEta * kCpEta * diam2) ? 1 : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int nn = 0, nl = nleaf;
      for (int i = 0; i < ncur; ++i) {
        const double u0 = cur[i][0], v0 = cur[i][1], s = cur[i][2];
        if (split[i]) {
          if (nn + 4 > kCpCur) { bad = 1; continue; }
          const double hs = 0.5 * s;
          const double cu[4] = {u0, u0 + hs, u0, u0 + hs}, cv[4] = {v0, v0, v0 + hs, v0 + hs};
          for (int j = 0; j < 4; ++j) { nxt[nn][0] = cu[j]; nxt[nn][1] = cv[j]; nxt[nn][2] = hs; ++nn; }
        } else {
          if (nl >= kCpLeaves) { bad = 1; continue; }
          L[3 * nl] = u0; L[3 * nl + 1] = v0; L[3 * nl + 2] = s;
          ++nl;
        }
      }
      for (int i = 0; i < nn; ++i) { cur[i][0] = nxt[i][0]; cur[i][1] = nxt[i][1]; cur[i][2] = nxt[i][2]; }
      ncur = nn;
      nleaf = nl;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && bad) atomicOr(A.flag, 1);
  // leaves in batches of 256 quadrature nodes
  const double c4 = 1.0 / (16.0 * atan(1.0));
  const int nnode = nleaf * q * q, stride = 1 + 2 * p;
  double acc = 0.0;
  const int ia = threadIdx.x % p, ib = threadIdx.x / p;
  for (int base = 0; base < nnode; base += 256) {
    const int nd = base + threadIdx.x;
    double* my = sh + (size_t)threadIdx.x * stride;
    if (nd < nnode) {
      const int leaf = nd / (q * q), r = nd % (q * q), ju = r % q, jv = r / q;
      const double u0 = L[3 * leaf], v0 = L[3 * leaf + 1], hs = 0.5 * L[3 * leaf + 2];
      const double u = u0 + hs * (A.gq[ju] + 1.0), v = v0 + hs * (A.gq[jv] + 1.0);
      double y[3], yu[3], yv[3];
      cp_geom(A, u, v, my + 1, my + 1 + p, y, yu, yv);
      const double n0 = yu[1] * yv[2] - yu[2] * yv[1], n1 = yu[2] * yv[0] - yu[0] * yv[2],
                   n2 = yu[0] * yv[1] - yu[1] * yv[0];
      const double d0 = x0 - y[0], d1 = x1 - y[1], d2 = x2 - y[2];
      const double rr = d0 * d0 + d1 * d1 + d2 * d2, rn = sqrt(rr);
      my[0] = c4 * A.gw[ju] * A.gw[jv] * hs * hs * (d0 * n0 + d1 * n1 + d2 * n2) / (rn * rr);
    } else {
      my[0] = 0.0;
      for (int i = 1; i < stride; ++i) my[i] = 0.0;
    }
    __syncthreads();
    if (threadIdx.x < h) {
      const int cnt = min(256, nnode - base);
      for (int i = 0; i < cnt; ++i) {
        const double* o = sh + (size_t)i * stride;
        acc += o[0] * o[1 + ia] * o[1 + p + ib];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < h) A.Ac[k + (size_t)A.m * threadIdx.x] = acc;
}

}  // namespace
}  // namespace qol

using namespace qol;

extern "C" void qol_lap3ddlp_closepanel_mex_(const int64_t* m_tgt, const double* t_x, const int64_t* npat,
                                             const double* s_x, const int64_t* order, const int64_t* ref,
                                             const double* if_adapt_d, double* Ac) {
  try {
    const int64_t p = *order, m = *m_tgt, h = p * p;
    if (p < 2 || p > kCpMaxP) throw std::runtime_error("qol_lap3ddlp_closepanel_mex_: order must be 2..16");
    if (*npat != h) throw std::runtime_error("qol_lap3ddlp_closepanel_mex_: npat must be order^2");
    // the reference implements the adaptive moments only (lap3d_close_mod.f90:234-245)
    if (*if_adapt_d == 0.0) throw std::runtime_error("qol_lap3ddlp_closepanel_mex_: if_adapt = false is not implemented "
                                                     "(as in the reference)");
    if (m < 0) throw std::runtime_error("qol_lap3ddlp_closepanel_mex_: m_tgt must be non-negative");
    if (m == 0) { set_status(0); return; }
    const int q = (int)(2 * p + 2 * (*ref > 0 ? *ref : 1));
    std::vector<double> t(p), w(p), wb(p), gq(q), gw(q);
    gl_nodes(p, t.data(), w.data(), nullptr);
    gl_nodes(q, gq.data(), gw.data(), nullptr);
    for (int64_t j = 0; j < p; ++j) {
      double pr = 1.0;
      for (int64_t k = 0; k < p; ++k)
        if (k != j) pr *= t[j] - t[k];
      wb[j] = 1.0 / pr;
    }
    DevBuf d_t, d_wb, d_gq, d_gw, d_sx, d_tx, d_leaf, d_Ac, d_flag;
    auto up = [](DevBuf& b, const void* src, size_t bytes) {
      b.alloc(bytes);
      HIPCHK(hipMemcpy(b.ptr, src, bytes, hipMemcpyHostToDevice));
    };
    up(d_t, t.data(), sizeof(double) * p);
    up(d_wb, wb.data(), sizeof(double) * p);
    up(d_gq, gq.data(), sizeof(double) * q);
    up(d_gw, gw.data(), sizeof(double) * q);
    up(d_sx, s_x, sizeof(double) * 3 * h);
    up(d_tx, t_x, sizeof(double) * 3 * m);
    d_leaf.alloc(sizeof(double) * 3 * (size_t)kCpLeaves * m);
    d_Ac.alloc(sizeof(double) * (size_t)m * h);
    d_flag.alloc(sizeof(int));
    const int zero = 0;
    HIPCHK(hipMemcpy(d_flag.ptr, &zero, sizeof(int), hipMemcpyHostToDevice));
    CpArgs A{(int)p, q, (int)m, (const double*)d_t.ptr, (const double*)d_wb.ptr, (const double*)d_gq.ptr,
             (const double*)d_gw.ptr, (const double*)d_sx.ptr, (const double*)d_tx.ptr, (double*)d_leaf.ptr,
             (double*)d_Ac.ptr, (int*)d_flag.ptr};
    const size_t lds = sizeof(double) * 256 * (1 + 2 * (size_t)p);
    if (lds > 64 * 1024)
      HIPCHK(hipFuncSetAttribute((const void*)closepanel_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(closepanel_kernel, dim3((unsigned)m), dim3(256), lds, 0, A);
    HIPCHK(hipGetLastError());
    int flag = 0;
    HIPCHK(hipMemcpy(&flag, d_flag.ptr, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(Ac, d_Ac.ptr, sizeof(double) * (size_t)m * h, hipMemcpyDeviceToHost));
    if (flag) throw std::runtime_error("qol_lap3ddlp_closepanel_mex_: a target's quadrature tree passed its size limit");
    set_status(0);
  } catch (const std::exception& e) {
    set_error(e.what());
    set_status(1);
  }
}
