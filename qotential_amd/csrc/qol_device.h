// Device-side math shared by the RRQ kernels (CDNA4 / gfx950, FP64).
//
// All FP64 code is compiled with -ffp-contract=off so that every product and sum
// rounds exactly as on the host; the ball tests that decide near lists and
// refinement levels rely on it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qol {

#define QOL_MAXP 16           // largest patch order the device kernels accept
#define QOL_MAXH 136          // QOL_MAXP*(QOL_MAXP+1)/2
#define QOL_MAXNQ 18          // QOL_MAXP + 2
#define QOL_MAXNCOEF 153      // (QOL_MAXP+1)(QOL_MAXP+2)/2

// result tile of v_mfma_f64_16x16x4 (4 doubles per lane)
typedef double qd4 __attribute__((ext_vector_type(4)));

struct cd {
  double re, im;
};
__device__ __forceinline__ cd cmk(double a, double b) { return cd{a, b}; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd csub(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cd cmul(cd a, cd b) { return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cd cscale(cd a, double s) { return cd{a.re * s, a.im * s}; }
__device__ __forceinline__ cd cconj(cd a) { return cd{a.re, -a.im}; }
__device__ __forceinline__ cd cdiv(cd a, cd b) {
  // Smith's algorithm
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, d = b.re + b.im * r;
    return cd{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  } else {
    double r = b.re / b.im, d = b.im + b.re * r;
    return cd{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
  }
}
__device__ __forceinline__ double cabs_(cd a) { return hypot(a.re, a.im); }
__device__ __forceinline__ cd csqrt_(cd z) {
  // principal square root
  double m = hypot(z.re, z.im);
  if (m == 0.0) return cd{0.0, 0.0};
  double t = sqrt(0.5 * (m + fabs(z.re)));
  if (z.re >= 0.0) return cd{t, z.im / (2.0 * t)};
  return cd{fabs(z.im) / (2.0 * t), z.im >= 0.0 ? t : -t};
}
// i*a
__device__ __forceinline__ cd cmuli(cd a) { return cd{-a.im, a.re}; }

// ---- Koornwinder basis on the triangle (same ordering and scaling as the host) ----
// f, fu, fv: arrays of (nmax+1)(nmax+2)/2
__device__ inline void koorn_eval(double u, double v, int nmax, double* f, double* fu, double* fv) {
  const int N = nmax + 1;
  double Q[QOL_MAXP + 1], Qu[QOL_MAXP + 1], Qv[QOL_MAXP + 1];
  const double a = 2.0 * u + v - 1.0, b = 1.0 - v, bb = b * b;
  Q[0] = 1.0; Qu[0] = 0.0; Qv[0] = 0.0;
  if (N > 1) { Q[1] = a; Qu[1] = 2.0; Qv[1] = 1.0; }
  for (int m = 1; m + 1 < N; ++m) {
    const double c1 = 2.0 * m + 1.0, c2 = (double)m, inv = 1.0 / (double)(m + 1);
    Q[m + 1] = (c1 * a * Q[m] - c2 * bb * Q[m - 1]) * inv;
    Qu[m + 1] = (c1 * (2.0 * Q[m] + a * Qu[m]) - c2 * bb * Qu[m - 1]) * inv;
    Qv[m + 1] = (c1 * (Q[m] + a * Qv[m]) - c2 * (bb * Qv[m - 1] - 2.0 * b * Q[m - 1])) * inv;
  }
  const double x = 2.0 * v - 1.0;
  int idx = 0;
  for (int d = 0; d < N; ++d) {
    for (int bd = 0; bd <= d; ++bd) {
      const int am = d - bd;
      const double al = 2.0 * am + 1.0;
      double j0 = 1.0, j1 = 0.0, jv0 = 0.0, jv1 = 0.0;  // P_bd and d/dv
      if (bd >= 1) {
        double pm = 1.0, p = 0.5 * ((al + 2.0) * x + al), dpm = 0.0, dp = al + 2.0;
        for (int k = 1; k < bd; ++k) {
          const double dk = (double)k;
          const double a1 = 2.0 * (dk + 1.0) * (dk + al + 1.0) * (2.0 * dk + al);
          const double a2 = (2.0 * dk + al + 1.0) * al * al;
          const double a3 = (2.0 * dk + al) * (2.0 * dk + al + 1.0) * (2.0 * dk + al + 2.0);
          const double a4 = 2.0 * (dk + al) * dk * (2.0 * dk + al + 2.0);
          const double pn = ((a2 + a3 * x) * p - a4 * pm) / a1;
          const double dpn = ((a2 + a3 * x) * dp + 2.0 * a3 * p - a4 * dpm) / a1;
          pm = p; p = pn; dpm = dp; dp = dpn;
        }
        j0 = p; jv0 = dp;
      } else {
        j0 = 1.0; jv0 = 0.0;
      }
      (void)j1; (void)jv1;
      const double sc = sqrt((2.0 * am + 1.0) * 2.0 * (double)(d + 1));
      f[idx] = sc * Q[am] * j0;
      if (fu) fu[idx] = sc * Qu[am] * j0;
      if (fv) fv[idx] = sc * (Qv[am] * j0 + Q[am] * jv0);
      ++idx;
    }
  }
}

// ---- Legendre series sum a_k P_k(s) and derivative (real and complex s) ----
__device__ __forceinline__ void leg_eval_r(const double* a, int n, double s, double& val, double& der) {
  double p0 = 1.0, p1 = s, d0 = 0.0, d1 = 1.0;
  val = a[0];
  der = 0.0;
  if (n > 1) { val += a[1] * p1; der += a[1]; }
  for (int k = 1; k + 1 < n; ++k) {
    const double dk = (double)k;
    const double p2 = ((2.0 * dk + 1.0) * s * p1 - dk * p0) / (dk + 1.0);
    const double d2 = ((2.0 * dk + 1.0) * (p1 + s * d1) - dk * d0) / (dk + 1.0);
    val += a[k + 1] * p2;
    der += a[k + 1] * d2;
    p0 = p1; p1 = p2; d0 = d1; d1 = d2;
  }
}

__device__ __forceinline__ void leg_eval_c(const double* a, int n, cd s, cd& val, cd& der) {
  cd p0 = cmk(1.0, 0.0), p1 = s, d0 = cmk(0.0, 0.0), d1 = cmk(1.0, 0.0);
  val = cmk(a[0], 0.0);
  der = cmk(0.0, 0.0);
  if (n > 1) { val = cadd(val, cscale(p1, a[1])); der = cadd(der, cmk(a[1], 0.0)); }
  for (int k = 1; k + 1 < n; ++k) {
    const double dk = (double)k, c1 = (2.0 * dk + 1.0) / (dk + 1.0), c2 = dk / (dk + 1.0);
    const cd p2 = csub(cscale(cmul(s, p1), c1), cscale(p0, c2));
    const cd d2 = csub(cscale(cadd(p1, cmul(s, d1)), c1), cscale(d0, c2));
    val = cadd(val, cscale(p2, a[k + 1]));
    der = cadd(der, cscale(d2, a[k + 1]));
    p0 = p1; p1 = p2; d0 = d1; d1 = d2;
  }
}

__device__ __forceinline__ double bernstein_rho(cd s) {
  cd r = cadd(s, cmul(csqrt_(cmk(s.re - 1.0, s.im)), csqrt_(cmk(s.re + 1.0, s.im))));
  double a = cabs_(r);
  return a >= 1.0 ? a : 1.0 / a;
}

// 1/sqrt(x) for finite x > 0: the hardware estimate refined by two Newton steps
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double t = x * y;
  y = fma(0.5 * y, fma(-t, y, 1.0), y);
  t = x * y;
  y = fma(0.5 * y, fma(-t, y, 1.0), y);
  return y;
}

// 1/x for finite nonzero x: the hardware estimate refined by two Newton steps
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(y, fma(-x, y, 1.0), y);
  y = fma(y, fma(-x, y, 1.0), y);
  return y;
}

// ---- solid harmonics H'_n^m (m = 0..n), polar axis = frame x, W = y + i z ----
__device__ __forceinline__ int hidx(int n, int m) { return n * (n + 1) / 2 + m; }

template <class Arr>
__device__ __forceinline__ cd hget(const Arr& H, int n, int m) {
  if (n < 0 || m > n || m < -n) return cmk(0.0, 0.0);
  if (m >= 0) return H[hidx(n, m)];
  cd v = cconj(H[hidx(n, -m)]);
  return ((-m) & 1) ? cmk(-v.re, -v.im) : v;
}

}  // namespace qol
