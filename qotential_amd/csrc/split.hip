// Split-precision direct S/D on the refinement levels: FP32 matrix cores.
//
// Same work and output as direct_mfma_kernel (levels.hip): for every pair
// served at level l, the weights As(c) = sum_i S(x, y_i) I(i, c) and
// Ad(c) = sum_i D(x, y_i) I(i, c) over the level's nodes y_i (the reference's
// direct block and dgemm with interpmat, src/patch_refine_mod.f90:569-602).
// Precision split:
//   - x - y_i and (x - y_i).n_i are formed in FP64 from the FP64 target, level
//     node and normal, then rounded to FP32, so each entry's relative error
//     stays at FP32 rounding however close or however tangential the target;
//   - the entries S = w/(4 pi r), D = w (x-y).n/(4 pi r^3) are FP32 (v_rsq_f32
//     and one Newton step);
//   - the contraction runs on v_mfma_f32_16x16x4f32 against the FP32 copy of
//     the padded interpolation rows, accumulating in FP32 over a chunk of CH
//     steps of 4 nodes; each chunk's sums are then added into FP64 registers,
//     which bounds the FP32 accumulation error by the chunk, not the level.
// The weights are written in FP64 into the same per-pair slots as the FP64
// kernel, so the apply and everything after it are unchanged.
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>

#include "qol_kernels.h"

namespace qol {

typedef float qf4 __attribute__((ext_vector_type(4)));

template <int NT, int CH>
__global__ void __launch_bounds__(64) direct_split_kernel(LevelArgs A, const int* __restrict__ lpairs,
                                                          const int* __restrict__ ltot,
                                                          const int* __restrict__ ppanel,
                                                          const int* __restrict__ tidx,
                                                          const double* __restrict__ tx,
                                                          const double* __restrict__ lgeo,
                                                          double* __restrict__ wout, int lo, int hi) {
  int b = blockIdx.x, l = -1;
  int start[4];
  {
    int s = 0;
    for (int q = 0; q < 4; ++q) { start[q] = s; s += ltot[q]; }
  }
  for (int q = hi; q >= lo; --q) {
    const int ntile = (ltot[q] + 15) >> 4;
    if (b < ntile) { l = q; break; }
    b -= ntile;
  }
  if (l < 0) return;
  const int h = A.h, ntot = A.ntot, pad = A.pad;
  const int first = start[l] + 16 * b;
  const int cnt = min(16, ltot[l] - 16 * b);
  const int lane = threadIdx.x, m = lane & 15, kq = lane >> 4;
  const bool live = m < cnt;
  const int pair = live ? lpairs[first + m] : 0;
  double x0 = 0.0, x1 = 0.0, x2 = 0.0;
  const double* G = lgeo;
  if (live) {
    const size_t t = (size_t)tidx[pair];
    x0 = tx[3 * t]; x1 = tx[3 * t + 1]; x2 = tx[3 * t + 2];
    G = lgeo + 8 * ((size_t)ppanel[pair] * ntot + A.lev_off[l]);
  }
  const int n = A.lev_n[l];
  const float* B = A.impf + (size_t)pad * A.lev_off[l];
  const float c4 = (float)(1.0 / (16.0 * atan(1.0)));
  qd4 accS[NT], accD[NT];
  qf4 cS[NT], cD[NT];
#pragma unroll
  for (int q = 0; q < NT; ++q) {
    accS[q] = qd4{0.0, 0.0, 0.0, 0.0};
    accD[q] = qd4{0.0, 0.0, 0.0, 0.0};
    cS[q] = qf4{0.0f, 0.0f, 0.0f, 0.0f};
    cD[q] = qf4{0.0f, 0.0f, 0.0f, 0.0f};
  }
  int step = 0;
  for (int k0 = 0; k0 < n; k0 += 4) {
    const int i = k0 + kq;
    const bool in = i < n;
    const int ic = in ? i : n - 1;
    const double2* g2 = (const double2*)(G + 8 * ic);
    const double2 ga = g2[0], gb = g2[1], gc = g2[2], gd = g2[3];
    float bq[NT];
#pragma unroll
    for (int q = 0; q < NT; ++q) bq[q] = B[(size_t)ic * pad + 16 * q + m];
    // the difference in FP64, then FP32
    const float fx = (float)(x0 - ga.x), fy = (float)(x1 - ga.y), fz = (float)(x2 - gb.x);
    const float rr = fx * fx + fy * fy + fz * fz;
    const bool ok = live && in && rr > 1.0e-28f;
    const float rs = ok ? rr : 1.0f;
    float y = __builtin_amdgcn_rsqf(rs);
    y = y * fmaf(-0.5f * rs * y, y, 1.5f);
    const float cw = (ok ? 1.0f : 0.0f) * (c4 * (float)gd.x);
    // (x - y).n in FP64: it cancels when x - y is nearly tangent to the patch
    const float dn = (float)((x0 - ga.x) * gb.y + (x1 - ga.y) * gc.x + (x2 - gb.x) * gc.y);
    const float s = cw * y;
    const float d = cw * dn * (y * y * y);
#pragma unroll
    for (int q = 0; q < NT; ++q) {
      cS[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(s, bq[q], cS[q], 0, 0, 0);
      cD[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(d, bq[q], cD[q], 0, 0, 0);
    }
    if (++step == CH) {
      step = 0;
#pragma unroll
      for (int q = 0; q < NT; ++q) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          accS[q][r] += (double)cS[q][r];
          accD[q][r] += (double)cD[q][r];
        }
        cS[q] = qf4{0.0f, 0.0f, 0.0f, 0.0f};
        cD[q] = qf4{0.0f, 0.0f, 0.0f, 0.0f};
      }
    }
  }
#pragma unroll
  for (int q = 0; q < NT; ++q)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      accS[q][r] += (double)cS[q][r];
      accD[q][r] += (double)cD[q][r];
    }
  // v_mfma_f32_16x16x4f32 result: lane l holds D[4 (l/16) + r][l%16] in slot
  // r, so this lane has pairs 4 kq + r of output column 16 q + m
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = 4 * kq + r;
    const int pr = __shfl(pair, row, 64);
    if (row < cnt) {
      double* o = wout + (size_t)pr * 2 * h;
#pragma unroll
      for (int q = 0; q < NT; ++q) {
        const int col = 16 * q + m;
        if (col < h) {
          o[col] = accS[q][r];
          o[h + col] = accD[q][r];
        }
      }
    }
  }
}

void launch_direct_split(LevelArgs A, const int* lpairs, const int* ltot, const int* ppanel, const int* tidx,
                         const double* tx, const double* lgeo, double* wout, int64_t nnz, hipStream_t s, int lo,
                         int hi) {
  if (lo > hi) return;
  const int grid = (int)((nnz + 15) / 16) + 4;
  constexpr int CH = 4;
#define QOL_SPLIT(NT)                                                                                          \
  case NT:                                                                                                     \
    hipLaunchKernelGGL((direct_split_kernel<NT, CH>), dim3(grid), dim3(64), 0, s, A, lpairs, ltot, ppanel, tidx, \
                       tx, lgeo, wout, lo, hi);                                                                \
    break;
  switch ((A.h + 15) / 16) {
    QOL_SPLIT(1) QOL_SPLIT(2) QOL_SPLIT(3) QOL_SPLIT(4) QOL_SPLIT(5) QOL_SPLIT(6) QOL_SPLIT(7) QOL_SPLIT(8)
    QOL_SPLIT(9)
    default: throw std::runtime_error("split direct kernel: unsupported patch order");
  }
#undef QOL_SPLIT
}

}  // namespace qol
