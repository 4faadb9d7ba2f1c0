// Refinement levels on the GPU: per-panel level geometry and sub-patch balls,
// pair classification, and direct S/D on the upsampled levels contracted with the
// interpolation matrix.
//
// Reference: build_patch_levels_r64   patch_refine_mod.f90:423-502
//            lap3dsdlpmat_levels_r64  patch_refine_mod.f90:516-615
//            lap3dsdlpmat_r64 (direct kernel; zero at coincident points)
#include "qol_kernels.h"

namespace qol {

// One workgroup per panel.  Level nodes are interpolated in the reference's
// summation order (sequential over the h panel nodes, no FMA) so that the ball
// centres and radii -- and hence every classification -- match the host oracle
// bit for bit.  Output: lgeo[panel][7][ntot] (x,y,z,nx,ny,nz,w), balls[panel][nsubtot][4].
__global__ void __launch_bounds__(256) level_geom_kernel(LevelArgs A, const double* __restrict__ sx,
                                                         const double* __restrict__ rts,
                                                         const double* __restrict__ rps,
                                                         double* __restrict__ lgeo, double* __restrict__ balls) {
  extern __shared__ double smem[];
  const int h = A.h, k = blockIdx.x;
  double* srcT = smem;              // [9][h]
  double* lx = srcT + 9 * h;        // [4][nmax] x,y,z,w of the current level
  const int nmax = A.lev_n[3] > A.lev_n[2] ? A.lev_n[3] : A.lev_n[2];
  for (int i = threadIdx.x; i < h; i += blockDim.x) {
    const size_t g = (size_t)k * h + i;
    for (int c = 0; c < 3; ++c) {
      srcT[c * h + i] = sx[3 * g + c];
      srcT[(3 + c) * h + i] = rts[3 * g + c];
      srcT[(6 + c) * h + i] = rps[3 * g + c];
    }
  }
  __syncthreads();
  double* G = lgeo + (size_t)k * 7 * A.ntot;
  double* B = balls + (size_t)k * 4 * A.nsubtot;
  for (int l = 0; l < 4; ++l) {
    const int n = A.lev_n[l], off = A.lev_off[l];
    const double* IM = A.interp + (size_t)h * off;  // column-major n x h
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      double geo[9];
      for (int c = 0; c < 9; ++c) geo[c] = 0.0;
      for (int q = 0; q < h; ++q) {
        const double v = IM[i + (size_t)n * q];
        for (int c = 0; c < 9; ++c) geo[c] = geo[c] + v * srcT[c * h + q];
      }
      const double n0 = geo[7] * geo[5] - geo[8] * geo[4];
      const double n1 = geo[8] * geo[3] - geo[6] * geo[5];
      const double n2 = geo[6] * geo[4] - geo[7] * geo[3];
      const double sp = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
      const double w = sp * A.wt[off + i];
      G[0 * A.ntot + off + i] = geo[0];
      G[1 * A.ntot + off + i] = geo[1];
      G[2 * A.ntot + off + i] = geo[2];
      G[3 * A.ntot + off + i] = n0 / sp;
      G[4 * A.ntot + off + i] = n1 / sp;
      G[5 * A.ntot + off + i] = n2 / sp;
      G[6 * A.ntot + off + i] = w;
      lx[i] = geo[0];
      lx[nmax + i] = geo[1];
      lx[2 * nmax + i] = geo[2];
      lx[3 * nmax + i] = w;
    }
    __syncthreads();
    const int ns = A.lev_nsub[l], sb = A.sub_base[l];
    for (int s = threadIdx.x; s < ns; s += blockDim.x) {
      const int i0 = A.suboff[sb + l + s], i1 = A.suboff[sb + l + s + 1];
      double swsum = 0.0;
      for (int i = i0; i < i1; ++i) swsum += lx[3 * nmax + i];
      double qp[3];
      for (int c = 0; c < 3; ++c) {
        double acc = 0.0;
        for (int i = i0; i < i1; ++i) acc += lx[c * nmax + i] * lx[3 * nmax + i];
        qp[c] = acc / swsum;
      }
      double rmax2 = 0.0;
      for (int i = i0; i < i1; ++i) {
        const double d0 = lx[i] - qp[0], d1 = lx[nmax + i] - qp[1], d2 = lx[2 * nmax + i] - qp[2];
        rmax2 = fmax(rmax2, d0 * d0 + d1 * d1 + d2 * d2);
      }
      const double r2 = A.subtyp[sb + s] == 1 ? A.dist2 * fmax(rmax2, 0.5 * swsum) : A.dist2 * fmax(rmax2, swsum);
      B[4 * (sb + s) + 0] = qp[0];
      B[4 * (sb + s) + 1] = qp[1];
      B[4 * (sb + s) + 2] = qp[2];
      B[4 * (sb + s) + 3] = r2;
    }
    __syncthreads();
  }
}

// Classify every (panel, target) pair: the first level whose sub-patch balls all
// exclude the target, else RRQ (label 5).  One workgroup per panel.  Also writes
// the pair's panel id and the per-panel count of RRQ pairs.
__global__ void __launch_bounds__(256) classify_kernel(LevelArgs A, const int* __restrict__ offs,
                                                       const int* __restrict__ tidx, const double* __restrict__ tx,
                                                       const double* __restrict__ balls, int8_t* __restrict__ label,
                                                       int* __restrict__ ppanel, int* __restrict__ nrrq,
                                                       int* __restrict__ lcount) {
  extern __shared__ double sb[];
  __shared__ int cnt[5];
  const int k = blockIdx.x;
  const double* B = balls + (size_t)k * 4 * A.nsubtot;
  for (int i = threadIdx.x; i < 4 * A.nsubtot; i += blockDim.x) sb[i] = B[i];
  if (threadIdx.x < 5) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int j0 = offs[k], j1 = offs[k + 1];
  for (int j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
    const int t = tidx[j];
    const double x0 = tx[3 * (size_t)t], x1 = tx[3 * (size_t)t + 1], x2 = tx[3 * (size_t)t + 2];
    int lab = 5;
    for (int l = 0; l < 4 && lab == 5; ++l) {
      bool outside = true;
      for (int s = A.sub_base[l]; s < A.sub_base[l] + A.lev_nsub[l]; ++s) {
        const double d0 = sb[4 * s] - x0, d1 = sb[4 * s + 1] - x1, d2 = sb[4 * s + 2] - x2;
        if (d0 * d0 + d1 * d1 + d2 * d2 < sb[4 * s + 3]) { outside = false; break; }
      }
      if (outside) lab = l + 1;
    }
    label[j] = (int8_t)lab;
    ppanel[j] = k;
    atomicAdd(&cnt[lab - 1], 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) nrrq[k] = cnt[4];
  if (threadIdx.x < 5) lcount[5 * k + threadIdx.x] = cnt[threadIdx.x];
}

// Direct S/D on level l for the pairs of one panel that classified to l, then
// the contraction with the level's interpolation matrix:
//   As(pair, j) = sum_i S(x, y_i) interp(i, j),  Ad likewise with the DLP kernel.
// One workgroup per panel; targets in tiles of kT, level nodes in chunks of kNC
// staged in LDS with the matching interpolation rows.
constexpr int kT = 32, kNC = 64, kJG = 8;

__global__ void __launch_bounds__(256) direct_level_kernel(LevelArgs A, int l, const int* __restrict__ offs,
                                                           const int* __restrict__ tidx,
                                                           const double* __restrict__ tx,
                                                           const int8_t* __restrict__ label,
                                                           const double* __restrict__ lgeo,
                                                           double* __restrict__ wout) {
  extern __shared__ double sm[];
  const int h = A.h, k = blockIdx.x;
  const int n = A.lev_n[l], off = A.lev_off[l];
  double* sg = sm;                        // [7][kNC] level geometry chunk
  double* sim = sg + 7 * kNC;             // [kNC][h] interpolation rows
  double* sS = sim + (size_t)kNC * h;     // [kT][kNC]
  double* sD = sS + kT * kNC;             // [kT][kNC]
  double* stx = sD + kT * kNC;            // [kT][3]
  __shared__ int plist[kT];
  __shared__ int ntile;
  const double pi4inv = 1.0 / (16.0 * atan(1.0));
  const double* G = lgeo + (size_t)k * 7 * A.ntot;
  const double* IM = A.interp + (size_t)h * off;
  const int j0 = offs[k], j1 = offs[k + 1];
  const int me_t = threadIdx.x / kJG, me_g = threadIdx.x % kJG;
  int scan = j0;
  while (true) {
    // gather the next tile of pairs with this level (in pair order)
    __syncthreads();
    if (threadIdx.x == 0) {
      int c = 0;
      while (scan < j1 && c < kT) {
        if (label[scan] == l + 1) plist[c++] = scan;
        ++scan;
      }
      ntile = c;
    }
    __syncthreads();
    const int nt = ntile;
    if (nt == 0) break;
    for (int i = threadIdx.x; i < 3 * nt; i += blockDim.x) stx[i] = tx[3 * (size_t)tidx[plist[i / 3]] + i % 3];
    double accS[QOL_MAXH / kJG + 1], accD[QOL_MAXH / kJG + 1];
    for (int q = 0; q < QOL_MAXH / kJG + 1; ++q) { accS[q] = 0.0; accD[q] = 0.0; }
    for (int c0 = 0; c0 < n; c0 += kNC) {
      const int nc = min(kNC, n - c0);
      __syncthreads();
      for (int i = threadIdx.x; i < 7 * kNC; i += blockDim.x) {
        const int c = i / kNC, ii = i % kNC;
        sg[i] = ii < nc ? G[(size_t)c * A.ntot + off + c0 + ii] : 0.0;
      }
      for (int i = threadIdx.x; i < kNC * h; i += blockDim.x) {
        const int ii = i % kNC, j = i / kNC;
        sim[ii * h + j] = ii < nc ? IM[(c0 + ii) + (size_t)n * j] : 0.0;
      }
      __syncthreads();
      for (int i = threadIdx.x; i < kT * kNC; i += blockDim.x) {
        const int t = i / kNC, ii = i % kNC;
        double s = 0.0, d = 0.0;
        if (t < nt && ii < nc) {
          const double dx = stx[3 * t] - sg[ii], dy = stx[3 * t + 1] - sg[kNC + ii], dz = stx[3 * t + 2] - sg[2 * kNC + ii];
          const double rr = dx * dx + dy * dy + dz * dz;
          if (rr > 1.0e-28) {
            const double r = sqrt(rr);
            const double w = sg[6 * kNC + ii];
            s = pi4inv * w / r;
            d = pi4inv * w * (dx * sg[3 * kNC + ii] + dy * sg[4 * kNC + ii] + dz * sg[5 * kNC + ii]) / (r * rr);
          }
        }
        sS[i] = s;
        sD[i] = d;
      }
      __syncthreads();
      if (me_t < nt) {
        for (int ii = 0; ii < nc; ++ii) {
          const double s = sS[me_t * kNC + ii], d = sD[me_t * kNC + ii];
          int q = 0;
          for (int j = me_g; j < h; j += kJG, ++q) {
            const double v = sim[ii * h + j];
            accS[q] = accS[q] + s * v;
            accD[q] = accD[q] + d * v;
          }
        }
      }
    }
    if (me_t < nt) {
      double* o = wout + (size_t)plist[me_t] * 2 * h;
      int q = 0;
      for (int j = me_g; j < h; j += kJG, ++q) {
        o[j] = accS[q];
        o[h + j] = accD[q];
      }
    }
  }
}

size_t direct_level_smem(int h) { return sizeof(double) * (7 * kNC + (size_t)kNC * h + 2 * kT * kNC + 3 * kT); }

}  // namespace qol
