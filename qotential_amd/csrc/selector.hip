// Pair classification: the reference's ball rule, and the learned level
// selector that may serve a pair at a coarser level than the rule.
//
// Ball rule (src/patch_refine_mod.f90:541-613): a pair is served at the first
// level whose sub-patch balls all exclude the target (|qpoint - x|^2 < qradii2
// means inside), else by RRQ (label 5).
//
// Learned selector (no counterpart in the reference): from 10 geometry
// features of the pair, an MLP 10 -> 64 -> 64 -> 7 (ReLU; BF16 weights, FP32
// biases and accumulation) picks the coarsest serving that still agrees with
// the ball rule's result to the training tolerance: class 0 the panel's own
// nodes (no correction), 1..4 direct on level 1..4, 5 RRQ, 6 direct on level 4
// where the rule asks for RRQ.  Below the confidence threshold the rule decides.
// Features (all dimensionless; s = sqrt(panel area)):
//   0 log10 |x - c| / s               (c: area-weighted centroid)
//   1 log10 |(x - c).n| / s           (n: mean unit normal; floored at 1e-14)
//   2 log10 min_i |x - y_i| / s       (nearest panel node)
//   3, 4 barycentric l1, l2 of x's projection on the plane of the panel's corners
//   5 min(l0, l1, l2)                 (signed distance to the nearest edge)
//   6 max_i |(y_i - c).n| / s         (out-of-plane bending)
//   7 max_i |n_i - n|                 (normal turning)
//   8 p / 12                          (patch order)
//   9 log10 |x - q1|^2 / r1^2          (margin of the level-1 ball)
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>

#include "qol_kernels.h"

namespace qol {

struct PanelFeat {
  double c[3], n[3], s, X0[3], e1[3], e2[3], i11, i12, i22, ka, kb, q1[3], r1;
};

__device__ __forceinline__ float bf16_to_f32(uint16_t v) { return __uint_as_float((uint32_t)v << 16); }

// panel quantities from the panel's nodes y (3h, in LDS); one thread
__device__ void panel_features(int h, const double* __restrict__ y, const double* __restrict__ snx,
                               const double* __restrict__ sw, const double* __restrict__ corner,
                               const double* __restrict__ ball1, PanelFeat& P) {
  double A = 0.0, c0 = 0.0, c1 = 0.0, c2 = 0.0, m0 = 0.0, m1 = 0.0, m2 = 0.0;
  for (int i = 0; i < h; ++i) {
    const double w = sw[i];
    A += w;
    c0 += w * y[3 * i]; c1 += w * y[3 * i + 1]; c2 += w * y[3 * i + 2];
    m0 += w * snx[3 * i]; m1 += w * snx[3 * i + 1]; m2 += w * snx[3 * i + 2];
  }
  P.c[0] = c0 / A; P.c[1] = c1 / A; P.c[2] = c2 / A;
  const double mn = sqrt(m0 * m0 + m1 * m1 + m2 * m2);
  P.n[0] = m0 / mn; P.n[1] = m1 / mn; P.n[2] = m2 / mn;
  P.s = sqrt(A);
  double X[3][3];
  for (int k = 0; k < 3; ++k)
    for (int d = 0; d < 3; ++d) {
      double a = 0.0;
      for (int i = 0; i < h; ++i) a += corner[k * h + i] * y[3 * i + d];
      X[k][d] = a;
    }
  double g11 = 0.0, g12 = 0.0, g22 = 0.0;
  for (int d = 0; d < 3; ++d) {
    P.X0[d] = X[0][d];
    P.e1[d] = X[1][d] - X[0][d];
    P.e2[d] = X[2][d] - X[0][d];
    g11 += P.e1[d] * P.e1[d]; g12 += P.e1[d] * P.e2[d]; g22 += P.e2[d] * P.e2[d];
  }
  const double det = g11 * g22 - g12 * g12;
  P.i11 = g22 / det; P.i12 = -g12 / det; P.i22 = g11 / det;
  double ka = 0.0, kb = 0.0;
  for (int i = 0; i < h; ++i) {
    const double dz = (y[3 * i] - P.c[0]) * P.n[0] + (y[3 * i + 1] - P.c[1]) * P.n[1] + (y[3 * i + 2] - P.c[2]) * P.n[2];
    ka = fmax(ka, fabs(dz));
    const double a0 = snx[3 * i] - P.n[0], a1 = snx[3 * i + 1] - P.n[1], a2 = snx[3 * i + 2] - P.n[2];
    kb = fmax(kb, sqrt(a0 * a0 + a1 * a1 + a2 * a2));
  }
  P.ka = ka / P.s;
  P.kb = kb;
  P.q1[0] = ball1[0]; P.q1[1] = ball1[1]; P.q1[2] = ball1[2];
  P.r1 = ball1[3];
}

__device__ void pair_feature_vector(int h, const PanelFeat& P, const double* __restrict__ y, double x0, double x1,
                                    double x2, float* f) {
  const int p = (int)((sqrt(8.0 * h + 1.0) - 1.0) * 0.5 + 0.5);  // h = p (p + 1) / 2
  const double d0 = x0 - P.c[0], d1 = x1 - P.c[1], d2 = x2 - P.c[2];
  const double dist = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  const double dn = d0 * P.n[0] + d1 * P.n[1] + d2 * P.n[2];
  double dmin2 = 1e300;
  for (int i = 0; i < h; ++i) {
    const double a0 = x0 - y[3 * i], a1 = x1 - y[3 * i + 1], a2 = x2 - y[3 * i + 2];
    dmin2 = fmin(dmin2, a0 * a0 + a1 * a1 + a2 * a2);
  }
  const double b0 = x0 - P.X0[0], b1 = x1 - P.X0[1], b2 = x2 - P.X0[2];
  const double u1 = P.e1[0] * b0 + P.e1[1] * b1 + P.e1[2] * b2;
  const double u2 = P.e2[0] * b0 + P.e2[1] * b1 + P.e2[2] * b2;
  const double l1 = P.i11 * u1 + P.i12 * u2, l2 = P.i12 * u1 + P.i22 * u2, l0 = 1.0 - l1 - l2;
  const double q0 = x0 - P.q1[0], q1 = x1 - P.q1[1], q2 = x2 - P.q1[2];
  const double s2 = P.s * P.s;
  f[0] = (float)log10(fmax(dist / P.s, 1e-30));
  f[1] = (float)log10(fmax(fabs(dn) / P.s, 1e-14));
  f[2] = (float)(0.5 * log10(fmax(dmin2 / s2, 1e-60)));
  f[3] = (float)l1;
  f[4] = (float)l2;
  f[5] = (float)fmin(l0, fmin(l1, l2));
  f[6] = (float)P.ka;
  f[7] = (float)P.kb;
  f[8] = (float)p / 12.0f;
  f[9] = (float)log10(fmax((q0 * q0 + q1 * q1 + q2 * q2) / P.r1, 1e-60));
}

// the MLP on normalised features: hidden layers h1, h2 and logits[7]; weights W
// (BF16) and biases B anywhere (LDS or global), FP32 accumulation in index order
__device__ __forceinline__ void selector_forward(const uint16_t* __restrict__ W, const float* __restrict__ B,
                                                 const float* xin, float* h1, float* h2, float* out) {
#pragma unroll
  for (int j = 0; j < kSelHid; ++j) {
    float a = B[j];
#pragma unroll
    for (int i = 0; i < kSelIn; ++i) a = fmaf(bf16_to_f32(W[j * kSelIn + i]), xin[i], a);
    h1[j] = fmaxf(a, 0.0f);
  }
#pragma unroll 4
  for (int j = 0; j < kSelHid; ++j) {
    float a = B[kSelHid + j];
#pragma unroll
    for (int i = 0; i < kSelHid; ++i) a = fmaf(bf16_to_f32(W[kSelW1 + j * kSelHid + i]), h1[i], a);
    h2[j] = fmaxf(a, 0.0f);
  }
#pragma unroll
  for (int k = 0; k < kSelOut; ++k) {
    float a = B[2 * kSelHid + k];
#pragma unroll
    for (int i = 0; i < kSelHid; ++i) a = fmaf(bf16_to_f32(W[kSelW1 + kSelW2 + k * kSelHid + i]), h2[i], a);
    out[k] = a;
  }
}

__device__ __forceinline__ void selector_logits(const uint16_t* __restrict__ W, const float* __restrict__ B,
                                                const float* xin, float* out) {
  float h1[kSelHid], h2[kSelHid];
  selector_forward(W, B, xin, h1, h2, out);
}

typedef float sqf4 __attribute__((ext_vector_type(4)));

// per-wave LDS of the classification's network: the features of 64 pairs (row
// stride kSelFS), a 16-pair hidden layer (row stride kSelHS = 4 mod 32, so the
// A-operand reads of rows l%16 start on different banks) and the 64 pairs'
// logits (row stride 8)
constexpr int kSelFS = 13, kSelHS = kSelHid + 4;
constexpr int kSelWave = 64 * kSelFS + 16 * kSelHS + 64 * 8;  // floats

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef __bf16 sbf8 __attribute__((ext_vector_type(8)));

// x = hi + mid + lo exactly, each part a BF16 (8 significant bits each, so the
// three carry the FP32 significand; each residual is exact in FP32)
__device__ __forceinline__ void split_bf16x3(float x, __bf16& hi, __bf16& mid, __bf16& lo) {
  hi = (__bf16)x;
  const float r1 = x - (float)hi;
  mid = (__bf16)r1;
  lo = (__bf16)(r1 - (float)mid);
}

__device__ __forceinline__ __bf16 bf16_bits(uint16_t u) { return __builtin_bit_cast(__bf16, u); }

// One 32-deep step of D += A B for the three BF16 parts of A: lane l supplies
// a[0..7] = A[l%16][8(l/16) + j] (FP32, split here) and b = B[8(l/16) + j][l%16].
__device__ __forceinline__ sqf4 mfma_bf16x3(const float* a, sbf8 b, sqf4 acc) {
  sbf8 ah, am, al;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    __bf16 x, y, z;
    split_bf16x3(a[j], x, y, z);
    ah[j] = x; am[j] = y; al[j] = z;
  }
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, b, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, b, acc, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, b, acc, 0, 0, 0);
}

// The network for the 64 pairs of one wave on the BF16 matrix cores
// (v_mfma_f32_16x16x32_bf16), ngroups groups of 16 pairs: F holds the normalised
// features (row stride kSelFS), the logits go to L[pair][8].  Every FP32
// activation enters as the sum of three BF16 parts and the weights are BF16, so
// every product is exact and the sums are FP32 (the order of the additions
// inside a step is the hardware's); the bias is added after the sum.  Every lane
// of the wave takes part.  Lane l supplies A[l%16][8(l/16)+j] and B[8(l/16)+j][l%16]
// of a 32-deep step and holds D[4(l/16)+r][l%16] of a tile.
__device__ void selector_mlp_wave(const uint16_t* __restrict__ W, const float* __restrict__ Bi, const float* F,
                                  float* Hb, float* L, int ngroups) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  for (int g = 0; g < ngroups; ++g) {
    const float* Fg = F + 16 * g * kSelFS;
    sqf4 acc[4];
    // layer 1: K = 10 in one step (columns 10..31 zero), N = 64 in four tiles
    {
      float a[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int kq = 8 * lk + j;
        a[j] = kq < kSelIn ? Fg[li * kSelFS + kq] : 0.0f;
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        sbf8 b;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int kq = 8 * lk + j;
          b[j] = bf16_bits(kq < kSelIn ? W[(16 * t + li) * kSelIn + kq] : (uint16_t)0);
        }
        acc[t] = mfma_bf16x3(a, b, sqf4{0.0f, 0.0f, 0.0f, 0.0f});
      }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) Hb[(4 * lk + r) * kSelHS + 16 * t + li] = fmaxf(acc[t][r] + Bi[16 * t + li], 0.0f);
    wave_sync();
    // layer 2: K = 64 in two steps, N = 64
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = sqf4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      float a[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) a[j] = Hb[li * kSelHS + 32 * s + 8 * lk + j];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        sbf8 b;
#pragma unroll
        for (int j = 0; j < 8; ++j) b[j] = bf16_bits(W[kSelW1 + (16 * t + li) * kSelHid + 32 * s + 8 * lk + j]);
        acc[t] = mfma_bf16x3(a, b, acc[t]);
      }
    }
    wave_sync();  // every lane has read the first hidden layer
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        Hb[(4 * lk + r) * kSelHS + 16 * t + li] = fmaxf(acc[t][r] + Bi[kSelHid + 16 * t + li], 0.0f);
    wave_sync();
    // layer 3: K = 64 in two steps, N = 7 in one tile (columns 7..15 zero)
    sqf4 o = sqf4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      float a[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) a[j] = Hb[li * kSelHS + 32 * s + 8 * lk + j];
      sbf8 b;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        b[j] = bf16_bits(li < kSelOut ? W[kSelW1 + kSelW2 + min(li, kSelOut - 1) * kSelHid + 32 * s + 8 * lk + j]
                                      : (uint16_t)0);
      o = mfma_bf16x3(a, b, o);
    }
    if (li < kSelOut)
#pragma unroll
      for (int r = 0; r < 4; ++r) L[(16 * g + 4 * lk + r) * 8 + li] = o[r] + Bi[2 * kSelHid + li];
    wave_sync();  // the hidden tile is rewritten by the next group
  }
}

size_t classify_smem(const LevelArgs& A, const SelectorArgs& S) {
  size_t b = sizeof(double) * 4 * (size_t)A.nsubtot;
  if (S.mode == 2) {
    b += sizeof(double) * 3 * (size_t)A.h;                       // panel nodes
    b += sizeof(float) * (kSelNB + 2 * kSelIn);                    // biases, normalisation
    b += sizeof(uint16_t) * kSelNW;                                // weights
    b += 64 + sizeof(PanelFeat);                                   // panel features (8-byte aligned)
    b += sizeof(float) * 4 * kSelWave;                             // the network's per-wave tiles
  }
  return b;
}

// The adaptive walk of one target (patch_refine_mod.f90:740-771, then
// one_level at level 4, :773-842) over the balls in LDS (ball s at b[4s]):
// label 1 when the level-1 cell is a leaf, 2 with the tree cut in `cut` when
// every leaf is at levels 2-3, else 4 (outside every level-4 ball) or 5 (RRQ).
// The walk's leaf test is d^2 > r^2 and one_level's "inside" is d^2 < r^2, as
// in the reference, so a target exactly on a ball is opened by the walk and
// served by level 4.
__device__ __forceinline__ int adaptive_label(const LevelArgs& A, const double* __restrict__ b, double x0, double x1,
                                              double x2, int& cut) {
  auto d2 = [&](int s) {
    const double e0 = b[4 * s] - x0, e1 = b[4 * s + 1] - x1, e2 = b[4 * s + 2] - x2;
    return e0 * e0 + e1 * e1 + e2 * e2;
  };
  cut = 0;
  const int s1 = A.sub_base[0];
  if (d2(s1) > b[4 * s1 + 3]) return 1;
  int open = 0;
  bool ok = true;
  for (int q = 0; q < 4 && ok; ++q) {
    const int s2 = A.sub_base[1] + q;
    if (d2(s2) > b[4 * s2 + 3]) continue;
    open |= 1 << q;
    for (int c = 0; c < 4; ++c) {
      const int s3 = A.sub_base[2] + 4 * q + c;
      if (!(d2(s3) > b[4 * s3 + 3])) { ok = false; break; }
    }
  }
  if (ok) { cut = 1 + open; return 2; }
  for (int s = A.sub_base[3]; s < A.sub_base[3] + A.lev_nsub[3]; ++s)
    if (d2(s) < b[4 * s + 3]) return 5;
  return 4;
}

// One workgroup per panel: every pair's label (ball rule, forced, or learned),
// the pair's panel id, the per-panel counts of labels 1..5 and of RRQ pairs.
// Label-0 pairs are counted in no list.  LEARNED instantiates the selector's
// features and network; the ball-rule instantiation carries none of it.
template <bool LEARNED>
__global__ void __launch_bounds__(256) classify_kernel(LevelArgs A, SelectorArgs S, const int* __restrict__ offs,
                                                       const int* __restrict__ tidx, const double* __restrict__ tx,
                                                       const double* __restrict__ sx, const double* __restrict__ snx,
                                                       const double* __restrict__ sw,
                                                       const double* __restrict__ balls, int8_t* __restrict__ label,
                                                       int* __restrict__ ppanel, int* __restrict__ nrrq,
                                                       int* __restrict__ lcount, int8_t* __restrict__ abucket) {
  extern __shared__ double sb[];
  __shared__ int cnt[5];
  // the panel's feature data lives past the network's parameters in dynamic LDS
  PanelFeat* pf = nullptr;
  const int k = blockIdx.x, h = A.h;
  const double* Bk = balls + (size_t)k * 4 * A.nsubtot;
  for (int i = threadIdx.x; i < 4 * A.nsubtot; i += blockDim.x) sb[i] = Bk[i];
  double* ynod = sb + 4 * A.nsubtot;
  float* fb = (float*)(ynod + 3 * h);
  uint16_t* wl = (uint16_t*)(fb + kSelNB + 2 * kSelIn);
  // past the biases, normalisation and weights, rounded up to whole doubles
  constexpr int kParamDoubles = (int)((sizeof(float) * (kSelNB + 2 * kSelIn) + sizeof(uint16_t) * kSelNW + 7) / 8);
  if (LEARNED) pf = (PanelFeat*)(ynod + 3 * h + kParamDoubles);
  if (LEARNED) {
    for (int i = threadIdx.x; i < 3 * h; i += blockDim.x) ynod[i] = sx[3 * (size_t)k * h + i];
    for (int i = threadIdx.x; i < kSelNB; i += blockDim.x) fb[i] = S.b[i];
    for (int i = threadIdx.x; i < 2 * kSelIn; i += blockDim.x) fb[kSelNB + i] = S.norm[i];
    for (int i = threadIdx.x; i < kSelNW; i += blockDim.x) wl[i] = S.w[i];
  }
  if (threadIdx.x < 5) cnt[threadIdx.x] = 0;
  __syncthreads();
  if (LEARNED && threadIdx.x == 0)
    panel_features(h, ynod, snx + 3 * (size_t)k * h, sw + (size_t)k * h, S.corner, sb + 4 * A.sub_base[0], pf[0]);
  __syncthreads();
  const int j0 = offs[k], j1 = offs[k + 1];
  if (LEARNED) {
    // the learned selector (launched for S.mode == 2 only): each wave takes 64
    // pairs at a time (a loop uniform over the wave, since the network runs on
    // the matrix cores with every lane)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float* F = (float*)(pf + 1) + (size_t)wv * kSelWave;
    float* Hb = F + 64 * kSelFS;
    float* LB = Hb + 16 * kSelHS;
    for (int base = j0 + 64 * wv; base < j1; base += blockDim.x) {
      const int j = base + lane;
      const bool live = j < j1;
      float f[kSelIn];
#pragma unroll
      for (int i = 0; i < kSelIn; ++i) f[i] = 0.0f;
      int lab = 5;
      if (live) {
        const int t = tidx[j];
        const double x0 = tx[3 * (size_t)t], x1 = tx[3 * (size_t)t + 1], x2 = tx[3 * (size_t)t + 2];
        for (int l = 0; l < 4 && lab == 5; ++l) {
          bool outside = true;
          for (int s = A.sub_base[l]; s < A.sub_base[l] + A.lev_nsub[l]; ++s) {
            const double d0 = sb[4 * s] - x0, d1 = sb[4 * s + 1] - x1, d2 = sb[4 * s + 2] - x2;
            if (d0 * d0 + d1 * d1 + d2 * d2 < sb[4 * s + 3]) { outside = false; break; }
          }
          if (outside) lab = l + 1;
        }
        pair_feature_vector(h, pf[0], ynod, x0, x1, x2, f);
#pragma unroll
        for (int i = 0; i < kSelIn; ++i) f[i] = (f[i] - fb[kSelNB + i]) * fb[kSelNB + kSelIn + i];
      }
#pragma unroll
      for (int i = 0; i < kSelIn; ++i) F[lane * kSelFS + i] = f[i];
      F[lane * kSelFS + kSelIn] = 0.0f;
      F[lane * kSelFS + kSelIn + 1] = 0.0f;
      wave_sync();
      selector_mlp_wave(wl, fb, F, Hb, LB, min(4, (j1 - base + 15) / 16));
      if (live) {
        const float* lg = LB + lane * 8;
        int best = 0;
        float mx = lg[0];
#pragma unroll
        for (int c = 1; c < kSelOut; ++c)
          if (lg[c] > mx) { mx = lg[c]; best = c; }
        float z = 0.0f;
#pragma unroll
        for (int c = 0; c < kSelOut; ++c) z += __expf(lg[c] - mx);
        if (1.0f / z >= S.threshold) lab = selector_label(best);
        label[j] = (int8_t)lab;
        ppanel[j] = k;
        if (lab > 0) atomicAdd(&cnt[lab - 1], 1);
      }
      wave_sync();  // the feature and logit tiles are rewritten next round
    }
  }
  for (int j = j0 + threadIdx.x; j < j1 && !LEARNED; j += blockDim.x) {
    const int t = tidx[j];
    const double x0 = tx[3 * (size_t)t], x1 = tx[3 * (size_t)t + 1], x2 = tx[3 * (size_t)t + 2];
    int lab = 5;
    if (!LEARNED && A.adaptive) {
      int cut = 0;
      lab = adaptive_label(A, sb, x0, x1, x2, cut);
      abucket[j] = (int8_t)cut;
    } else {
      for (int l = 0; l < 4 && lab == 5; ++l) {
        bool outside = true;
        for (int s = A.sub_base[l]; s < A.sub_base[l] + A.lev_nsub[l]; ++s) {
          const double d0 = sb[4 * s] - x0, d1 = sb[4 * s + 1] - x1, d2 = sb[4 * s + 2] - x2;
          if (d0 * d0 + d1 * d1 + d2 * d2 < sb[4 * s + 3]) { outside = false; break; }
        }
        if (outside) lab = l + 1;
      }
    }
    if (S.mode == 1) {
      lab = S.forced;
    } else if (S.mode == 3) {
      if (lab <= 4) lab = min(4, lab + S.forced);
    }
    label[j] = (int8_t)lab;
    ppanel[j] = k;
    if (lab > 0) atomicAdd(&cnt[lab - 1], 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) nrrq[k] = cnt[4];
  if (threadIdx.x < 5) lcount[5 * k + threadIdx.x] = cnt[threadIdx.x];
}

void launch_classify(const LevelArgs& A, const SelectorArgs& S, int npan, const int* offs, const int* tidx,
                     const double* tx, const double* sx, const double* snx, const double* sw, const double* balls,
                     int8_t* label, int* ppanel, int* nrrq, int* lcount, int8_t* abucket, hipStream_t s) {
  if (A.adaptive && (S.mode != 0 || !abucket))
    throw std::runtime_error("adaptive levels: the ball-rule walk only (no selector, forced or shifted labels)");
  const size_t lds = classify_smem(A, S);
  if (S.mode == 2) {
    if (lds > 64 * 1024)
      (void)hipFuncSetAttribute((const void*)classify_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(classify_kernel<true>, dim3(npan), dim3(256), lds, s, A, S, offs, tidx, tx, sx, snx, sw, balls,
                       label, ppanel, nrrq, lcount, abucket);
  } else {
    hipLaunchKernelGGL(classify_kernel<false>, dim3(npan), dim3(256), lds, s, A, S, offs, tidx, tx, sx, snx, sw,
                       balls, label, ppanel, nrrq, lcount, abucket);
  }
}

__global__ void __launch_bounds__(256) pair_features_kernel(LevelArgs A, const double* __restrict__ corner,
                                                            const int* __restrict__ offs, const int* __restrict__ tidx,
                                                            const double* __restrict__ tx,
                                                            const double* __restrict__ sx,
                                                            const double* __restrict__ snx,
                                                            const double* __restrict__ sw,
                                                            const double* __restrict__ balls,
                                                            float* __restrict__ feat) {
  extern __shared__ double ynod[];
  __shared__ PanelFeat pf;
  const int k = blockIdx.x, h = A.h;
  for (int i = threadIdx.x; i < 3 * h; i += blockDim.x) ynod[i] = sx[3 * (size_t)k * h + i];
  __syncthreads();
  if (threadIdx.x == 0)
    panel_features(h, ynod, snx + 3 * (size_t)k * h, sw + (size_t)k * h, corner,
                   balls + (size_t)k * 4 * A.nsubtot + 4 * A.sub_base[0], pf);
  __syncthreads();
  for (int j = offs[k] + threadIdx.x; j < offs[k + 1]; j += blockDim.x) {
    const int t = tidx[j];
    float f[kSelIn];
    pair_feature_vector(h, pf, ynod, tx[3 * (size_t)t], tx[3 * (size_t)t + 1], tx[3 * (size_t)t + 2], f);
    for (int i = 0; i < kSelIn; ++i) feat[(size_t)j * kSelIn + i] = f[i];
  }
}


// ---------------------------------------------------------------------------
// Training of the selector on the GPU: forward with the same arithmetic as the
// classification (BF16 weights, FP32 accumulation), class-weighted
// cross-entropy, backward, per-parameter gradient sums over the batch in a
// fixed order (deterministic), Adam on FP32 master weights that are rounded to
// BF16 (nearest even) after every step.

__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// batch rows b < n: the sample perm[first + b] of the training set
__global__ void __launch_bounds__(256) mlp_forward_kernel(int n, int first, const int* __restrict__ perm,
                                                          const float* __restrict__ feat, const int* __restrict__ lab,
                                                          const float* __restrict__ cw, const uint16_t* __restrict__ W,
                                                          const float* __restrict__ B, const float* __restrict__ norm,
                                                          float inv_n, float* __restrict__ X, float* __restrict__ H1,
                                                          float* __restrict__ H2, float* __restrict__ dlog,
                                                          float* __restrict__ loss_part) {
  __shared__ float red[256];
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  float loss = 0.0f;
  if (b < n) {
    const int r = perm ? perm[first + b] : first + b;
    float x[kSelIn], h1[kSelHid], h2[kSelHid], lg[kSelOut];
#pragma unroll
    for (int i = 0; i < kSelIn; ++i) x[i] = (feat[(size_t)r * kSelIn + i] - norm[i]) * norm[kSelIn + i];
    selector_forward(W, B, x, h1, h2, lg);
    float mx = lg[0];
#pragma unroll
    for (int c = 1; c < kSelOut; ++c) mx = fmaxf(mx, lg[c]);
    float z = 0.0f;
#pragma unroll
    for (int c = 0; c < kSelOut; ++c) z += expf(lg[c] - mx);
    const int y = lab[r];
    const float wy = cw[y];
    loss = wy * (logf(z) + mx - lg[y]);
#pragma unroll
    for (int c = 0; c < kSelOut; ++c) {
      const float pc = expf(lg[c] - mx) / z;
      dlog[(size_t)b * kSelOut + c] = wy * (pc - (c == y ? 1.0f : 0.0f)) * inv_n;
    }
#pragma unroll
    for (int i = 0; i < kSelIn; ++i) X[(size_t)b * kSelIn + i] = x[i];
#pragma unroll
    for (int j = 0; j < kSelHid; ++j) {
      H1[(size_t)b * kSelHid + j] = h1[j];
      H2[(size_t)b * kSelHid + j] = h2[j];
    }
  }
  red[threadIdx.x] = loss;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// per sample: dH2 = (W3^T dlog) * [H2 > 0], dH1 = (W2^T dH2) * [H1 > 0]
__global__ void __launch_bounds__(256) mlp_backward_kernel(int n, const uint16_t* __restrict__ W,
                                                           const float* __restrict__ H1, const float* __restrict__ H2,
                                                           const float* __restrict__ dlog, float* __restrict__ dH1,
                                                           float* __restrict__ dH2) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  float d2[kSelHid];
#pragma unroll
  for (int j = 0; j < kSelHid; ++j) {
    float a = 0.0f;
#pragma unroll
    for (int c = 0; c < kSelOut; ++c)
      a = fmaf(bf16_to_f32(W[kSelW1 + kSelW2 + c * kSelHid + j]), dlog[(size_t)b * kSelOut + c], a);
    d2[j] = H2[(size_t)b * kSelHid + j] > 0.0f ? a : 0.0f;
    dH2[(size_t)b * kSelHid + j] = d2[j];
  }
  for (int i = 0; i < kSelHid; ++i) {
    float a = 0.0f;
#pragma unroll
    for (int j = 0; j < kSelHid; ++j) a = fmaf(bf16_to_f32(W[kSelW1 + j * kSelHid + i]), d2[j], a);
    dH1[(size_t)b * kSelHid + i] = H1[(size_t)b * kSelHid + i] > 0.0f ? a : 0.0f;
  }
}

// one thread per parameter, summing over the batch in sample order
__global__ void __launch_bounds__(256) mlp_grad_kernel(int n, const float* __restrict__ X,
                                                       const float* __restrict__ H1, const float* __restrict__ H2,
                                                       const float* __restrict__ dH1, const float* __restrict__ dH2,
                                                       const float* __restrict__ dlog, float* __restrict__ grad) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= kSelNP) return;
  const float *A, *Bv;
  int sa, sb, ia, ib;
  if (q < kSelW1) {
    A = dH1; sa = kSelHid; ia = q / kSelIn; Bv = X; sb = kSelIn; ib = q % kSelIn;
  } else if (q < kSelW1 + kSelW2) {
    const int r = q - kSelW1;
    A = dH2; sa = kSelHid; ia = r / kSelHid; Bv = H1; sb = kSelHid; ib = r % kSelHid;
  } else if (q < kSelNW) {
    const int r = q - kSelW1 - kSelW2;
    A = dlog; sa = kSelOut; ia = r / kSelHid; Bv = H2; sb = kSelHid; ib = r % kSelHid;
  } else {
    const int r = q - kSelNW;  // biases: b1, b2, b3
    if (r < kSelHid) { A = dH1; sa = kSelHid; ia = r; }
    else if (r < 2 * kSelHid) { A = dH2; sa = kSelHid; ia = r - kSelHid; }
    else { A = dlog; sa = kSelOut; ia = r - 2 * kSelHid; }
    Bv = nullptr; sb = 0; ib = 0;
  }
  float g = 0.0f;
  if (Bv) {
    for (int b = 0; b < n; ++b) g = fmaf(A[(size_t)b * sa + ia], Bv[(size_t)b * sb + ib], g);
  } else {
    for (int b = 0; b < n; ++b) g += A[(size_t)b * sa + ia];
  }
  grad[q] = g;
}

__global__ void __launch_bounds__(256) mlp_adam_kernel(float* __restrict__ master, const float* __restrict__ grad,
                                                       float* __restrict__ m, float* __restrict__ v, float lr,
                                                       float beta1, float beta2, float eps, float bc1, float bc2,
                                                       uint16_t* __restrict__ wbf) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= kSelNP) return;
  const float g = grad[q];
  const float mq = beta1 * m[q] + (1.0f - beta1) * g;
  const float vq = beta2 * v[q] + (1.0f - beta2) * g * g;
  m[q] = mq;
  v[q] = vq;
  const float w = master[q] - lr * (mq / bc1) / (sqrtf(vq / bc2) + eps);
  master[q] = w;
  if (q < kSelNW) wbf[q] = f32_to_bf16(w);
}

__global__ void __launch_bounds__(256) mlp_round_kernel(const float* __restrict__ master, uint16_t* __restrict__ wbf) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < kSelNW) wbf[q] = f32_to_bf16(master[q]);
}

__global__ void __launch_bounds__(256) mlp_predict_kernel(int n, const float* __restrict__ feat,
                                                          const uint16_t* __restrict__ W, const float* __restrict__ B,
                                                          const float* __restrict__ norm, float* __restrict__ prob) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  float x[kSelIn], lg[kSelOut];
#pragma unroll
  for (int i = 0; i < kSelIn; ++i) x[i] = (feat[(size_t)b * kSelIn + i] - norm[i]) * norm[kSelIn + i];
  selector_logits(W, B, x, lg);
  float mx = lg[0];
#pragma unroll
  for (int c = 1; c < kSelOut; ++c) mx = fmaxf(mx, lg[c]);
  float z = 0.0f;
#pragma unroll
  for (int c = 0; c < kSelOut; ++c) z += __expf(lg[c] - mx);
#pragma unroll
  for (int c = 0; c < kSelOut; ++c) prob[(size_t)b * kSelOut + c] = __expf(lg[c] - mx) / z;
}

}  // namespace qol

// ---------------------------------------------------------------------------
// Host side of the selector: the model object and its C-ABI.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "qol_engine.h"
#include "qotential_hip.h"

namespace qol {
void set_status(int64_t s);
void set_error(const std::string& s);

struct Mlp {
  DevBuf master, wbf, m, v, grad, norm, cw;
  DevBuf feat, lab, perm, X, H1, H2, dH1, dH2, dlog, lossp;
  int64_t t = 0;
  int64_t cap = 0;  // batch rows the work buffers hold
  hipStream_t s = nullptr;
  Mlp() {
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    master.alloc(sizeof(float) * kSelNP);
    wbf.alloc(sizeof(uint16_t) * kSelNW);
    m.alloc(sizeof(float) * kSelNP);
    v.alloc(sizeof(float) * kSelNP);
    grad.alloc(sizeof(float) * kSelNP);
    norm.alloc(sizeof(float) * 2 * kSelIn);
    cw.alloc(sizeof(float) * kSelOut);
    HIPCHK(hipMemsetAsync(master.ptr, 0, sizeof(float) * kSelNP, s));
    HIPCHK(hipMemsetAsync(m.ptr, 0, sizeof(float) * kSelNP, s));
    HIPCHK(hipMemsetAsync(v.ptr, 0, sizeof(float) * kSelNP, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  ~Mlp() {
    if (s) (void)hipStreamDestroy(s);
  }
  void reserve(int64_t rows) {
    if (rows <= cap) return;
    X.alloc(sizeof(float) * kSelIn * rows);
    H1.alloc(sizeof(float) * kSelHid * rows);
    H2.alloc(sizeof(float) * kSelHid * rows);
    dH1.alloc(sizeof(float) * kSelHid * rows);
    dH2.alloc(sizeof(float) * kSelHid * rows);
    dlog.alloc(sizeof(float) * kSelOut * rows);
    lossp.alloc(sizeof(float) * ((rows + 255) / 256));
    cap = rows;
  }
  void upload_set(int64_t n, const float* f, const int64_t* y) {
    std::vector<int> yi(n);
    for (int64_t i = 0; i < n; ++i) {
      if (y[i] < 0 || y[i] >= kSelOut) throw std::runtime_error("selector label out of range 0..6");
      yi[i] = (int)y[i];
    }
    feat.alloc(sizeof(float) * kSelIn * std::max<int64_t>(n, 1));
    lab.alloc(sizeof(int) * std::max<int64_t>(n, 1));
    HIPCHK(hipMemcpyAsync(feat.ptr, f, sizeof(float) * kSelIn * n, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(lab.ptr, yi.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  // forward + backward + gradient over rows first..first+nb of the (permuted) set; returns the loss sum
  double step_grad(int64_t first, int64_t nb, const int* dperm) {
    const int g = (int)((nb + 255) / 256);
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(g), dim3(256), 0, s, (int)nb, (int)first, dperm,
                       (const float*)feat.ptr, (const int*)lab.ptr, (const float*)cw.ptr, (const uint16_t*)wbf.ptr,
                       (const float*)master.ptr + kSelNW, (const float*)norm.ptr, 1.0f / (float)nb, (float*)X.ptr,
                       (float*)H1.ptr, (float*)H2.ptr, (float*)dlog.ptr, (float*)lossp.ptr);
    hipLaunchKernelGGL(mlp_backward_kernel, dim3(g), dim3(256), 0, s, (int)nb, (const uint16_t*)wbf.ptr,
                       (const float*)H1.ptr, (const float*)H2.ptr, (const float*)dlog.ptr, (float*)dH1.ptr,
                       (float*)dH2.ptr);
    hipLaunchKernelGGL(mlp_grad_kernel, dim3((kSelNP + 255) / 256), dim3(256), 0, s, (int)nb, (const float*)X.ptr,
                       (const float*)H1.ptr, (const float*)H2.ptr, (const float*)dH1.ptr, (const float*)dH2.ptr,
                       (const float*)dlog.ptr, (float*)grad.ptr);
    HIPCHK(hipGetLastError());
    std::vector<float> lp(g);
    HIPCHK(hipMemcpyAsync(lp.data(), lossp.ptr, sizeof(float) * g, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double tot = 0.0;
    for (float x : lp) tot += x;
    return tot;
  }
};

Mlp* M(void* h) {
  if (!h) throw std::runtime_error("null selector handle");
  return static_cast<Mlp*>(h);
}

SelectorArgs selector_args(void* mlp, double threshold) {
  SelectorArgs S;
  if (!mlp) return S;
  Mlp* m = M(mlp);
  S.mode = 2;
  S.threshold = (float)threshold;
  S.w = (const uint16_t*)m->wbf.ptr;
  S.b = (const float*)m->master.ptr + kSelNW;
  S.norm = (const float*)m->norm.ptr;
  return S;
}

}  // namespace qol

using namespace qol;

namespace {
template <class F>
int guarded_sel(F&& f) {
  try {
    f();
    set_status(0);
    return 0;
  } catch (const std::exception& e) {
    set_error(e.what());
    set_status(1);
    return 1;
  }
}
}  // namespace

extern "C" {

int qol_selector_create(void** out) {
  return guarded_sel([&] {
    *out = nullptr;
    *out = new Mlp();
  });
}

int qol_selector_destroy(void* h) {
  return guarded_sel([&] { delete M(h); });
}

int qol_selector_set_params(void* h, const float* params, const float* norm) {
  return guarded_sel([&] {
    Mlp* m = M(h);
    HIPCHK(hipMemcpyAsync(m->master.ptr, params, sizeof(float) * kSelNP, hipMemcpyHostToDevice, m->s));
    HIPCHK(hipMemcpyAsync(m->norm.ptr, norm, sizeof(float) * 2 * kSelIn, hipMemcpyHostToDevice, m->s));
    hipLaunchKernelGGL(mlp_round_kernel, dim3((kSelNW + 255) / 256), dim3(256), 0, m->s, (const float*)m->master.ptr,
                       (uint16_t*)m->wbf.ptr);
    HIPCHK(hipMemsetAsync(m->m.ptr, 0, sizeof(float) * kSelNP, m->s));
    HIPCHK(hipMemsetAsync(m->v.ptr, 0, sizeof(float) * kSelNP, m->s));
    m->t = 0;
    HIPCHK(hipStreamSynchronize(m->s));
  });
}

int qol_selector_get_params(void* h, float* params, float* norm) {
  return guarded_sel([&] {
    Mlp* m = M(h);
    HIPCHK(hipMemcpyAsync(params, m->master.ptr, sizeof(float) * kSelNP, hipMemcpyDeviceToHost, m->s));
    if (norm) HIPCHK(hipMemcpyAsync(norm, m->norm.ptr, sizeof(float) * 2 * kSelIn, hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipStreamSynchronize(m->s));
  });
}

int qol_selector_grad(void* h, int64_t n, const float* feat, const int64_t* labels, const float* class_w,
                      float* grad, double* loss) {
  return guarded_sel([&] {
    Mlp* m = M(h);
    if (n <= 0) throw std::runtime_error("qol_selector_grad: n must be positive");
    m->upload_set(n, feat, labels);
    HIPCHK(hipMemcpyAsync(m->cw.ptr, class_w, sizeof(float) * kSelOut, hipMemcpyHostToDevice, m->s));
    m->reserve(n);
    const double l = m->step_grad(0, n, nullptr);
    if (loss) *loss = l / (double)n;
    HIPCHK(hipMemcpyAsync(grad, m->grad.ptr, sizeof(float) * kSelNP, hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipStreamSynchronize(m->s));
  });
}

int qol_selector_train(void* h, int64_t n, const float* feat, const int64_t* labels, const float* class_w,
                       int64_t epochs, int64_t batch, double lr, uint64_t seed, double* loss_epoch) {
  return guarded_sel([&] {
    Mlp* m = M(h);
    if (n <= 0 || batch <= 0 || epochs <= 0) throw std::runtime_error("qol_selector_train: bad sizes");
    m->upload_set(n, feat, labels);
    HIPCHK(hipMemcpyAsync(m->cw.ptr, class_w, sizeof(float) * kSelOut, hipMemcpyHostToDevice, m->s));
    m->reserve(std::min(batch, n));
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 rng(seed);
    m->perm.alloc(sizeof(int) * n);
    for (int64_t e = 0; e < epochs; ++e) {
      std::shuffle(perm.begin(), perm.end(), rng);
      HIPCHK(hipMemcpyAsync(m->perm.ptr, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, m->s));
      double tot = 0.0;
      for (int64_t first = 0; first < n; first += batch) {
        const int64_t nb = std::min(batch, n - first);
        tot += m->step_grad(first, nb, (const int*)m->perm.ptr);
        m->t += 1;
        const float bc1 = 1.0f - std::pow(0.9f, (float)m->t), bc2 = 1.0f - std::pow(0.999f, (float)m->t);
        hipLaunchKernelGGL(mlp_adam_kernel, dim3((kSelNP + 255) / 256), dim3(256), 0, m->s, (float*)m->master.ptr,
                           (const float*)m->grad.ptr, (float*)m->m.ptr, (float*)m->v.ptr, (float)lr, 0.9f, 0.999f,
                           1e-8f, bc1, bc2, (uint16_t*)m->wbf.ptr);
      }
      HIPCHK(hipStreamSynchronize(m->s));
      if (loss_epoch) loss_epoch[e] = tot / (double)n;
    }
  });
}

int qol_selector_predict(void* h, int64_t n, const float* feat, float* prob) {
  return guarded_sel([&] {
    Mlp* m = M(h);
    if (n <= 0) return;
    DevBuf f, p;
    f.alloc(sizeof(float) * kSelIn * n);
    p.alloc(sizeof(float) * kSelOut * n);
    HIPCHK(hipMemcpyAsync(f.ptr, feat, sizeof(float) * kSelIn * n, hipMemcpyHostToDevice, m->s));
    hipLaunchKernelGGL(mlp_predict_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->s, (int)n,
                       (const float*)f.ptr, (const uint16_t*)m->wbf.ptr, (const float*)m->master.ptr + kSelNW,
                       (const float*)m->norm.ptr, (float*)p.ptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(prob, p.ptr, sizeof(float) * kSelOut * n, hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipStreamSynchronize(m->s));
  });
}

}  // extern "C"
