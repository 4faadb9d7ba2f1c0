// Refinement levels on the GPU: per-panel level geometry and sub-patch balls,
// pair classification, and direct S/D on the upsampled levels contracted with the
// interpolation matrix on the FP64 matrix cores.
//
// Reference: build_patch_levels_r64   patch_refine_mod.f90:423-502
//            lap3dsdlpmat_levels_r64  patch_refine_mod.f90:516-615
//            lap3dsdlpmat_r64 (direct kernel; zero at coincident points)
#include <stdexcept>
#include <vector>

#include "qol_kernels.h"

namespace qol {

// kLP panels per workgroup, so each interpolation-matrix entry is loaded once
// for all of them.  Level nodes are interpolated as sequential fused multiply-adds
// over the h panel nodes, the same operations in the same order as the host
// oracle, so the ball centres and radii -- and hence every classification --
// match it bit for bit.
// Work goes by chunks of whole sub-patches of one level (at most kChunk nodes,
// one node per thread; the host builds the chunk table): the chunk's nodes are
// written to HBM (plain stores: streaming ones measured slower) and kept in LDS,
// and the balls of its sub-patches are summed from LDS in node order, so nothing
// is read back.
// Output: lgeo[panel][ntot][8] (x,y,z,nx,ny,nz,w, pad: 64 bytes per node, so one
// node step of the direct kernel reads two cache lines), balls[panel][nsubtot][4].
constexpr int kLP = 4;
constexpr int kChunk = 256;
// LDS stride of one panel's chunk buffer (4 x kChunk doubles): 8 doubles of
// padding put the kLP panels of one node in different LDS banks (an unpadded
// stride, a multiple of the bank count, serialised their accesses 4 ways)
constexpr int kCP = 4 * kChunk + 8;

size_t level_geom_smem(int h) { return sizeof(double) * (size_t)kLP * kCP; }
int level_geom_blocks(int npan) { return (npan + kLP - 1) / kLP; }

std::vector<int> level_chunks(const LevelArgs& A, const int* suboff_host) {
  std::vector<int> ch;
  for (int l = 0; l < 4; ++l) {
    const int* so = suboff_host + A.sub_base[l] + l;  // nsub + 1 offsets of level l
    int s0 = 0;
    while (s0 < A.lev_nsub[l]) {
      int s1 = s0 + 1;
      if (so[s1] - so[s0] > kChunk) throw std::runtime_error("level_chunks: a sub-patch exceeds the chunk size");
      while (s1 < A.lev_nsub[l] && so[s1 + 1] - so[s0] <= kChunk) ++s1;
      ch.insert(ch.end(), {l, so[s0], so[s1], s0, s1});
      s0 = s1;
    }
  }
  return ch;
}

// balls of a chunk's sub-patches (node range [c0, ...) of level l, sub-patches
// [sb0, sb1)) from the chunk's x, y, z, w in LDS, summed in node order
__device__ __forceinline__ void chunk_balls(const LevelArgs& A, const double* cn, int k0, int np_, int l, int c0,
                                            int sb0, int sb1, double* __restrict__ balls) {
  const int nsc = sb1 - sb0;
  for (int e = threadIdx.x; e < np_ * nsc; e += blockDim.x) {
    const int pp = e / nsc, s = sb0 + e % nsc;
    const int s_all = A.sub_base[l] + s;
    const int i0 = A.suboff[A.sub_base[l] + l + s] - c0, i1 = A.suboff[A.sub_base[l] + l + s + 1] - c0;
    const double* L = cn + (size_t)pp * kCP;
    // four independent sums, each in node order (the oracle's), in one pass; the
    // LDS loads go in batches of kB nodes ahead of the additions, which keep
    // their order (one thread serialises a sub-patch, so load latency, not the
    // adds, bounded the unbatched loop)
    constexpr int kB = 8;
    double swsum = 0.0, a0 = 0.0, a1 = 0.0, a2 = 0.0;
    int j = i0;
    for (; j + kB <= i1; j += kB) {
      double bx[kB], by[kB], bz[kB], bw[kB];
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        bx[u] = L[j + u];
        by[u] = L[kChunk + j + u];
        bz[u] = L[2 * kChunk + j + u];
        bw[u] = L[3 * kChunk + j + u];
      }
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        swsum += bw[u];
        a0 += bx[u] * bw[u];
        a1 += by[u] * bw[u];
        a2 += bz[u] * bw[u];
      }
    }
    for (; j < i1; ++j) {
      const double w = L[3 * kChunk + j];
      swsum += w;
      a0 += L[j] * w;
      a1 += L[kChunk + j] * w;
      a2 += L[2 * kChunk + j] * w;
    }
    const double qp[3] = {a0 / swsum, a1 / swsum, a2 / swsum};
    // a maximum is exact in any order: kB partial maxima over batches
    double rm[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) rm[u] = 0.0;
    j = i0;
    for (; j + kB <= i1; j += kB) {
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        const double d0 = L[j + u] - qp[0], d1 = L[kChunk + j + u] - qp[1], d2 = L[2 * kChunk + j + u] - qp[2];
        rm[u] = fmax(rm[u], d0 * d0 + d1 * d1 + d2 * d2);
      }
    }
    for (; j < i1; ++j) {
      const double d0 = L[j] - qp[0], d1 = L[kChunk + j] - qp[1], d2 = L[2 * kChunk + j] - qp[2];
      rm[0] = fmax(rm[0], d0 * d0 + d1 * d1 + d2 * d2);
    }
    double rmax2 = rm[0];
#pragma unroll
    for (int u = 1; u < kB; ++u) rmax2 = fmax(rmax2, rm[u]);
    const double r2 = A.subtyp[s_all] == 1 ? A.dist2 * fmax(rmax2, 0.5 * swsum) : A.dist2 * fmax(rmax2, swsum);
    double* B = balls + (size_t)(k0 + pp) * 4 * A.nsubtot + 4 * s_all;
    B[0] = qp[0];
    B[1] = qp[1];
    B[2] = qp[2];
    B[3] = r2;
  }
}

__global__ void __launch_bounds__(256) level_geom_kernel(LevelArgs A, int npan, const double* __restrict__ sx,
                                                         const double* __restrict__ rts,
                                                         const double* __restrict__ rps,
                                                         double* __restrict__ lgeo, double* __restrict__ balls) {
  extern __shared__ double smem[];
  const int h = A.h, k0 = blockIdx.x * kLP;
  const int np_ = min(kLP, npan - k0);
  const size_t nt_ = A.ntot;
  double* cn = smem;  // [kLP][kCP]: x, y, z, w of the chunk's nodes (4 x kChunk) and padding
  for (int ch = 0; ch < A.nchunk; ++ch) {
    const int* C = A.chunks + 5 * ch;
    const int l = C[0], c0 = C[1], c1 = C[2], sb0 = C[3], sb1 = C[4];
    const int n = A.lev_n[l], off = A.lev_off[l];
    const double* IM = A.interp + (size_t)h * off;  // column-major n x h
    const int i = c0 + (int)threadIdx.x;
    if (i < c1) {
      double geo[kLP][9];
#pragma unroll
      for (int pp = 0; pp < kLP; ++pp)
#pragma unroll
        for (int c = 0; c < 9; ++c) geo[pp][c] = 0.0;
      // sequential fma in q for every accumulator (the oracle's order); the next
      // interpolation entry is fetched one step ahead.  The panel values of a
      // step are the same for the whole workgroup: read with uniform addresses
      // they come through the scalar cache and feed the FMAs from scalar
      // registers (no LDS traffic).
      double vnext = IM[i];
      for (int q = 0; q < h; ++q) {
        const double v = vnext;
        if (q + 1 < h) vnext = IM[i + (size_t)n * (q + 1)];
#pragma unroll
        for (int pp = 0; pp < kLP; ++pp) {
          const size_t g = (size_t)(k0 + (pp < np_ ? pp : 0)) * h + q;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            geo[pp][c] = fma(v, sx[3 * g + c], geo[pp][c]);
            geo[pp][3 + c] = fma(v, rts[3 * g + c], geo[pp][3 + c]);
            geo[pp][6 + c] = fma(v, rps[3 * g + c], geo[pp][6 + c]);
          }
        }
      }
      const int li = i - c0;
#pragma unroll
      for (int pp = 0; pp < kLP; ++pp) {
        if (pp >= np_) break;
        const double* gp = geo[pp];
        const double n0 = gp[7] * gp[5] - gp[8] * gp[4];
        const double n1 = gp[8] * gp[3] - gp[6] * gp[5];
        const double n2 = gp[6] * gp[4] - gp[7] * gp[3];
        const double sp = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        const double w = sp * A.wt[off + i];
        double2* G = (double2*)(lgeo + 8 * ((size_t)(k0 + pp) * nt_ + off + i));
        G[0] = double2{gp[0], gp[1]};
        G[1] = double2{gp[2], n0 / sp};
        G[2] = double2{n1 / sp, n2 / sp};
        G[3] = double2{w, 0.0};
        double* L = cn + (size_t)pp * kCP + li;
        L[0] = gp[0];
        L[kChunk] = gp[1];
        L[2 * kChunk] = gp[2];
        L[3 * kChunk] = w;
      }
    }
    __syncthreads();
    chunk_balls(A, cn, k0, np_, l, c0, sb0, sb1, balls);
    __syncthreads();  // the chunk buffer is rewritten by the next chunk
  }
}

// Level geometry on the FP64 matrix cores (h up to 80, orders up to 12, in KS = ceil(h/4) blocks of 4):
// per workgroup kLP panels, per wave tiles of 16 level nodes,
//   D(node, 9*pp + c) = sum_q IM(node, q) S_pp(q, c)   (c: x, y, z, x_u, x_v)
// as chained v_mfma_f64_16x16x4 over blocks of 4 panel nodes.  The matrix core
// forms each block as fma(a3,b3, fma(a2,b2, fma(a1,b1, fma(a0,b0, c)))) (checked
// bit for bit on gfx950), so the chain is the oracle's sequential fma over q and
// the level nodes -- and the balls -- are bit-identical to it.  The panel values
// (B operand) stay in registers for the whole workgroup; a tile's results pass
// through a small per-wave LDS stage to one lane per (node, panel).
constexpr int kLV = 4 * 9;  // output values per node (kLP panels x 9)
size_t level_geom_mfma_smem() { return sizeof(double) * ((size_t)kLP * kCP + (size_t)4 * 16 * kLV); }

template <int KS>
__global__ void __launch_bounds__(256) level_geom_mfma_kernel(LevelArgs A, int npan, const double* __restrict__ sx,
                                                              const double* __restrict__ rts,
                                                              const double* __restrict__ rps,
                                                              double* __restrict__ lgeo, double* __restrict__ balls) {
  static_assert(kLP * 9 == kLV && kLV <= 48, "three 16-column tiles");
  extern __shared__ double smem[];
  const int h = A.h, k0 = blockIdx.x * kLP;
  const int np_ = min(kLP, npan - k0);
  const size_t nt_ = A.ntot;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
  double* cn = smem;                                           // [kLP][kCP]
  double* stg = smem + (size_t)kLP * kCP + (size_t)wv * 16 * kLV;  // [16 nodes][kLV], this wave
  // B operand: lane supplies B[lk][li] of block s, tile t = S(q = 4s + lk, column 16t + li)
  double bf[KS][3];
#pragma unroll
  for (int s = 0; s < KS; ++s)
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      const int col = 16 * t + li, pp = col / 9, c = col % 9, q = 4 * s + lk;
      double v = 0.0;
      if ((col < kLV) & (pp < np_) & (q < h)) {
        const size_t g = (size_t)(k0 + pp) * h + q;
        v = c < 3 ? sx[3 * g + c] : (c < 6 ? rts[3 * g + c - 3] : rps[3 * g + c - 6]);
      }
      bf[s][t] = v;
    }
  for (int ch = 0; ch < A.nchunk; ++ch) {
    const int* C = A.chunks + 5 * ch;
    const int l = C[0], c0 = C[1], c1 = C[2], sb0 = C[3], sb1 = C[4];
    const int off = A.lev_off[l];
    const int ntile = (c1 - c0 + 15) / 16;
    for (int tl = wv; tl < ntile; tl += 4) {
      // A operand: lane supplies A[li][lk] of block s = IM(node c0 + 16 tl + li, q = 4s + lk),
      // from the column-major copy, so one load reads 16 consecutive nodes of 4
      // columns (4 cache lines instead of the 16 rows of the row-major copy);
      // columns past h are zero
      const int nd = min(c0 + 16 * tl + li, c1 - 1);
      const int nl = A.lev_n[l], st = 4 * nl;  // 4 columns
      const double* cb = A.interp + (size_t)h * off + nd;  // column 0 of this lane's node
      const double* col = cb + (size_t)nl * lk;
      double af[KS];
#pragma unroll
      for (int s = 0; s < KS - 1; ++s) af[s] = col[s * st];
      {
        // only the last block can pass column h: the load stays in the table
        const int q = 4 * (KS - 1) + lk;
        const double v = cb[(size_t)nl * min(q, h - 1)];
        af[KS - 1] = q < h ? v : 0.0;
      }
      qd4 acc[3];
#pragma unroll
      for (int t = 0; t < 3; ++t) acc[t] = qd4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int t = 0; t < 3; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[s], bf[s][t], acc[t], 0, 0, 0);
      // this lane holds D[4j + lk][16t + li] in slot j of tile t
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        const int col = 16 * t + li;
        if (col < kLV) {
#pragma unroll
          for (int j = 0; j < 4; ++j) stg[(4 * j + lk) * kLV + col] = acc[t][j];
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // one lane per (node, panel)
      {
        const int nl = lane >> 2, pp = lane & 3;
        const int i = c0 + 16 * tl + nl;
        if ((i < c1) & (pp < np_)) {
          const double* gp = stg + nl * kLV + 9 * pp;
          const double n0 = gp[7] * gp[5] - gp[8] * gp[4];
          const double n1 = gp[8] * gp[3] - gp[6] * gp[5];
          const double n2 = gp[6] * gp[4] - gp[7] * gp[3];
          const double sp = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
          const double w = sp * A.wt[off + i];
          double2* G = (double2*)(lgeo + 8 * ((size_t)(k0 + pp) * nt_ + off + i));
          G[0] = double2{gp[0], gp[1]};
          G[1] = double2{gp[2], n0 / sp};
          G[2] = double2{n1 / sp, n2 / sp};
          G[3] = double2{w, 0.0};
          double* L = cn + (size_t)pp * kCP + (i - c0);
          L[0] = gp[0];
          L[kChunk] = gp[1];
          L[2 * kChunk] = gp[2];
          L[3 * kChunk] = w;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __syncthreads();
    chunk_balls(A, cn, k0, np_, l, c0, sb0, sb1, balls);
    __syncthreads();  // the chunk buffer is rewritten by the next chunk
  }
}

void launch_level_geom(const LevelArgs& A, int npan, const double* sx, const double* rts, const double* rps,
                       double* lgeo, double* balls, hipStream_t s) {
  const int ks = (A.h + 3) / 4;
  const dim3 grid(level_geom_blocks(npan)), blk(256);
  if (ks <= 20 && 4 * ks <= A.pad) {
    const size_t lds = level_geom_mfma_smem();
#define QOL_LEVEL_MFMA(KS_)                                                                                    \
  case KS_:                                                                                                    \
    hipLaunchKernelGGL(level_geom_mfma_kernel<KS_>, grid, blk, lds, s, A, npan, sx, rts, rps, lgeo, balls);  \
    return;
    switch (ks) {
      QOL_LEVEL_MFMA(1) QOL_LEVEL_MFMA(2) QOL_LEVEL_MFMA(3) QOL_LEVEL_MFMA(4) QOL_LEVEL_MFMA(5)
      QOL_LEVEL_MFMA(6) QOL_LEVEL_MFMA(7) QOL_LEVEL_MFMA(8) QOL_LEVEL_MFMA(9) QOL_LEVEL_MFMA(10)
      QOL_LEVEL_MFMA(11) QOL_LEVEL_MFMA(12) QOL_LEVEL_MFMA(13) QOL_LEVEL_MFMA(14) QOL_LEVEL_MFMA(15)
      QOL_LEVEL_MFMA(16) QOL_LEVEL_MFMA(17) QOL_LEVEL_MFMA(18) QOL_LEVEL_MFMA(19) QOL_LEVEL_MFMA(20)
      default: break;
    }
#undef QOL_LEVEL_MFMA
  }
  const size_t lds = level_geom_smem(A.h);
  if (lds > 64 * 1024)
    (void)hipFuncSetAttribute((const void*)level_geom_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(level_geom_kernel, grid, blk, lds, s, A, npan, sx, rts, rps, lgeo, balls);
}

// Per-label pair lists, stable (panel order, then pair order inside a panel):
// lstart[l][k] = first slot of panel k's label-(l+1) pairs inside list l,
// ltot[l] = length of list l.  One block per label.
__global__ void __launch_bounds__(1024) level_scan_kernel(const int* __restrict__ lcount, int npan,
                                                          int* __restrict__ lstart, int* __restrict__ ltot) {
  __shared__ int part[1024];
  const int l = blockIdx.x;
  const int per = (npan + blockDim.x - 1) / blockDim.x;
  const int a = threadIdx.x * per, b = min(npan, a + per);
  int s = 0;
  for (int k = a; k < b; ++k) s += lcount[5 * k + l];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int i = 0; i < (int)blockDim.x; ++i) { const int v = part[i]; part[i] = acc; acc += v; }
    ltot[l] = acc;
  }
  __syncthreads();
  int acc = part[threadIdx.x];
  for (int k = a; k < b; ++k) { lstart[(size_t)l * npan + k] = acc; acc += lcount[5 * k + l]; }
}

// Scatter every pair index into its label's list (lists concatenated in label
// order: list l starts at ltot[0] + ... + ltot[l-1]).  One workgroup per panel;
// order inside a panel kept by wave ballots.
__global__ void __launch_bounds__(256) compact_kernel(const int* __restrict__ offs, const int8_t* __restrict__ label,
                                                      const int* __restrict__ lstart, const int* __restrict__ ltot,
                                                      int npan, int* __restrict__ lpairs) {
  __shared__ int wcnt[4][5];
  __shared__ int run[5];
  const int k = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (threadIdx.x < 5) {
    int base = 0;
    for (int q = 0; q < (int)threadIdx.x; ++q) base += ltot[q];
    run[threadIdx.x] = base + lstart[(size_t)threadIdx.x * npan + k];
  }
  __syncthreads();
  const int j0 = offs[k], j1 = offs[k + 1];
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int c = j0; c < j1; c += blockDim.x) {
    const int j = c + threadIdx.x;
    const int lab = j < j1 ? (int)label[j] : 0;
    int rank = 0;
    for (int q = 1; q <= 5; ++q) {
      const unsigned long long bal = __ballot(lab == q);
      if (lab == q) rank = __popcll(bal & below);
      if (lane == 0) wcnt[wv][q - 1] = __popcll(bal);
    }
    __syncthreads();
    if (lab > 0) {
      int pos = run[lab - 1] + rank;
      for (int w = 0; w < wv; ++w) pos += wcnt[w][lab - 1];
      lpairs[pos] = j;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
      int add = 0;
      for (int w = 0; w < 4; ++w) add += wcnt[w][threadIdx.x];
      run[threadIdx.x] += add;
    }
    __syncthreads();
  }
}

// Direct S/D on the refinement levels with FP64 matrix cores.  One wave per
// tile of 16 pairs of one level (tiles of the heaviest level first); for each
// step of 4 level nodes every lane forms one kernel entry,
//   S(pair m, node i) = w_i / (4 pi |x - y_i|),  D = w_i (x - y_i).n_i / (4 pi |x - y_i|^3)
// (zero at coincident points), which is the A operand of
// v_mfma_f64_16x16x4 against 4 rows of the level's interpolation matrix
// (row-major, padded to 16*NT columns).  Accumulators: 16 pairs x 16*NT
// outputs for S and for D.

// One step of the direct contraction: this lane's pair (A row m of the tile)
// against node ic (the step's row kq), S and D entries on the vector ALUs, then
// accS/D[q] += entry x row ic of the padded interpolation table B (columns
// 16q + m) on the matrix cores.  use = false gives exactly zero entries (the
// operands stay finite: clamped index, r^2 = 1), so no branch is needed.
template <int NT>
__device__ __forceinline__ void direct_step(const double* __restrict__ G, const double* __restrict__ B, int pad,
                                            int m, int ic, bool use, double x0, double x1, double x2, double c4,
                                            qd4* accS, qd4* accD) {
  const double2* g2 = (const double2*)(G + 8 * ic);
  const double2 ga = g2[0], gb = g2[1], gc = g2[2], gd = g2[3];
  const double gx = ga.x, gy = ga.y, gz = gb.x, nx = gb.y, ny = gc.x, nz = gc.y, gw = gd.x;
  double bq[NT];
#pragma unroll
  for (int q = 0; q < NT; ++q) bq[q] = B[(size_t)ic * pad + 16 * q + m];
  const double dx = x0 - gx, dy = x1 - gy, dz = x2 - gz;
  const double rr = dx * dx + dy * dy + dz * dz;
  const bool ok = use && rr > 1.0e-28;
  const double rs = ok ? rr : 1.0;
  // 1/r from the hardware estimate and two Newton steps
  double y = __builtin_amdgcn_rsq(rs);
  double t = rs * y;
  y = fma(0.5 * y, fma(-t, y, 1.0), y);
  t = rs * y;
  y = fma(0.5 * y, fma(-t, y, 1.0), y);
  const double msk = ok ? 1.0 : 0.0;
  const double cw = msk * (c4 * gw);
  const double dn = dx * nx + dy * ny + dz * nz;
  const double s = cw * y;
  const double d = cw * dn * (y * y * y);
  // past a segment's last node the A entries are exactly zero and the B row
  // (the clamped node's) is finite, so B needs no mask
#pragma unroll
  for (int q = 0; q < NT; ++q) {
    accS[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(s, bq[q], accS[q], 0, 0, 0);
    accD[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(d, bq[q], accD[q], 0, 0, 0);
  }
}

template <int NT>
__global__ void __launch_bounds__(64) direct_mfma_kernel(LevelArgs A, const int* __restrict__ lpairs,
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
  const double* B = A.imp + (size_t)pad * A.lev_off[l];
  const double c4 = 1.0 / (16.0 * atan(1.0));
  qd4 accS[NT], accD[NT];
#pragma unroll
  for (int q = 0; q < NT; ++q) {
    accS[q] = qd4{0.0, 0.0, 0.0, 0.0};
    accD[q] = qd4{0.0, 0.0, 0.0, 0.0};
  }
  // branch-free body (clamped indices, selects) so the accumulators stay put
  for (int k0 = 0; k0 < n; k0 += 4) {
    const int i = k0 + kq;
    const bool in = i < n;
    direct_step<NT>(G, B, pad, m, in ? i : n - 1, live && in, x0, x1, x2, c4, accS, accD);
  }
  // v_mfma_f64_16x16x4 operands and result: lane l supplies A[l%16][l/16] and
  // B[l/16][l%16] and holds D[4r + l/16][l%16] in slot r, so this lane has
  // pairs 4r + kq of output column 16q + m
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = 4 * r + kq;
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

void launch_direct_mfma(LevelArgs A, const int* lpairs, const int* ltot, const int* ppanel, const int* tidx,
                        const double* tx, const double* lgeo, double* wout, int64_t nnz, hipStream_t s, int lo,
                        int hi) {
  if (lo > hi) return;
  const int grid = (int)((nnz + 15) / 16) + 4;
  switch ((A.h + 15) / 16) {
    case 1: hipLaunchKernelGGL(direct_mfma_kernel<1>, dim3(grid), dim3(64), 0, s, A, lpairs, ltot, ppanel, tidx, tx, lgeo, wout, lo, hi); break;
    case 2: hipLaunchKernelGGL(direct_mfma_kernel<2>, dim3(grid), dim3(64), 0, s, A, lpairs, ltot, ppanel, tidx, tx, lgeo, wout, lo, hi); break;
    case 3: hipLaunchKernelGGL(direct_mfma_kernel<3>, dim3(grid), dim3(64), 0, s, A, lpairs, ltot, ppanel, tidx, tx, lgeo, wout, lo, hi); break;
    case 4: hipLaunchKernelGGL(direct_mfma_kernel<4>, dim3(grid), dim3(64), 0, s, A, lpairs, ltot, ppanel, tidx, tx, lgeo, wout, lo, hi); break;
    case 5: hipLaunchKernelGGL(direct_mfma_kernel<5>, dim3(grid), dim3(64), 0, s, A, lpairs, ltot, ppanel, tidx, tx, lgeo, wout, lo, hi); break;
    case 6: hipLaunchKernelGGL(direct_mfma_kernel<6>, dim3(grid), dim3(64), 0, s, A, lpairs, ltot, ppanel, tidx, tx, lgeo, wout, lo, hi); break;
    case 7: hipLaunchKernelGGL(direct_mfma_kernel<7>, dim3(grid), dim3(64), 0, s, A, lpairs, ltot, ppanel, tidx, tx, lgeo, wout, lo, hi); break;
    case 8: hipLaunchKernelGGL(direct_mfma_kernel<8>, dim3(grid), dim3(64), 0, s, A, lpairs, ltot, ppanel, tidx, tx, lgeo, wout, lo, hi); break;
    case 9: hipLaunchKernelGGL(direct_mfma_kernel<9>, dim3(grid), dim3(64), 0, s, A, lpairs, ltot, ppanel, tidx, tx, lgeo, wout, lo, hi); break;
    default: throw std::runtime_error("direct level kernel: unsupported patch order");
  }
}

// ---------------------------------------------------------------------------
// The reference's adaptive quadtree (PR_ADAPTIVE: adaptive_pass, walk, cut_gI,
// patch_refine_mod.f90:617-771) on triangle levels.  walk descends from the
// level-1 cell: a cell whose ball the target is outside of (strict >) is a
// leaf; otherwise its four children are tested, and a level-3 cell the target
// is inside of sends the pair to one_level at level 4.  With three tree levels
// a cut is fixed by which level-2 cells are opened, so the 17 possible cuts
// are buckets; the reference's gI (cut_gI :721-738) is the leaf cells'
// interpolation rows, which here stay in the levels' padded table.

// per cut: the leaf cells' first nodes in the concatenated level arrays, in the
// reference's walk order (level-2 leaves by cell, then the level-3 leaves)
std::vector<int> adaptive_segments(const LevelArgs& A) {
  std::vector<int> seg((size_t)kAdaptBuckets * (kAdaptSegs + 1), 0);
  const int n1 = A.lev_n[0];
  seg[0] = 1;
  seg[1] = A.lev_off[0];
  for (int m = 0; m < 16; ++m) {
    int* g = seg.data() + (size_t)(1 + m) * (kAdaptSegs + 1);
    int c = 0;
    for (int q = 0; q < 4; ++q)
      if (!((m >> q) & 1)) g[1 + c++] = A.lev_off[1] + q * n1;
    for (int q = 0; q < 4; ++q)
      if ((m >> q) & 1)
        for (int b = 0; b < 4; ++b) g[1 + c++] = A.lev_off[2] + (4 * q + b) * n1;
    g[0] = c;
  }
  return seg;
}

// acnt: [0, 17) pairs per cut, [17, 34) first slot of each cut, [34, 51) cursors
__global__ void __launch_bounds__(256) abucket_count_kernel(const int* __restrict__ ltot,
                                                            const int* __restrict__ lpairs,
                                                            const int8_t* __restrict__ abucket,
                                                            int* __restrict__ acnt) {
  __shared__ int lh[kAdaptBuckets];
  if (threadIdx.x < kAdaptBuckets) lh[threadIdx.x] = 0;
  __syncthreads();
  const int n = ltot[1], base = ltot[0];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(lh + abucket[lpairs[base + i]], 1);
  __syncthreads();
  if (threadIdx.x < kAdaptBuckets && lh[threadIdx.x]) atomicAdd(acnt + threadIdx.x, lh[threadIdx.x]);
}

__global__ void abucket_offsets_kernel(int* __restrict__ acnt) {
  if (threadIdx.x != 0) return;
  int s = 0;
  for (int b = 0; b < kAdaptBuckets; ++b) {
    acnt[kAdaptBuckets + b] = s;
    acnt[2 * kAdaptBuckets + b] = s;
    s += acnt[b];
  }
}

// one global range per non-empty cut per workgroup, ranks inside it from LDS;
// the order inside a cut does not matter (every pair's weights depend on that
// pair alone)
__global__ void __launch_bounds__(256) abucket_scatter_kernel(const int* __restrict__ ltot,
                                                              const int* __restrict__ lpairs,
                                                              const int8_t* __restrict__ abucket,
                                                              int* __restrict__ acnt, int* __restrict__ apairs) {
  __shared__ int lh[kAdaptBuckets], lb[kAdaptBuckets];
  if (threadIdx.x < kAdaptBuckets) lh[threadIdx.x] = 0;
  __syncthreads();
  const int n = ltot[1], base = ltot[0];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int pair = -1, b = 0, r = 0;
  if (i < n) {
    pair = lpairs[base + i];
    b = abucket[pair];
    r = atomicAdd(lh + b, 1);
  }
  __syncthreads();
  if (threadIdx.x < kAdaptBuckets)
    lb[threadIdx.x] = lh[threadIdx.x] ? atomicAdd(acnt + 2 * kAdaptBuckets + threadIdx.x, lh[threadIdx.x]) : 0;
  __syncthreads();
  if (pair >= 0) apairs[lb[b] + r] = pair;
}

// one wave per 16 pairs of one cut: direct_mfma_kernel's step over the cut's
// leaf cells (n1 nodes each) instead of one level's nodes
template <int NT>
__global__ void __launch_bounds__(64) direct_adaptive_kernel(LevelArgs A, const int* __restrict__ acnt,
                                                             const int* __restrict__ apairs,
                                                             const int* __restrict__ ppanel,
                                                             const int* __restrict__ tidx,
                                                             const double* __restrict__ tx,
                                                             const double* __restrict__ lgeo,
                                                             double* __restrict__ wout) {
  int b = blockIdx.x, bk = -1;
  for (int q = 0; q < kAdaptBuckets; ++q) {
    const int ntile = (acnt[q] + 15) / 16;
    if (b < ntile) { bk = q; break; }
    b -= ntile;
  }
  if (bk < 0) return;
  const int h = A.h, n1 = A.n1, pad = A.pad;
  const int first = acnt[kAdaptBuckets + bk] + 16 * b;
  const int cnt = min(16, acnt[bk] - 16 * b);
  const int lane = threadIdx.x, m = lane & 15, kq = lane >> 4;
  const bool live = m < cnt;
  const int pair = live ? apairs[first + m] : 0;
  double x0 = 0.0, x1 = 0.0, x2 = 0.0;
  const double* G = lgeo;
  if (live) {
    const size_t t = (size_t)tidx[pair];
    x0 = tx[3 * t]; x1 = tx[3 * t + 1]; x2 = tx[3 * t + 2];
    G = lgeo + 8 * (size_t)ppanel[pair] * A.ntot;
  }
  const int* sg = A.aseg + bk * (kAdaptSegs + 1);
  const int nseg = sg[0];
  const double c4 = 1.0 / (16.0 * atan(1.0));
  qd4 accS[NT], accD[NT];
#pragma unroll
  for (int q = 0; q < NT; ++q) {
    accS[q] = qd4{0.0, 0.0, 0.0, 0.0};
    accD[q] = qd4{0.0, 0.0, 0.0, 0.0};
  }
  for (int c = 0; c < nseg; ++c) {
    const int o = sg[1 + c];
    for (int k0 = 0; k0 < n1; k0 += 4) {
      const int i = k0 + kq;
      const bool in = i < n1;
      direct_step<NT>(G, A.imp, pad, m, o + (in ? i : n1 - 1), live && in, x0, x1, x2, c4, accS, accD);
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = 4 * r + kq;
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

void launch_direct_adaptive(LevelArgs A, const int* ltot, const int* lpairs, const int8_t* abucket, int* acnt,
                            int* apairs, const int* ppanel, const int* tidx, const double* tx, const double* lgeo,
                            double* wout, int64_t nnz, hipStream_t s) {
  if (nnz <= 0) return;
  if (!A.adaptive || !A.aseg) throw std::runtime_error("adaptive direct kernel without adaptive level tables");
  const dim3 gs((unsigned)((nnz + 255) / 256)), blk(256);
  hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, s, acnt, 3 * kAdaptBuckets);
  hipLaunchKernelGGL(abucket_count_kernel, gs, blk, 0, s, ltot, lpairs, abucket, acnt);
  hipLaunchKernelGGL(abucket_offsets_kernel, dim3(1), dim3(64), 0, s, acnt);
  hipLaunchKernelGGL(abucket_scatter_kernel, gs, blk, 0, s, ltot, lpairs, abucket, acnt, apairs);
  const int grid = (int)((nnz + 15) / 16) + kAdaptBuckets;
#define QOL_ADAPT(NT_)                                                                                       \
  case NT_:                                                                                                  \
    hipLaunchKernelGGL(direct_adaptive_kernel<NT_>, dim3(grid), dim3(64), 0, s, A, (const int*)acnt,         \
                       (const int*)apairs, ppanel, tidx, tx, lgeo, wout);                                    \
    break;
  switch ((A.h + 15) / 16) {
    QOL_ADAPT(1) QOL_ADAPT(2) QOL_ADAPT(3) QOL_ADAPT(4) QOL_ADAPT(5) QOL_ADAPT(6) QOL_ADAPT(7) QOL_ADAPT(8)
    QOL_ADAPT(9)
    default: throw std::runtime_error("adaptive direct kernel: unsupported patch order");
  }
#undef QOL_ADAPT
}

}  // namespace qol
