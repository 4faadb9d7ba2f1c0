// Near list by the reference's ball rule, CSR by panel with targets ascending.
//
// Reference: stellarator_grf_core_mod.f90:652-695 (count, then fill) and
// test/stellarator_bench.f90:88-94: ball centre = mean of the panel's nodes,
// radius 1.75*sqrt(sum sw), target kept when |x - c|^2 < r^2.  With a diameter
// factor f > 0 the radius is f h instead, h the largest distance between the
// corners of the panel's interpolant (the "radius 2h" near list, f = 2; h as the
// target generator defines the patch diameter).
// Brute force over (panel, target) like the reference, but kPB panels per
// workgroup share every target load, and the fill keeps ascending target order
// with wavefront ballots so the CSR is bit-identical to the serial loop.
#include "qol_kernels.h"

namespace qol {

constexpr int kPB = 16;  // panels per workgroup

// ball[k] = (c0, c1, c2, r^2), summed in the reference's order; corner: the 3 x h
// rows interpolating a panel at its corners (used when diam > 0)
__global__ void nearlist_prep_kernel(int npan, int h, const double* __restrict__ sx, const double* __restrict__ sw,
                                     const double* __restrict__ corner, double diam, double* __restrict__ ball) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npan) return;
  double q0 = 0.0, q1 = 0.0, q2 = 0.0, s = 0.0;
  for (int i = 0; i < h; ++i) {
    const size_t g = (size_t)k * h + i;
    q0 += sx[3 * g];
    q1 += sx[3 * g + 1];
    q2 += sx[3 * g + 2];
  }
  for (int i = 0; i < h; ++i) s += sw[(size_t)k * h + i];
  double qr = 1.75 * sqrt(s);
  if (diam > 0.0) {
    double v[3][3];
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < 3; ++d) {
        double a = 0.0;
        for (int i = 0; i < h; ++i) a += corner[c * h + i] * sx[3 * ((size_t)k * h + i) + d];
        v[c][d] = a;
      }
    double hd = 0.0;
    for (int a = 0; a < 3; ++a)
      for (int b = a + 1; b < 3; ++b) {
        const double e0 = v[a][0] - v[b][0], e1 = v[a][1] - v[b][1], e2 = v[a][2] - v[b][2];
        hd = fmax(hd, sqrt(e0 * e0 + e1 * e1 + e2 * e2));
      }
    qr = diam * hd;
  }
  ball[4 * k] = q0 / (double)h;
  ball[4 * k + 1] = q1 / (double)h;
  ball[4 * k + 2] = q2 / (double)h;
  ball[4 * k + 3] = qr * qr;
}

__device__ __forceinline__ bool in_ball(const double* b, double x0, double x1, double x2) {
  const double d0 = x0 - b[0], d1 = x1 - b[1], d2 = x2 - b[2];
  return d0 * d0 + d1 * d1 + d2 * d2 < b[3];
}

__global__ void __launch_bounds__(256) nearlist_count_kernel(int npan, int nt, const double* __restrict__ tx,
                                                             const double* __restrict__ ball, int* __restrict__ count) {
  __shared__ double sb[4 * kPB];
  __shared__ int sc[kPB];
  const int k0 = blockIdx.x * kPB;
  if (threadIdx.x < 4 * kPB) sb[threadIdx.x] = (k0 + threadIdx.x / 4 < npan) ? ball[4 * k0 + threadIdx.x] : 0.0;
  if (threadIdx.x < kPB) sc[threadIdx.x] = 0;
  __syncthreads();
  int my[kPB];
  for (int q = 0; q < kPB; ++q) my[q] = 0;
  for (int t = blockIdx.y * blockDim.x + threadIdx.x; t < nt; t += gridDim.y * blockDim.x) {
    const double x0 = tx[3 * (size_t)t], x1 = tx[3 * (size_t)t + 1], x2 = tx[3 * (size_t)t + 2];
    for (int q = 0; q < kPB; ++q) my[q] += in_ball(sb + 4 * q, x0, x1, x2) ? 1 : 0;
  }
  for (int q = 0; q < kPB; ++q) {
    int v = my[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(&sc[q], v);
  }
  __syncthreads();
  if (threadIdx.x < kPB && k0 + threadIdx.x < npan && sc[threadIdx.x]) atomicAdd(&count[k0 + threadIdx.x], sc[threadIdx.x]);
}

// one workgroup per kPB panels, targets in ascending chunks of blockDim
__global__ void __launch_bounds__(256) nearlist_fill_kernel(int npan, int nt, const double* __restrict__ tx,
                                                            const double* __restrict__ ball,
                                                            const int* __restrict__ offs, int* __restrict__ idx) {
  __shared__ double sb[4 * kPB];
  __shared__ int base[kPB];
  __shared__ int wcnt[kPB][4];
  const int k0 = blockIdx.x * kPB;
  if (threadIdx.x < 4 * kPB) sb[threadIdx.x] = (k0 + threadIdx.x / 4 < npan) ? ball[4 * k0 + threadIdx.x] : 0.0;
  if (threadIdx.x < kPB) base[threadIdx.x] = (k0 + threadIdx.x < npan) ? offs[k0 + threadIdx.x] : 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int t0 = 0; t0 < nt; t0 += blockDim.x) {
    const int t = t0 + threadIdx.x;
    double x0 = 0.0, x1 = 0.0, x2 = 0.0;
    if (t < nt) { x0 = tx[3 * (size_t)t]; x1 = tx[3 * (size_t)t + 1]; x2 = tx[3 * (size_t)t + 2]; }
    uint64_t bal[kPB];
    for (int q = 0; q < kPB; ++q) {
      const bool in = (t < nt) && (k0 + q < npan) && in_ball(sb + 4 * q, x0, x1, x2);
      bal[q] = __ballot(in);
      if (lane == 0) wcnt[q][wv] = __popcll(bal[q]);
    }
    __syncthreads();
    for (int q = 0; q < kPB; ++q) {
      if ((bal[q] >> lane) & 1ull) {
        int pos = base[q];
        for (int w = 0; w < wv; ++w) pos += wcnt[q][w];
        pos += __popcll(bal[q] & ((1ull << lane) - 1ull));
        idx[pos] = t;
      }
    }
    __syncthreads();
    if (threadIdx.x < kPB) {
      int s = 0;
      for (int w = 0; w < 4; ++w) s += wcnt[threadIdx.x][w];
      base[threadIdx.x] += s;
    }
    __syncthreads();
  }
}

// exclusive scan of n ints into out[0..n] with a single workgroup of 1024 threads
__global__ void __launch_bounds__(1024) scan_kernel(const int* __restrict__ in, int n, int* __restrict__ out) {
  __shared__ int part[1024];
  const int per = (n + blockDim.x - 1) / blockDim.x;
  const int a = threadIdx.x * per, b = min(n, a + per);
  int s = 0;
  for (int i = a; i < b; ++i) s += in[i];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int i = 0; i < (int)blockDim.x; ++i) { const int v = part[i]; part[i] = acc; acc += v; }
    out[n] = acc;
  }
  __syncthreads();
  int acc = part[threadIdx.x];
  for (int i = a; i < b; ++i) { out[i] = acc; acc += in[i]; }
}


// ---------------------------------------------------------------------------
// The same near list through a uniform grid of cells no smaller than the
// largest panel ball (the spatial search the reference's MATLAB caller does
// with a k-d tree, test/torus_fmm3dbie_rrq_comparison.m:81-99): targets are
// binned by cell, each panel tests only the targets of the cells its ball's
// bounding box covers (at most 3 x 3 x 3), and each panel's list is sorted, so
// the CSR equals the brute-force one bit for bit (the same in_ball test).

__device__ __forceinline__ int grid_coord(double x, double lo, double inv_c, int n) {
  const double f = floor((x - lo) * inv_c);
  return f < 0.0 ? 0 : (f > (double)(n - 1) ? n - 1 : (int)f);
}

__global__ void grid_bin_kernel(int nt, const double* __restrict__ tx, GridArgs G, int* __restrict__ cell_of,
                                int* __restrict__ cell_count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int i = grid_coord(tx[3 * (size_t)t], G.lo[0], G.inv_c, G.n[0]);
  const int j = grid_coord(tx[3 * (size_t)t + 1], G.lo[1], G.inv_c, G.n[1]);
  const int k = grid_coord(tx[3 * (size_t)t + 2], G.lo[2], G.inv_c, G.n[2]);
  const int c = (k * G.n[1] + j) * G.n[0] + i;
  cell_of[t] = c;
  atomicAdd(&cell_count[c], 1);
}

__global__ void grid_scatter_kernel(int nt, const int* __restrict__ cell_of, const int* __restrict__ cell_start,
                                    int* __restrict__ cursor, int* __restrict__ cell_targets) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int c = cell_of[t];
  cell_targets[cell_start[c] + atomicAdd(&cursor[c], 1)] = t;
}

// one wave per panel; fill == 0 counts, fill == 1 writes the hits (unordered)
__global__ void __launch_bounds__(64) grid_panel_kernel(int npan, const double* __restrict__ ball, GridArgs G,
                                                        const int* __restrict__ cell_start,
                                                        const int* __restrict__ cell_targets,
                                                        const double* __restrict__ tx, int fill,
                                                        int* __restrict__ count, const int* __restrict__ offs,
                                                        int* __restrict__ cursor, int* __restrict__ idx) {
  const int k = blockIdx.x;
  if (k >= npan) return;
  const double* b = ball + 4 * (size_t)k;
  const double r = sqrt(b[3]);
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = grid_coord(b[d] - r, G.lo[d], G.inv_c, G.n[d]);
    hi[d] = grid_coord(b[d] + r, G.lo[d], G.inv_c, G.n[d]);
  }
  int cnt = 0;
  for (int kk = lo[2]; kk <= hi[2]; ++kk)
    for (int jj = lo[1]; jj <= hi[1]; ++jj)
      for (int ii = lo[0]; ii <= hi[0]; ++ii) {
        const int c = (kk * G.n[1] + jj) * G.n[0] + ii;
        for (int s = cell_start[c] + threadIdx.x; s < cell_start[c + 1]; s += 64) {
          const int t = cell_targets[s];
          if (in_ball(b, tx[3 * (size_t)t], tx[3 * (size_t)t + 1], tx[3 * (size_t)t + 2])) {
            if (fill) idx[offs[k] + atomicAdd(&cursor[k], 1)] = t;
            else ++cnt;
          }
        }
      }
  if (!fill) {
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if (threadIdx.x == 0) count[k] = cnt;
  }
}

// each panel's list in ascending target order: bitonic sort in LDS up to
// kSortMax entries, one thread's insertion sort beyond
constexpr int kSortMax = 4096;
__global__ void __launch_bounds__(256) segment_sort_kernel(int npan, const int* __restrict__ offs, int* __restrict__ idx) {
  __shared__ int s[kSortMax];
  const int k = blockIdx.x;
  const int a = offs[k], n = offs[k + 1] - a;
  if (n <= 1) return;
  if (n > kSortMax) {
    if (threadIdx.x == 0)
      for (int i = a + 1; i < a + n; ++i) {
        const int v = idx[i];
        int j = i - 1;
        while (j >= a && idx[j] > v) { idx[j + 1] = idx[j]; --j; }
        idx[j + 1] = v;
      }
    return;
  }
  int m = 1;
  while (m < n) m <<= 1;
  for (int i = threadIdx.x; i < m; i += blockDim.x) s[i] = i < n ? idx[a + i] : 0x7fffffff;
  __syncthreads();
  for (int size = 2; size <= m; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < m; i += blockDim.x) {
        const int j = i ^ stride;
        if (j > i) {
          const bool up = (i & size) == 0;
          const int x = s[i], y = s[j];
          if ((x > y) == up) { s[i] = y; s[j] = x; }
        }
      }
      __syncthreads();
    }
  for (int i = threadIdx.x; i < n; i += blockDim.x) idx[a + i] = s[i];
}

}  // namespace qol
