// Host engine: device buffers, per-order tables, kernel launches and the C-ABI
// of the batched close-evaluation path.
//
// One engine holds one triangulated surface of patch order p (geometry resident
// in HBM), a target set and its near list, and the per-pair close-evaluation
// weights.  Every stage runs on the engine's HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "qol_engine.h"
#include "qol_host.h"
#include "qol_kernels.h"
#include "qotential_hip.h"

namespace qol {

void set_status(int64_t s);
void set_error(const std::string& s);


void add_timeinfo(const double* ms, double* timeinfo) {
  timeinfo[0] += 1e-3 * ms[0];
  timeinfo[1] += 1e-3 * ms[3];
  timeinfo[2] += 1e-3 * (ms[1] + ms[2]);
  timeinfo[3] += 1e-3 * ms[6];
  timeinfo[4] += 1e-3 * (ms[5] + ms[8]);
  timeinfo[5] += 1e-3 * (ms[7] + ms[9]);
  timeinfo[7] += 1e-3 * ms[4];
}

Arena*& thread_arena() {
  static thread_local Arena* a = nullptr;
  return a;
}

DevBuf::~DevBuf() {
  if (ptr && owned) (void)hipFree(ptr);
}
void DevBuf::alloc(size_t b) {
  if (b <= bytes && ptr) return;
  if (ptr && owned) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
  owned = false;
  if (b == 0) return;
  if (Arena* a = thread_arena()) {
    const size_t off = (a->used + 255) & ~(size_t)255;
    if (off > a->cap || b > a->cap - off)
      throw std::runtime_error("device scratch exhausted: " + std::to_string(off + b) + " bytes needed, " +
                               std::to_string(a->cap) + " available");
    ptr = a->base + off;
    bytes = b;
    a->used = off + b;
    return;
  }
  HIPCHK(hipMalloc(&ptr, b));
  bytes = b;
  owned = true;
}

void StreamSet::create_one_stream() {
  HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  s1 = s2 = s3 = stream;
  for (auto& st : tev)
    for (auto& e : st) HIPCHK(hipEventCreate(&e));
  for (hipEvent_t* e : {&ev_start, &ev_lists, &ev_direct, &ev_setup, &ev_system, &ev_solve, &ev_edge, &ev_proj})
    HIPCHK(hipEventCreateWithFlags(e, hipEventDisableTiming));
}

namespace {

template <class T>
void upload(DevBuf& d, const T* h, size_t n, hipStream_t s) {
  d.alloc(n * sizeof(T));
  if (n) HIPCHK(hipMemcpyAsync(d.ptr, h, n * sizeof(T), hipMemcpyHostToDevice, s));
}

int blocks(size_t n, int b) { return (int)((n + b - 1) / b); }

void set_lds(const void* f, size_t bytes) {
  if (bytes > 64 * 1024) HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  if (bytes > 160 * 1024) throw std::runtime_error("kernel needs more LDS than a CU has");
}

}  // namespace

// ---------------------------------------------------------------------------
// the four levels' offsets and sizes in the concatenated node and sub-patch
// arrays, and the concatenated suboff / subtyp tables
static void level_layout(const OrderTables& T, LevelArgs& LA, std::vector<int>& suboff, std::vector<int>& subtyp) {
  LA.h = (int)T.h; LA.ntot = (int)T.ntot; LA.nsubtot = (int)T.nsubtot;
  int off = 0, sb = 0;
  for (int l = 0; l < 4; ++l) {
    const LevelTable& L = T.lev[l];
    LA.lev_n[l] = (int)L.n;
    LA.lev_off[l] = off;
    LA.lev_nsub[l] = (int)L.nsub;
    LA.sub_base[l] = sb;
    for (i64 q = 0; q <= L.nsub; ++q) suboff.push_back((int)L.suboff[q]);
    for (i64 q = 0; q < L.nsub; ++q) subtyp.push_back((int)L.subtyp[q]);
    off += (int)L.n;
    sb += (int)L.nsub;
  }
}

DeviceTables::DeviceTables(i64 p, bool square, hipStream_t s) {
  const OrderTables& T = order_tables(p, square);
  h = (int)T.h;
  ntot = (int)T.ntot;
  nsubtot = (int)T.nsubtot;
  std::vector<double> interp((size_t)T.h * T.ntot), wt(T.ntot);
  std::vector<int> suboff, subtyp;
  level_layout(T, LA, suboff, subtyp);
  for (int l = 0; l < 4; ++l) {
    const LevelTable& L = T.lev[l];
    const int off = LA.lev_off[l];
    std::memcpy(interp.data() + (size_t)T.h * off, L.interpmat.data(), sizeof(double) * L.n * T.h);
    std::memcpy(wt.data() + off, L.wt.data(), sizeof(double) * L.n);
  }
  // row-major copy padded to whole 16-column MFMA tiles
  const int pad = 16 * ((h + 15) / 16);
  std::vector<double> imp((size_t)pad * T.ntot, 0.0);
  for (int l = 0; l < 4; ++l) {
    const int n = LA.lev_n[l], o = LA.lev_off[l];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < h; ++j) imp[(size_t)(o + i) * pad + j] = interp[(size_t)h * o + i + (size_t)n * j];
  }
  upload(d_imp, imp.data(), imp.size(), s);
  LA.imp = (const double*)d_imp.ptr;
  const std::vector<float> impf(imp.begin(), imp.end());
  upload(d_impf, impf.data(), impf.size(), s);
  LA.impf = (const float*)d_impf.ptr;
  LA.pad = pad;
  upload(d_interp, interp.data(), interp.size(), s);
  upload(d_wt, wt.data(), wt.size(), s);
  upload(d_suboff, suboff.data(), suboff.size(), s);
  upload(d_subtyp, subtyp.data(), subtyp.size(), s);
  LA.interp = (const double*)d_interp.ptr;
  LA.wt = (const double*)d_wt.ptr;
  LA.suboff = (const int*)d_suboff.ptr;
  LA.subtyp = (const int*)d_subtyp.ptr;
  LA.dist2 = 1.4 * 1.4;
  const std::vector<int> chunks = level_chunks(LA, suboff.data());
  upload(d_chunks, chunks.data(), chunks.size(), s);
  LA.chunks = (const int*)d_chunks.ptr;
  LA.nchunk = (int)chunks.size() / 5;
  LA.n1 = LA.lev_n[0];
  if (!square) {
    const std::vector<int> seg = adaptive_segments(LA);
    upload(d_aseg, seg.data(), seg.size(), s);
    LA.aseg = (const int*)d_aseg.ptr;
  }
  upload(d_tgl, T.tgl.data(), T.tgl.size(), s);
  upload(d_wgl, T.wgl.data(), T.wgl.size(), s);
  upload(d_Leg, T.Legmat.data(), T.Legmat.size(), s);
  upload(d_umatr, T.umatr.data(), T.umatr.size(), s);
  upload(d_uvn, T.uvn.data(), T.uvn.size(), s);
  RT.p = (int)T.p; RT.h = h; RT.nq = (int)T.nq; RT.ncoef = (int)T.ncoef;
  RT.tgl = (const double*)d_tgl.ptr;
  RT.wgl = (const double*)d_wgl.ptr;
  RT.Legmat = (const double*)d_Leg.ptr;
  // [k * nq + j]: the coefficient of s^j in the Legendre polynomial P_k,
  // (-1)^m C(k, m) C(2k - 2m, k) / 2^k for j = k - 2m: integers over a power of
  // two, exact in double for the orders here (k <= 13)
  {
    const int nq = (int)T.nq;
    std::vector<double> l2m((size_t)nq * nq, 0.0);
    auto binom = [](int n, int r) {
      double b = 1.0;
      for (int i = 1; i <= r; ++i) b = b * (double)(n - r + i) / (double)i;
      return b;
    };
    for (int kq = 0; kq < nq; ++kq)
      for (int m = 0; 2 * m <= kq; ++m)
        l2m[(size_t)kq * nq + kq - 2 * m] =
            ((m & 1) ? -1.0 : 1.0) * binom(kq, m) * binom(2 * kq - 2 * m, kq) * std::ldexp(1.0, -kq);
    upload(d_l2m, l2m.data(), l2m.size(), s);
    RT.leg2mon = (const double*)d_l2m.ptr;
  }
  RT.umatr = (const double*)d_umatr.ptr;
  RT.uvn = (const double*)d_uvn.ptr;
  const std::vector<double> kt = koorn_eval_table((int)T.p);
  upload(d_ktab, kt.data(), kt.size(), s);
  RT.ktab = (const double*)d_ktab.ptr;
  // the geometry's interpolant at the corners (0,0), (1,0), (0,1), as the level
  // rows are formed: Koornwinder values times umatr
  std::vector<double> corner(3 * (size_t)T.h), pol(T.h);
  const double cuv[3][2] = {{0.0, 0.0}, {1.0, 0.0}, {0.0, 1.0}};
  for (int c = 0; c < 3; ++c) {
    koornwinder(cuv[c][0], cuv[c][1], T.p - 1, pol.data(), nullptr, nullptr);
    for (i64 j = 0; j < T.h; ++j) {
      double a = 0.0;
      for (i64 q = 0; q < T.h; ++q) a += pol[q] * T.umatr[q + T.h * j];
      corner[(size_t)c * T.h + j] = a;
    }
  }
  upload(d_corner, corner.data(), corner.size(), s);
  HIPCHK(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
Engine::Engine(i64 p_, i64 npan_, const double* sx, const double* snx, const double* sw, const double* rts,
               const double* rps, i64 ichart_, const double* sxbd, const double* rv, bool device_ptrs,
               const StreamSet* borrow, bool square_levels)
    : p((int)p_), npan((int)npan_), ichart((int)ichart_), square(square_levels) {
  if (p < 2 || p > QOL_MAXP) throw std::runtime_error("patch order out of range for the device kernels");
  if (npan <= 0) throw std::runtime_error("npan must be positive");
  if (borrow) {
    borrowed = true;
    stream = borrow->stream; s1 = borrow->s1; s2 = borrow->s2; s3 = borrow->s3;
    for (int i = 0; i < 10; ++i)
      for (int j = 0; j < 2; ++j) tev[i][j] = borrow->tev[i][j];
    ev_start = borrow->ev_start; ev_lists = borrow->ev_lists; ev_direct = borrow->ev_direct;
    ev_setup = borrow->ev_setup; ev_system = borrow->ev_system; ev_solve = borrow->ev_solve; ev_edge = borrow->ev_edge;
    ev_proj = borrow->ev_proj;
  } else {
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // QOL_ONE_STREAM=1 runs every stage on the main stream (kernel times without
    // overlap, for profiling); by default the stages overlap on three streams
    const char* one = std::getenv("QOL_ONE_STREAM");
    if (one && one[0] == '1') {
      s1 = s2 = s3 = stream;
    } else {
      // three streams: the per-pair edge chain (s3 role) runs on the main
      // stream, which has nothing else to do between the pair lists and the
      // moments.  A fourth stream would share a hardware queue with one of the
      // others (HIP spreads streams over the process's few queues in creation
      // order), and the edge chain then waited behind the solve.
      HIPCHK(hipStreamCreateWithFlags(&s1, hipStreamNonBlocking));
      HIPCHK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
      s3 = stream;
    }
    for (auto& st : tev)
      for (auto& e : st) HIPCHK(hipEventCreate(&e));
    for (hipEvent_t* e : {&ev_start, &ev_lists, &ev_direct, &ev_setup, &ev_system, &ev_solve, &ev_edge, &ev_proj})
      HIPCHK(hipEventCreateWithFlags(e, hipEventDisableTiming));
  }
  tab = std::make_unique<DeviceTables>(p, square, stream);
  h = tab->h;
  nq = p + 2;
  ncoef = (p + 1) * (p + 2) / 2;
  nsrc = (size_t)npan * h;
  set_surface(sx, snx, sw, rts, rps, sxbd, rv, device_ptrs);
}

void Engine::set_surface(const double* sx, const double* snx, const double* sw, const double* rts, const double* rps,
                         const double* sxbd, const double* rv, bool device_ptrs) {
  const hipMemcpyKind kind = device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  auto put = [&](DevBuf& d, const double* a, size_t n) {
    d.alloc(n * sizeof(double));
    HIPCHK(hipMemcpyAsync(d.ptr, a, n * sizeof(double), kind, stream));
  };
  put(d_sx, sx, 3 * nsrc);
  put(d_snx, snx, 3 * nsrc);
  put(d_sw, sw, nsrc);
  put(d_rts, rts, 3 * nsrc);
  put(d_rps, rps, 3 * nsrc);
  if (ichart == 1) {
    if (!sxbd || !rv) throw std::runtime_error("ichart=1 needs the chart boundary and vertices");
    put(d_sxbd, sxbd, (size_t)npan * 9 * nq);
    put(d_rv, rv, (size_t)npan * 9);
  }
  HIPCHK(hipStreamSynchronize(stream));
  // everything derived from the old surface is stale
  levels_built = false;
  evaluated = false;
  have_omega = false;
  density_set = false;
  val_current = false;
  nnz = -1;
  tlists_nnz = -1;
}

Engine::~Engine() {
  if (borrowed) return;
  for (auto& st : tev)
    for (auto& e : st) (void)hipEventDestroy(e);
  for (hipEvent_t e : {ev_start, ev_lists, ev_direct, ev_setup, ev_system, ev_solve, ev_edge, ev_proj})
    (void)hipEventDestroy(e);
  if (s1 && s1 != stream) (void)hipStreamDestroy(s1);
  if (s2 && s2 != stream) (void)hipStreamDestroy(s2);
  if (s3 && s3 != stream) (void)hipStreamDestroy(s3);
  if (stream) (void)hipStreamDestroy(stream);
}

size_t Engine::one_panel_bytes(i64 p, i64 m, i64 ichart, bool square) {
  const OrderTables& T = order_tables(p, square);
  LevelArgs LA{};
  std::vector<int> suboff, subtyp;
  level_layout(T, LA, suboff, subtyp);
  const size_t h = (size_t)T.h, ntot = (size_t)T.ntot, nsubtot = (size_t)T.nsubtot, nq = (size_t)p + 2;
  const size_t nb = 3 * nq, ncoef = (size_t)(p + 1) * (p + 2) / 2, pad = 16 * ((h + 15) / 16);
  const size_t mm = (size_t)(m > 0 ? m : 1), D = sizeof(double), I = sizeof(int);
  size_t tot = 0;
  auto add = [&](size_t b) { tot += (b + 255) & ~(size_t)255; };
  // per-order tables (DeviceTables)
  for (size_t b : {pad * ntot * D, pad * ntot * sizeof(float), h * ntot * D, ntot * D, suboff.size() * I, subtyp.size() * I,
                   level_chunks(LA, suboff.data()).size() * I, T.tgl.size() * D, T.wgl.size() * D,
                   T.Legmat.size() * D, T.Legmat.size() * D, T.umatr.size() * D, T.uvn.size() * D,
                   koorn_eval_table((int)p).size() * D,
                   3 * h * D})
    add(b);
  if (!square) add((size_t)kAdaptBuckets * (kAdaptSegs + 1) * I);
  // the panel (constructor), targets and pairs
  for (size_t b : {3 * h * D, 3 * h * D, h * D, 3 * h * D, 3 * h * D}) add(b);
  if (ichart == 1) { add(9 * nq * D); add(9 * D); }
  add(3 * mm * D);                  // tx
  add(2 * I); add(mm * I);          // offs, tidx
  // alloc_pairs
  for (size_t b : {mm, mm * I, 2 * h * D * mm, mm * D, (size_t)I, 5 * (size_t)I, 5 * (size_t)I, 8 * (size_t)I, mm * I,
                   mm, mm * I, 3 * (size_t)kAdaptBuckets * I})
    add(b);
  // alloc_rrq
  for (size_t b : {16 * D, 9 * nq * D, 6 * nb * D, 3 * h * D, 2 * nb * (size_t)(p * (p + 1) / 2) * D, 3 * h * h * D, h * h * D,
                   h * h * D, 10 * h * h * D, h * h * D})
    add(b);
  if (!rrq_solve_g1_in_lds((int)h)) add(4 * h * h * D);
  // eval: Omega, Ias, the per-pair RRQ data, the level geometry
  for (size_t b : {4 * h * mm * D, mm * D, 4 * mm * D, 3 * nq * mm * D, 3 * mm * D, 9 * mm * D,
                   rrq_edge_order_ints((int64_t)mm) * I, (size_t)16, 5 * h * mm * D, 8 * ntot * D, 4 * nsubtot * D})
    add(b);
  return tot;
}

void Engine::set_targets(i64 nt_, const double* tx, bool device_ptr) {
  nt = (int)nt_;
  d_tx.alloc(sizeof(double) * 3 * (size_t)(nt > 0 ? nt : 1));
  if (nt > 0)
    HIPCHK(hipMemcpyAsync(d_tx.ptr, tx, sizeof(double) * 3 * nt, device_ptr ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                          stream));
  nnz = -1;
  evaluated = false;
  tlists_nnz = -1;  // the per-target pair lists belong to the old targets
}

void Engine::alloc_pairs() {
  tlists_nnz = -1;  // a new pair list: the per-target lists are rebuilt on the next apply
  val_current = false;
  d_label.alloc((size_t)(nnz > 0 ? nnz : 1));
  d_ppanel.alloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
  d_wout.alloc(sizeof(double) * 2 * (size_t)h * (nnz > 0 ? nnz : 1));
  d_val.alloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
  d_nrrq.alloc(sizeof(int) * npan);
  d_lcount.alloc(sizeof(int) * 5 * npan);
  d_lstart.alloc(sizeof(int) * 5 * npan);
  d_ltot.alloc(sizeof(int) * 8);
  d_lpairs.alloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
  d_abucket.alloc((size_t)(nnz > 0 ? nnz : 1));
  d_apairs.alloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
  d_acnt.alloc(sizeof(int) * 3 * kAdaptBuckets);
}

// The reference's PR_ADAPTIVE and PR_SQUARE (patch_refine_mod.f90:51-55, 109-117):
// square levels (the default), triangle levels, or triangle levels served by the
// adaptive walk.  New level tables are uploaded when the level form changes.
void Engine::set_level_mode(bool adaptive_, bool square_) {
  if (adaptive_ && square_) throw std::runtime_error("PR_SQUARE and PR_ADAPTIVE cannot be combined");
  if (square_ != square) {
    HIPCHK(hipStreamSynchronize(stream));
    tab = std::make_unique<DeviceTables>(p, square_, stream);
    square = square_;
    levels_built = false;
    evaluated = false;
    val_current = false;
  }
  adaptive = adaptive_;
  tab->LA.adaptive = adaptive ? 1 : 0;
}

// classification and per-label pair lists (main stream)
void Engine::classify_lists() {
  SelectorArgs S = sel;
  S.corner = (const double*)tab->d_corner.ptr;
  launch_classify(tab->LA, S, npan, (const int*)d_offs.ptr, (const int*)d_tidx.ptr, (const double*)d_tx.ptr,
                  (const double*)d_sx.ptr, (const double*)d_snx.ptr, (const double*)d_sw.ptr,
                  (const double*)d_balls.ptr, (int8_t*)d_label.ptr, (int*)d_ppanel.ptr, (int*)d_nrrq.ptr,
                  (int*)d_lcount.ptr, (int8_t*)d_abucket.ptr, stream);
  // pairs the selector serves with the panel's own nodes: naive weights, and a
  // zero correction when eval makes the corrections itself
  if (S.mode == 2 || (S.mode == 1 && S.forced == 0))
    hipLaunchKernelGGL(naive_fill_kernel, dim3(blocks(nnz, 256)), dim3(256), 0, stream, (int)nnz, h,
                       (const int8_t*)d_label.ptr, (const int*)d_tidx.ptr, (const int*)d_ppanel.ptr,
                       (const double*)d_tx.ptr, (const double*)d_sx.ptr, (const double*)d_snx.ptr,
                       (const double*)d_sw.ptr, (double*)d_wout.ptr, density_set ? (double*)d_val.ptr : nullptr);
  hipLaunchKernelGGL(level_scan_kernel, dim3(5), dim3(1024), 0, stream, (const int*)d_lcount.ptr, npan,
                     (int*)d_lstart.ptr, (int*)d_ltot.ptr);
  hipLaunchKernelGGL(compact_kernel, dim3(npan), dim3(256), 0, stream, (const int*)d_offs.ptr,
                     (const int8_t*)d_label.ptr, (const int*)d_lstart.ptr, (const int*)d_ltot.ptr, npan,
                     (int*)d_lpairs.ptr);
}

// direct S/D of every level-served pair: levels 1..split_levels on the FP32
// matrix cores (split precision), the rest on the FP64 matrix cores
void Engine::direct(hipStream_t s) {
  if (adaptive) {
    // level 1 whole, the tree cuts (list 1), level 4 for what the tree cannot serve
    if (split_levels > 0) throw std::runtime_error("split precision is not built for adaptive levels");
    launch_direct_mfma(tab->LA, (const int*)d_lpairs.ptr, (const int*)d_ltot.ptr, (const int*)d_ppanel.ptr,
                       (const int*)d_tidx.ptr, (const double*)d_tx.ptr, (const double*)d_lgeo.ptr, (double*)d_wout.ptr,
                       nnz, s, 0, 0);
    launch_direct_adaptive(tab->LA, (const int*)d_ltot.ptr, (const int*)d_lpairs.ptr, (const int8_t*)d_abucket.ptr,
                           (int*)d_acnt.ptr, (int*)d_apairs.ptr, (const int*)d_ppanel.ptr, (const int*)d_tidx.ptr,
                           (const double*)d_tx.ptr, (const double*)d_lgeo.ptr, (double*)d_wout.ptr, nnz, s);
    launch_direct_mfma(tab->LA, (const int*)d_lpairs.ptr, (const int*)d_ltot.ptr, (const int*)d_ppanel.ptr,
                       (const int*)d_tidx.ptr, (const double*)d_tx.ptr, (const double*)d_lgeo.ptr, (double*)d_wout.ptr,
                       nnz, s, 3, 3);
    return;
  }
  const int ns = std::max(0, std::min(4, split_levels));
  launch_direct_split(tab->LA, (const int*)d_lpairs.ptr, (const int*)d_ltot.ptr, (const int*)d_ppanel.ptr,
                      (const int*)d_tidx.ptr, (const double*)d_tx.ptr, (const double*)d_lgeo.ptr, (double*)d_wout.ptr,
                      nnz, s, 0, ns - 1);
  launch_direct_mfma(tab->LA, (const int*)d_lpairs.ptr, (const int*)d_ltot.ptr, (const int*)d_ppanel.ptr,
                     (const int*)d_tidx.ptr, (const double*)d_tx.ptr, (const double*)d_lgeo.ptr, (double*)d_wout.ptr,
                     nnz, s, ns, 3);
}

i64 Engine::nearlist() {
  if (nt < 0) throw std::runtime_error("targets not set");
  d_nball.alloc(sizeof(double) * 4 * npan);
  d_cnt.alloc(sizeof(int) * npan);
  d_offs.alloc(sizeof(int) * (npan + 1));
  HIPCHK(hipMemsetAsync(d_cnt.ptr, 0, sizeof(int) * npan, stream));
  hipLaunchKernelGGL(nearlist_prep_kernel, dim3(blocks(npan, 256)), dim3(256), 0, stream, npan, h,
                     (const double*)d_sx.ptr, (const double*)d_sw.ptr, (const double*)tab->d_corner.ptr, near_diam,
                     (double*)d_nball.ptr);
  const int pb = 16;
  const int gy = nt > 0 ? std::min(64, blocks(nt, 256)) : 1;
  if (nt > 0)
    hipLaunchKernelGGL(nearlist_count_kernel, dim3(blocks(npan, pb), gy), dim3(256), 0, stream, npan, nt,
                       (const double*)d_tx.ptr, (const double*)d_nball.ptr, (int*)d_cnt.ptr);
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, stream, (const int*)d_cnt.ptr, npan, (int*)d_offs.ptr);
  int tot = 0;
  HIPCHK(hipMemcpyAsync(&tot, (int*)d_offs.ptr + npan, sizeof(int), hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  nnz = tot;
  d_tidx.alloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
  if (nt > 0 && nnz > 0)
    hipLaunchKernelGGL(nearlist_fill_kernel, dim3(blocks(npan, pb)), dim3(256), 0, stream, npan, nt,
                       (const double*)d_tx.ptr, (const double*)d_nball.ptr, (const int*)d_offs.ptr, (int*)d_tidx.ptr);
  HIPCHK(hipGetLastError());
  alloc_pairs();
  evaluated = false;
  return nnz;
}

// The same near list through a uniform grid of targets (cells at least as
// large as the largest panel ball, at most 2^24 cells): bit-identical CSR.
i64 Engine::nearlist_grid() {
  if (nt < 0) throw std::runtime_error("targets not set");
  d_nball.alloc(sizeof(double) * 4 * npan);
  d_cnt.alloc(sizeof(int) * npan);
  d_offs.alloc(sizeof(int) * (npan + 1));
  hipLaunchKernelGGL(nearlist_prep_kernel, dim3(blocks(npan, 256)), dim3(256), 0, stream, npan, h,
                     (const double*)d_sx.ptr, (const double*)d_sw.ptr, (const double*)tab->d_corner.ptr, near_diam,
                     (double*)d_nball.ptr);
  std::vector<double> bl(4 * (size_t)npan);
  HIPCHK(hipMemcpyAsync(bl.data(), d_nball.ptr, sizeof(double) * bl.size(), hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, rmax = 0.0;
  for (int k = 0; k < npan; ++k) {
    const double r = std::sqrt(bl[4 * k + 3]);
    rmax = std::max(rmax, r);
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], bl[4 * k + d] - r);
      hi[d] = std::max(hi[d], bl[4 * k + d] + r);
    }
  }
  double c = rmax > 0.0 ? rmax : 1.0;
  GridArgs G{};
  for (;;) {
    double cells = 1.0;
    for (int d = 0; d < 3; ++d) {
      G.n[d] = std::max(1, (int)std::ceil((hi[d] - lo[d]) / c));
      cells *= G.n[d];
    }
    if (cells <= (double)(1 << 24)) break;
    c *= 1.26;
  }
  for (int d = 0; d < 3; ++d) G.lo[d] = lo[d];
  G.inv_c = 1.0 / c;
  const int ncell = G.n[0] * G.n[1] * G.n[2];
  DevBuf d_cell_of, d_ccount, d_cstart, d_ccur, d_ctarg, d_pcur;
  d_cell_of.alloc(sizeof(int) * (size_t)std::max(nt, 1));
  d_ccount.alloc(sizeof(int) * ncell);
  d_cstart.alloc(sizeof(int) * (ncell + 1));
  d_ccur.alloc(sizeof(int) * ncell);
  d_ctarg.alloc(sizeof(int) * (size_t)std::max(nt, 1));
  d_pcur.alloc(sizeof(int) * npan);
  HIPCHK(hipMemsetAsync(d_ccount.ptr, 0, sizeof(int) * ncell, stream));
  HIPCHK(hipMemsetAsync(d_ccur.ptr, 0, sizeof(int) * ncell, stream));
  HIPCHK(hipMemsetAsync(d_pcur.ptr, 0, sizeof(int) * npan, stream));
  if (nt > 0) {
    hipLaunchKernelGGL(grid_bin_kernel, dim3(blocks(nt, 256)), dim3(256), 0, stream, nt, (const double*)d_tx.ptr, G,
                       (int*)d_cell_of.ptr, (int*)d_ccount.ptr);
  }
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, stream, (const int*)d_ccount.ptr, ncell, (int*)d_cstart.ptr);
  if (nt > 0)
    hipLaunchKernelGGL(grid_scatter_kernel, dim3(blocks(nt, 256)), dim3(256), 0, stream, nt, (const int*)d_cell_of.ptr,
                       (const int*)d_cstart.ptr, (int*)d_ccur.ptr, (int*)d_ctarg.ptr);
  hipLaunchKernelGGL(grid_panel_kernel, dim3(npan), dim3(64), 0, stream, npan, (const double*)d_nball.ptr, G,
                     (const int*)d_cstart.ptr, (const int*)d_ctarg.ptr, (const double*)d_tx.ptr, 0, (int*)d_cnt.ptr,
                     (const int*)nullptr, (int*)nullptr, (int*)nullptr);
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, stream, (const int*)d_cnt.ptr, npan, (int*)d_offs.ptr);
  int tot = 0;
  HIPCHK(hipMemcpyAsync(&tot, (int*)d_offs.ptr + npan, sizeof(int), hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  nnz = tot;
  d_tidx.alloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
  if (nnz > 0) {
    hipLaunchKernelGGL(grid_panel_kernel, dim3(npan), dim3(64), 0, stream, npan, (const double*)d_nball.ptr, G,
                       (const int*)d_cstart.ptr, (const int*)d_ctarg.ptr, (const double*)d_tx.ptr, 1, (int*)nullptr,
                       (const int*)d_offs.ptr, (int*)d_pcur.ptr, (int*)d_tidx.ptr);
    hipLaunchKernelGGL(segment_sort_kernel, dim3(npan), dim3(256), 0, stream, npan, (const int*)d_offs.ptr,
                       (int*)d_tidx.ptr);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(stream));
  alloc_pairs();
  evaluated = false;
  return nnz;
}

void Engine::set_pairs(const i64* offs, const i64* idx1) {
  std::vector<int> o(npan + 1);
  for (int k = 0; k <= npan; ++k) o[k] = (int)offs[k];
  if (o[0] != 0) throw std::runtime_error("pair offsets must start at 0");
  for (int k = 0; k < npan; ++k)
    if (o[k + 1] < o[k]) throw std::runtime_error("pair offsets must be non-decreasing");
  nnz = o[npan];
  std::vector<int> t((size_t)(nnz > 0 ? nnz : 1));
  for (i64 j = 0; j < nnz; ++j) {
    if (idx1[j] < 1 || idx1[j] > nt) throw std::runtime_error("pair target index out of range");
    t[j] = (int)(idx1[j] - 1);
  }
  upload(d_offs, o.data(), o.size(), stream);
  upload(d_tidx, t.data(), t.size(), stream);
  HIPCHK(hipStreamSynchronize(stream));
  alloc_pairs();
  evaluated = false;
}

void Engine::get_pairs(i64* offs, i64* idx1) {
  std::vector<int> o(npan + 1), t((size_t)(nnz > 0 ? nnz : 1));
  HIPCHK(hipMemcpyAsync(o.data(), d_offs.ptr, sizeof(int) * (npan + 1), hipMemcpyDeviceToHost, stream));
  if (nnz > 0) HIPCHK(hipMemcpyAsync(t.data(), d_tidx.ptr, sizeof(int) * nnz, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  for (int k = 0; k <= npan; ++k) offs[k] = o[k];
  for (i64 j = 0; j < nnz; ++j) idx1[j] = (i64)t[j] + 1;
}

void Engine::alloc_levels() {
  d_lgeo.alloc(sizeof(double) * 8 * (size_t)tab->ntot * npan);
  d_balls.alloc(sizeof(double) * 4 * (size_t)tab->nsubtot * npan);
}

void Engine::build_levels() {
  alloc_levels();
  launch_level_geom(tab->LA, npan, (const double*)d_sx.ptr, (const double*)d_rts.ptr, (const double*)d_rps.ptr,
                    (double*)d_lgeo.ptr, (double*)d_balls.ptr, stream);
  HIPCHK(hipGetLastError());
  levels_built = true;
}

void Engine::alloc_rrq() {
  if (rrq_alloc) return;
  const size_t np_ = npan, nb = 3 * (size_t)nq;
  d_frame.alloc(sizeof(double) * 16 * np_);
  d_edge.alloc(sizeof(double) * 9 * nq * np_);
  d_bX.alloc(sizeof(double) * 6 * nb * np_);
  d_cx.alloc(sizeof(double) * 3 * (size_t)h * np_);
  d_basis.alloc(sizeof(double) * 2 * nb * (size_t)(p * (p + 1) / 2) * np_);  // boundary harmonics of degree < p
  d_Fq.alloc(sizeof(double) * 3 * (size_t)h * h * np_);
  d_Msl.alloc(sizeof(double) * (size_t)h * h * np_);
  d_Fn.alloc(sizeof(double) * (size_t)h * h * np_);
  d_G.alloc(sizeof(double) * 10 * (size_t)h * h * np_);
  d_Binv.alloc(sizeof(double) * (size_t)h * h * np_);
  if (!rrq_solve_g1_in_lds(h)) d_scr.alloc(sizeof(double) * 4 * (size_t)h * h * np_);
  RP.frame = (double*)d_frame.ptr; RP.edge = (double*)d_edge.ptr; RP.bX = (double*)d_bX.ptr;
  RP.cx = (double*)d_cx.ptr; RP.basis = (double*)d_basis.ptr; RP.Fq = (double*)d_Fq.ptr;
  RP.Msl = (double*)d_Msl.ptr; RP.Fn = (double*)d_Fn.ptr;
  RP.G = (double*)d_G.ptr; RP.Binv = (double*)d_Binv.ptr; RP.scr = rrq_solve_g1_in_lds(h) ? nullptr : (double*)d_scr.ptr;
  rrq_alloc = true;
}

void Engine::eval_levels_only() {
  if (nnz < 0) throw std::runtime_error("pairs not set (call nearlist or set_pairs)");
  build_levels();
  if (nnz > 0) {
    classify_lists();
    direct(stream);
  }
  HIPCHK(hipGetLastError());
  evaluated = true;
  have_omega = false;
  projected = false;
}

// Stages and streams (all at the default priority).  The per-panel RRQ set-up
// (frame kernel, then system kernel) and solve need only the surface, so they
// run for every panel on s2 from the start, beside the level geometry and
// classification on the main stream.  After the pair lists: the per-pair
// projection, root/SSQ and solid angle in the s3 role (the main stream itself
// for an engine that makes its own streams; after the frame kernel), and the
// direct levels on s1 (matrix cores) once the projection is done; the moments
// join the edge chain and the system kernel, the weights the solve, and the
// main stream joins s1 last.
// tev[i] brackets stage i: levels, classify+lists, direct, rrq setup, rrq solve,
// projection, root + SSQ (solid angle from tev[8][0]), moments, -, weights.
void Engine::eval(int exterior, double side_tol, bool want_omega) {
  if (nnz < 0) throw std::runtime_error("pairs not set (call nearlist or set_pairs)");
  if (p > 12) throw std::runtime_error("the RRQ kernels support patch orders up to 12");
  // stage timing events are left out of a captured graph (timing = false)
  auto rec = [&](int i, int j, hipStream_t s) {
    if (timing) HIPCHK(hipEventRecord(tev[i][j], s));
  };
  // the stream roles (s3 is the main stream unless borrowed), or the main stream for all of them
  hipStream_t s1 = one_stream_run ? stream : this->s1, s2 = one_stream_run ? stream : this->s2,
              s3 = one_stream_run ? stream : this->s3;
  if (nnz <= 0) {
    rec(0, 0, stream);
    build_levels();
    rec(0, 1, stream);
    HIPCHK(hipMemsetAsync(d_nrrq.ptr, 0, sizeof(int) * npan, stream));
    HIPCHK(hipMemsetAsync(d_lcount.ptr, 0, sizeof(int) * 5 * npan, stream));
    for (int i = 1; i < 10; ++i) { rec(i, 0, stream); rec(i, 1, stream); }
    HIPCHK(hipGetLastError());
    val_current = false;
    evaluated = true;
    have_omega = want_omega;
    projected = false;
    return;
  }
  alloc_rrq();
  if (want_omega) {
    d_omega.alloc(sizeof(double) * 4 * (size_t)h * nnz);
    d_ias.alloc(sizeof(double) * (size_t)nnz);
    HIPCHK(hipMemsetAsync(d_omega.ptr, 0, sizeof(double) * 4 * (size_t)h * nnz, stream));
    HIPCHK(hipMemsetAsync(d_ias.ptr, 0, sizeof(double) * (size_t)nnz, stream));
  }
  d_pq.alloc(sizeof(double) * 4 * (size_t)nnz);
  d_w1.alloc(sizeof(double) * 3 * (size_t)nq * nnz);
  d_oe.alloc(sizeof(double) * 3 * (size_t)nnz);
  d_root.alloc(sizeof(double) * 9 * (size_t)nnz);
  d_eorder.alloc(sizeof(int) * rrq_edge_order_ints(nnz));
  d_pfail.alloc(16);
  d_vt.alloc(sizeof(double) * 5 * (size_t)h * nnz);
  alloc_levels();  // every buffer is in place before the first launch
  HIPCHK(hipEventRecord(ev_start, stream));
  // s2: RRQ set-up and solve for every panel
  HIPCHK(hipStreamWaitEvent(s2, ev_start, 0));
  rec(3, 0, s2);
  const size_t lds1 = rrq_frame_smem(h, nq);
  set_lds((const void*)rrq_panel_frame_kernel, lds1);
  // one wave per panel: the kernel's loops run over the h nodes or the boundary
  // nodes, so wider workgroups leave lanes idle while fewer panels share a CU
  hipLaunchKernelGGL(rrq_panel_frame_kernel, dim3(npan), dim3(64), lds1, s2, tab->RT, RP, npan, (const int*)nullptr,
                     (const double*)d_sx.ptr, (const double*)d_snx.ptr, (const double*)d_sw.ptr, ichart,
                     (const double*)d_sxbd.ptr, (const double*)d_rv.ptr);
  HIPCHK(hipEventRecord(ev_setup, s2));
  const size_t lds2 = rrq_system_smem(h, nq, ncoef);
  set_lds((const void*)rrq_panel_system_kernel, lds2);
  hipLaunchKernelGGL(rrq_panel_system_kernel, dim3(npan), dim3(256), lds2, s2, tab->RT, RP, npan, (const int*)nullptr,
                     (const double*)d_sx.ptr, (const double*)d_snx.ptr);
  rec(3, 1, s2);
  HIPCHK(hipEventRecord(ev_system, s2));
  rec(4, 0, s2);
  launch_rrq_panel_solve(tab->RT, RP, npan, nullptr, s2, solve_variant);
  rec(4, 1, s2);
  HIPCHK(hipEventRecord(ev_solve, s2));
  // main: levels, classification, pair lists
  rec(0, 0, stream);
  build_levels();
  rec(0, 1, stream);
  rec(1, 0, stream);
  classify_lists();
  rec(1, 1, stream);
  HIPCHK(hipEventRecord(ev_lists, stream));
  // s3: projection, root + SSQ, solid angle (needs the lists and the set-up)
  HIPCHK(hipStreamWaitEvent(s3, ev_lists, 0));
  HIPCHK(hipStreamWaitEvent(s3, ev_setup, 0));
  rec(5, 0, s3);
  launch_rrq_project(tab->RT, RP, (const int*)d_lpairs.ptr, (const int*)d_ltot.ptr, (const int*)d_ppanel.ptr,
                     (const int*)d_tidx.ptr, (const double*)d_tx.ptr, (const double*)d_sx.ptr, exterior, side_tol,
                     (double*)d_pq.ptr, (int*)d_eorder.ptr, (int*)d_pfail.ptr, nnz, s3);
  rec(5, 1, s3);
  HIPCHK(hipEventRecord(ev_proj, s3));
  rec(6, 0, s3);
  launch_rrq_edge(tab->RT, RP, (const int*)d_lpairs.ptr, (const int*)d_ltot.ptr, (const int*)d_ppanel.ptr,
                  (const int*)d_tidx.ptr, (const double*)d_tx.ptr, (const double*)d_pq.ptr, (double*)d_w1.ptr,
                  (double*)d_root.ptr, (double*)d_oe.ptr, (int*)d_eorder.ptr, nnz, s3, timing ? tev[8][0] : nullptr);
  rec(6, 1, s3);
  HIPCHK(hipEventRecord(ev_edge, s3));
  // s1: direct levels, after the edge chain (projection, roots, solid angle): that
  // chain heads the path the moments wait for, the direct kernel's 121k waves
  // would crowd it out, and the direct kernel then runs beside the moments
  HIPCHK(hipStreamWaitEvent(s1, ev_edge, 0));
  rec(2, 0, s1);
  direct(s1);
  rec(2, 1, s1);
  // with the density resident, the level-served pairs' corrections follow at once
  const bool fuse = density_set;
  if (fuse) launch_correction(0, 3, s1);
  HIPCHK(hipEventRecord(ev_direct, s1));
  // main: moments once the edges and the boundary harmonics are in, weights once the
  // solve is in too
  HIPCHK(hipStreamWaitEvent(stream, ev_edge, 0));
  HIPCHK(hipStreamWaitEvent(stream, ev_system, 0));
  rec(7, 0, stream);
  const size_t lds3 = rrq_target_smem(p, h, nq, ncoef);
  set_lds((const void*)rrq_target_kernel, lds3);
  hipLaunchKernelGGL(rrq_target_kernel, dim3(npan), dim3(256), lds3, stream, tab->RT, RP, (const int*)d_lpairs.ptr,
                     (const int*)d_lstart.ptr, (const int*)d_ltot.ptr, npan, (const int*)d_tidx.ptr,
                     (const double*)d_tx.ptr, (const int*)d_nrrq.ptr, (const double*)d_w1.ptr,
                     (const double*)d_oe.ptr, (double*)d_vt.ptr, want_omega ? (double*)d_omega.ptr : nullptr,
                     want_omega ? (double*)d_ias.ptr : nullptr);
  rec(7, 1, stream);
  HIPCHK(hipStreamWaitEvent(stream, ev_solve, 0));
  rec(9, 0, stream);
  launch_rrq_weights(tab->RT, RP, (const int*)d_lpairs.ptr, (const int*)d_lstart.ptr, (const int*)d_ltot.ptr, npan,
                     (const int*)d_nrrq.ptr, (const double*)d_vt.ptr, (double*)d_wout.ptr, stream);
  rec(9, 1, stream);
  if (fuse) launch_correction(4, 4, stream);
  // join: everything after eval on the main stream sees every weight
  HIPCHK(hipStreamWaitEvent(stream, ev_direct, 0));
  val_current = fuse;
  HIPCHK(hipGetLastError());
  evaluated = true;
  have_omega = want_omega;
  projected = true;
}

i64 Engine::projection_unconverged(double* max_step) {
  if (!evaluated) throw std::runtime_error("not evaluated");
  if (max_step) *max_step = 0.0;
  if (!projected) return 0;
  HIPCHK(hipStreamSynchronize(stream));
  int stat[4] = {0, 0, 0, 0};
  HIPCHK(hipMemcpy(stat, d_pfail.ptr, sizeof(stat), hipMemcpyDeviceToHost));
  if (max_step) std::memcpy(max_step, stat + 2, sizeof(double));
  return stat[0];
}

void Engine::stage_ms(double* ms) {
  HIPCHK(hipStreamSynchronize(stream));
  auto el = [&](hipEvent_t a, hipEvent_t b) {
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, a, b));
    return (double)t;
  };
  for (int i = 0; i < 8; ++i) ms[i] = el(tev[i][0], tev[i][1]);
  ms[9] = el(tev[9][0], tev[9][1]);
  // the edge stage is two kernels: root + SSQ, then the solid angle (tev[8][0] between)
  if (nnz > 0) {
    ms[6] = el(tev[6][0], tev[8][0]);
    ms[8] = el(tev[8][0], tev[6][1]);
  } else {
    ms[8] = 0.0;
  }
}

void Engine::pair_features(float* feat) {
  if (nnz < 0) throw std::runtime_error("pairs not set (call nearlist or set_pairs)");
  build_levels();
  if (nnz > 0) {
    DevBuf d_feat;
    d_feat.alloc(sizeof(float) * kSelIn * (size_t)nnz);
    hipLaunchKernelGGL(pair_features_kernel, dim3(npan), dim3(256), sizeof(double) * 3 * (size_t)h, stream, tab->LA,
                       (const double*)tab->d_corner.ptr, (const int*)d_offs.ptr, (const int*)d_tidx.ptr,
                       (const double*)d_tx.ptr, (const double*)d_sx.ptr, (const double*)d_snx.ptr,
                       (const double*)d_sw.ptr, (const double*)d_balls.ptr, (float*)d_feat.ptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(feat, d_feat.ptr, sizeof(float) * kSelIn * (size_t)nnz, hipMemcpyDeviceToHost, stream));
  }
  HIPCHK(hipStreamSynchronize(stream));
}

void Engine::get_weights(double* As, double* Ad, i64* label, double* Omega, double* Ias) {
  if (!evaluated) throw std::runtime_error("weights not evaluated");
  HIPCHK(hipStreamSynchronize(stream));
  if ((As || Ad) && nnz > 0) {
    std::vector<double> w(2 * (size_t)h * nnz);
    HIPCHK(hipMemcpy(w.data(), d_wout.ptr, sizeof(double) * 2 * h * (size_t)nnz, hipMemcpyDeviceToHost));
    // column-major (nnz, h)
    for (i64 j = 0; j < nnz; ++j)
      for (int c = 0; c < h; ++c) {
        if (As) As[j + (size_t)nnz * c] = w[(size_t)j * 2 * h + c];
        if (Ad) Ad[j + (size_t)nnz * c] = w[(size_t)j * 2 * h + h + c];
      }
  }
  if (label && nnz > 0) {
    std::vector<int8_t> l(nnz);
    HIPCHK(hipMemcpy(l.data(), d_label.ptr, nnz, hipMemcpyDeviceToHost));
    for (i64 j = 0; j < nnz; ++j) label[j] = l[j];
  }
  if (Omega && nnz > 0) {
    if (!have_omega) throw std::runtime_error("Omega was not requested at eval");
    HIPCHK(hipMemcpy(Omega, d_omega.ptr, sizeof(double) * 4 * (size_t)h * nnz, hipMemcpyDeviceToHost));
  }
  if (Ias && nnz > 0) {
    if (!have_omega) throw std::runtime_error("Ias was not requested at eval");
    HIPCHK(hipMemcpy(Ias, d_ias.ptr, sizeof(double) * (size_t)nnz, hipMemcpyDeviceToHost));
  }
}

void Engine::apply(const double* ub, const double* ubn, double* u, double* val, bool device_ptrs) {
  if (!evaluated) throw std::runtime_error("weights not evaluated");
  const hipMemcpyKind kin = device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  const hipMemcpyKind kout = device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  d_ub.alloc(sizeof(double) * nsrc);
  d_ubn.alloc(sizeof(double) * nsrc);
  d_u.alloc(sizeof(double) * (size_t)(nt > 0 ? nt : 1));
  HIPCHK(hipMemcpyAsync(d_ub.ptr, ub, sizeof(double) * nsrc, kin, stream));
  HIPCHK(hipMemcpyAsync(d_ubn.ptr, ubn, sizeof(double) * nsrc, kin, stream));
  if (nt > 0) HIPCHK(hipMemcpyAsync(d_u.ptr, u, sizeof(double) * nt, kin, stream));
  density_set = true;
  val_current = false;  // the density may have changed since the last eval
  apply_device();
  if (nt > 0) HIPCHK(hipMemcpyAsync(u, d_u.ptr, sizeof(double) * nt, kout, stream));
  if (val && nnz > 0) HIPCHK(hipMemcpyAsync(val, d_val.ptr, sizeof(double) * nnz, kout, stream));
  HIPCHK(hipStreamSynchronize(stream));
}

void Engine::build_target_lists() {
  if (tlists_nnz == nnz) return;
  d_tcount.alloc(sizeof(int) * (size_t)(nt > 0 ? nt : 1));
  d_toffs.alloc(sizeof(int) * (size_t)(nt + 1));
  d_tcur.alloc(sizeof(int) * (size_t)(nt > 0 ? nt : 1));
  d_tpairs.alloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
  HIPCHK(hipMemsetAsync(d_tcount.ptr, 0, sizeof(int) * nt, stream));
  HIPCHK(hipMemsetAsync(d_tcur.ptr, 0, sizeof(int) * nt, stream));
  hipLaunchKernelGGL(target_count_kernel, dim3(blocks(nnz, 256)), dim3(256), 0, stream, (int)nnz,
                     (const int*)d_tidx.ptr, (int*)d_tcount.ptr);
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, stream, (const int*)d_tcount.ptr, nt, (int*)d_toffs.ptr);
  hipLaunchKernelGGL(target_fill_kernel, dim3(blocks(nnz, 256)), dim3(256), 0, stream, (int)nnz,
                     (const int*)d_tidx.ptr, (const int*)d_toffs.ptr, (int*)d_tcur.ptr, (int*)d_tpairs.ptr);
  HIPCHK(hipGetLastError());
  tlists_nnz = nnz;
}

// corrections of the pairs in per-label lists first..last (after their weights)
void Engine::launch_correction(int first, int last, hipStream_t s) {
  hipLaunchKernelGGL(naive_apply_list_kernel, dim3(blocks(4 * (size_t)nnz, 256)), dim3(256), 0, s,
                     (const int*)d_lpairs.ptr, (const int*)d_ltot.ptr, first, last, h, (const int*)d_tidx.ptr,
                     (const int*)d_ppanel.ptr, (const double*)d_tx.ptr, (const double*)d_sx.ptr,
                     (const double*)d_snx.ptr, (const double*)d_sw.ptr, (const double*)d_ub.ptr,
                     (const double*)d_ubn.ptr, (const double*)d_wout.ptr, (double*)d_val.ptr);
}

// near-correction values (unless the last eval already made them from the
// resident density) and the fixed-order per-target sum into u
void Engine::apply_device() {
  if (nnz <= 0) return;
  if (!val_current)
    hipLaunchKernelGGL(naive_apply_kernel, dim3(blocks(4 * (size_t)nnz, 256)), dim3(256), 0, stream, (int)nnz, h,
                       (const int*)d_tidx.ptr, (const int*)d_ppanel.ptr, (const double*)d_tx.ptr,
                       (const double*)d_sx.ptr, (const double*)d_snx.ptr, (const double*)d_sw.ptr,
                       (const double*)d_ub.ptr, (const double*)d_ubn.ptr, (const double*)d_wout.ptr,
                       (double*)d_val.ptr);
  build_target_lists();
  hipLaunchKernelGGL(target_sum_kernel, dim3(blocks(nt, 256)), dim3(256), 0, stream, nt, (const int*)d_toffs.ptr,
                     (int*)d_tpairs.ptr, (const double*)d_val.ptr, (double*)d_u.ptr);
  HIPCHK(hipGetLastError());
}

void Engine::get_values(double* val, double* u) {
  HIPCHK(hipStreamSynchronize(stream));
  if (val && nnz > 0) HIPCHK(hipMemcpy(val, d_val.ptr, sizeof(double) * nnz, hipMemcpyDeviceToHost));
  if (u && nt > 0) HIPCHK(hipMemcpy(u, d_u.ptr, sizeof(double) * nt, hipMemcpyDeviceToHost));
}

void Engine::far(const double* ub, const double* ubn, double* u) {
  d_ub.alloc(sizeof(double) * nsrc);
  d_ubn.alloc(sizeof(double) * nsrc);
  d_u.alloc(sizeof(double) * (size_t)(nt > 0 ? nt : 1));
  HIPCHK(hipMemcpyAsync(d_ub.ptr, ub, sizeof(double) * nsrc, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(d_ubn.ptr, ubn, sizeof(double) * nsrc, hipMemcpyHostToDevice, stream));
  density_set = true;
  val_current = false;
  if (nt > 0)
    hipLaunchKernelGGL(far_direct_kernel, dim3(blocks(nt, 256)), dim3(256), 0, stream, nt, (const double*)d_tx.ptr,
                       (int)nsrc, (const double*)d_sx.ptr, (const double*)d_snx.ptr, (const double*)d_sw.ptr,
                       (const double*)d_ub.ptr, (const double*)d_ubn.ptr, (double*)d_u.ptr);
  HIPCHK(hipGetLastError());
  if (nt > 0) HIPCHK(hipMemcpyAsync(u, d_u.ptr, sizeof(double) * nt, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
}

void Engine::level_data(double* lgeo, double* balls) {
  build_levels();
  HIPCHK(hipStreamSynchronize(stream));
  if (lgeo) {
    // device layout: 8 doubles per (panel, node): x, y, z, nx, ny, nz, w, pad;
    // the caller's: [panel][7][ntot]
    const size_t nt_ = (size_t)tab->ntot, n = nt_ * npan;
    std::vector<double> g(8 * n);
    HIPCHK(hipMemcpy(g.data(), d_lgeo.ptr, sizeof(double) * 8 * n, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < (size_t)npan; ++k)
      for (size_t i = 0; i < nt_; ++i)
        for (int c = 0; c < 7; ++c) lgeo[(k * 7 + c) * nt_ + i] = g[8 * (k * nt_ + i) + c];
  }
  if (balls) HIPCHK(hipMemcpy(balls, d_balls.ptr, sizeof(double) * 4 * (size_t)tab->nsubtot * npan, hipMemcpyDeviceToHost));
}

void Engine::triplets(const i64* trow, i64* si, i64* sj, double* vs, double* vd) {
  if (!evaluated) throw std::runtime_error("weights not evaluated");
  std::vector<int> tr((size_t)(nt > 0 ? nt : 1), 0);
  for (int t = 0; t < nt; ++t) tr[t] = (int)trow[t];
  DevBuf d_tr, d_si, d_sj, d_vs, d_vd;
  upload(d_tr, tr.data(), tr.size(), stream);
  const size_t n = (size_t)nnz * h;
  d_si.alloc(sizeof(int64_t) * (n ? n : 1)); d_sj.alloc(sizeof(int64_t) * (n ? n : 1));
  d_vs.alloc(sizeof(double) * (n ? n : 1)); d_vd.alloc(sizeof(double) * (n ? n : 1));
  if (nnz > 0)
    hipLaunchKernelGGL(triplet_kernel, dim3(blocks(nnz, 256)), dim3(256), 0, stream, (int)nnz, h,
                       (const int*)d_tidx.ptr, (const int*)d_ppanel.ptr, (const int*)d_tr.ptr, (const int*)d_offs.ptr,
                       (const double*)d_tx.ptr, (const double*)d_sx.ptr, (const double*)d_snx.ptr,
                       (const double*)d_sw.ptr, (const double*)d_wout.ptr, (int64_t*)d_si.ptr, (int64_t*)d_sj.ptr,
                       (double*)d_vs.ptr, (double*)d_vd.ptr);
  HIPCHK(hipGetLastError());
  if (n) {
    HIPCHK(hipMemcpyAsync(si, d_si.ptr, sizeof(int64_t) * n, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(sj, d_sj.ptr, sizeof(int64_t) * n, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(vs, d_vs.ptr, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(vd, d_vd.ptr, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
  }
  HIPCHK(hipStreamSynchronize(stream));
}

}  // namespace qol
