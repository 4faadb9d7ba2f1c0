// C-ABI of libqotential_hip.so: the batched engine and the reference drop-ins.
//
// Drop-in entry points keep the reference's MEX/Fortran symbols and argument
// lists (gfortran mangling, everything by reference):
//   qol_rrq_mex_                     src/lap3d_close_mex.f90:18-45
//   qol_simplex_precomp_mex_         src/lap3d_close_mex.f90:48-59
//   qol_getnearquad_lap_rrq_mex_     src/lap3d_close_mex.f90:62-81
//   qol_build_patch_levels_mex_      src/patch_refine_mex.f90:198-208
//   qol_lap3dsdlpmat_levels_mex_     src/patch_refine_mex.f90:211-223
//   qol_clear_patch_levels_mex_      src/patch_refine_mex.f90:226-233
//   qol_clear_patch_levels_all_mex_  src/patch_refine_mex.f90:236-242
// Each runs on the GPU through a short-lived engine; the batched engine API
// (qol_engine_*) keeps geometry, targets and weights resident across calls.
#include <hip/hip_runtime.h>

#include <atomic>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "qol_engine.h"
#include "qol_host.h"
#include "qotential_hip.h"

namespace qol {
void set_status(int64_t s);
void set_error(const std::string& s);
void set_unconverged(int64_t n, double max_step);
}  // namespace qol

using namespace qol;

namespace {

template <class F>
int guarded(F&& f) {
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

Engine* E(void* h) {
  if (!h) throw std::runtime_error("null engine handle");
  return static_cast<Engine*>(h);
}

// the reference's fixed relation between order, nquad and orderff on the RRQ path
void check_orders(int64_t order, int64_t nquad, int64_t orderff, double distff) {
  if (nquad != order + 2 || orderff != order + 2)
    throw std::runtime_error("this library evaluates nquad = orderff = order + 2 (the reference drivers' choice)");
  if (distff != 1.4) throw std::runtime_error("distff must be 1.4 (the reference's level radius factor)");
}

// One-panel engine per calling thread for the per-panel drop-ins, kept across
// calls (streams, per-order tables and buffers are made once per thread and
// order; each call uploads only the panel, its targets and pair list).  The
// cache is deliberately never destroyed: thread-exit destructors could run
// after the HIP runtime has shut down.
struct PanelEngineCache {
  std::unique_ptr<Engine> e;
  int64_t p = -1, ichart = -1;
};

// the reference's PR_ADAPTIVE / PR_SQUARE module variables (process-wide, as
// Fortran module variables are; patch_refine_mod.f90:51-55)
std::atomic<int> g_adaptive{0}, g_square{1};

Engine& panel_engine(int64_t p, int64_t ichart, const double* sx, const double* snx, const double* sw,
                     const double* rts, const double* rps, const double* sxbd, const double* rv) {
  static thread_local PanelEngineCache* cache = nullptr;
  if (!cache) cache = new PanelEngineCache;
  if (!cache->e || cache->p != p || cache->ichart != ichart) {
    cache->e.reset();
    cache->e.reset(new Engine(p, 1, sx, snx, sw, rts, rps, ichart, sxbd, rv, false));
    cache->p = p;
    cache->ichart = ichart;
  } else {
    cache->e->set_surface(sx, snx, sw, rts, rps, sxbd, rv, false);
  }
  bool ad = false, sq = true;
  level_mode(&ad, &sq);
  cache->e->set_level_mode(ad, sq);
  return *cache->e;
}

}  // namespace

namespace qol {
void level_mode(bool* adaptive, bool* square) {
  *adaptive = g_adaptive.load() != 0;
  *square = g_square.load() != 0;
}
}  // namespace qol

extern "C" {

// ---------------------------------------------------------------- engine API
int qol_engine_create(int64_t p, int64_t npan, const double* sx, const double* snx, const double* sw,
                      const double* rts, const double* rps, int64_t ichart, const double* sxbd,
                      const double* rv, int64_t device_ptrs, void** out) {
  return guarded([&] {
    *out = nullptr;
    *out = new Engine(p, npan, sx, snx, sw, rts, rps, ichart, sxbd, rv, device_ptrs != 0);
  });
}

int qol_engine_set_level_mode(void* h, int64_t adaptive, int64_t square) {
  return guarded([&] { E(h)->set_level_mode(adaptive != 0, square != 0); });
}

int qol_set_level_mode(int64_t adaptive, int64_t square) {
  return guarded([&] {
    g_adaptive.store(adaptive != 0 ? 1 : 0);
    g_square.store(square != 0 ? 1 : 0);
  });
}

int qol_get_level_mode(int64_t* adaptive, int64_t* square) {
  return guarded([&] {
    *adaptive = g_adaptive.load();
    *square = g_square.load();
  });
}

int qol_engine_destroy(void* h) {
  return guarded([&] { delete E(h); });
}

int qol_engine_set_targets(void* h, int64_t nt, const double* tx, int64_t device_ptr) {
  return guarded([&] { E(h)->set_targets(nt, tx, device_ptr != 0); });
}

int qol_engine_nearlist(void* h, int64_t* nnz) {
  return guarded([&] { *nnz = E(h)->nearlist(); });
}

int qol_engine_nearlist_grid(void* h, int64_t* nnz) {
  return guarded([&] { *nnz = E(h)->nearlist_grid(); });
}

int qol_engine_set_near_radius(void* h, double diam_factor) {
  return guarded([&] {
    if (!(diam_factor >= 0.0) || !std::isfinite(diam_factor))
      throw std::runtime_error("near radius: the diameter factor must be 0 (the reference ball) or positive");
    E(h)->near_diam = diam_factor;
  });
}

int qol_engine_set_pairs(void* h, const int64_t* offs, const int64_t* idx) {
  return guarded([&] { E(h)->set_pairs(offs, idx); });
}

int qol_engine_get_pairs(void* h, int64_t* offs, int64_t* idx) {
  return guarded([&] { E(h)->get_pairs(offs, idx); });
}

int qol_engine_eval(void* h, int64_t exterior, double side_tol, int64_t want_omega) {
  return guarded([&] { E(h)->eval((int)exterior, side_tol, want_omega != 0); });
}

int qol_engine_sync(void* h) {
  return guarded([&] {
    if (hipStreamSynchronize(E(h)->stream) != hipSuccess) throw std::runtime_error("stream synchronize failed");
  });
}

int qol_engine_stage_ms(void* h, double* ms) {
  return guarded([&] { E(h)->stage_ms(ms); });
}

int qol_engine_projection_unconverged(void* h, int64_t* count, double* max_step) {
  return guarded([&] { *count = E(h)->projection_unconverged(max_step); });
}

int qol_engine_get_weights(void* h, double* As, double* Ad, int64_t* label, double* Omega, double* Ias) {
  return guarded([&] { E(h)->get_weights(As, Ad, label, Omega, Ias); });
}

int qol_engine_apply(void* h, const double* ub, const double* ubn, double* u, double* val) {
  return guarded([&] { E(h)->apply(ub, ubn, u, val, false); });
}

int qol_engine_apply_resident(void* h) {
  return guarded([&] { E(h)->apply_device(); });
}

int qol_engine_get_values(void* h, double* val, double* u) {
  return guarded([&] { E(h)->get_values(val, u); });
}

int qol_engine_set_potential(void* h, const double* u) {
  return guarded([&] {
    Engine* e = E(h);
    if (e->nt <= 0) return;
    e->d_u.alloc(sizeof(double) * (size_t)e->nt);
    HIPCHK(hipMemcpyAsync(e->d_u.ptr, u, sizeof(double) * e->nt, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
  });
}

int qol_engine_far(void* h, const double* ub, const double* ubn, double* u) {
  return guarded([&] { E(h)->far(ub, ubn, u); });
}

int qol_engine_level_info(void* h, int64_t* info) {
  return guarded([&] {
    const LevelArgs& A = E(h)->tab->LA;
    info[0] = A.ntot; info[1] = A.nsubtot;
    for (int l = 0; l < 4; ++l) { info[2 + l] = A.lev_n[l]; info[6 + l] = A.lev_nsub[l]; }
    info[10] = E(h)->square ? 1 : 0;
    info[11] = E(h)->adaptive ? 1 : 0;
  });
}

int qol_engine_level_data(void* h, double* lgeo, double* balls) {
  return guarded([&] { E(h)->level_data(lgeo, balls); });
}

int qol_engine_triplets(void* h, const int64_t* trow, int64_t* si, int64_t* sj, double* vs, double* vd) {
  return guarded([&] { E(h)->triplets(trow, si, sj, vs, vd); });
}

int qol_engine_set_precision(void* h, int64_t split_levels) {
  return guarded([&] {
    if (split_levels < 0 || split_levels > 4) throw std::runtime_error("split_levels must be 0..4");
    E(h)->split_levels = (int)split_levels;
  });
}

int qol_engine_set_selector(void* h, void* selector, double threshold) {
  return guarded([&] {
    if (threshold < 0.0 || threshold > 1.0) throw std::runtime_error("threshold must be in [0, 1]");
    E(h)->sel = selector_args(selector, threshold);
  });
}

int qol_engine_force_label(void* h, int64_t label) {
  return guarded([&] {
    if (label < -1 || label > 5) throw std::runtime_error("forced label must be -1 (ball rule) or 0..5");
    SelectorArgs S;
    if (label >= 0) { S.mode = 1; S.forced = (int)label; }
    E(h)->sel = S;
  });
}

int qol_engine_refine_levels(void* h, int64_t shift) {
  return guarded([&] {
    if (shift < 0 || shift > 3) throw std::runtime_error("level shift must be 0..3");
    SelectorArgs S;
    if (shift > 0) { S.mode = 3; S.forced = (int)shift; }
    E(h)->sel = S;
  });
}

int qol_engine_pair_features(void* h, float* feat) {
  return guarded([&] { E(h)->pair_features(feat); });
}

int qol_engine_stream(void* h, void** stream) {
  return guarded([&] { *stream = (void*)E(h)->stream; });
}

int qol_geo_build_gpu(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode, const double* mnrz,
                      int64_t iterate_aspect, double* sx, double* snx, double* sw, double* rts, double* rps,
                      double* charts) {
  return guarded([&] {
    Surface s;
    s.nfp = nfp;
    for (int64_t k = 0; k < nmode; ++k) {
      s.m.push_back(mnrz[4 * k]);
      s.n.push_back(mnrz[4 * k + 1]);
      s.rc.push_back(mnrz[4 * k + 2]);
      s.zs.push_back(mnrz[4 * k + 3]);
    }
    GeometryPlan P;
    plan_geometry(s, mp, np, p, iterate_aspect != 0, P);
    build_geometry_gpu(s, P, sx, snx, sw, rts, rps);
    if (charts)
      for (size_t k = 0; k < P.charts.size(); ++k) {
        const Chart& c = P.charts[k];
        const double v[6] = {c.a1, c.b1, c.a2, c.b2, c.jj, c.ii};
        for (int j = 0; j < 6; ++j) charts[6 * k + j] = v[j];
      }
  });
}

int qol_geo_build_charts_gpu(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode, const double* mnrz,
                             int64_t nchart, const double* charts, double* sx, double* snx, double* sw, double* rts,
                             double* rps) {
  return guarded([&] {
    if (nchart < 0 || (nchart > 0 && !charts)) throw std::runtime_error("qol_geo_build_charts_gpu: bad chart list");
    Surface s;
    s.nfp = nfp;
    for (int64_t k = 0; k < nmode; ++k) {
      s.m.push_back(mnrz[4 * k]);
      s.n.push_back(mnrz[4 * k + 1]);
      s.rc.push_back(mnrz[4 * k + 2]);
      s.zs.push_back(mnrz[4 * k + 3]);
    }
    std::vector<Chart> L(nchart);
    for (int64_t k = 0; k < nchart; ++k)
      L[k] = Chart{charts[6 * k], charts[6 * k + 1], charts[6 * k + 2], charts[6 * k + 3], charts[6 * k + 4],
                   charts[6 * k + 5]};
    GeometryPlan P;
    plan_geometry(s, mp, np, p, L, P);
    build_geometry_gpu(s, P, sx, snx, sw, rts, rps);
  });
}

int64_t qol_geo_charts_gpu(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode, const double* mnrz,
                           double restol, int64_t cap, double* charts) {
  int64_t n = -2;
  const int st = guarded([&] {
    Surface s;
    s.nfp = nfp;
    for (int64_t k = 0; k < nmode; ++k) {
      s.m.push_back(mnrz[4 * k]);
      s.n.push_back(mnrz[4 * k + 1]);
      s.rc.push_back(mnrz[4 * k + 2]);
      s.zs.push_back(mnrz[4 * k + 3]);
    }
    bool over = false;
    const std::vector<Chart> L = build_charts_refined(s, mp, np, p, restol, cap, &over, true);
    if (over) { n = -1; return; }
    if (charts)
      for (size_t k = 0; k < L.size(); ++k) {
        const Chart& c = L[k];
        const double v[6] = {c.a1, c.b1, c.a2, c.b2, c.jj, c.ii};
        for (int j = 0; j < 6; ++j) charts[6 * k + j] = v[j];
      }
    n = (int64_t)L.size();
  });
  return st == 0 ? n : -2;
}

int qol_geo_curvature_tails_gpu(int64_t mp, int64_t np, int64_t p, int64_t nfp, int64_t nmode, const double* mnrz,
                                int64_t nchart, const double* charts, double* tails) {
  return guarded([&] {
    if (nchart < 0 || (nchart > 0 && (!charts || !tails)))
      throw std::runtime_error("qol_geo_curvature_tails_gpu: bad chart list");
    Surface s;
    s.nfp = nfp;
    for (int64_t k = 0; k < nmode; ++k) {
      s.m.push_back(mnrz[4 * k]);
      s.n.push_back(mnrz[4 * k + 1]);
      s.rc.push_back(mnrz[4 * k + 2]);
      s.zs.push_back(mnrz[4 * k + 3]);
    }
    std::vector<Chart> L(nchart);
    for (int64_t k = 0; k < nchart; ++k)
      L[k] = Chart{charts[6 * k], charts[6 * k + 1], charts[6 * k + 2], charts[6 * k + 3], charts[6 * k + 4],
                   charts[6 * k + 5]};
    const std::vector<double> t = curvature_tails_gpu(s, L, mp, np, p);
    for (int64_t k = 0; k < nchart; ++k) tails[k] = t[k];
  });
}

int qol_hbm_copy(int64_t nbytes, double* ms) {
  return guarded([&] {
    if (nbytes <= 0 || nbytes % 16) throw std::runtime_error("qol_hbm_copy: nbytes must be a positive multiple of 16");
    void *src = nullptr, *dst = nullptr;
    hipEvent_t e0, e1;
    HIPCHK(hipMalloc(&src, (size_t)nbytes));
    HIPCHK(hipMalloc(&dst, (size_t)nbytes));
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const size_t n = (size_t)nbytes / 16;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
    HIPCHK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(hbm_copy_kernel, dim3(grid), dim3(256), 0, 0, n, (const double2*)src, (double2*)dst);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e1, 0));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    HIPCHK(hipEventDestroy(e0));
    HIPCHK(hipEventDestroy(e1));
    HIPCHK(hipFree(src));
    HIPCHK(hipFree(dst));
  });
}

int qol_order_info(int64_t p, int64_t* info) {
  return guarded([&] {
    const OrderTables& T = order_tables(p, g_square.load() != 0);
    info[0] = T.h; info[1] = T.nq; info[2] = T.ncoef; info[3] = T.ntot; info[4] = T.nsubtot;
    for (int l = 0; l < 4; ++l) { info[5 + l] = T.lev[l].n; info[9 + l] = T.lev[l].nsub; }
  });
}

int qol_order_level(int64_t p, int64_t lev, double* uv, double* wt, double* interp, int64_t* suboff,
                    int64_t* subtyp) {
  return guarded([&] {
    const OrderTables& T = order_tables(p, g_square.load() != 0);
    if (lev < 1 || lev > 4) throw std::runtime_error("level must be 1..4");
    const LevelTable& L = T.lev[lev - 1];
    std::memcpy(uv, L.uv.data(), sizeof(double) * 2 * L.n);
    std::memcpy(wt, L.wt.data(), sizeof(double) * L.n);
    std::memcpy(interp, L.interpmat.data(), sizeof(double) * L.n * T.h);
    for (int64_t q = 0; q <= L.nsub; ++q) suboff[q] = L.suboff[q];
    for (int64_t q = 0; q < L.nsub; ++q) subtyp[q] = L.subtyp[q];
  });
}

// ---------------------------------------------------------------- drop-ins
void qol_simplex_precomp_mex_(const int64_t* nquad, const int64_t* korder, const int64_t* kpols, double* tgl,
                              double* wgl, double* Dgl, double* w_bclag, double* Legmat, double* umatr,
                              double* vmatr) {
  guarded([&] {
    const int64_t p = *korder + 1;
    if (*kpols != p * (p + 1) / 2) throw std::runtime_error("kpols must be (korder+1)(korder+2)/2");
    const OrderTables& T = order_tables(p);
    const int64_t nq = *nquad;
    std::vector<double> t(nq), w(nq), D(nq * nq);
    gl_nodes(nq, t.data(), w.data(), D.data());
    std::memcpy(tgl, t.data(), sizeof(double) * nq);
    std::memcpy(wgl, w.data(), sizeof(double) * nq);
    std::memcpy(Dgl, D.data(), sizeof(double) * nq * nq);
    for (int64_t j = 0; j < nq; ++j) {
      double pr = 1.0;
      for (int64_t k = 0; k < nq; ++k) if (k != j) pr *= t[j] - t[k];
      w_bclag[j] = 1.0 / pr;
    }
    for (int64_t i = 0; i < nq; ++i) {
      double pkm1 = 0.0, pk = 1.0;
      for (int64_t k = 0; k < nq; ++k) {
        Legmat[k + nq * i] = pk * w[i] * (2.0 * k + 1.0) * 0.5;
        const double pn = ((2.0 * k + 1.0) * t[i] * pk - (double)k * pkm1) / (double)(k + 1);
        pkm1 = pk;
        pk = pn;
      }
    }
    std::memcpy(umatr, T.umatr.data(), sizeof(double) * T.h * T.h);
    std::memcpy(vmatr, T.vmatr.data(), sizeof(double) * T.h * T.h);
  });
}

// rrq_r64 for m targets against one panel of n = order(order+1)/2 nodes.  The
// per-order tables (tgl ... vmatr) are rebuilt internally; the caller's copies
// are accepted for interface compatibility (they come from qol_simplex_precomp).
// Targets within 1e-10 of the patch take the side given by exterior; every other
// target is classified by its own side, so off-surface targets on either side
// are served.  Omega holds the moment right-hand side scaled by 1/(4 pi alpha),
// IalphaAsvestas the solid-angle moment int z.n/|z|^3 dS.
void qol_rrq_mex_(const int64_t* m, const double* tx, const int64_t* n, const double* sx, const double* snx,
                  const double* sw, const double* rts, const double* rps, const int64_t* order, const int64_t* nquad,
                  const int64_t* orderff, const double* distff, const int64_t* exterior, const int64_t* isimd,
                  const int64_t* ichart, const double* sxbd_chart, const double* rv_chart, const double* tgl,
                  const double* wgl, const double* Dgl, const double* w_bclag, const double* Legmat,
                  const double* umatr, const double* vmatr, double* As, double* Ad, double* Omega,
                  double* IalphaAsvestas, double* timeinfo) {
  (void)isimd; (void)tgl; (void)wgl; (void)Dgl; (void)w_bclag; (void)Legmat; (void)umatr; (void)vmatr;
  guarded([&] {
    set_unconverged(0, 0.0);
    const int64_t p = *order, h = p * (p + 1) / 2;
    if (*n != h) throw std::runtime_error("n must be order(order+1)/2");
    check_orders(p, *nquad, *orderff, *distff);
    const int64_t mm = *m;
    std::memset(As, 0, sizeof(double) * mm * h);
    std::memset(Ad, 0, sizeof(double) * mm * h);
    std::memset(Omega, 0, sizeof(double) * 4 * h * mm);
    std::memset(IalphaAsvestas, 0, sizeof(double) * mm);
    if (mm == 0) return;
    Engine& e = panel_engine(p, *ichart, sx, snx, sw, rts, rps, sxbd_chart, rv_chart);
    e.set_targets(mm, tx, false);
    std::vector<int64_t> offs = {0, mm}, idx(mm);
    for (int64_t j = 0; j < mm; ++j) idx[j] = j + 1;
    e.set_pairs(offs.data(), idx.data());
    e.eval(*exterior != 0 ? 1 : 0, 1e-10, true);
    e.get_weights(As, Ad, nullptr, Omega, IalphaAsvestas);
    {
      double st = 0.0;
      const int64_t n = e.projection_unconverged(&st);
      set_unconverged(n, st);
    }
    double ms[10];
    e.stage_ms(ms);
    add_timeinfo(ms, timeinfo);
  });
}

void qol_getnearquad_lap_rrq_mex_(const int64_t* npan, const int64_t* nterms, const int64_t* hdim,
                                  const int64_t* nquad, const int64_t* nsrc, const double* sx, const double* snx,
                                  const double* sw, const double* rts, const double* rps, const int64_t* ntcx,
                                  const double* tcx, const int64_t* tcxrow, const int64_t* tcxi,
                                  const int64_t* orderff, const double* distff, const int64_t* iside,
                                  const int64_t* isimd, const int64_t* nnz, int64_t* sparsei, int64_t* sparsej,
                                  double* sparsevs, double* sparsevd, double* timeinfo) {
  (void)isimd;
  guarded([&] {
    set_unconverged(0, 0.0);
    const int64_t p = *nterms, h = *hdim;
    if (h != p * (p + 1) / 2) throw std::runtime_error("hdim must be nterms(nterms+1)/2");
    if (*nsrc != *npan * h) throw std::runtime_error("nsrc must be npan*hdim");
    check_orders(p, *nquad, *orderff, *distff);
    const int64_t np_ = *npan, nt = *ntcx;
    if (tcxi[0] != 1) throw std::runtime_error("tcxi(1) must be 1");
    if ((tcxi[np_] - 1) * h != *nnz) throw std::runtime_error("nnz must equal (tcxi(npan+1)-1)*hdim");
    bool ad = false, sq = true;
    level_mode(&ad, &sq);
    Engine e(p, np_, sx, snx, sw, rts, rps, 0, nullptr, nullptr, false, nullptr, sq);
    e.set_level_mode(ad, sq);
    e.set_targets(nt, tcx, false);
    std::vector<int64_t> offs(np_ + 1), idx(tcxi[np_] - 1);
    for (int64_t k = 0; k <= np_; ++k) offs[k] = tcxi[k] - 1;
    for (size_t j = 0; j < idx.size(); ++j) idx[j] = (int64_t)j + 1;
    e.set_pairs(offs.data(), idx.data());
    e.eval((*iside % 10) == 0 ? 1 : 0, 1e-10, false);
    e.triplets(tcxrow, sparsei, sparsej, sparsevs, sparsevd);
    {
      double st = 0.0;
      const int64_t n = e.projection_unconverged(&st);
      set_unconverged(n, st);
    }
    double ms[10];
    e.stage_ms(ms);
    add_timeinfo(ms, timeinfo);
  });
}

}  // extern "C"

// ---------------------------------------------------------------- level handles
// One slot with a generation-tagged handle, as the reference's registry
// (patch_refine_mex.f90:8-195): handle = generation*1 + slot; ier 0 ok,
// 1 no free slot, 2 bad handle, 3 npols mismatch, 5 unsupported level set.
namespace {
struct LevelSlot {
  bool live = false;
  int64_t generation = 1, npols = 0, order = 0;
  std::vector<double> sx, rts, rps;
};
std::mutex g_slot_mu;
LevelSlot g_slot;
}  // namespace

extern "C" {

void qol_build_patch_levels_mex_(const int64_t* korder, const int64_t* npols, const double* sx, const double* rts,
                                 const double* rps, const int64_t* order, const double* dist, const int64_t* nlevel,
                                 int64_t* lv, int64_t* ier) {
  std::lock_guard<std::mutex> g(g_slot_mu);
  *lv = 0;
  const int64_t p = *korder + 1;
  if (*npols != p * (p + 1) / 2 || *order != p + 2 || *nlevel != 4 || *dist != 1.4 || p > 12) {
    *ier = 5;
    set_error("levels are built for nlevel=4, order=korder+3, dist=1.4, korder<=11");
    return;
  }
  if (g_slot.live) { *ier = 1; return; }
  const int64_t h = *npols;
  g_slot.sx.assign(sx, sx + 3 * h);
  g_slot.rts.assign(rts, rts + 3 * h);
  g_slot.rps.assign(rps, rps + 3 * h);
  g_slot.npols = h;
  g_slot.order = p;
  g_slot.live = true;
  *lv = g_slot.generation * 1 + 1;
  *ier = 0;
}

void qol_lap3dsdlpmat_levels_mex_(const int64_t* m, const double* tx, const int64_t* npols, const int64_t* lv,
                                  const int64_t* isimd, double* As, double* Ad, int64_t* idxs, int64_t* ms,
                                  int64_t* ier) {
  (void)isimd;
  std::lock_guard<std::mutex> g(g_slot_mu);
  if (*lv <= 0 || !g_slot.live || (*lv - 1) / 1 != g_slot.generation) { *ms = 0; *ier = 2; return; }
  if (*npols != g_slot.npols) { *ms = 0; *ier = 3; return; }
  const int rc = guarded([&] {
    const int64_t h = g_slot.npols, mm = *m;
    // unit normals/weights are not needed for the level path; pass the tangents
    std::vector<double> snx(3 * h, 0.0), sw(h, 1.0);
    Engine& e = panel_engine(g_slot.order, 0, g_slot.sx.data(), snx.data(), sw.data(), g_slot.rts.data(),
                             g_slot.rps.data(), nullptr, nullptr);
    e.set_targets(mm, tx, false);
    std::vector<int64_t> offs = {0, mm}, idx(mm);
    for (int64_t j = 0; j < mm; ++j) idx[j] = j + 1;
    e.set_pairs(offs.data(), idx.data());
    e.eval_levels_only();
    std::vector<double> a(mm * h), d(mm * h);
    std::vector<int64_t> lab(mm);
    e.get_weights(a.data(), d.data(), lab.data(), nullptr, nullptr);
    int64_t k = 0;
    for (int64_t j = 0; j < mm; ++j) {
      if (lab[j] <= 4) {
        for (int64_t c = 0; c < h; ++c) { As[j + mm * c] = a[j + mm * c]; Ad[j + mm * c] = d[j + mm * c]; }
      } else {
        idxs[k++] = j + 1;
      }
    }
    *ms = k;
  });
  *ier = rc == 0 ? 0 : 6;
}

void qol_clear_patch_levels_mex_(const int64_t* lv, int64_t* ier) {
  std::lock_guard<std::mutex> g(g_slot_mu);
  if (*lv <= 0 || !g_slot.live || (*lv - 1) / 1 != g_slot.generation) { *ier = 2; return; }
  g_slot.live = false;
  g_slot.generation += 1;
  g_slot.sx.clear(); g_slot.rts.clear(); g_slot.rps.clear();
  *ier = 0;
}

void qol_clear_patch_levels_all_mex_(int64_t* ier) {
  std::lock_guard<std::mutex> g(g_slot_mu);
  if (g_slot.live) {
    g_slot.live = false;
    g_slot.generation += 1;
  }
  *ier = 0;
}

}  // extern "C"
