// Host-only driver for the AddressSanitizer / UndefinedBehaviorSanitizer run of
// the library's host code (geometry, per-order tables, workload C-ABI) and of
// the oracle.  Built and run by tests/test_asan_host.py with g++; it makes no
// GPU call (the one GPU entry the host geometry code can reach, the batched
// curvature tails, is given a stand-in that refuses).
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "qol_host.h"
#include "qotential_hip.h"
#include "qol_oracle.h"

extern "C" {
i64 orc_nearlist_count_mt(i64 npan, i64 hdim, const double* sx, const double* sw, i64 nt, const double* tx, i64* offs,
                          i64 nthreads);
void orc_nearlist_fill_mt(i64 npan, i64 hdim, const double* sx, const double* sw, i64 nt, const double* tx,
                          const i64* offs, i64* idx, i64 nthreads);
void orc_labels(i64 npan, i64 order, const double* sx, const double* rts, const double* rps, const double* tx,
                const i64* offs, const i64* idx, i64* label, i64* cut, i64 nthreads);
i64 orc_set_level_mode(i64 adaptive, i64 square);
}

namespace qol {
std::vector<double> curvature_tails_gpu(const Surface&, const std::vector<Chart>&, i64, i64, i64) {
  throw std::runtime_error("no GPU in the sanitizer build");
}
}  // namespace qol

int main() {
  // per-order tables for every supported order
  for (int p = 2; p <= 12; ++p)
    for (bool square : {true, false}) {
      const qol::OrderTables& T = qol::order_tables(p, square);
      if (T.h != p * (p + 1) / 2) return 1;
    }
  // geometry: built-in surface, VMEC surface, refined W7-X-like charts
  const double modes[3][4] = {{0, 0, 1.0, 0.0}, {1, 0, 0.3, 0.3}, {1, 1, 0.06, -0.06}};
  for (int p : {4, 8}) {
    const int64_t n = qol_geo_count(12, 36, p, 5, 3, &modes[0][0], 1);
    if (n <= 0) return 2;
    const int64_t h = p * (p + 1) / 2, ns = n * h;
    std::vector<double> sx(3 * ns), snx(3 * ns), sw(ns), rts(3 * ns), rps(3 * ns), ch(6 * (n / 2));
    if (qol_geo_build(12, 36, p, 5, 3, &modes[0][0], 1, sx.data(), snx.data(), sw.data(), rts.data(), rps.data(),
                      ch.data()))
      return 3;
    std::vector<double> sxbd(3 * 3 * (p + 2) * n), rv(9 * n);
    if (qol_geo_chart_boundary(12, 36, p, 5, 3, &modes[0][0], 1, p + 2, sxbd.data(), rv.data())) return 4;
    const int64_t nt = 500;
    std::vector<double> tx(3 * nt), dist(nt);
    std::vector<int64_t> tri(nt);
    if (qol_make_targets(p, n, sx.data(), snx.data(), nt, -6.0, -1.0, 0, 7, tx.data(), tri.data(), dist.data()))
      return 5;
    // the oracle on a few panels of this surface
    std::vector<int64_t> offs(n + 1);
    orc_nearlist_count_mt(n, h, sx.data(), sw.data(), nt, tx.data(), offs.data(), 2);
    std::vector<int64_t> idx(offs[n] > 0 ? offs[n] : 1);
    orc_nearlist_fill_mt(n, h, sx.data(), sw.data(), nt, tx.data(), offs.data(), idx.data(), 2);
    std::vector<int64_t> lab(idx.size()), cut(idx.size());
    // square levels, then the adaptive walk on triangle levels
    for (int mode = 0; mode < 2; ++mode) {
    orc_set_level_mode(mode, 1 - mode);
    orc_labels(n, p, sx.data(), rts.data(), rps.data(), tx.data(), offs.data(), idx.data(), lab.data(), cut.data(), 2);
    for (int64_t k = 0; k < n; k += n / 5) {
      const int64_t m = offs[k + 1] - offs[k];
      if (m == 0) continue;
      std::vector<double> t3(3 * m), As(m * h), Ad(m * h), Om(4 * h * m), Ia(m), ti(20, 0.0);
      std::vector<int64_t> lb(m);
      for (int64_t j = 0; j < m; ++j)
        for (int c = 0; c < 3; ++c) t3[3 * j + c] = tx[3 * (idx[offs[k] + j] - 1) + c];
      const int64_t nq = p + 2;
      std::vector<double> tgl(nq), wgl(nq), Dgl(nq * nq), wb(nq), Leg(nq * nq), um(h * h), vm(h * h);
      orc::simplex_precomp(nq, p - 1, h, tgl.data(), wgl.data(), Dgl.data(), wb.data(), Leg.data(), um.data(), vm.data());
      orc::rrq(m, t3.data(), h, sx.data() + 3 * k * h, snx.data() + 3 * k * h, sw.data() + k * h,
               rts.data() + 3 * k * h, rps.data() + 3 * k * h, p, nq, p + 2, 1.4, true, 1,
               sxbd.data() + 9 * nq * k, rv.data() + 9 * k, tgl.data(), wgl.data(), Dgl.data(), wb.data(),
               Leg.data(), um.data(), vm.data(), As.data(), Ad.data(), Om.data(), Ia.data(), ti.data(), lb.data());
      for (double v : As)
        if (!std::isfinite(v)) return 6;
    }
    }
    orc_set_level_mode(0, 1);
  }
  // adaptive chart refinement on the host
  std::vector<double> cl(6 * 200000);
  if (qol_geo_charts(8, 24, 6, 5, 3, &modes[0][0], 0.05, 0, cl.data()) <= 0) return 7;
  // tensor-patch close DLP
  const int p = 5;
  std::vector<double> t(p), w(p), sxq(3 * p * p), txq = {0.01, 0.02, 0.05, 0.0, 0.0, 1e-6}, Ac(2 * p * p);
  orc::gauss(p, t.data(), w.data(), nullptr);
  for (int b = 0; b < p; ++b)
    for (int a = 0; a < p; ++a) {
      sxq[3 * (b * p + a)] = 0.1 * t[a];
      sxq[3 * (b * p + a) + 1] = 0.1 * t[b];
      sxq[3 * (b * p + a) + 2] = 0.0;
    }
  orc::closepanel_dlp(2, txq.data(), p, sxq.data(), 1, Ac.data());
  std::printf("asan driver done\n");
  return 0;
}
