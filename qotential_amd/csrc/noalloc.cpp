// qol_rrq_noalloc_: qol_rrq_mex_ for callers that cannot allocate device memory.
//
// Every device buffer of the call (per-order tables, the panel, targets, pair
// lists, level geometry, RRQ set-up and per-pair data) is sliced from `scratch`,
// a file-scope device array of kScratchBytes per GPU that lives until the
// process ends; the call itself makes no hipMalloc / hipFree.  Set-up happens
// once per device, on the first call with that device current: the array's
// device address (hipGetSymbolAddress) and one stream with its events.
// Calls on one GPU share the array, so overlapping calls would race on it;
// the library serialises them with a per-device lock.  The array keeps the
// last call's contents; the kernels initialise what they read before reading
// it, as they do for freshly allocated buffers.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "qol_engine.h"
#include "qol_host.h"
#include "qotential_hip.h"

namespace qol {
void set_status(int64_t s);
void set_unconverged(int64_t n, double max_step);
void set_error(const std::string& s);
}  // namespace qol

using namespace qol;

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;  // 64 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) char scratch[kScratchBytes];

struct DeviceScratch {
  std::mutex mu;     // serialises calls on this device
  bool ready = false;
  char* ptr = nullptr;  // device address of this GPU's copy of `scratch`
  StreamSet ss;
};

std::mutex g_setup_mu;
DeviceScratch g_dev[kMaxDevices];

DeviceScratch& device_scratch() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
  DeviceScratch& d = g_dev[dev];
  std::lock_guard<std::mutex> g(g_setup_mu);
  if (!d.ready) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    d.ptr = static_cast<char*>(p);
    d.ss.create_one_stream();
    d.ready = true;
  }
  return d;
}

// the calling thread's device buffers come from `a` while this object lives
struct ArenaScope {
  Arena* prev;
  explicit ArenaScope(Arena* a) : prev(thread_arena()) { thread_arena() = a; }
  ~ArenaScope() { thread_arena() = prev; }
};

}  // namespace

extern "C" void qol_rrq_noalloc_(const int64_t* m, const double* tx, const int64_t* n, const double* sx,
                                 const double* snx, const double* sw, const double* rts, const double* rps,
                                 const int64_t* order, const int64_t* nquad, const int64_t* orderff,
                                 const double* distff, const int64_t* exterior, const int64_t* isimd,
                                 const int64_t* ichart, const double* sxbd_chart, const double* rv_chart,
                                 const double* tgl, const double* wgl, const double* Dgl, const double* w_bclag,
                                 const double* Legmat, const double* umatr, const double* vmatr, double* As,
                                 double* Ad, double* Omega, double* IalphaAsvestas, double* timeinfo) {
  (void)isimd; (void)tgl; (void)wgl; (void)Dgl; (void)w_bclag; (void)Legmat; (void)umatr; (void)vmatr;
  try {
    set_unconverged(0, 0.0);
    const int64_t p = *order, h = p * (p + 1) / 2;
    if (p < 2 || p > 12) throw std::runtime_error("qol_rrq_noalloc_: order must be 2..12");
    if (*n != h) throw std::runtime_error("n must be order(order+1)/2");
    if (*nquad != p + 2 || *orderff != p + 2)
      throw std::runtime_error("this library evaluates nquad = orderff = order + 2 (the reference drivers' choice)");
    if (*distff != 1.4) throw std::runtime_error("distff must be 1.4 (the reference's level radius factor)");
    const int64_t mm = *m;
    if (mm < 0) throw std::runtime_error("m must be non-negative");
    bool ad = false, sq = true;
    level_mode(&ad, &sq);
    if (ad && sq) throw std::runtime_error("PR_SQUARE and PR_ADAPTIVE cannot be combined");
    const size_t need = Engine::one_panel_bytes(p, mm, *ichart, sq);
    if (need > kScratchBytes)
      throw std::runtime_error("qol_rrq_noalloc_: " + std::to_string(mm) + " targets at order " + std::to_string(p) +
                               " need " + std::to_string(need) + " bytes of device scratch; the library holds " +
                               std::to_string(kScratchBytes) + " (use qol_rrq_mex_ for larger calls)");
    std::memset(As, 0, sizeof(double) * mm * h);
    std::memset(Ad, 0, sizeof(double) * mm * h);
    std::memset(Omega, 0, sizeof(double) * 4 * h * mm);
    std::memset(IalphaAsvestas, 0, sizeof(double) * mm);
    if (mm > 0) {
      DeviceScratch& d = device_scratch();
      std::lock_guard<std::mutex> g(d.mu);
      Arena a{d.ptr, kScratchBytes, 0};
      ArenaScope scope(&a);
      Engine e(p, 1, sx, snx, sw, rts, rps, *ichart, sxbd_chart, rv_chart, false, &d.ss, sq);
      e.set_level_mode(ad, sq);
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
    }
    set_status(0);
  } catch (const std::exception& ex) {
    set_error(ex.what());
    set_status(1);
  }
}
