// Mixed-order batches: the close-evaluation step (eval + corrections + the
// per-target sums) of several engines -- one per patch order of a mixed-order
// surface -- launched together from one call, and a per-order autotuner over
// the step's schedule.  Capturing the step into one HIP graph is built but
// switched off: the one capture run so far ended its process with a host-side
// segmentation fault whose cause has not been found, so qol_batch_capture
// refuses until it is.
//
// The reference evaluates a mixed-order surface panel by panel, each rrq_r64
// call taking that panel's own order (src/lap3d_close_mod.f90:622-649 inside the
// OpenMP loop of stellarator_grf_core_mod.f90:759-888).  Here each order's
// panels form one engine (work bucketed by order, and inside an engine by
// level label); one call launches every bucket's kernels with their stream
// dependencies, so the buckets overlap on the device with no host wait.
//
// Autotuner: for each engine, the candidates are the three-stream and the
// one-stream schedule and the RRQ solve variants (register-resident inverses
// and the fused kernel at every order, the matrix-core form up to order 8); each
// candidate's single-engine step is timed over a few steps and the fastest
// kept; then the order in which the engines are launched (ascending or
// descending patch order) is timed the same way.  The schedule choice does
// not change any result; the solve variants agree to the conditioning of the
// quaternion system.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "qol_engine.h"
#include "qotential_hip.h"

namespace qol {
void set_status(int64_t s);
void set_error(const std::string& s);

struct Batch {
  std::vector<Engine*> eng;
  std::vector<int> order;  // engines in graph order
  hipStream_t origin = nullptr;
  hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
  std::vector<hipEvent_t> join;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int exterior = 1;
  double side_tol = 1e-10;

  explicit Batch(const std::vector<Engine*>& e) : eng(e) {
    if (eng.empty()) throw std::runtime_error("batch: no engines");
    for (Engine* x : eng) {
      if (!x) throw std::runtime_error("batch: null engine");
      if (!x->evaluated || !x->density_set || x->nnz < 0 || x->tlists_nnz != x->nnz)
        throw std::runtime_error("batch: every engine needs targets, pairs, a resident density and one eager "
                                 "eval + apply before capture");
    }
    HIPCHK(hipStreamCreateWithFlags(&origin, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    HIPCHK(hipEventCreate(&t0));
    HIPCHK(hipEventCreate(&t1));
    join.resize(eng.size());
    for (auto& j : join) HIPCHK(hipEventCreateWithFlags(&j, hipEventDisableTiming));
    for (int i = 0; i < (int)eng.size(); ++i) order.push_back(i);
  }
  ~Batch() {
    release();
    for (auto& j : join) (void)hipEventDestroy(j);
    if (fork) (void)hipEventDestroy(fork);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (origin) (void)hipStreamDestroy(origin);
  }
  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  // every engine's device work must be finished before a capture starts
  void sync_engines() {
    for (Engine* x : eng) HIPCHK(hipStreamSynchronize(x->stream));
  }
  // one step of the given engines, launched directly: each engine's stages on
  // its own streams, all joined into the origin stream
  void step(const std::vector<int>& which) {
    HIPCHK(hipEventRecord(fork, origin));
    for (int i : which) {
      Engine* e = eng[i];
      HIPCHK(hipStreamWaitEvent(e->stream, fork, 0));
      e->eval(exterior, side_tol, false);
      e->apply_device();
      HIPCHK(hipEventRecord(join[i], e->stream));
      HIPCHK(hipStreamWaitEvent(origin, join[i], 0));
    }
  }
  void capture(const std::vector<int>& which) {
    throw std::runtime_error("batch graph capture is switched off: its first run ended the process with a host-side "
                             "segmentation fault whose cause is not yet found");
    release();
    sync_engines();
    HIPCHK(hipStreamBeginCapture(origin, hipStreamCaptureModeRelaxed));
    try {
      HIPCHK(hipEventRecord(fork, origin));
      for (int i : which) {
        Engine* e = eng[i];
        HIPCHK(hipStreamWaitEvent(e->stream, fork, 0));
        e->timing = false;
        e->eval(exterior, side_tol, false);
        e->apply_device();
        e->timing = true;
        HIPCHK(hipEventRecord(join[i], e->stream));
        HIPCHK(hipStreamWaitEvent(origin, join[i], 0));
      }
    } catch (...) {
      for (Engine* e : eng) e->timing = true;
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(origin, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(origin, &graph));
    HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  void launch() {
    if (exec) HIPCHK(hipGraphLaunch(exec, origin));
    else step(order);
  }
  // mean ms per step of the given engines over reps steps (after one untimed step)
  double timed(const std::vector<int>& which, int reps) {
    step(which);
    HIPCHK(hipEventRecord(t0, origin));
    for (int r = 0; r < reps; ++r) step(which);
    HIPCHK(hipEventRecord(t1, origin));
    HIPCHK(hipEventSynchronize(t1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, t0, t1));
    return (double)ms / reps;
  }
};

Batch* B(void* h) {
  if (!h) throw std::runtime_error("null batch handle");
  return static_cast<Batch*>(h);
}

}  // namespace qol

using namespace qol;

namespace {
template <class F>
int guarded_batch(F&& f) {
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

int qol_batch_create(int64_t n, void** engines, void** out) {
  return guarded_batch([&] {
    *out = nullptr;
    std::vector<Engine*> e(n);
    for (int64_t i = 0; i < n; ++i) e[i] = static_cast<Engine*>(engines[i]);
    *out = new Batch(e);
  });
}

int qol_batch_destroy(void* h) {
  return guarded_batch([&] { delete B(h); });
}

int qol_batch_capture(void* h, int64_t exterior, double side_tol) {
  return guarded_batch([&] {
    Batch* b = B(h);
    b->exterior = (int)exterior;
    b->side_tol = side_tol;
    b->capture(b->order);
  });
}

int qol_batch_launch(void* h) {
  return guarded_batch([&] { B(h)->launch(); });
}

int qol_batch_sync(void* h) {
  return guarded_batch([&] { HIPCHK(hipStreamSynchronize(B(h)->origin)); });
}

int qol_batch_autotune(void* h, int64_t reps, double* ms, int64_t* choice) {
  return guarded_batch([&] {
    Batch* b = B(h);
    const int n = (int)b->eng.size(), r = (int)std::max<int64_t>(1, reps);
    for (int i = 0; i < n; ++i) {
      Engine* e = b->eng[i];
      double best = 1e300;
      int bs = 0, bv = 0;
      for (int one = 0; one < 2; ++one)
        for (int v = 0; v < 3; ++v) {
          double t = -1.0;
          if (v < 2 || e->p <= 8) {
            e->one_stream_run = one != 0;
            e->solve_variant = v;
            t = b->timed({i}, r);
            if (t < best) { best = t; bs = one; bv = v; }
          }
          if (ms) ms[(size_t)i * 6 + one * 3 + v] = t;
        }
      e->one_stream_run = bs != 0;
      e->solve_variant = bv;
      if (choice) { choice[2 * i] = bs; choice[2 * i + 1] = bv; }
    }
    // the engines' order in the graph: ascending or descending patch order
    std::vector<int> up(n);
    for (int i = 0; i < n; ++i) up[i] = i;
    std::sort(up.begin(), up.end(), [&](int a, int c) { return b->eng[a]->p < b->eng[c]->p; });
    std::vector<int> down(up.rbegin(), up.rend());
    const double tu = b->timed(up, r);
    const double td = b->timed(down, r);
    b->order = td < tu ? down : up;
    if (choice) choice[2 * n] = td < tu ? 1 : 0;
    if (ms) { ms[(size_t)n * 6] = tu; ms[(size_t)n * 6 + 1] = td; }
  });
}

}  // extern "C"
