// pt_ensemble.cpp -- ensembles of independent cases advanced together.
//
// The reference runs one case per process (PS/main.f90:245-300); an
// ensemble of cases with different grids and parameters is run here in one
// process on one GPU.  Each case is a full context (its own device arrays
// and HIP stream); pt_ensemble_run_steps hands the cases to a pool of host
// threads, each thread stepping its cases in turn, so that the small,
// latency-bound kernels of different cases (coarse multigrid levels, filter
// passes, reductions) run concurrently on the GPU instead of one case at a
// time.  A case's arithmetic does not depend on the pool, so every case's
// results are bit-identical to the same case run alone.
//
// Autotuning: with workers <= 0 the first steps are run with each candidate
// pool size in turn and the rest with the fastest (the tuning steps are real
// steps: they advance the cases exactly as any other steps do).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <vector>
#include "pt_internal.hpp"

namespace pt {

struct Ensemble {
  std::vector<void*> cases;
  std::vector<int> status;    // per case: PT_OK or the error that stopped it
  int workers = 0;            // pool size of the last run (chosen by the tuner when asked)
  std::vector<double> tune_s; // seconds per ensemble step of each tried pool size
  std::vector<int> tune_w;
};

// one ensemble step with `w` threads: thread t steps cases t, t+w, ...
static void step_all(Ensemble& e, int w) {
  const int n = (int)e.cases.size();
  std::vector<std::thread> pool;
  pool.reserve(w);
  for (int t = 0; t < w; t++) {
    pool.emplace_back([&e, t, w, n]() {
      for (int i = t; i < n; i += w) {
        if (e.status[i] != PT_OK) continue;
        (void)hipSetDevice(static_cast<Context*>(e.cases[i])->device);
        e.status[i] = pt_time_step(e.cases[i]);
      }
    });
  }
  for (std::thread& th : pool) th.join();
}

}  // namespace pt

using namespace pt;

extern "C" {

int pt_ensemble_create(int32_t n, const pt_params* params, void** ens) {
  if (!params || !ens || n < 1) { set_error("pt_ensemble_create: n >= 1 cases required"); return PT_ERR_INVALID; }
  *ens = nullptr;
  Ensemble* e = new Ensemble();
  for (int i = 0; i < n; i++) {
    void* c = nullptr;
    const int rc = pt_create(&params[i], &c);
    if (rc != PT_OK) {
      for (void* q : e->cases) pt_destroy(q);
      delete e;
      return rc;
    }
    e->cases.push_back(c);
  }
  e->status.assign(n, PT_OK);
  *ens = e;
  return PT_OK;
}

int pt_ensemble_destroy(void* ens) {
  Ensemble* e = static_cast<Ensemble*>(ens);
  if (!e) return PT_OK;
  for (void* c : e->cases) pt_destroy(c);
  delete e;
  return PT_OK;
}

int pt_ensemble_case(void* ens, int32_t i, void** ctx) {
  Ensemble* e = static_cast<Ensemble*>(ens);
  if (!e || !ctx || i < 0 || i >= (int)e->cases.size()) return PT_ERR_INVALID;
  *ctx = e->cases[i];
  return PT_OK;
}

int pt_ensemble_run_steps(void* ens, int32_t nsteps, int32_t workers, int32_t* status) {
  Ensemble* e = static_cast<Ensemble*>(ens);
  if (!e || nsteps < 0) return PT_ERR_INVALID;
  const int n = (int)e->cases.size();
  int step = 0;
  if (workers <= 0) {
    // tune the pool size on the first steps
    e->tune_w.clear();
    e->tune_s.clear();
    for (int w : {1, 2, 4, 8, 16}) {
      if (w > n || step >= nsteps) break;
      const auto t0 = std::chrono::steady_clock::now();
      step_all(*e, w);
      for (void* c : e->cases) (void)pt_synchronize(c);
      const auto t1 = std::chrono::steady_clock::now();
      e->tune_w.push_back(w);
      e->tune_s.push_back(std::chrono::duration<double>(t1 - t0).count());
      step++;
    }
    int best = 1;
    double bt = 1e300;
    for (size_t k = 0; k < e->tune_w.size(); k++)
      if (e->tune_s[k] < bt) { bt = e->tune_s[k]; best = e->tune_w[k]; }
    e->workers = best;
  } else {
    e->workers = std::min(workers, n);
  }
  for (; step < nsteps; step++) step_all(*e, e->workers);
  for (void* c : e->cases) (void)pt_synchronize(c);
  if (status)
    for (int i = 0; i < n; i++) status[i] = e->status[i];
  return PT_OK;
}

int pt_ensemble_workers(void* ens, int32_t* workers) {
  Ensemble* e = static_cast<Ensemble*>(ens);
  if (!e || !workers) return PT_ERR_INVALID;
  *workers = e->workers;
  return PT_OK;
}

}  // extern "C"
