// pt_api.hip -- C ABI entry points (include/projectracer_gpu.h).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "pt_internal.hpp"

namespace pt {

static thread_local std::string g_err;
void set_error(const std::string& s) { g_err = s; }

// Per-kernel HIP-event timing on the context stream.  Events are recorded
// around launches and folded into the total once the ring fills.
static void timer_flush(KTimer& t) {
  if (t.n == 0) return;
  (void)hipEventSynchronize(t.end[t.n - 1]);
  for (int j = 0; j < t.n; j++) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.beg[j], t.end[j]) == hipSuccess) t.ms += ms;
  }
  t.launches += t.n;
  t.n = 0;
}
void timer_begin(Context& c, int which, int64_t units) {
  KTimer& t = c.kt[which];
  if (t.n == 1024) timer_flush(t);
  t.work += units;
  (void)hipEventRecord(t.beg[t.n], c.stream);
}
void timer_end(Context& c, int which) {
  KTimer& t = c.kt[which];
  (void)hipEventRecord(t.end[t.n], c.stream);
  t.n++;
}

static int check_supported(const pt_params& p) {
  if (p.nx < 8 || p.nz < 8 || p.grids < 2 || p.grids > MAXL || p.tpc < 1 || p.mingrid < 2) {
    set_error("nx, nz >= 8, 2 <= grids <= 6, tpc >= 1, mingrid >= 2 required");
    return PT_ERR_INVALID;
  }
  if (p.nstageS < 2 || p.nstageS > 60 || p.Nmulti < 1 || p.max_cycles < 1 || !(p.tolerance > 0.0)) {
    set_error("2 <= nstageS <= 60, Nmulti >= 1, max_cycles >= 1, tolerance > 0 required");
    return PT_ERR_INVALID;
  }
  if (p.ivis < 0 || p.ivis > 2 || p.RaT == 0.0 || (p.Vbc != 0 && p.Vbc != 1) || (p.Tbot != 0 && p.Tbot != 1)) {
    set_error("supported: ivis in {0,1,2}, RaT != 0, Vbc and Tbot in {0,1}");
    return PT_ERR_UNSUPPORTED;
  }
  if (p.precision != PT_PREC_FP64 && p.precision != PT_PREC_MIXED) {
    set_error("precision must be PT_PREC_FP64 or PT_PREC_MIXED");
    return PT_ERR_INVALID;
  }
  if (p.guess != PT_GUESS_WARM && p.guess != PT_GUESS_LEARNED) {
    set_error("guess must be PT_GUESS_WARM or PT_GUESS_LEARNED");
    return PT_ERR_INVALID;
  }
  if (p.iterate != 0 && p.iterate != 1) {
    set_error("iterate must be 0 (direct) or 1 (multigrid)");
    return PT_ERR_INVALID;
  }
  if (p.iterate == 0 && (p.ivis != 0 || (p.Database != 1 && p.Database != 2))) {
    set_error("the direct solver (iterate=0) is built for ivis=0 and Database 1 or 2");
    return PT_ERR_UNSUPPORTED;
  }
  if ((p.iC_continuous != 0 && p.iC_continuous != 1) || (p.equalize != 0 && p.equalize != 1)) {
    set_error("iC_continuous and equalize must be 0 or 1");
    return PT_ERR_INVALID;
  }
  if (p.equalize == 1 && (p.iC_continuous != 0 || p.original != 1 || p.buffer < 1 || !(p.dist_factor >= 0.0))) {
    // the reference requires discrete types for equalize (PS/main.f90:92);
    // the threaded equalize_tracers (original=0) is not built
    set_error("equalize=1 is built for iC_continuous=0, original=1, buffer >= 1, dist_factor >= 0");
    return PT_ERR_UNSUPPORTED;
  }
  if (p.RKC1 != 0 && p.RKC1 != 1) {
    set_error("RKC1 must be 0 (RK1) or 1 (RKC1)");
    return PT_ERR_INVALID;
  }
  if (p.RKC1 == 0 && p.iterate == 1 && p.precision != PT_PREC_FP64) {
    // the mixed path's FP32 cycle is built on RKC1 sweeps
    set_error("the RK1 smoother (RKC1=0) is built for precision = PT_PREC_FP64");
    return PT_ERR_UNSUPPORTED;
  }
  return PT_OK;
}

static Context* ctx_of(void* p) { return static_cast<Context*>(p); }

static double* field_ptr(Context& c, int id, size_t* n) {
  size_t n3 = (size_t)(c.T.nx + 7) * (c.T.nz + 7), n5 = (size_t)(c.T.nx + 11) * (c.T.nz + 11);
  switch (id) {
    case PT_FIELD_SF: *n = n3; return c.SF;
    case PT_FIELD_T: *n = n3; return c.Tratio;
    case PT_FIELD_TVIS: *n = n3; return c.Tvis;
    case PT_FIELD_TBUOY: *n = n3; return c.Tbuoy;
    case PT_FIELD_U: *n = n3; return c.u;
    case PT_FIELD_W: *n = n3; return c.w;
    case PT_FIELD_CNEW: *n = n5; return c.Cnew;
    case PT_FIELD_CVIS: *n = n5; return c.Cvis;
    case PT_FIELD_CBUOY: *n = n5; return c.Cbuoy;
    case PT_FIELD_CONDUCT: *n = n5; return c.cond;
    case PT_FIELD_VIS: *n = n3; return c.L[1].v.vis;
    case PT_FIELD_RESIDUAL: *n = (size_t)(c.T.nx - 1) * (c.T.nz - 1); return c.L[1].res;
    case PT_FIELD_SOURCE: *n = n3; return c.src;
    default: *n = 0; return nullptr;
  }
}

}  // namespace pt

using namespace pt;

extern "C" {

const char* pt_last_error(void) { return g_err.c_str(); }

void pt_default_params(pt_params* p) {
  std::memset(p, 0, sizeof *p);
  p->aspect = 1.0; p->nx = 64; p->nz = 64; p->grids = 6; p->mingrid = 4;
  p->tolerance = 1e-8; p->max_cycles = 1000; p->Nmulti = 1; p->eig_ratio = 1.0; p->vis_iter = 12;
  p->nstageS = 35; p->dampingS = 1000.0; p->c_2 = 0.95; p->c_1 = 0.9;
  p->courant = 0.5; p->courantT = 0.99; p->courantC = 0.99; p->courant_stream = 0.99;
  p->tpc = 20; p->Tbot = 0; p->Vbc = 0; p->RaT = 1e5; p->RaC = 0.0;
  p->H = 0.0; p->conduct_factor = 1.0; p->Htype = 0.0;
  p->ivis = 0; p->visP = 1.0; p->visT = 1e3; p->delta_visT = 2.0; p->eig_ratio_min = 2.0;
  p->visC = 10.0; p->delta_visC = 2.0; p->precision = PT_PREC_FP64; p->guess = PT_GUESS_WARM;
  p->dent = 0.5; p->iterate = 1; p->Database = 2;
  p->iC_continuous = 1; p->equalize = 0; p->original = 1; p->buffer = 10; p->dist_factor = 1e-4;
  p->RKC1 = 1;
}

int pt_create(const pt_params* p, void** out) {
  if (!p || !out) { set_error("null argument"); return PT_ERR_INVALID; }
  *out = nullptr;
  PT_TRY(check_supported(*p));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("no HIP device visible");
    return PT_ERR_NO_DEVICE;
  }
  hipDeviceProp_t prop;
  int dev = 0;
  PT_HIP(hipGetDevice(&dev));
  PT_HIP(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_error(std::string("device is ") + prop.gcnArchName + ", this build targets gfx950");
    return PT_ERR_NO_DEVICE;
  }
  Context* c = new Context();
  std::memset(&c->L, 0, sizeof c->L);
  c->P = *p;
  c->device = dev;
  const char* graph_env = std::getenv("PT_GRAPH");
  c->use_graph = !(graph_env && graph_env[0] == '0');
  const char* gs_env = std::getenv("PT_GUESS_SMOOTH");
  if (gs_env) c->guess_smooth = std::max(0, std::atoi(gs_env));
  const char* gj_env = std::getenv("PT_GUESS_JUDGE");
  if (gj_env) c->guess_judge = std::min(2, std::max(0, std::atoi(gj_env)));
  const char* filter_env = std::getenv("PT_FILTER_HOST");
  c->filter_on_device = !(filter_env && filter_env[0] == '1');
  const char* radix_env = std::getenv("PT_PROJECT_RADIX");
  c->project_radix_only = radix_env && radix_env[0] == '1';
  int rc = build_tables(c->P, c->T);
  if (rc != PT_OK) { delete c; return rc; }
  c->P.nx = c->T.nx; c->P.nz = c->T.nz; c->P.grids = c->T.grids;
  std::memcpy(c->cyc, c->T.multi, sizeof c->cyc);
  c->cyc_n = c->T.nsteps;
  if (c->P.precision == PT_PREC_MIXED) {
    // the mixed path's cycle: a V-cycle whose coarse corrections start from
    // zero at each descent and whose levels >= 2 are smoothed on the way up
    // only (V(0,1) below level 1).  PT_MIXED_CYCLE=W|F|V,
    // PT_MIXED_COARSE_ZERO=0|1 and PT_MIXED_PRESMOOTH_FROM=l (smooth on the
    // way down too below level l; 99: at every level) select others, e.g. the
    // reference's W-cycle with warm-started coarse levels
    const char* ce = std::getenv("PT_MIXED_CYCLE");
    const char kind = (ce && (ce[0] == 'W' || ce[0] == 'F' || ce[0] == 'V')) ? ce[0] : 'V';
    c->cyc_n = cycle_schedule_kind(c->T.grids, kind, c->cyc);
    const char* cz = std::getenv("PT_MIXED_COARSE_ZERO");
    c->coarse_zero = !(cz && cz[0] == '0');
    const char* ps = std::getenv("PT_MIXED_PRESMOOTH_FROM");
    c->presmooth_from = ps ? std::max(2, std::atoi(ps)) : 2;
    // its coarse levels' sweeps: 15 RKC1 stages (the reference's recurrence
    // and damping; dt_stream for that stage count); PT_MIXED_COARSE_STAGES=s
    // selects another count, nstageS or more the reference's everywhere
    const char* cs = std::getenv("PT_MIXED_COARSE_STAGES");
    const int sc = cs ? std::atoi(cs) : 15;
    if (sc >= 2 && sc < c->P.nstageS) rkc1_coarse_tables(c->T, sc, c->P.dampingS);
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    delete c;
    return PT_ERR_HIP;
  }
  for (auto& t : c->kt)
    for (int j = 0; j < 1024; j++) {
      (void)hipEventCreate(&t.beg[j]);
      (void)hipEventCreate(&t.end[j]);
    }
  rc = stats_scratch_setup(*c);
  if (rc == PT_OK) rc = solver_alloc(*c);
  if (rc == PT_OK) rc = tracers_alloc(*c);
  if (rc == PT_OK && c->P.iterate == 0) rc = direct_alloc(*c);
  if (rc != PT_OK) {
    solver_free(*c);
    tracers_free(*c);
    direct_free(*c);
    delete c;
    return rc;
  }
  *out = c;
  return PT_OK;
}

int pt_destroy(void* p) {
  Context* c = ctx_of(p);
  if (!c) return PT_OK;
  (void)hipStreamSynchronize(c->stream);
  solver_free(*c);
  tracers_free(*c);
  direct_free(*c);
  for (auto& t : c->kt)
    for (int j = 0; j < 1024; j++) {
      (void)hipEventDestroy(t.beg[j]);
      (void)hipEventDestroy(t.end[j]);
    }
  (void)hipStreamDestroy(c->stream);
  delete c;
  return PT_OK;
}

int pt_get_dims(void* p, int32_t* nx, int32_t* nz, int32_t* grids, int64_t* ntr) {
  Context* c = ctx_of(p);
  if (!c) return PT_ERR_INVALID;
  if (nx) *nx = c->T.nx;
  if (nz) *nz = c->T.nz;
  if (grids) *grids = c->T.grids;
  if (ntr) *ntr = c->T.ntr;
  return PT_OK;
}

int pt_load_state(void* p, const double* T_nodes, double Tbot_iso, double Ttop_iso, const double* rtr,
                  const double* str, const double* Ttr, const double* Ctr) {
  Context* c = ctx_of(p);
  if (!c || !T_nodes || !rtr || !str || !Ttr || !Ctr) { set_error("null argument"); return PT_ERR_INVALID; }
  if (c->P.iC_continuous == 0)
    for (int64_t j = 0; j < c->T.ntr; j++)
      if (Ctr[j] != 0.0 && Ctr[j] != 1.0) {
        set_error("discrete tracer types (iC_continuous=0): Ctr must be 0 or 1, tracer " + std::to_string(j + 1) +
                  " has " + std::to_string(Ctr[j]));
        return PT_ERR_INVALID;
      }
  return load_state(*c, T_nodes, Tbot_iso, Ttop_iso, rtr, str, Ttr, Ctr);
}

int pt_multigrid(void* p) {
  Context* c = ctx_of(p);
  if (!c) return PT_ERR_INVALID;
  if (c->P.iterate != 1) {
    set_error("multigrid needs a context created with iterate=1");
    return PT_ERR_INVALID;
  }
  PT_TRY(multigrid(*c));
  PT_HIP(hipStreamSynchronize(c->stream));
  return PT_OK;
}

int pt_compute_velocities(void* p) {
  Context* c = ctx_of(p);
  if (!c) return PT_ERR_INVALID;
  PT_TRY(compute_velocities(*c));
  PT_HIP(hipStreamSynchronize(c->stream));
  return PT_OK;
}

int pt_update_velocities(void* p) {
  Context* c = ctx_of(p);
  if (!c) return PT_ERR_INVALID;
  PT_TRY(solve_stream(*c));
  PT_TRY(compute_velocities(*c));
  PT_HIP(hipStreamSynchronize(c->stream));
  return PT_OK;
}

int pt_build_sf_matrix(void* p) {
  Context* c = ctx_of(p);
  if (!c) return PT_ERR_INVALID;
  PT_TRY(build_sf_matrix(*c));
  PT_HIP(hipStreamSynchronize(c->stream));
  return PT_OK;
}

int pt_solve_sf_equation(void* p) {
  Context* c = ctx_of(p);
  if (!c) return PT_ERR_INVALID;
  PT_TRY(solve_sf_equation(*c));
  PT_HIP(hipStreamSynchronize(c->stream));
  return PT_OK;
}

int pt_energy_time(void* p) {
  Context* c = ctx_of(p);
  if (!c) return PT_ERR_INVALID;
  PT_TRY(energy_time(*c));
  PT_HIP(hipStreamSynchronize(c->stream));
  return PT_OK;
}

int pt_tracers_to_corners(void* p) {
  Context* c = ctx_of(p);
  if (!c) return PT_ERR_INVALID;
  size_t b3 = sizeof(double) * (size_t)(c->T.nx + 7) * (c->T.nz + 7);
  PT_HIP(hipMemcpyAsync(c->T0, c->Tratio, b3, hipMemcpyDeviceToDevice, c->stream));
  PT_TRY(tracers_to_corners(*c, c->r, c->s, c->Tt));
  PT_HIP(hipStreamSynchronize(c->stream));
  return PT_OK;
}

static int time_step_async(Context& c) {
  PT_TRY(solve_stream(c));
  PT_TRY(compute_velocities(c));
  PT_TRY(energy_time(c));
  c.time = c.time + c.dt;
  return PT_OK;
}

int pt_time_step(void* p) {
  Context* c = ctx_of(p);
  if (!c) return PT_ERR_INVALID;
  PT_TRY(time_step_async(*c));
  PT_HIP(hipStreamSynchronize(c->stream));
  return PT_OK;
}

int pt_run_steps(void* p, int32_t n) {
  Context* c = ctx_of(p);
  if (!c || n < 0) return PT_ERR_INVALID;
  for (int j = 0; j < n; j++) PT_TRY(time_step_async(*c));
  PT_HIP(hipStreamSynchronize(c->stream));
  return PT_OK;
}

int64_t pt_field_extent(void* p, int32_t id) {
  Context* c = ctx_of(p);
  if (!c) return -1;
  size_t n;
  return field_ptr(*c, id, &n) ? (int64_t)n : -1;
}

int pt_get_field(void* p, int32_t id, double* host) {
  Context* c = ctx_of(p);
  if (!c || !host) return PT_ERR_INVALID;
  size_t n;
  double* d = field_ptr(*c, id, &n);
  if (!d) { set_error("unknown field id"); return PT_ERR_INVALID; }
  if (id == PT_FIELD_RESIDUAL) {
    // level-1 residual is stored with the level pitch; return (1:nx-1,1:nz-1)
    int nx = c->T.nx, nz = c->T.nz, pitch = nx + 7;
    PT_HIP(hipMemcpy2DAsync(host, sizeof(double) * (nx - 1), d + (long)4 * pitch + 4, sizeof(double) * pitch,
                            sizeof(double) * (nx - 1), nz - 1, hipMemcpyDeviceToHost, c->stream));
  } else {
    PT_HIP(hipMemcpyAsync(host, d, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  }
  PT_HIP(hipStreamSynchronize(c->stream));
  return PT_OK;
}

int pt_set_field(void* p, int32_t id, const double* host) {
  Context* c = ctx_of(p);
  if (!c || !host || id == PT_FIELD_RESIDUAL) return PT_ERR_INVALID;
  size_t n;
  double* d = field_ptr(*c, id, &n);
  if (!d) { set_error("unknown field id"); return PT_ERR_INVALID; }
  PT_HIP(hipMemcpyAsync(d, host, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  PT_HIP(hipStreamSynchronize(c->stream));
  return PT_OK;
}

int pt_get_tracers(void* p, double* rtr, double* str, double* Ttr, double* Ctr) {
  Context* c = ctx_of(p);
  if (!c) return PT_ERR_INVALID;
  size_t b = sizeof(double) * (size_t)c->T.ntr;
  double *dr = nullptr, *ds = nullptr, *dT = nullptr, *dC = nullptr;
  PT_TRY(tracers_in_id_order(*c, rtr ? &dr : nullptr, str ? &ds : nullptr, Ttr ? &dT : nullptr,
                             Ctr ? &dC : nullptr));
  if (rtr) PT_HIP(hipMemcpyAsync(rtr, dr, b, hipMemcpyDeviceToHost, c->stream));
  if (str) PT_HIP(hipMemcpyAsync(str, ds, b, hipMemcpyDeviceToHost, c->stream));
  if (Ttr) PT_HIP(hipMemcpyAsync(Ttr, dT, b, hipMemcpyDeviceToHost, c->stream));
  if (Ctr) PT_HIP(hipMemcpyAsync(Ctr, dC, b, hipMemcpyDeviceToHost, c->stream));
  PT_HIP(hipStreamSynchronize(c->stream));
  return PT_OK;
}

int pt_get_nempty(void* p, int32_t* out) {
  Context* c = ctx_of(p);
  if (!c || !out) return PT_ERR_INVALID;
  PT_HIP(hipMemcpyAsync(out, c->nempty, sizeof(int32_t) * (size_t)(c->T.nx + 1) * (c->T.nz + 1),
                        hipMemcpyDeviceToHost, c->stream));
  PT_HIP(hipStreamSynchronize(c->stream));
  return PT_OK;
}

int pt_get_info(void* p, pt_solve_info* info) {
  Context* c = ctx_of(p);
  if (!c || !info) return PT_ERR_INVALID;
  std::memset(info, 0, sizeof *info);
  info->q = c->last_q; info->residual = c->residual1; info->res_ref = c->res_ref1;
  info->dt = c->dt; info->time = c->time; info->solves = c->solves; info->cycles = c->cycles;
  info->n_eta = c->n_eta; info->n_temp = c->n_temp; info->n_comp = c->n_comp; info->eig_fine = c->eig_fine;
  info->cycle_graphs = c->graph_state == 2 ? (int32_t)c->cycle_graphs.size() : 0;
  return PT_OK;
}

int pt_stats(void* p, pt_stats_row* row) {
  Context* c = ctx_of(p);
  if (!c || !row) return PT_ERR_INVALID;
  return stats(*c, row);
}

int pt_stats_noalloc(void* p, pt_stats_row* row) {
  Context* c = ctx_of(p);
  if (!c || !row) return PT_ERR_INVALID;
  return stats_noalloc(*c, row);
}

int pt_write_restart(void* p, const char* dir) {
  Context* c = ctx_of(p);
  if (!c || !dir) return PT_ERR_INVALID;
  return write_restart(*c, dir);
}

int pt_write_snapshot(void* p, const char* dir, int32_t fcount) {
  Context* c = ctx_of(p);
  if (!c || !dir) return PT_ERR_INVALID;
  return write_snapshot(*c, dir, fcount);
}

int pt_read_restart(void* p, const char* dir, double Tbot_iso, double Ttop_iso) {
  Context* c = ctx_of(p);
  if (!c || !dir) return PT_ERR_INVALID;
  return read_restart(*c, dir, Tbot_iso, Ttop_iso);
}

int pt_synchronize(void* p) {
  Context* c = ctx_of(p);
  if (!c) return PT_ERR_INVALID;
  PT_HIP(hipStreamSynchronize(c->stream));
  return PT_OK;
}

int pt_enable_kernel_timing(void* p, int32_t on) {
  Context* c = ctx_of(p);
  if (!c) return PT_ERR_INVALID;
  PT_HIP(hipStreamSynchronize(c->stream));
  for (auto& t : c->kt) { timer_flush(t); t.ms = 0.0; t.launches = 0; t.work = 0; }
  c->timing = on != 0;
  return PT_OK;
}

int pt_kernel_time(void* p, int32_t which, double* ms, int64_t* launches) {
  Context* c = ctx_of(p);
  if (!c || which < 0 || which >= PT_K_COUNT) return PT_ERR_INVALID;
  PT_HIP(hipStreamSynchronize(c->stream));
  timer_flush(c->kt[which]);
  if (ms) *ms = c->kt[which].ms;
  if (launches) *launches = c->kt[which].launches;
  return PT_OK;
}

int pt_set_guess_net(void* p, void* net, int32_t guess, int32_t capture) {
  Context* c = ctx_of(p);
  if (!c || (guess != PT_GUESS_WARM && guess != PT_GUESS_LEARNED)) return PT_ERR_INVALID;
  if (c->P.iterate != 1) {
    set_error("the learned guess starts the multigrid solve (iterate=1)");
    return PT_ERR_INVALID;
  }
  if (!net && (guess == PT_GUESS_LEARNED || capture)) {
    set_error("a network is needed for the learned guess or for capture");
    return PT_ERR_INVALID;
  }
  c->guess_net = net;
  c->guess_on = guess == PT_GUESS_LEARNED;
  c->capture = capture != 0;
  c->P.guess = guess;
  return PT_OK;
}

int pt_projection_counts(void* p, int64_t* windowed, int64_t* radix) {
  Context* c = ctx_of(p);
  if (!c || !windowed || !radix) return PT_ERR_INVALID;
  *windowed = c->proj_window;
  *radix = c->proj_fallback;
  return PT_OK;
}

int pt_equalize_counts(void* p, int64_t* nequalize, int64_t* moved_last, int64_t* moved_total) {
  Context* c = ctx_of(p);
  if (!c || !nequalize || !moved_last || !moved_total) return PT_ERR_INVALID;
  *nequalize = c->nequalize;
  *moved_last = c->eq_moved_last;
  *moved_total = c->eq_moved_total;
  return PT_OK;
}

int pt_get_rng_state(void* p, int32_t* state) {
  Context* c = ctx_of(p);
  if (!c || !state) return PT_ERR_INVALID;
  for (int q = 0; q < 4; q++) state[q] = c->rng[q];
  return PT_OK;
}

int pt_set_rng_state(void* p, const int32_t* state) {
  Context* c = ctx_of(p);
  if (!c || !state) return PT_ERR_INVALID;
  for (int q = 0; q < 3; q++)
    if (state[q] < 0 || state[q] >= 2147483579) {
      // every state the reference's seeds lead to keeps seed_i, seed_j,
      // seed_k in [0, 2147483579) (PS/basics.f90:321-331, composition.f90:12-16)
      set_error("mzran state: seed_i, seed_j, seed_k must lie in [0, 2147483579)");
      return PT_ERR_INVALID;
    }
  for (int q = 0; q < 4; q++) c->rng[q] = state[q];
  return PT_OK;
}

int pt_guess_counts(void* p, int64_t* tried, int64_t* used) {
  Context* c = ctx_of(p);
  if (!c || !tried || !used) return PT_ERR_INVALID;
  *tried = c->guess_tried;
  *used = c->guess_used;
  return PT_OK;
}

int pt_guess_last(void* p, double* l1_warm, double* l1_guess, double* alpha) {
  Context* c = ctx_of(p);
  if (!c || !l1_warm || !l1_guess || !alpha) return PT_ERR_INVALID;
  *l1_warm = c->lg_l1_warm;
  *l1_guess = c->lg_l1_guess;
  *alpha = c->lg_alpha;
  return PT_OK;
}

int pt_kernel_work(void* p, int32_t which, int64_t* units) {
  Context* c = ctx_of(p);
  if (!c || !units || which < 0 || which >= PT_K_COUNT) return PT_ERR_INVALID;
  *units = c->kt[which].work;
  return PT_OK;
}

}  // extern "C"
