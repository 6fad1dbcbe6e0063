// pt_internal.hpp -- device context of the gfx950 hot path.
//
// Layout in HBM (all FP64 unless noted; every 2-D field is column-major,
// i fastest, exactly the Fortran array with its halo so exchanges with the
// host are plain copies):
//   node fields, halo 3 : pitch nx+7  (SF, Tratio, T0, Tvis, Tbuoy, u, w, source,
//                                       conduct_r, conduct_s)
//   node fields, halo 5 : pitch nx+11 (Cnew, C0, Cvis, Cbuoy, conduct, Textend)
//   multigrid level l   : pitch n_l+7, halo 3 for every per-level array
//                         (error, RKC scratch, vis_* set, RHS, residual, dt_stream)
//   interpolation table : (i, k, field, d) with d = {f, d/dr, d2/dr2, d3/dr3},
//                         k fastest so the 7 rows a tracer reads are contiguous
//   tracers             : SoA, ID order for r/s/T/C and the RK4 registers
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include "../../include/projectracer_gpu.h"

namespace pt {

constexpr int MAXL = 6;
constexpr int H3 = 3;   // span / span_interp
constexpr int H5 = 5;   // spanT

// ---- host-side tables computed once at create (pt_setup.cpp) -------------
struct Tables {
  int nx, nz, grids, nsteps, multi[52];
  double aspect, dr, ds, dr2, ds2, dr3, ds3, dr4, ds4;
  double factorial[5], dr_power[5], ds_power[5];
  int nxg[MAXL + 1], nzg[MAXL + 1];
  double drg[MAXL + 1], dsg[MAXL + 1], dr2g[MAXL + 1], ds2g[MAXL + 1], dr3g[MAXL + 1], ds3g[MAXL + 1],
      dr4g[MAXL + 1], ds4g[MAXL + 1], drdsg[MAXL + 1], drds2g[MAXL + 1], dr2dsg[MAXL + 1], dr2ds2g[MAXL + 1];
  double D1[7], D2[7], D3[7], D4[7];
  double Drs[7][7], Drss[7][7], Drrs[7][7], Drrss[7][7];   // [k+3][i+3]
  double D[5][7];                                          // D(m,j) = D[m][j+3]
  double Tr_base, Trr_base, Trrr_base, Trrrr_base, Trs_base, Trrs_base, Trrss_base, Tr_base_so, Trr_base_so;
  double SR1[MAXL + 1], SR2[MAXL + 1], SR4[MAXL + 1], SR5[MAXL + 1], SR6[MAXL + 1];
  double SI1[MAXL + 1], SI2[MAXL + 1], SI4[MAXL + 1], SI5[MAXL + 1], SI6[MAXL + 1];
  double SM1[7][7], SM2[7][7], SM4[7][7], SM5[7][7], SM6[7][7];   // [dk+3][di+3]
  double mewt[64], mew[64], vv[64], cc[64], RKC1_stability;
  // the mixed path's coarse levels with s_coarse stages (0: nstageS, the tables above)
  double mewt_c[64], mew_c[64], vv_c[64], cc_c[64], RKC1_stability_c;
  int s_coarse = 0;
  double delta_T, delta_C;
  double dts;                    // dts_array value (uniform grid => constant)
  double dt_vis[MAXL + 1];       // dt_vis per level (constant per level)
  int kl, ku, ngc;
  int dn, dkl, dku;               // direct solver (iterate=0): unknowns, band widths
  int64_t ntr;
};
int build_tables(const pt_params& p, Tables& t);   // returns PT_* status
void rkc1_coarse_tables(Tables& t, int s, double damping);
int cycle_schedule_kind(int grids, char kind, int* multi);   // 'W', 'V' or 'F'; returns nsteps

// ---- stencil constants passed to kernels by value ----------------------
struct Stencil {
  double D1[5], D2[5];           // offsets -2..2
  double D3[7], D4[7];           // offsets -3..3
  double Drs[5][5], Drss[5][5], Drrs[5][5], Drrss[5][5];   // [b+2][a+2]
};

struct LevelGeom {
  int n, m, pitch;               // n_l, m_l, pitch = n_l + 7
  double dr, ds, dr2, ds2, dr3, ds3, dr4, ds4, drds, drds2, dr2ds, dr2ds2;
};

// viscosity field set of one level (primary or scratch)
struct VisSet {
  double *vis, *vx, *vz, *vxx, *vzz, *vxz, *vxxpzz, *vxxmzz, *vsmooth;
};

struct Level {
  LevelGeom g;
  double *err, *bufB, *bufC;      // psi correction and RKC rotating buffers
  double *rhs, *res, *dts, *ratio;
  double *vis_static;
  VisSet v, vs;                   // primary / scratch viscosity sets
  // FP32 mirrors for the mixed-precision smoother
  float *ferr, *fB, *fC, *fD, *fE, *frhs, *fdts, *fvis, *fvxm, *fvxz, *fvx, *fvz;
  size_t nelem;                   // (n+7)*(m+7)
};

struct KTimer {
  hipEvent_t beg[1024], end[1024];
  int n = 0;
  double ms = 0.0;
  int64_t launches = 0;
  int64_t work = 0;   // units of work of the timed launches (interior points of stage launches)
};

struct Context {
  pt_params P;
  Tables T;
  Stencil st;
  hipStream_t stream = nullptr;
  int device = 0;
  Level L[MAXL + 1];
  // node fields
  double *SF, *Tratio, *T0, *Tvis, *Tbuoy, *u, *w, *src, *cond_r, *cond_s;
  double* dtab;                   // D(1:3,-3:3) interpolation weights on the device
  double *Cnew, *C0, *Cvis, *Cbuoy, *cond, *Textend;
  double* itab;                   // interpolation table (u, w, source)
  double* ctab;                   // per-node interpolation polynomial coefficients
  // coarse grid: band LU (DGBTF2) and the dense operator G = A^-1 diag(RowB)
  // restricted to interior unknowns, built from the LU once per factorization
  double *Mc, *RowB, *Bc; int32_t* ipiv; int32_t* d_info;
  double* Gc; int nci;
  // direct solver (iterate=0): equilibrated band LU of the whole grid, row
  // factors, right-hand side / solution vector
  double *dAB = nullptr, *dRow = nullptr, *dB = nullptr; int32_t* dipiv = nullptr;
  int32_t *dju = nullptr, *d_lustate = nullptr;   // LU panels: ju per step; carried ju, info
  bool direct_built = false;
  // tracers, SoA in cell order (re-sorted after every projection of the step's
  // final positions); tid maps a storage slot to the tracer's reference ID
  double *r, *s, *Tt, *Ct, *r0, *s0, *T0t, *r1, *s1, *T1t, *r2, *s2, *T2t;
  int32_t *tid, *tid_tmp;
  // projection work: cell keys, stable (cell, index) sort, per-cell partial sums
  int32_t *cellkey, *skey, *iota, *perm, *cstart, *cend, *nempty;
  void* sort_tmp; size_t sort_bytes; int sort_bits;
  // the storage order's cells: okey per slot and each cell's slot range
  // (ostart, oend), valid once the tracers have been reordered; ccount /
  // bsum: per-cell counts and per-block totals of the windowed counting sort
  int32_t *okey = nullptr, *ostart = nullptr, *oend = nullptr, *ccount = nullptr, *bsum = nullptr;
  bool order_valid = false;
  bool project_radix_only = false;              // PT_PROJECT_RADIX=1 at pt_create
  int window_skip = 0;                          // projections left before the windowed sort is tried again
  int64_t proj_window = 0, proj_fallback = 0;   // projections by windowed counting sort / by radix sort
  // tracer redistribution (equalize_tracers_original): mzran stream 0
  // (seed_i, seed_j, seed_k, seed_n), counters, and the work buffers
  // allocated on the first call (eq_node: per-node counts, excess and the
  // lottery's cell list; eq_id: per-ID donor class and slot; eq_save: the
  // donor slots in IDsave order; eq_dev: device scalars)
  int32_t rng[4] = {521288629, 362436069, 16163801, 1131199299};
  int64_t nequalize = 0, eq_moved_last = 0, eq_moved_total = 0;
  int32_t *eq_node = nullptr, *eq_id = nullptr, *eq_save = nullptr, *eq_dev = nullptr, *eq_bsum = nullptr;
  double* cellsum;          // 27 per cell: (T*w, C*w, w) for the 3x3 neighbour nodes
  double *TxS, *CxS;        // halo-5 scratch fields for the filters
  int32_t* d_flags;         // [0]=tracer outside domain, [1]=tracer more than one cell from its slot's cell
  // device-driven filter loops: control (done flag, pass count), its pinned
  // host mirror, and the ring of per-check smoothness maxima (FP64 bit
  // patterns); filter_on_device is false when PT_FILTER_HOST=1 at pt_create
  int32_t* d_fctl = nullptr;
  int32_t* h_fctl = nullptr;
  unsigned long long* d_fsm = nullptr;
  bool filter_on_device = true;
  int hint_comp = 0, hint_temp = 0;   // pass counts of the previous C / T filter loops
  // reductions
  double* red;              // block partials
  double* red_out;          // device scalars (pinned mirror below)
  double* h_red;            // pinned host mirror
  size_t red_cap;
  // state
  double Tbot_iso = 0, Ttop_iso = 0, time = 0, dt = 0;
  double res_ref1 = 0, residual1 = 0;
  int last_q = 0; int64_t solves = 0, cycles = 0;
  int n_eta = 0, n_temp = 0, n_comp = 0; double eig_fine = 0;
  bool first_solve = true;
  // the multigrid cycle as captured HIP graphs, one per arrangement of the
  // rotating level buffers it starts from (`before`; it leaves them as
  // `after`): graph_state 0 = not yet run, 1 = run eagerly once, 2 =
  // capturing/replaying, -1 = eager only; use_graph is false when PT_GRAPH=0
  // at pt_create
  struct CycleGraph {
    std::vector<void*> before, after;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  bool use_graph = true;
  int graph_state = 0;
  // cycle schedule the solve follows: the reference's W-cycle in the FP64
  // path; the mixed path may take a cheaper one (V or F, the reference's own
  // alternatives in PS/multigrid_cycle.f90), since it is judged by the same
  // FP64 stop test.  coarse_zero: the mixed path starts each coarse-level
  // correction from zero when the cycle descends to it.
  int cyc[52] = {0};
  int cyc_n = 0;
  bool coarse_zero = false;
  int presmooth_from = 99;   // the mixed path's cycle smooths levels >= this on the way up only (2 by default)
  std::vector<CycleGraph> cycle_graphs;
  // mixed precision work
  double* dres64;           // FP64 residual for defect correction
  // stats: the reference's module T is the initial field until the first
  // energy_time; entrainment is normalised by the first call's C integral
  bool t_initial = true;
  double mass = 0.0; bool mass_set = false;
  double* stats_rows = nullptr;
  double* stats_scratch = nullptr;   // this device's `scratch` array (pt_stats_noalloc)
  double* strain_buf = nullptr;   // strain-rate invariant at the nodes (snapshots)
  // learned initial guess: the attached network (pt_set_guess_net), whether
  // solves propose its guess (guess_on) and/or add (inputs, correction) pairs
  // to its dataset (capture); level-1 work fields; counters of proposals made
  // and kept; the dataset slot and residual scale of the solve in flight
  void* guess_net = nullptr;
  bool guess_on = false, capture = false;
  double *lg_res = nullptr, *lg_res2 = nullptr, *lg_psi = nullptr;
  int64_t guess_tried = 0, guess_used = 0;
  int lg_slot = -1;
  int guess_smooth = 2;   // binomial smoothing passes of the proposed correction
  int guess_judge = 0;    // 0: keep a proposal with a lower L1 residual; 1: step by the tiles' smooth residual
  double lg_sr = 0.0, lg_l1_warm = 0.0, lg_l1_guess = 0.0, lg_alpha = 0.0;
  // timing
  bool timing = false;
  KTimer kt[PT_K_COUNT];
};

// error reporting
void set_error(const std::string& s);
#define PT_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      ::pt::set_error(std::string(#call) + ": " + hipGetErrorString(e_));              \
      return PT_ERR_HIP;                                                               \
    }                                                                                  \
  } while (0)
#define PT_TRY(call)              \
  do {                            \
    int rc_ = (call);             \
    if (rc_ != PT_OK) return rc_; \
  } while (0)

// ---- solver (pt_solver.hip) ------------------------------------------------
int solver_alloc(Context& c);
int solver_free(Context& c);
int multigrid(Context& c);
int compute_velocities(Context& c);
int compute_rhs1(Context& c);
int launch_equilibrate(Context& c, double* ab, double* row, int n, int kl, int ku);
// ---- direct solver (pt_direct.hip) --------------------------------------------
int direct_alloc(Context& c);
int direct_free(Context& c);
int build_sf_matrix(Context& c);
int solve_sf_equation(Context& c);
int solve_direct(Context& c);
// the solve of update_velocities: direct (iterate=0) or multigrid
inline int solve_stream(Context& c) { return c.P.iterate == 0 ? solve_direct(c) : multigrid(c); }
// ---- tracers (pt_tracers.hip) ------------------------------------------------
int tracers_alloc(Context& c);
int tracers_free(Context& c);
int tracers_to_corners(Context& c, const double* r, const double* s, const double* Tt, bool filters = true);
int energy_time(Context& c);
// equalize_tracers_original (pt_equalize.hip); frees its buffers with the context
int equalize_tracers(Context& c);
void equalize_free(Context& c);
// its helpers in the tracer pipeline: storage in the current cell order, Ttr
// at listed slots interpolated from Tratio, the closing projection
int ensure_cell_order(Context& c);
int interpolate_T_at(Context& c, const int32_t* slots, int64_t n);
int equalize_projection(Context& c, bool moved);
int load_state(Context& c, const double* T_nodes, double Tbot, double Ttop, const double* rtr,
               const double* str, const double* Ttr, const double* Ctr);
// tracer state in reference ID order into device buffers (any may be null)
int tracers_in_id_order(Context& c, double** r, double** s, double** Tt, double** Ct);
int restart_state(Context& c, const double* T3, const double* C5, const double* SF3, double Tbot, double Ttop,
                  const double* rtr, const double* str, const double* Ttr, const double* Ctr);
int stream_bcs(Context& c, double* S);
int strain_field(Context& c, double* out);
int write_snapshot(Context& c, const char* dir, int fcount);
int write_restart(Context& c, const char* dir);
int read_restart(Context& c, const char* dir, double Tbot, double Ttop);
// ---- diagnostics (pt_stats.hip) -------------------------------------------------
int stats(Context& c, pt_stats_row* row);
int stats_noalloc(Context& c, pt_stats_row* row);   // no allocation: the file-scope scratch array
int stats_scratch_setup(Context& c);                 // once per device, at pt_create
// ---- sorting (pt_sort.hip) ---------------------------------------------------
int sort_pairs_temp_bytes(int64_t n, int bits, size_t* bytes);
int sort_pairs(void* tmp, size_t bytes, const int32_t* kin, int32_t* kout, const int32_t* vin, int32_t* vout,
               int64_t n, int bits, hipStream_t s);
// ---- reductions (pt_solver.hip) ----------------------------------------------
enum RedOp { RED_SUM_ABS = 0, RED_MAX = 1, RED_MIN = 2, RED_SUM_SQ = 3 };
int level1_residual_of(Context& c, const double* psi, double* res, double* l1);
// ---- learned initial guess (pt_unet.hip) ----------------------------------------
int learned_before_solve(Context& c);   // network proposal and/or dataset capture
int learned_after_solve(Context& c);    // the solve's correction as a training target
int reduce_to_host(Context& c, const double* d, int n, int m, int pitch, int off, RedOp op, double* out);
__global__ void k_reduce_final(const double* partial, int nb, int op, double* out);

// timing helpers
void timer_begin(Context& c, int which, int64_t units = 0);
void timer_end(Context& c, int which);

}  // namespace pt
