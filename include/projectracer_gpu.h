/*
 * projectracer_gpu.h -- C ABI of the MI355X (gfx950) implementation of
 * ProjecTracer's per-timestep hot path.
 *
 * The reference program is Fortran 90: its hot-path subroutines take no
 * arguments and work on module globals (`use basics; use arrays`).  This
 * ABI keeps that shape: one opaque context owns the device mirrors of the
 * module arrays and is driven by argument-free calls that stand in for the
 * reference subroutines.  Everything crossing the boundary is a plain C
 * type (int32/int64/double and pointers to host memory laid out exactly as
 * the Fortran arrays, column-major with the reference's halo widths), so a
 * Fortran `interface ... bind(C)` block binds it directly (the module in
 * projectracer_amd/fortran/projectracer_gpu.f90 is that block).
 *
 * Every entry point returns PT_OK (0) or one of the PT_ERR_* codes below.
 * The reference has no return codes: it prints a message and calls `stop`
 * (the cited lines); the Fortran shim reproduces those messages from the
 * code.
 *
 * Supported subset (what the hot path and its configurations use):
 * uniform grid (gridX=gridZ=0), order=4, Database=2, iterate=1, MG=1,
 * RKC1=1 or 0 (FP64), splitting=0, non_Newtonian=0, plastic=0, comp=1, ntype=1,
 * iC_continuous=1 or 0, shape_function=1, tracer=1, torder=4, RK_update=1,
 * vel_update=1, local_time=0, Chebyshev=0, CFL=0, equalize=0 (or 1 with
 * original=1 and iC_continuous=0), iheattype=0.  pt_create rejects anything
 * else with PT_ERR_UNSUPPORTED.
 */
#ifndef PROJECTRACER_GPU_H
#define PROJECTRACER_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define PT_OK                   0
#define PT_ERR_TRACER_OUTSIDE   1  /* PS/composition.f90:1007-1010 "Tracer advected outside of domain" */
#define PT_ERR_MAX_CYCLES       2  /* PS/iterative_solver.f90:325-334 "Too Many Multigrid Cycles" */
#define PT_ERR_HIGH_RESIDUAL    3  /* PS/iterative_solver.f90:438-447 "Abnormally High Residual" */
#define PT_ERR_SINGULAR_COARSE  4  /* DGBTRF INFO>0 at PS/iterative_solver.f90:153 (ignored by the reference) */
#define PT_ERR_HIP              5  /* HIP runtime error (no reference counterpart) */
#define PT_ERR_INVALID          6  /* bad argument */
#define PT_ERR_UNSUPPORTED      7  /* option outside the supported subset */
#define PT_ERR_NO_DEVICE        8  /* no gfx950 device visible */

/* ---- solver precision / initial-guess modes ---------------------------- */
#define PT_PREC_FP64   0  /* plain path: every operation in FP64, as the reference */
#define PT_PREC_MIXED  1  /* FP32 smoothing/restriction/prolongation inside an FP64
                             defect-correction loop; FP64 residual, coarse LU and psi */

#define PT_GUESS_WARM    0  /* warm start from the previous psi (PS/iterative_solver.f90:134) */
#define PT_GUESS_LEARNED 1  /* learned initial guess for psi (no reference counterpart); takes
                               effect once a network is attached (pt_set_guess_net) */

/* ---- namelist values (PS/basics.f90:54-66, read at PS/main.f90:11-21) --- */
typedef struct pt_params {
  /* &inputs: grid */
  double aspect;          /* aspect ratio lambda */
  int32_t nx, nz;         /* cells; rewritten by optimize_grid (PS/iterative_solver.f90:1-61) */
  int32_t grids;          /* max multigrid levels (2..6); rewritten by optimize_grid */
  int32_t mingrid;        /* min coarsest-grid size */
  /* &inputs: velocity solver */
  double  tolerance;      /* L1-mean relative residual target */
  int32_t max_cycles;
  int32_t Nmulti;         /* smoother iterations per level visit */
  double  eig_ratio;      /* kept for completeness; unused by the supported path */
  int32_t vis_iter;       /* viscosity smoother iterations before restriction */
  int32_t nstageS;        /* RKC1 stages */
  double  dampingS;       /* RKC1 damping */
  double  c_2, c_1;       /* RKC1 Courant range */
  /* &inputs: time stepping */
  double  courant, courantT, courantC, courant_stream;
  /* &inputs: tracers */
  int32_t tpc;            /* tracers per grid node */
  /* &inputs: boundary conditions */
  int32_t Tbot;           /* 0 isothermal, 1 insulating bottom */
  int32_t Vbc;            /* 0 free-slip, 1 rigid top/bottom */
  /* &Rayleigh_numbers */
  double  RaT, RaC;
  /* &heating_conductivity (iheattype=0) */
  double  H, conduct_factor, Htype;
  /* &viscosity */
  int32_t ivis;           /* 0 isoviscous, 1 T/z dependent, 2 T/C/z dependent */
  double  visP, visT, delta_visT, eig_ratio_min, visC, delta_visC;
  /* path selection (no reference counterpart) */
  int32_t precision;      /* PT_PREC_* */
  int32_t guess;          /* PT_GUESS_* */
  /* &inputs: diagnostics */
  double  dent;           /* reference height of the entrainment integral */
  /* &inputs: stream-function solver */
  int32_t iterate;        /* 0 direct banded LU (PS/momentum.f90:172-653), 1 multigrid */
  int32_t Database;       /* direct solver unknown order (PS/momentum.f90:238-253):
                             1 p = (nx+1) k + i + 1, 2 p = (nz+1) i + k + 1 */
  /* &inputs: tracer composition and redistribution */
  int32_t iC_continuous;  /* 1: continuous C carried by the tracers (Ctr); 0: discrete
                             types Ttype in {0,1} (PS/composition.f90:1090-1276),
                             exchanged as Ctr = 0.0 / 1.0 */
  int32_t equalize;       /* 1: redistribute tracers at the start of every energy_time
                             (PS/energy.f90:198-204); needs iC_continuous = 0 */
  int32_t original;       /* 1: equalize_tracers_original (PS/composition.f90:84-263),
                             the only redistribution built */
  int32_t buffer;         /* half-width of the single-type neighbourhood test (10) */
  double  dist_factor;    /* wall distance of redistributed tracers, in cells (1e-4) */
  /* &inputs: fine grid smoother */
  int32_t RKC1;           /* 1: RKC1 sweeps of nstageS stages; 0: RK1, one explicit step
                             per iteration with local Chebyshev-ordered steps
                             (PS/iterative_solver.f90:1468-1485, :1503-1513); RK1 is
                             built for precision = PT_PREC_FP64 */
} pt_params;

/* Fill *p with the reference defaults used by this build (inputs.nml values
 * for the supported path; tolerance 1e-8, 6 levels, RKC1 35 stages,
 * dent 0.5). */
void pt_default_params(pt_params* p);

/* ---- field identifiers for pt_get_field / pt_set_field ----------------- */
/* Each field is exchanged with its Fortran bounds (column-major, i fastest):
 *   PT_FIELD_SF, PT_FIELD_T(ratio), PT_FIELD_TVIS, PT_FIELD_TBUOY: (-3:nx+3,-3:nz+3)
 *   PT_FIELD_U, PT_FIELD_W: (-3:nx+3,-3:nz+3)   (span_interp halo)
 *   PT_FIELD_CNEW, PT_FIELD_CVIS, PT_FIELD_CBUOY, PT_FIELD_CONDUCT: (-5:nx+5,-5:nz+5)
 *   PT_FIELD_VIS: level-1 vis_grid (-3:nx+3,-3:nz+3)
 *   PT_FIELD_RESIDUAL: level-1 residual (1:nx-1,1:nz-1)
 *   PT_FIELD_SOURCE: tracer_space_array (-3:nx+3,-3:nz+3)
 * pt_field_extent() returns the element count so callers can size buffers. */
enum {
  PT_FIELD_SF = 0, PT_FIELD_T = 1, PT_FIELD_TVIS = 2, PT_FIELD_TBUOY = 3,
  PT_FIELD_U = 4, PT_FIELD_W = 5, PT_FIELD_CNEW = 6, PT_FIELD_CVIS = 7,
  PT_FIELD_CBUOY = 8, PT_FIELD_CONDUCT = 9, PT_FIELD_VIS = 10,
  PT_FIELD_RESIDUAL = 11, PT_FIELD_SOURCE = 12
};

/* ---- per-solve / per-step counters ------------------------------------ */
typedef struct pt_solve_info {
  int32_t q;              /* cycles of the last solve (PS/iterative_solver.f90:161-335) */
  double  residual;       /* residual_mag(1) of the last solve */
  double  res_ref;        /* res_ref(1) of the last solve */
  double  dt;             /* last time step (PS/energy.f90:756-810) */
  double  time;           /* model time */
  int64_t solves;         /* solves since pt_create */
  int64_t cycles;         /* multigrid cycles since pt_create */
  int32_t n_eta;          /* viscosity filter passes of the last level-1 smoother_vis */
  int32_t n_temp;         /* T filter passes of the last projection */
  int32_t n_comp;         /* C filter passes of the last projection */
  double  eig_fine;       /* ratio_min of the last level-1 smoother_vis */
  int32_t cycle_graphs;   /* multigrid cycles captured as HIP graphs and replayed
                             (one per starting buffer arrangement; 0: launched
                             one kernel at a time) */
} pt_solve_info;

/* ---- per-step diagnostics (PS/basics.f90:363-462) ---------------------
 * The values `stats` writes to stats.dat (comp=1, tracer=1 row, format at
 * PS/basics.f90:421-422) and viscosity_smoother.dat (:453-455), computed on
 * the device from the current state.  The first call after pt_load_state
 * plays the reference's tstep=0 call: it fixes `mass` (the initial C
 * integral) that entrainment is normalised by. */
typedef struct pt_stats_row {
  double  time, Tavg, vrms, Nu, Nu_bot, RaT_avg, mobility, strain_max, q1, q2, q3, q4;
  int32_t empty_cells, tracer_min, tracer_max;
  double  Cavg, ent;
  double  vis_min, vis_max;   /* viscosity_smoother.dat */
  int32_t n_eta, n_temp, n_comp;
  double  eig_fine;
} pt_stats_row;

/* ---- lifecycle ---------------------------------------------------------
 * pt_create: compute_smoothing_parameters, multigrid_setup (optimize_grid),
 *   compute_basics, storage, Chebyshev_prep_stream, FD_coefficients,
 *   build_grid, derivative_eigenvalues, stability_RK4_smoother,
 *   eigenvalue_prep, stream_space_matrix_prep, stability_vis_smoother
 *   (PS/main.f90:168-203).  *p is copied; nx/nz/grids are rewritten in the
 *   context as optimize_grid does (read them back with pt_get_dims). */
int pt_create(const pt_params* p, void** ctx);
int pt_destroy(void* ctx);
int pt_get_dims(void* ctx, int32_t* nx, int32_t* nz, int32_t* grids, int64_t* ntr);

/* pt_load_state: the tail of initial_temperature + initialize_tracers
 * (PS/basics.f90:352-359, PS/composition.f90:848) and PS/main.f90:228-229:
 * T node values (0:nx,0:nz), the isothermal boundary values, and tracer
 * positions/values in ID order (ntr each); projects the tracers
 * (tracers_to_corners) and zeroes psi and the coarse-level corrections.
 * With discrete types (iC_continuous=0) Ctr carries Ttype: every value must
 * be 0.0 or 1.0 (PT_ERR_INVALID otherwise). */
int pt_load_state(void* ctx, const double* T_nodes, double Tbot_iso, double Ttop_iso,
                  const double* rtr, const double* str, const double* Ttr,
                  const double* Ctr);

/* ---- hot path: drop-in replacements for the reference subroutines ------ */
int pt_multigrid(void* ctx);            /* PS/iterative_solver.f90:124-341 (iterate=1) */
int pt_compute_velocities(void* ctx);   /* PS/momentum.f90:655-710 + velocityBCs */
/* PS/momentum.f90:173-189: the direct solve when the context was created with
 * iterate=0 (building the matrix on its first call, and on every call for
 * ivis >= 1, as PS/main.f90:247-251 does), else multigrid; then velocities */
int pt_update_velocities(void* ctx);
/* Direct solver (iterate=0 contexts), PS/momentum.f90:319-653:
 * build_SF_matrix assembles the whole-grid banded operator (Database 1 or 2
 * ordering), scales its rows by DGBEQU's factors and factors it (band LU with
 * partial pivoting, DGBTF2's column order); solve_SF_equation forms the
 * buoyancy vector from Tbuoy/Cbuoy, runs DGBTRS and writes psi with its ghost
 * images (Stream_grid).  pt_solve_sf_equation before any build returns
 * PT_ERR_INVALID; a zero pivot returns PT_ERR_SINGULAR_COARSE. */
int pt_build_sf_matrix(void* ctx);      /* PS/momentum.f90:321-488 */
int pt_solve_sf_equation(void* ctx);    /* PS/momentum.f90:621-653 */
int pt_energy_time(void* ctx);          /* PS/energy.f90:181-556 (RK4, 3 solves, 4 projections) */
int pt_tracers_to_corners(void* ctx);   /* PS/composition.f90:923-1087 on (rtr,str,Ttr) */
/* One iteration of the time loop at PS/main.f90:245-299 (without stats and
 * snapshots): multigrid, compute_velocities, energy_time, time += dt. */
int pt_time_step(void* ctx);
/* Run n time steps back to back (device-resident; one host sync at the end). */
int pt_run_steps(void* ctx, int32_t n);

/* ---- host <-> device exchange (when a host consumer needs the data) ----- */
int64_t pt_field_extent(void* ctx, int32_t field_id);
int pt_get_field(void* ctx, int32_t field_id, double* host);
int pt_set_field(void* ctx, int32_t field_id, const double* host);
/* Tracers in reference ID order (the device stores them in ID order; the
 * projection works through a cell-sorted permutation it keeps separately). */
int pt_get_tracers(void* ctx, double* rtr, double* str, double* Ttr, double* Ctr);
/* nempty(0:nx,0:nz): tracers binned to each node's dual cell (INT32). */
int pt_get_nempty(void* ctx, int32_t* nempty);
int pt_get_info(void* ctx, pt_solve_info* info);
/* stats (PS/basics.f90:363-462) of the current state; see pt_stats_row.
 * Its partial-sum buffer is allocated on the first call. */
int pt_stats(void* ctx, pt_stats_row* row);
/* pt_stats without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same kernels
 * and results as pt_stats; the partial sums live in a file-scope device
 * array of the library (one per GPU, 147,712 bytes, held until the process
 * ends), so grids are limited to nz <= 2047 (PT_ERR_UNSUPPORTED beyond).
 * All calls on one GPU share that array: calls that overlap on one GPU
 * (several host threads, several streams) race. */
int pt_stats_noalloc(void* ctx, pt_stats_row* row);

/* ---- checkpoint / resume in the reference's files --------------------
 * pt_write_restart: print_restart (PS/main.f90:457-496) into directory dir:
 *   Trestart        xg(-6:nx+6), zg(-6:nz+6), T(-3:nx+3,-3:nz+3),
 *                   Cnew(1,-5:nx+5,-5:nz+5), SF(-3:nx+3,-3:nz+3)
 *   restart_tracers rtr, str, Ctr, Ttr in tracer ID order (discrete types,
 *                   iC_continuous=0: rtr, str, Ttype as INTEGER*4, Ttr)
 * each one Fortran unformatted sequential record (4-byte length before and
 * after), the files of a RaT /= 0, comp=1, tracer=1 run.
 * T is the module T (the initial field before the first energy_time).
 * pt_read_restart: the restart=1 start (PS/grid.f90:55-103, PS/main.f90:
 *   204-209): reads both files, re-projects the tracers onto T and Cnew
 *   (file values kept at nodes without tracers), keeps SF as the initial
 *   guess, resets time to 0.  The walls take Tbot_iso/Ttop_iso from the
 *   caller: the reference leaves them at 0 on restart.
 * PT_ERR_INVALID when a file is missing, has another record length than
 * this context's grid and tracer count, or another grid.  Records above
 * 2^31-1 bytes are PT_ERR_UNSUPPORTED. */
int pt_write_restart(void* ctx, const char* dir);
/* snapshots (PS/main.f90:377-424) into directory dir: T<10000+fcount> with,
 * per node (i outer, k inner), xg, zg, T, SF, u, w, 2*vis*strain, vis, Tvis
 * (g14.5E3), nempty (i6), Cnew, Cvis (g14.5E3) and a blank record after each
 * i; P<10000+fcount> with zg and the lateral means of T, strain and
 * 2*vis*strain per row.  strain is compute_strain_rate_invariant_direct
 * (PS/momentum.f90:73-93) of the current psi.  fcount 0..89999. */
int pt_write_snapshot(void* ctx, const char* dir, int32_t fcount);
int pt_read_restart(void* ctx, const char* dir, double Tbot_iso, double Ttop_iso);
int pt_synchronize(void* ctx);

/* ---- timing hooks (no reference counterpart; PS/basics.f90:27-31 hold
 * the reference's omp_get_wtime accumulators) -------------------------- */
/* Enable per-kernel HIP-event timing of the dominant kernels on the
 * context's stream.  While enabled, pt_kernel_time reports the accumulated
 * device time (ms) and launch count of kernel `which` (PT_K_*). */
/*   PT_K_RKC_STAGE_L1     level-1 RKC1 stage launches of the solver
 *   PT_K_TRACER_STAGE     the four RK4 tracer stage kernels of energy_time
 *   PT_K_PROJECT          one projection (binning, sort, sums, node combine)
 *   PT_K_TRACER_RK2       the second RK4 tracer stage kernel alone
 *   PT_K_RKC_STAGE_COARSE RKC1 stage launches on levels >= 2 (the levels
 *                         swept one launch per stage)
 *   PT_K_FILTER           one filter loop of a projection (smoother_time or
 *                         smoother_time_T, its batches and waits included)
 *   PT_K_SOLVE            one multigrid call (set-up and cycles)
 *   PT_K_EQUALIZE         one call of the tracer redistribution (equalize=1):
 *                         the counts, and when it redistributes the donors,
 *                         lottery, T interpolation and closing projection,
 *                         host waits included */
enum { PT_K_RKC_STAGE_L1 = 0, PT_K_TRACER_STAGE = 1, PT_K_PROJECT = 2, PT_K_TRACER_RK2 = 3,
       PT_K_RKC_STAGE_COARSE = 4, PT_K_FILTER = 5, PT_K_SOLVE = 6, PT_K_EQUALIZE = 7, PT_K_COUNT = 8 };
int pt_enable_kernel_timing(void* ctx, int32_t on);
int pt_kernel_time(void* ctx, int32_t which, double* ms, int64_t* launches);
/* Units of work of the timed launches of `which`: interior grid points for
 * the stage launches (PT_K_RKC_STAGE_*), 0 for the others. */
int pt_kernel_work(void* ctx, int32_t which, int64_t* units);
/* projections whose stable (cell, slot) sort was the windowed counting sort
 * (tracers stored in an earlier projection's cell order, none more than one
 * cell away) and those that took the radix sort (no such order yet, a
 * farther move or one in the last 8 projections, or PT_PROJECT_RADIX=1 at
 * pt_create); both give the same order, element for element */
int pt_projection_counts(void* ctx, int64_t* windowed, int64_t* radix);

/* ---- tracer redistribution (equalize=1, original=1, iC_continuous=0) -----
 * equalize_tracers_original (PS/composition.f90:84-263) runs on the device
 * at the start of every energy_time.  nequalize: calls that redistributed
 * (the reference's counter, :151-155); moved_last / moved_total: tracers
 * moved by the last such call and by all of them. */
int pt_equalize_counts(void* ctx, int64_t* nequalize, int64_t* moved_last, int64_t* moved_total);
/* mzran stream 0 (PS/composition.f90:7-20), the stream the redistribution's
 * lottery draws from: state = seed_i(0), seed_j(0), seed_k(0), seed_n(0).
 * pt_create starts it at PS/basics.f90:321-324's values; a host that drew
 * the initial cloud from stream 0 (one OpenMP thread) sets the state it
 * left.  seed_i, seed_j and seed_k lie in [0, 2147483579) in every state the
 * reference's seeds lead to; pt_set_rng_state rejects others (PT_ERR_INVALID). */
int pt_get_rng_state(void* ctx, int32_t* state);
int pt_set_rng_state(void* ctx, const int32_t* state);

/* ---- learned initial guess (no reference counterpart) -------------------
 * A four-level convolutional encoder-decoder (3x3 convolutions, channels
 * 32-64-128-256, skip connections; BF16 weights on MFMA, FP32 accumulation)
 * that proposes a correction to the warm start of a solve.  The reference
 * warm-starts every solve from the previous psi (PS/iterative_solver.f90:134);
 * with guess = PT_GUESS_LEARNED and a network attached (pt_set_guess_net) the
 * solve evaluates the network on the level-1 fields and keeps its proposal
 * only when the proposal's FP64 L1 residual is below the warm start's; the
 * cycles then run to the same stop test either way.
 * Network input, one pixel per node (i, k), 0 <= i < nx, 0 <= k < nz
 * (nx, nz multiples of 8): Tbuoy, Cbuoy, ln(eta), psi / rms(psi) and the warm
 * start's residual r / rms(r); output: the correction in units of
 * rms(r) * out_scale. */
int pt_net_create(int32_t seed, void** net);   /* He-initialised weights from `seed` */
int pt_net_destroy(void* net);
int pt_net_num_params(void* net, int64_t* count);
int pt_net_get_params(void* net, float* host);          /* FP32 master weights, flat */
int pt_net_set_params(void* net, const float* host);
/* forward pass on host inputs x (N, H, W, 5) FP32 -> y (N, H, W) FP32 */
int pt_net_forward(void* net, int32_t N, int32_t H, int32_t W, const float* x, float* y);
/* masked mean-square loss of the forward pass against t under mask (both
 * (N, H, W)), and its gradient with respect to the flat parameters */
int pt_net_loss_grad(void* net, int32_t N, int32_t H, int32_t W, const float* x, const float* t, const float* mask,
                     double* loss, float* grad);
/* one AdamW step (beta1 0.9, beta2 0.999, eps 1e-8) with the last gradient */
int pt_net_adamw_step(void* net, double lr, double weight_decay);
/* Training set on the device: `capacity` (inputs, correction) pairs of
 * nz x nx solves (H = nz, W = nx); a new call empties it. */
int pt_net_dataset(void* net, int32_t H, int32_t W, int32_t capacity);
int pt_net_dataset_size(void* net, int32_t* count);
/* stored pair `slot`: inputs x (H, W, 5) and target t (H, W) = the solve's
 * correction / rms(residual) */
int pt_net_dataset_get(void* net, int32_t slot, float* x, float* t);
/* `iters` AdamW steps on random batches of the stored pairs (targets
 * divided by out_scale, the pairs' rms correction, fixed at the first call);
 * *loss: the masked mean-square loss of the last batch */
int pt_net_train(void* net, int32_t iters, int32_t batch, double lr, double weight_decay, uint64_t seed,
                 double* loss);
int pt_net_out_scale(void* net, double* scale);
/* weight lam of the loss's curvature term: loss = (sum m e^2 + lam sum
 * (m lap e)^2) / sum m with e = y - t and lap the 5-point Laplacian on each
 * image (default 0) */
int pt_net_set_loss(void* net, double lap_weight);
/* Attach a network to a context (null detaches): guess = PT_GUESS_LEARNED
 * makes every solve propose the network's guess (kept only when its FP64
 * L1 residual is below the warm start's); capture != 0 adds every solve's
 * (inputs, correction) pair to the network's training set while it has
 * room.  The network must outlive the attachment. */
int pt_set_guess_net(void* ctx, void* net, int32_t guess, int32_t capture);
/* solves that evaluated the network, and those that kept its proposal */
int pt_guess_counts(void* ctx, int64_t* tried, int64_t* used);
/* the last proposal: the warm start's and the proposal's L1 residual sums,
 * and the step along the network's (smoothed) correction that minimises the
 * proposal's L2 residual */
int pt_guess_last(void* ctx, double* l1_warm, double* l1_guess, double* alpha);

/* ---- ensembles of independent cases (no reference counterpart: the
 * reference runs one case per process, PS/main.f90:245-300) --------------
 * n cases, each a full context with its own stream (params: an array of n,
 * grids and parameters free per case).  pt_ensemble_run_steps advances
 * every case by nsteps time steps (pt_time_step), the cases handed to a pool
 * of `workers` host threads so that different cases' kernels overlap on the
 * GPU; workers <= 0 tunes the pool size on the first steps (1, 2, 4, 8, 16
 * threads, one step each) and keeps the fastest.  A case's results do not
 * depend on the pool.  status[i] (optional, n entries) receives each case's
 * code: a case that fails stops there while the others go on.  The cases'
 * contexts (pt_ensemble_case) take every other call of this header. */
int pt_ensemble_create(int32_t n, const pt_params* params, void** ens);
int pt_ensemble_destroy(void* ens);
int pt_ensemble_case(void* ens, int32_t i, void** ctx);
int pt_ensemble_run_steps(void* ens, int32_t nsteps, int32_t workers, int32_t* status);
int pt_ensemble_workers(void* ens, int32_t* workers);   /* pool size of the last run */

/* Last error message (thread-local, never NULL). */
const char* pt_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* PROJECTRACER_GPU_H */
