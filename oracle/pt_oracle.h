/*
 * pt_oracle.h -- CPU oracle for the ProjecTracer hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a single-threaded plain-C restatement
 * of the reference Fortran (seantrim/ProjecTracer): the oracle the HIP
 * implementation is compared with.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as that oracle or the
 * timed CPU baseline -- never as a product code path.
 *
 * Parameters and field layouts are the ones of include/projectracer_gpu.h
 * so a test can run the same case through both and compare.
 */
#ifndef PT_ORACLE_H
#define PT_ORACLE_H
#include <stdint.h>
#include "../include/projectracer_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pto_case pto_case;

int  pto_create(const pt_params* p, pto_case** out);
void pto_destroy(pto_case* c);
void pto_dims(pto_case* c, int32_t* nx, int32_t* nz, int32_t* grids, int64_t* ntr);
int  pto_load_state(pto_case* c, const double* T_nodes, double Tbot_iso, double Ttop_iso,
                    const double* rtr, const double* str, const double* Ttr, const double* Ctr);
int  pto_multigrid(pto_case* c);
void pto_compute_velocities(pto_case* c);
int  pto_update_velocities(pto_case* c);
int  pto_solve_direct(pto_case* c);
int  pto_direct_matrix(pto_case* c, double* ab, double* row, int32_t* kl, int32_t* ku);
void pto_direct_dims(pto_case* c, int32_t* ngrid, int32_t* kl, int32_t* ku);
int  pto_energy_time(pto_case* c);
int  pto_tracers_to_corners(pto_case* c);
int  pto_time_step(pto_case* c);

int64_t pto_field_extent(pto_case* c, int32_t field_id);
int  pto_get_field(pto_case* c, int32_t field_id, double* out);
int  pto_set_field(pto_case* c, int32_t field_id, const double* in);
void pto_get_tracers(pto_case* c, double* rtr, double* str, double* Ttr, double* Ctr);
void pto_get_nempty(pto_case* c, int32_t* out);
void pto_get_info(pto_case* c, pt_solve_info* info);
/* stats (PS/basics.f90:363-462); the first call after pto_load_state fixes
 * the entrainment normalisation as the reference's tstep=0 call does */
int  pto_stats(pto_case* c, pt_stats_row* row);

/* tracer redistribution (equalize=1): the counters of pt_equalize_counts,
 * and mzran stream 0's state (get and/or set, either may be null) */
void pto_equalize_counts(pto_case* c, int64_t* nequalize, int64_t* moved_last, int64_t* moved_total);
void pto_rng_state(pto_case* c, int32_t* get, const int32_t* set);

/* unit-level hooks for closed-form tests */
/* stencil tables: D1..D4 (-3:3) and the interpolation table D(1:4,-3:3) */
void pto_fd_tables(pto_case* c, double* D1, double* D2, double* D3, double* D4, double* Dtab);
/* RKC1 coefficients mewt(1:s), mew(2:s), v(2:s), cc(2:s) (index 0 unused
 * where the reference array starts at 2), and RKC1_stability */
void pto_rkc_coefficients(pto_case* c, double* mewt, double* mew, double* v, double* cc,
                          double* stability);
/* derivative eigenvalue bases (derivative_eigenvalues): out[0..9] =
 * Tr, Trr, Trrr, Trrrr, Trs, Trrs, Trrss, Tr_so, Trr_so, and dt */
void pto_eigen_bases(pto_case* c, double* out);
/* psi (level numbering as the reference) given as (0:nxl,0:nzl) with
 * boundary values forced to zero and free-slip ghosts applied; writes
 * out(1:nxl-1,1:nzl-1) = L psi (the operator without RHS and sign flip). */
int  pto_apply_operator(pto_case* c, int level, const double* psi_nodes, double* out);
/* interpolate field (PT_FIELD_U, PT_FIELD_W or PT_FIELD_SOURCE, current
 * device-equivalent state) at n points using compute_derivatives+interpolate */
int  pto_interpolate(pto_case* c, int32_t field_id, int64_t n, const double* r,
                     const double* s, double* out);
/* coarse band LU on an arbitrary band matrix (dgbtf2/dgbtrs restatement):
 * ab is (2kl+ku+1) x n column-major, overwritten; b overwritten with x */
int  pto_band_lu_solve(int n, int kl, int ku, double* ab, int32_t* ipiv, double* b);
/* mzran stream 0 (PS/composition.f90:7-20 with seeds PS/basics.f90:321-324) */
void pto_mzran(int64_t n, double* out);
/* initialize_tracers positions as the reference draws them with nthreads
 * OpenMP threads (PS/composition.f90:786-847); ntr = (nx+1)(nz+1)tpc */
void pto_initial_positions(pto_case* c, int nthreads, double* rtr, double* str);
/* node coordinates of a uniform grid with the halos print_restart writes
 * (build_grid, PS/grid.f90:46-116): xg(-6:nx+6), zg(-6:nz+6) */
void pto_grid_coords(int nx, int nz, double aspect, double* xg, double* zg);

#ifdef __cplusplus
}
#endif
#endif
