!! projectracer_gpu.f90 -- ISO_C_BINDING interface to libprojectracer_gpu.so
!! (include/projectracer_gpu.h), the MI355X implementation of ProjecTracer's
!! per-timestep hot path.
!!
!! The reference program keeps its subroutine interfaces; their bodies call
!! the functions below on one module-level context (pt_ctx).  Arrays cross
!! with their Fortran bounds (column-major, halos included): e.g. SF, u, w,
!! Tratio, Tvis, Tbuoy, vis_grid(:,:,1) as (-3:nx+3,-3:nz+3) and Cnew(1,:,:),
!! Cvis(1,:,:), Cbuoy(1,:,:), conduct as (-5:nx+5,-5:nz+5).
!!
!! pt_check reproduces the reference's messages before `stop`:
!!   PT_ERR_TRACER_OUTSIDE  PS/composition.f90:1007-1010, PS/energy.f90 stage loops
!!   PT_ERR_MAX_CYCLES      PS/iterative_solver.f90:325-334
!!   PT_ERR_HIGH_RESIDUAL   PS/iterative_solver.f90:438-447
module projectracer_gpu
 use, intrinsic :: iso_c_binding
 implicit none
 private :: c_strlen

 integer(c_int), parameter :: PT_OK = 0, PT_ERR_TRACER_OUTSIDE = 1, PT_ERR_MAX_CYCLES = 2, &
  PT_ERR_HIGH_RESIDUAL = 3, PT_ERR_SINGULAR_COARSE = 4, PT_ERR_HIP = 5, PT_ERR_INVALID = 6, &
  PT_ERR_UNSUPPORTED = 7, PT_ERR_NO_DEVICE = 8
 integer(c_int32_t), parameter :: PT_PREC_FP64 = 0, PT_PREC_MIXED = 1
 integer(c_int32_t), parameter :: PT_GUESS_WARM = 0, PT_GUESS_LEARNED = 1
 integer(c_int32_t), parameter :: PT_FIELD_SF = 0, PT_FIELD_T = 1, PT_FIELD_TVIS = 2, &
  PT_FIELD_TBUOY = 3, PT_FIELD_U = 4, PT_FIELD_W = 5, PT_FIELD_CNEW = 6, PT_FIELD_CVIS = 7, &
  PT_FIELD_CBUOY = 8, PT_FIELD_CONDUCT = 9, PT_FIELD_VIS = 10, PT_FIELD_RESIDUAL = 11, &
  PT_FIELD_SOURCE = 12
 integer(c_int32_t), parameter :: PT_K_RKC_STAGE_L1 = 0, PT_K_TRACER_STAGE = 1, PT_K_PROJECT = 2, &
  PT_K_TRACER_RK2 = 3, PT_K_RKC_STAGE_COARSE = 4, PT_K_FILTER = 5, PT_K_SOLVE = 6, PT_K_EQUALIZE = 7

 !! namelist values (PS/basics.f90:54-66); layout of struct pt_params
 type, bind(C) :: pt_params
  real(c_double) :: aspect
  integer(c_int32_t) :: nx, nz, grids, mingrid
  real(c_double) :: tolerance
  integer(c_int32_t) :: max_cycles, Nmulti
  real(c_double) :: eig_ratio
  integer(c_int32_t) :: vis_iter, nstageS
  real(c_double) :: dampingS, c_2, c_1
  real(c_double) :: courant, courantT, courantC, courant_stream
  integer(c_int32_t) :: tpc, Tbot, Vbc
  real(c_double) :: RaT, RaC
  real(c_double) :: H, conduct_factor, Htype
  integer(c_int32_t) :: ivis
  real(c_double) :: visP, visT, delta_visT, eig_ratio_min, visC, delta_visC
  integer(c_int32_t) :: precision, guess
  real(c_double) :: dent
  integer(c_int32_t) :: iterate, Database
  integer(c_int32_t) :: iC_continuous, equalize, original, buffer
  real(c_double) :: dist_factor
  integer(c_int32_t) :: RKC1
 end type pt_params

 !! stats.dat / viscosity_smoother.dat values (PS/basics.f90:363-462)
 type, bind(C) :: pt_stats_row
  real(c_double) :: time, Tavg, vrms, Nu, Nu_bot, RaT_avg, mobility, strain_max, q1, q2, q3, q4
  integer(c_int32_t) :: empty_cells, tracer_min, tracer_max
  real(c_double) :: Cavg, ent
  real(c_double) :: vis_min, vis_max
  integer(c_int32_t) :: n_eta, n_temp, n_comp
  real(c_double) :: eig_fine
 end type pt_stats_row

 type, bind(C) :: pt_solve_info
  integer(c_int32_t) :: q
  real(c_double) :: residual, res_ref, dt, time
  integer(c_int64_t) :: solves, cycles
  integer(c_int32_t) :: n_eta, n_temp, n_comp
  real(c_double) :: eig_fine
  integer(c_int32_t) :: cycle_graphs
 end type pt_solve_info

 type(c_ptr), save :: pt_ctx = c_null_ptr

 interface
  subroutine pt_default_params(p) bind(C, name="pt_default_params")
   import :: pt_params
   type(pt_params), intent(out) :: p
  end subroutine
  integer(c_int) function pt_create(p, ctx) bind(C, name="pt_create")
   import :: pt_params, c_int, c_ptr
   type(pt_params), intent(in) :: p
   type(c_ptr), intent(out) :: ctx
  end function
  integer(c_int) function pt_destroy(ctx) bind(C, name="pt_destroy")
   import :: c_int, c_ptr
   type(c_ptr), value :: ctx
  end function
  integer(c_int) function pt_get_dims(ctx, nx, nz, grids, ntr) bind(C, name="pt_get_dims")
   import :: c_int, c_ptr, c_int32_t, c_int64_t
   type(c_ptr), value :: ctx
   integer(c_int32_t), intent(out) :: nx, nz, grids
   integer(c_int64_t), intent(out) :: ntr
  end function
  integer(c_int) function pt_load_state(ctx, T_nodes, Tbot_iso, Ttop_iso, rtr, str, Ttr, Ctr) &
    bind(C, name="pt_load_state")
   import :: c_int, c_ptr, c_double
   type(c_ptr), value :: ctx
   real(c_double), intent(in) :: T_nodes(*)
   real(c_double), value :: Tbot_iso, Ttop_iso
   real(c_double), intent(in) :: rtr(*), str(*), Ttr(*), Ctr(*)
  end function
  integer(c_int) function pt_multigrid(ctx) bind(C, name="pt_multigrid")
   import :: c_int, c_ptr
   type(c_ptr), value :: ctx
  end function
  integer(c_int) function pt_compute_velocities(ctx) bind(C, name="pt_compute_velocities")
   import :: c_int, c_ptr
   type(c_ptr), value :: ctx
  end function
  integer(c_int) function pt_update_velocities(ctx) bind(C, name="pt_update_velocities")
   import :: c_int, c_ptr
   type(c_ptr), value :: ctx
  end function
  integer(c_int) function pt_build_sf_matrix(ctx) bind(C, name="pt_build_sf_matrix")
   import :: c_int, c_ptr
   type(c_ptr), value :: ctx
  end function
  integer(c_int) function pt_solve_sf_equation(ctx) bind(C, name="pt_solve_sf_equation")
   import :: c_int, c_ptr
   type(c_ptr), value :: ctx
  end function
  integer(c_int) function pt_energy_time(ctx) bind(C, name="pt_energy_time")
   import :: c_int, c_ptr
   type(c_ptr), value :: ctx
  end function
  integer(c_int) function pt_tracers_to_corners(ctx) bind(C, name="pt_tracers_to_corners")
   import :: c_int, c_ptr
   type(c_ptr), value :: ctx
  end function
  integer(c_int) function pt_time_step(ctx) bind(C, name="pt_time_step")
   import :: c_int, c_ptr
   type(c_ptr), value :: ctx
  end function
  integer(c_int) function pt_run_steps(ctx, n) bind(C, name="pt_run_steps")
   import :: c_int, c_ptr, c_int32_t
   type(c_ptr), value :: ctx
   integer(c_int32_t), value :: n
  end function
  integer(c_int64_t) function pt_field_extent(ctx, field_id) bind(C, name="pt_field_extent")
   import :: c_ptr, c_int32_t, c_int64_t
   type(c_ptr), value :: ctx
   integer(c_int32_t), value :: field_id
  end function
  integer(c_int) function pt_get_field(ctx, field_id, host) bind(C, name="pt_get_field")
   import :: c_int, c_ptr, c_int32_t, c_double
   type(c_ptr), value :: ctx
   integer(c_int32_t), value :: field_id
   real(c_double), intent(out) :: host(*)
  end function
  integer(c_int) function pt_set_field(ctx, field_id, host) bind(C, name="pt_set_field")
   import :: c_int, c_ptr, c_int32_t, c_double
   type(c_ptr), value :: ctx
   integer(c_int32_t), value :: field_id
   real(c_double), intent(in) :: host(*)
  end function
  integer(c_int) function pt_get_tracers(ctx, rtr, str, Ttr, Ctr) bind(C, name="pt_get_tracers")
   import :: c_int, c_ptr, c_double
   type(c_ptr), value :: ctx
   real(c_double), intent(out) :: rtr(*), str(*), Ttr(*), Ctr(*)
  end function
  integer(c_int) function pt_get_nempty(ctx, nempty) bind(C, name="pt_get_nempty")
   import :: c_int, c_ptr, c_int32_t
   type(c_ptr), value :: ctx
   integer(c_int32_t), intent(out) :: nempty(*)
  end function
  integer(c_int) function pt_get_info(ctx, info) bind(C, name="pt_get_info")
   import :: c_int, c_ptr, pt_solve_info
   type(c_ptr), value :: ctx
   type(pt_solve_info), intent(out) :: info
  end function
  integer(c_int) function pt_stats(ctx, row) bind(C, name="pt_stats")
   import :: c_int, c_ptr, pt_stats_row
   type(c_ptr), value :: ctx
   type(pt_stats_row), intent(out) :: row
  end function
  !! pt_stats without any allocation (scratch in the library; nz <= 2047)
  integer(c_int) function pt_stats_noalloc(ctx, row) bind(C, name="pt_stats_noalloc")
   import :: c_int, c_ptr, pt_stats_row
   type(c_ptr), value :: ctx
   type(pt_stats_row), intent(out) :: row
  end function
  integer(c_int) function pt_write_restart(ctx, dir) bind(C, name="pt_write_restart")
   import :: c_int, c_ptr, c_char
   type(c_ptr), value :: ctx
   character(kind=c_char), intent(in) :: dir(*)
  end function
  integer(c_int) function pt_write_snapshot(ctx, dir, fcount) bind(C, name="pt_write_snapshot")
   import :: c_int, c_ptr, c_char, c_int32_t
   type(c_ptr), value :: ctx
   character(kind=c_char), intent(in) :: dir(*)
   integer(c_int32_t), value :: fcount
  end function
  integer(c_int) function pt_read_restart(ctx, dir, Tbot_iso, Ttop_iso) bind(C, name="pt_read_restart")
   import :: c_int, c_ptr, c_char, c_double
   type(c_ptr), value :: ctx
   character(kind=c_char), intent(in) :: dir(*)
   real(c_double), value :: Tbot_iso, Ttop_iso
  end function
  integer(c_int) function pt_synchronize(ctx) bind(C, name="pt_synchronize")
   import :: c_int, c_ptr
   type(c_ptr), value :: ctx
  end function
  integer(c_int) function pt_enable_kernel_timing(ctx, on) bind(C, name="pt_enable_kernel_timing")
   import :: c_int, c_ptr, c_int32_t
   type(c_ptr), value :: ctx
   integer(c_int32_t), value :: on
  end function
  integer(c_int) function pt_kernel_time(ctx, which, ms, launches) bind(C, name="pt_kernel_time")
   import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
   type(c_ptr), value :: ctx
   integer(c_int32_t), value :: which
   real(c_double), intent(out) :: ms
   integer(c_int64_t), intent(out) :: launches
  end function
  integer(c_int) function pt_kernel_work(ctx, which, units) bind(C, name="pt_kernel_work")
   import :: c_int, c_ptr, c_int32_t, c_int64_t
   type(c_ptr), value :: ctx
   integer(c_int32_t), value :: which
   integer(c_int64_t), intent(out) :: units
  end function
  integer(c_int) function pt_projection_counts(ctx, windowed, radix) bind(C, name="pt_projection_counts")
   import :: c_int, c_int64_t, c_ptr
   type(c_ptr), value :: ctx
   integer(c_int64_t), intent(out) :: windowed, radix
  end function
  !! tracer redistribution (equalize_tracers_original, PS/composition.f90:84-263)
  integer(c_int) function pt_equalize_counts(ctx, nequalize, moved_last, moved_total) &
    bind(C, name="pt_equalize_counts")
   import :: c_int, c_int64_t, c_ptr
   type(c_ptr), value :: ctx
   integer(c_int64_t), intent(out) :: nequalize, moved_last, moved_total
  end function
  integer(c_int) function pt_get_rng_state(ctx, state) bind(C, name="pt_get_rng_state")
   import :: c_int, c_int32_t, c_ptr
   type(c_ptr), value :: ctx
   integer(c_int32_t), intent(out) :: state(4)
  end function
  integer(c_int) function pt_set_rng_state(ctx, state) bind(C, name="pt_set_rng_state")
   import :: c_int, c_int32_t, c_ptr
   type(c_ptr), value :: ctx
   integer(c_int32_t), intent(in) :: state(4)
  end function
  !! learned initial guess (include/projectracer_gpu.h, "learned initial guess")
  integer(c_int) function pt_net_create(seed, net) bind(C, name="pt_net_create")
   import :: c_int, c_int32_t, c_ptr
   integer(c_int32_t), value :: seed
   type(c_ptr), intent(out) :: net
  end function
  integer(c_int) function pt_net_destroy(net) bind(C, name="pt_net_destroy")
   import :: c_int, c_ptr
   type(c_ptr), value :: net
  end function
  integer(c_int) function pt_net_dataset(net, H, W, capacity) bind(C, name="pt_net_dataset")
   import :: c_int, c_int32_t, c_ptr
   type(c_ptr), value :: net
   integer(c_int32_t), value :: H, W, capacity
  end function
  integer(c_int) function pt_net_dataset_size(net, count) bind(C, name="pt_net_dataset_size")
   import :: c_int, c_int32_t, c_ptr
   type(c_ptr), value :: net
   integer(c_int32_t), intent(out) :: count
  end function
  integer(c_int) function pt_net_train(net, iters, batch, lr, weight_decay, seed, loss) bind(C, name="pt_net_train")
   import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
   type(c_ptr), value :: net
   integer(c_int32_t), value :: iters, batch
   real(c_double), value :: lr, weight_decay
   integer(c_int64_t), value :: seed
   real(c_double), intent(out) :: loss
  end function
  integer(c_int) function pt_set_guess_net(ctx, net, guess, capture) bind(C, name="pt_set_guess_net")
   import :: c_int, c_int32_t, c_ptr
   type(c_ptr), value :: ctx, net
   integer(c_int32_t), value :: guess, capture
  end function
  integer(c_int) function pt_guess_counts(ctx, tried, used) bind(C, name="pt_guess_counts")
   import :: c_int, c_int64_t, c_ptr
   type(c_ptr), value :: ctx
   integer(c_int64_t), intent(out) :: tried, used
  end function
  !! ensembles of independent cases on one GPU
  integer(c_int) function pt_ensemble_create(n, params, ens) bind(C, name="pt_ensemble_create")
   import :: c_int, c_int32_t, c_ptr, pt_params
   integer(c_int32_t), value :: n
   type(pt_params), intent(in) :: params(*)
   type(c_ptr), intent(out) :: ens
  end function
  integer(c_int) function pt_ensemble_destroy(ens) bind(C, name="pt_ensemble_destroy")
   import :: c_int, c_ptr
   type(c_ptr), value :: ens
  end function
  integer(c_int) function pt_ensemble_case(ens, i, ctx) bind(C, name="pt_ensemble_case")
   import :: c_int, c_int32_t, c_ptr
   type(c_ptr), value :: ens
   integer(c_int32_t), value :: i
   type(c_ptr), intent(out) :: ctx
  end function
  integer(c_int) function pt_ensemble_run_steps(ens, nsteps, workers, status) bind(C, name="pt_ensemble_run_steps")
   import :: c_int, c_int32_t, c_ptr
   type(c_ptr), value :: ens
   integer(c_int32_t), value :: nsteps, workers
   integer(c_int32_t), intent(out) :: status(*)
  end function
  integer(c_int) function pt_ensemble_workers(ens, workers) bind(C, name="pt_ensemble_workers")
   import :: c_int, c_int32_t, c_ptr
   type(c_ptr), value :: ens
   integer(c_int32_t), intent(out) :: workers
  end function
  type(c_ptr) function pt_last_error() bind(C, name="pt_last_error")
   import :: c_ptr
  end function
  integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
   import :: c_size_t, c_ptr
   type(c_ptr), value :: s
  end function
 end interface

contains

 !! text of pt_last_error() as a Fortran string
 function pt_error_message() result(msg)
  character(len=:), allocatable :: msg
  type(c_ptr) :: p
  character(kind=c_char), pointer :: chars(:)
  integer :: n, j
  p = pt_last_error()
  n = int(c_strlen(p))
  call c_f_pointer(p, chars, [n])
  allocate(character(len=n) :: msg)
  do j = 1, n
   msg(j:j) = chars(j)
  end do
 end function

 !! stop with the reference's message when a call fails
 subroutine pt_check(ierr, where)
  integer(c_int), intent(in) :: ierr
  character(len=*), intent(in) :: where
  if (ierr .eq. PT_OK) return
  select case (ierr)
  case (PT_ERR_TRACER_OUTSIDE)
   write(*,*) "Tracer advected outside of domain - stopping"
  case (PT_ERR_MAX_CYCLES)
   write(*,*) "Multigrid Convergence Problem: Too Many Multigrid Cycles Needed"
  case (PT_ERR_HIGH_RESIDUAL)
   write(*,*) "Multigrid Convergence Problem: Abnormally High Residual Detected -- stopping"
  end select
  write(*,*) trim(where), ": ", pt_error_message()
  stop
 end subroutine

end module projectracer_gpu
