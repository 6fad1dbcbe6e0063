!! pt_steps.f90 -- the reference's time loop with its hot path on the GPU,
!! driven from Fortran through projectracer_gpu.f90.
!!
!! The program a reference maintainer gets by changing PS/main.f90 the way
!! the interface module describes: the namelists are read as
!! PS/main.f90:11-21 reads them (same file, groups and names, declared with
!! the kinds of PS/basics.f90:3-66), their values fill pt_params, and each
!! iteration of the time loop (PS/main.f90:245-300: velocity solve, RK4
!! energy step, tracer projection) is one pt_time_step on the GPU.
!!
!! Inputs, in the working directory:
!!   inputs.nml  the reference's namelist file (groups inputs,
!!               Rayleigh_numbers, heating_conductivity, viscosity); one
!!               composition type (comp=1, ntype=1)
!!   steps.nml   &steps nrun=<time steps>, precision=<0 FP64, 1 mixed> /
!!   inject.bin  stream file: Tbot_iso, Ttop_iso, T(0:nx,0:nz), then rtr,
!!               str, Ttr, Ctr (ntr values each): the state the reference
!!               holds after initial_temperature and initialize_tracers
!! Output: state.bin, stream: nx, nz, grids (int32), ntr (int64), the cycle
!!   count of the last solve of each time step and of one more solve after
!!   the loop (nrun+1 int32), the model time (double), then SF, T, U, W
!!   each as (extent int64, values with halos), nempty(0:nx,0:nz) (int32)
!!   and rtr, str, Ttr, Ctr (ntr doubles each).
program pt_steps
 use projectracer_gpu
 implicit none
 !! the reference's namelist variables (PS/basics.f90:54-66)
 real*8 :: aspect, thickness, spaceX, spaceZ, tolerance, eig_ratio, tol_coarse, dampingS, c_2, c_1
 real*8 :: tend, tsnap, damping, aRK2, tmovie_snap, dt_manual, courant, courantT, courantC, courant_stream
 real*8 :: dist_factor, delta_visT_init, dlayer, dent, klayer
 integer*4 :: nx, nz, Database, order, gridX, gridZ, nbound, iterate, MG, non_Newtonian, Nmulti, mingrid, grids
 integer*4 :: vis_iter, max_cycles, fp_max, splitting, RKC1, nstageS, Chebyshev, nstage, torder, movie
 integer*4 :: Nlocal, Nsnap, local_time, RK_update, vel_update, CFL, comp, ntype, iC_continuous
 integer*4 :: boundary_tracers, original, shape_function, buffer, equalize, tracer, tpc, restart
 integer*4 :: initial_temp, initial_comp, initial_temp_smooth, add_layer, Tbot, Vbc
 real*8 :: RaT, H, visP, visT, delta_visT, eig_ratio_min, yield, yield_gradient, vis_star, mpla, Cpla, alpha
 integer*4 :: iheattype, plastic, ivis
 real*8 :: RaC(1), conduct_factor(1), Htype(1), visC(1), delta_visC(1)
 namelist /inputs/ aspect,nx,nz,Database,order,gridX,gridZ,thickness,nbound,spaceX,spaceZ,&
  &iterate,MG,non_Newtonian,tolerance,Nmulti,mingrid,grids,eig_ratio,vis_iter,max_cycles,tol_coarse,fp_max,&
  &splitting,RKC1,nstageS,dampingS,c_2,c_1,&
  &tend,tsnap,Chebyshev,nstage,damping,aRK2,torder,movie,tmovie_snap,Nlocal,Nsnap,local_time,dt_manual,&
  &courant,courantT,courantC,courant_stream,RK_update,vel_update,CFL,&
  &comp,ntype,iC_continuous,boundary_tracers,original,shape_function,buffer,equalize,tracer,tpc,dist_factor,&
  &restart,initial_temp,initial_comp,initial_temp_smooth,delta_visT_init,add_layer,dlayer,dent,klayer,Tbot,Vbc
 namelist /Rayleigh_numbers/ RaT,RaC
 namelist /heating_conductivity/ iheattype,H,conduct_factor,Htype
 namelist /viscosity/ plastic,ivis,visP,visT,delta_visT,eig_ratio_min,visC,delta_visC,yield,yield_gradient,&
  &vis_star,mpla,Cpla,alpha
 integer*4 :: nrun, precision
 namelist /steps/ nrun, precision

 type(pt_params) :: p
 type(pt_solve_info) :: info
 type(c_ptr) :: ctx
 integer(c_int32_t) :: gnx, gnz, ggrids
 integer(c_int64_t) :: ntr, ext
 integer(c_int32_t), allocatable :: qs(:), nempty(:)
 real(c_double), allocatable :: Tn(:), rtr(:), str(:), Ttr(:), Ctr(:), f(:)
 real(c_double) :: Tbot_iso, Ttop_iso
 integer(c_int32_t), parameter :: out_fields(4) = [PT_FIELD_SF, PT_FIELD_T, PT_FIELD_U, PT_FIELD_W]
 integer :: step, j
 integer(c_int64_t) :: neq, moved_last, moved_total
 integer(c_int32_t) :: seeds(4)

 ntype = 1; comp = 1
 !! the supported values of the options pt_params does not carry, for a file that omits them
 order = 4; gridX = 0; gridZ = 0; MG = 1; non_Newtonian = 0; splitting = 0; RKC1 = 1; Chebyshev = 0
 torder = 4; local_time = 0; CFL = 0; RK_update = 1; vel_update = 1; tracer = 1; shape_function = 1
 plastic = 0; iheattype = 0
 open(unit=501, file="inputs.nml", status="old")
 read(nml=inputs, unit=501)
 read(nml=Rayleigh_numbers, unit=501)
 read(nml=heating_conductivity, unit=501)
 read(nml=viscosity, unit=501)
 close(501)
 if (comp .ne. 1 .or. ntype .ne. 1) then
  write(*,*) "pt_steps: one composition type (comp=1, ntype=1) is supported"
  stop 2
 end if
 !! options outside the subset the GPU path is built for (include/projectracer_gpu.h); pt_params
 !! has no field for them, so they are refused here rather than ignored
 if (order .ne. 4 .or. gridX .ne. 0 .or. gridZ .ne. 0 .or. MG .ne. 1 .or. non_Newtonian .ne. 0 .or. &
     & (RKC1 .eq. 1 .and. splitting .ne. 0) .or. Chebyshev .ne. 0 .or. torder .ne. 4 .or. &
     & local_time .ne. 0 .or. CFL .ne. 0 .or. RK_update .ne. 1 .or. vel_update .ne. 1 .or. &
     & tracer .ne. 1 .or. shape_function .ne. 1 .or. plastic .ne. 0 .or. iheattype .ne. 0) then
  write(*,*) "pt_steps: a namelist option outside the supported subset: order=4, gridX=gridZ=0, MG=1, ", &
     & "non_Newtonian=0, splitting=0, Chebyshev=0, torder=4, local_time=0, CFL=0, RK_update=1, ", &
     & "vel_update=1, tracer=1, shape_function=1, plastic=0, iheattype=0"
  stop 2
 end if
 nrun = 1; precision = PT_PREC_FP64
 open(unit=502, file="steps.nml", status="old")
 read(nml=steps, unit=502)
 close(502)

 !! namelist values -> pt_params (the field order of type pt_params)
 call pt_default_params(p)
 p%aspect = aspect; p%nx = nx; p%nz = nz; p%grids = grids; p%mingrid = mingrid
 p%tolerance = tolerance; p%max_cycles = max_cycles; p%Nmulti = Nmulti; p%eig_ratio = eig_ratio
 p%vis_iter = vis_iter; p%nstageS = nstageS; p%dampingS = dampingS; p%c_2 = c_2; p%c_1 = c_1
 p%courant = courant; p%courantT = courantT; p%courantC = courantC; p%courant_stream = courant_stream
 p%tpc = tpc; p%Tbot = Tbot; p%Vbc = Vbc
 p%RaT = RaT; p%RaC = RaC(1)
 p%H = H; p%conduct_factor = conduct_factor(1); p%Htype = Htype(1)
 p%ivis = ivis; p%visP = visP; p%visT = visT; p%delta_visT = delta_visT; p%eig_ratio_min = eig_ratio_min
 p%visC = visC(1); p%delta_visC = delta_visC(1)
 p%precision = precision; p%guess = PT_GUESS_WARM
 p%dent = dent; p%iterate = iterate; p%Database = Database
 p%iC_continuous = iC_continuous; p%equalize = equalize; p%original = original; p%buffer = buffer
 p%dist_factor = dist_factor
 p%RKC1 = RKC1

 call pt_check(pt_create(p, ctx), "pt_create")
 call pt_check(pt_get_dims(ctx, gnx, gnz, ggrids, ntr), "pt_get_dims")
 allocate(Tn((gnx+1)*(gnz+1)), rtr(ntr), str(ntr), Ttr(ntr), Ctr(ntr), nempty((gnx+1)*(gnz+1)), qs(nrun+1))
 open(unit=601, file="inject.bin", access="stream", form="unformatted", status="old")
 read(601) Tbot_iso, Ttop_iso
 read(601) Tn
 read(601) rtr, str, Ttr, Ctr
 close(601)
 call pt_check(pt_load_state(ctx, Tn, Tbot_iso, Ttop_iso, rtr, str, Ttr, Ctr), "pt_load_state")

 !! the time loop: each iteration solves for psi, takes the RK4 energy step
 !! (three more solves inside) and projects the tracers
 do step = 1, nrun
  call pt_check(pt_time_step(ctx), "pt_time_step")
  call pt_check(pt_get_info(ctx, info), "pt_get_info")
  qs(step) = info%q
 end do
 !! the next iteration's velocity solve, for the state to compare
 call pt_check(pt_multigrid(ctx), "pt_multigrid")
 call pt_check(pt_compute_velocities(ctx), "pt_compute_velocities")
 call pt_check(pt_get_info(ctx, info), "pt_get_info")
 qs(nrun+1) = info%q

 open(unit=701, file="state.bin", access="stream", form="unformatted", status="replace")
 write(701) gnx, gnz, ggrids, ntr
 write(701) qs
 write(701) info%time
 do j = 1, size(out_fields)
  ext = pt_field_extent(ctx, out_fields(j))
  allocate(f(ext))
  call pt_check(pt_get_field(ctx, out_fields(j), f), "pt_get_field")
  write(701) ext
  write(701) f
  deallocate(f)
 end do
 call pt_check(pt_get_nempty(ctx, nempty), "pt_get_nempty")
 write(701) nempty
 call pt_check(pt_get_tracers(ctx, rtr, str, Ttr, Ctr), "pt_get_tracers")
 write(701) rtr, str, Ttr, Ctr
 close(701)
 if (equalize .eq. 1) then
  !! the reference's nequalize (PS/composition.f90:151-155) and mzran stream 0
  call pt_check(pt_equalize_counts(ctx, neq, moved_last, moved_total), "pt_equalize_counts")
  call pt_check(pt_get_rng_state(ctx, seeds), "pt_get_rng_state")
  write(*,'(a,3(i12),4(i12))') "pt_steps: nequalize, moved, moved total, seeds ", neq, moved_last, &
   moved_total, seeds
 end if
 call pt_check(pt_destroy(ctx), "pt_destroy")
 write(*,'(a,i4,a,i8)') "pt_steps: ", nrun, " steps, last cycle count ", qs(nrun+1)
end program pt_steps
