!! ref_driver.f90 -- harness program for the flang build of the reference.
!!
!! TEST INFRASTRUCTURE ONLY.  Links against the reference's own objects
!! (compiled from its .f90 sources by oracle/Makefile) and drives them in the
!! order of PS/main.f90:168-300, with three differences that make the
!! reference usable as an oracle:
!!   * initial T and tracers can be injected from inject.bin (the configs'
!!     seeded initial conditions are not expressible in PS/initial_conditions.f90),
!!     with Tbot_iso/Ttop_iso set explicitly (restart=1 never sets them,
!!     PS/basics.f90:352-353);
!!   * the time loop runs a fixed number of steps (controls in driver.nml);
!!   * the state is written in full FP64 to state_NNNNN.bin (stream access)
!!     instead of the reference's snapshots (dump_every<0: no state files;
!!     each loop iteration's wall time is printed for timing runs).
!! With iterate=0 the set-up and the velocity solve follow main.f90's direct
!! branches (PS/main.f90:189-203, 224-226, 247-251).  With discrete tracer
!! types (iC_continuous=0) the injected and written C values are Ttype as
!! 0.0 / 1.0; with equalize=1 every step also prints nequalize and the
!! state of mzran stream 0.
!! Nothing in the reference sources is modified.
program ref_driver
 use basics
 use arrays
 implicit none
 integer*4 :: nrun, inject, dump_every, level, i, k, step_dump
 real*8 :: t1, t2, t_solve0, t_it0, t_it1
 real*8, allocatable :: Tn(:,:), Cin(:)
 character(len=32) :: fname
 namelist /driver/ nrun, inject, dump_every

 nrun=1; inject=0; dump_every=0
 open(unit=501,file="inputs.nml")
 read(nml=inputs,unit=501)
 if (comp.eq.1) allocate(RaC(1:ntype))
 read(nml=Rayleigh_numbers,unit=501)
 if (comp.eq.1) allocate(conduct_factor(1:ntype),Htype(1:ntype))
 read(nml=heating_conductivity,unit=501)
 if (comp.eq.1) allocate(visC(1:ntype),delta_visC(1:ntype))
 read(nml=viscosity,unit=501)
 close(501)
 open(unit=502,file="driver.nml")
 read(nml=driver,unit=502)
 close(502)

 call compute_smoothing_parameters
 if (iterate.eq.1) call multigrid_setup
 call compute_basics
 call storage
 if (RKC1.eq.1) call Chebyshev_prep_stream
 call FD_coefficients
 call build_grid
 call derivative_eigenvalues
 call stability_RK4_smoother
 if (iterate.eq.1) then
  call eigenvalue_prep
  call stream_space_matrix_prep
  do level=1,grids
   call stability_vis_smoother(level)
  end do
 else
  grids=1
  level=1
  call stability_vis_smoother(level)
 end if
 call write_init

 if (inject.eq.0) then
  call initial_temperature
  call initialize_tracers
 else
  allocate(Tn(0:nx,0:nz))
  open(unit=601,file="inject.bin",access='stream',form='unformatted',status='old')
  read(601) Tbot_iso,Ttop_iso
  read(601) Tn
  if (iC_continuous.eq.1) then
   read(601) rtr,str,Ttr,Ctr
  else
   allocate(Cin(ntr))
   read(601) rtr,str,Ttr,Cin
   Ttype=nint(Cin)
   deallocate(Cin)
  end if
  close(601)
  T(0:nx,0:nz)=Tn
  T(:,0)=Tbot_iso
  T(:,nz)=Ttop_iso
  call enforceBCs(T)
  call tracers_to_corners(rtr,str,Ttr,T)
 end if
 tstep=-1
 if (iterate.eq.0) then
  call build_SF_matrix
 else
  SF=0.d0
  error(:,:,2:grids)=0.d0
 end if
 time=0.d0
 write(*,*) "driver: initialised nx,nz,grids,ntr=",nx,nz,grids,ntr
 t_solve0=0.d0
 call cpu_time(t1)
!$ t1=omp_get_wtime_d()
 do tstep=0,nrun
  t_it0=omp_get_wtime_d()
  if (iterate.eq.0) then
   if ((ivis.ge.1).or.(tstep.eq.0)) call build_SF_matrix
   call solve_SF_equation
  else
   call multigrid
  end if
  call compute_velocities
  step_dump=0
  if ((tstep.eq.nrun).and.(dump_every.ge.0)) step_dump=1
  if ((dump_every.gt.0)) then
   if (mod(tstep,dump_every).eq.0) step_dump=1
  end if
  if (step_dump.eq.1) call write_state(tstep)
  if (tstep.eq.nrun) exit
  call energy_time
  if (equalize.eq.1) write(*,'(a,i8,4(i12))') " driver: nequalize,seeds=",nequalize,seed_i(0),seed_j(0),&
   &seed_k(0),seed_n(0)
  time=time+dt
  t_it1=omp_get_wtime_d()
  write(*,'(a,i6,3(g25.16))') " driver: step done",tstep,time,dt,t_it1-t_it0
 end do
 call cpu_time(t2)
!$ t2=omp_get_wtime_d()
 write(*,'(a,g25.16)') " driver: loop seconds=",t2-t1
 write(*,'(a,4(g25.16))') " driver: timers multigrid,energy,advect,convert=",tmultigrid,tenergy,tadvect_tracer,tconvert
 if (equalize.eq.1) write(*,'(a,g25.16,i8)') " driver: equalize seconds, calls=",tequalize,nequalize

contains

 real*8 function omp_get_wtime_d()
!$ use OMP_LIB
  omp_get_wtime_d=0.d0
!$ omp_get_wtime_d=omp_get_wtime()
 end function

 subroutine write_init
  integer*4 :: i_rk
  open(unit=602,file="init.bin",access='stream',form='unformatted',status='replace')
  write(602) nx,nz,grids,ntr,nstageS
  write(602) D1,D2,D3,D4,D
  write(602) Tr_base,Trr_base,Trrr_base,Trrrr_base,Trs_base,Trrs_base,Trrss_base,Tr_base_so,Trr_base_so
  if (RKC1.eq.1) then
   write(602) mewtChebS(1:nstageS),mewChebS(2:nstageS),vChebS(2:nstageS),cChebS(2:nstageS),RKC1_stability
  else
   !! RK1 (RKC1=0): the reference allocates no RKC1 tables; zeros keep the layout
   write(602) (0.d0, i_rk = 1, 4*nstageS-2)
  end if
  write(602) dts_array(0,0),delta_T,delta_C(1)
  close(602)
 end subroutine

 subroutine write_state(step)
  integer*4 :: step
  write(fname,'(a,i5.5,a)') "state_",step,".bin"
  open(unit=603,file=trim(fname),access='stream',form='unformatted',status='replace')
  write(603) nx,nz,ntr,step
  write(603) dt,time,Tbot_iso,Ttop_iso
  write(603) SF,Tratio,u,w,Tvis,Tbuoy
  write(603) Cnew(1,:,:),Cvis(1,:,:),conduct
  if (iterate.eq.0) then
   write(603) vis
  else
   write(603) vis_grid(:,:,1)
  end if
  write(603) nempty(0:nx,0:nz)
  if (iC_continuous.eq.1) then
   write(603) rtr,str,Ttr,Ctr
  else
   write(603) rtr,str,Ttr,real(Ttype,8)
  end if
  close(603)
 end subroutine

end program ref_driver
