!*************************************************************************
!> Fortran host driver of the MI355X per-timestep path: the main loop of
!! program/main.F90:100-164 (transform, nonlinear, predictor or corrector,
!! tim_check_cgce) on fixed dt, with the physics and transforms on the GPU
!! through the C-ABI (module lsd_c).
!!
!! Reads one namelist file (argument 1):
!!   &lsd_run_nml  L, M, Mp, N, Th, Ph, cod, comp, mag, precision,
!!                 rratio, Ra, Ra_comp, Pr, Sc, Pm, E, dt, implicit, dterr,
!!                 source, source_comp, nsteps, mode, state_in, state_out /
!! mode = 'step'   : lsd_step (the library's own loop, fused step path)
!! mode = 'phases' : the loop below, one C-ABI call per phase
!! state files: stream-unformatted FP64, for each of the six fields
!! (vel tor, vel pol, cod, comp, mag tor, mag pol) Re(N,0:H-1) then Im.
!*************************************************************************
 program lsd_run
   use lsd_c
   implicit none
   integer :: L = 32, M = 32, Mp = 1, N = 32, Th = 0, Ph = 0
   integer :: cod = 1, comp = 0, mag = 1, precision = 0
   double precision :: rratio = 0.35d0, Ra = 6.5d4, Ra_comp = 1d0, Pr = 1d0, Sc = 1d0, Pm = 5d0
   double precision :: E = 1d-3, dt = 1d-4, implicit = 0.5d0, dterr = 0.025d0
   double precision :: source = 0d0, source_comp = 0d0
   integer :: nsteps = 1
   character(len=16) :: mode = 'step'
   character(len=512) :: state_in = '', state_out = ''
   namelist /lsd_run_nml/ L, M, Mp, N, Th, Ph, cod, comp, mag, precision, rratio, Ra, Ra_comp, &
                          Pr, Sc, Pm, E, dt, implicit, dterr, source, source_comp, nsteps, mode, &
                          state_in, state_out

   type(lsd_config) :: cfg
   type(lsd_step_stats) :: stats
   type(c_ptr) :: ctx
   integer(c_int) :: H, iTh, iPh
   character(len=512) :: nml
   double precision, allocatable :: re(:,:), im(:,:)
   ! timestep state (modules/timestep.F90:12-20)
   integer :: tim_it, tim_step, evals
   double precision :: tim_t, tim_dterr, lasterr, corr_dt, d
   integer :: f, u

   call get_command_argument(1, nml)
   open(newunit=u, file=trim(nml), status='old')
   read(u, nml=lsd_run_nml)
   close(u)

   cfg%L = L; cfg%M = M; cfg%Mp = Mp; cfg%N = N; cfg%Th = Th; cfg%Ph = Ph
   cfg%vel = 1; cfg%cod = cod; cfg%comp = comp; cfg%mag = mag
   cfg%vel_bc = 1; cfg%cod_bc = 1; cfg%comp_bc = 1; cfg%precision = precision
   cfg%rratio = rratio; cfg%Ra = Ra; cfg%Ra_comp = Ra_comp; cfg%Pr = Pr; cfg%Sc = Sc
   cfg%Pm = Pm; cfg%E = E; cfg%dt = dt; cfg%implicit = implicit; cfg%dterr = dterr
   cfg%source = source; cfg%source_comp = source_comp; cfg%step_control = 2
   cfg%Nic = 1; cfg%rot = 0

   call lsd_check(lsd_config_sizes(cfg, H, iTh, iPh), 'par_precompute')
   call lsd_check(lsd_create(cfg, 0_c_int, ctx), 'initialise')
   allocate(re(N, 0:H-1), im(N, 0:H-1))

   open(newunit=u, file=trim(state_in), access='stream', form='unformatted', status='old')
   do f = 0, 5
      read(u) re
      read(u) im
      call lsd_check(lsd_set_coll(ctx, int(f, c_int), re, im), 'io_load_state')
   end do
   close(u)

   if (trim(mode) == 'step') then
      call lsd_check(lsd_step(ctx, int(nsteps, c_int), stats), 'lsd_step')
      evals = stats%evaluations
      tim_step = stats%steps
   else
      tim_it = 0; tim_step = 0; tim_t = 0d0; tim_dterr = 0d0; evals = 0
      lasterr = 1d99; corr_dt = 0d0
      do while (tim_it /= 0 .or. tim_step < nsteps)
         call lsd_check(lsd_transform(ctx), 'transform')
         call lsd_check(lsd_nonlinear(ctx), 'nonlinear')
         evals = evals + 1
         if (tim_it == 0) then
            call lsd_check(lsd_predictor(ctx), 'predictor')
            tim_it = 1
         else
            call lsd_check(lsd_corrector(ctx, d), 'corrector')
            tim_dterr = tim_dterr + d
            call check_cgce()
         end if
         if (tim_it == 0) then
            tim_t = tim_t + dt
            tim_step = tim_step + 1
         end if
      end do
   end if
   call lsd_check(lsd_sync(ctx), 'sync')

   open(newunit=u, file=trim(state_out), access='stream', form='unformatted', status='replace')
   do f = 0, 5
      call lsd_check(lsd_get_coll(ctx, int(f, c_int), re, im), 'io_save_state')
      write(u) re
      write(u) im
   end do
   close(u)
   print '(a,i0,a,i0,a)', 'lsd_run: ', tim_step, ' steps, ', evals, ' nonlinear evaluations'
   call lsd_destroy(ctx)

 contains

   !> tim_check_cgce, i_time_step_control = 2 (modules/timestep.F90:791-827)
   subroutine check_cgce()
      if (tim_it == 1) then
         corr_dt = dt * sqrt(dterr / tim_dterr)
         lasterr = 1d99
      end if
      if (dt < 1d-12 .and. tim_step > 30) then
         stop 'tim_check_cgce: dt small'
      else if (tim_it > 10) then
         stop 'tim_check_cgce: too many its'
      else if (tim_dterr > lasterr) then
         if (tim_dterr > 2d0 * dterr) stop 'tim_check_cgce: increasing error'
         tim_it = 0
      else if (tim_dterr > dterr) then
         lasterr = tim_dterr
         tim_it = tim_it + 1
      else
         tim_it = 0
      end if
      tim_dterr = 0d0
   end subroutine check_cgce

 end program lsd_run
