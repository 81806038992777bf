!-------------------------------------------------------------------------
! ref_step -- test infrastructure only (never part of the product).
!
! Drives the reference's own, unmodified physics modules, compiled from
! their sources by oracle/ref_build.py, through the main loop of
! program/main.F90:100-164 without any file I/O, so the repository can
! write golden per-step fixtures and time the reference's FFTW/LAPACK CPU
! path.  Serial build (_Nth = _Nr = 1), so every MPI branch is compiled
! out.
!
! Usage: ref_step <state.in> <dump.out> <nsteps> <dump_every> [time]
!   state.in : int32 N, int32 H, then uT, uP, C, X, BT, BP as Re(N,0:H-1)
!              followed by Im(N,0:H-1), float64 (the reference coll layout);
!              with a conducting inner core (i_Nic > 1) or the inner-core
!              rotation (b_rot_tstep) then icBT, icBP as Re(Nic,0:H-1),
!              Im(Nic,0:H-1) and rot_omega
!   dump.out : a record per dumped step: int32 step, int32 evaluations of
!              the step, float64 tim_t, float64 tim_dterr of the accepted
!              corrector, then the six fields as in state.in (and icBT,
!              icBP, rot_omega, rot_velTorq, rot_magTorq in the inner-core
!              builds); step 0 also carries the mesh arrays r, dr(1..3),
!              radLap, r1dr (7 x N band storage) and intr2dr (and the
!              inner-core radii), and the tables leg_gp, leg_gw, leg_P,
!              leg_gwP.
!   time     : if present, steps 2..nsteps are timed and the seconds per
!              step printed as "sec/step <value>" (program/main.F90:414
!              prints the same figure).
!
! Initialisation follows program/main.F90:220-400 with io_load_state
! replaced by the raw read above and io_load_radial_txtfiles by its
! Boussinesq branch (program/io.F90:508-550: eta = 1, d eta/dr = 0).
! With d_timestep < 0 each step starts with non_maxtstep, tim_new_tstep
! and the *_matrices rebuild on a new dt (program/main.F90:124-135).  The
! inner-core rotation follows program/main.F90:108, 136, 148.
!-------------------------------------------------------------------------
#include "../parallel.h"
 program ref_step
   use meshs, only : rdom, mes_precompute
   use variables, only : coll, phys, var_precompute
   use parameters
   use legendre, only : leg_precompute, leg_gp, leg_gw, leg_P, leg_gwP
   use transform, only : tra_precompute
   use timestep, only : tim_it, tim_t, tim_dt, tim_step, tim_dterr, tim_check_cgce, tim_precompute, &
                        tim_new_tstep, tim_new_dt
   use codensity, only : cod_precompute, cod_transform, cod_matrices, cod_predictor, cod_corrector
   use composition, only : comp_precompute, comp_transform, comp_matrices, comp_predictor, comp_corrector
   use velocity, only : vel_precompute, vel_angmomclr, vel_transform, vel_matrices, vel_predictor, &
                        vel_corrector
   use magnetic, only : mag_precompute, mag_transform, mag_matrices, mag_predictor, mag_corrector
   use nonlinear, only : non_precompute, non_codensity, non_composition, non_magnetic, non_velocity, &
                         non_maxtstep, non_rotation
   use rotation, only : rot_precompute, rot_predictor, rot_corrector
   implicit none

   type(rdom) :: mes_oc, mes_ic
   type(coll) :: vel_tor, vel_pol, vel_ntor, vel_npol
   type(phys) :: vel_r, vel_t, vel_p, vel_curlr, vel_curlt, vel_curlp
   type(phys) :: vel_dur, vel_qv1, vel_dutr, vel_dupr, vel_qv2
   type(coll) :: mag_tor, mag_pol, mag_ntor, mag_npol, mag_icTor, mag_icPol, mag_icNTor, mag_icNPol
   type(phys) :: mag_r, mag_t, mag_p, mag_curlr, mag_curlt, mag_curlp
   double precision :: mag_bcRe(2,0:i_pH1), mag_bcIm(2,0:i_pH1)
   double precision :: mag_bc_PolRe(2,0:i_pH1), mag_bc_PolIm(2,0:i_pH1)
   type(coll) :: cod_C, cod_NC
   type(phys) :: cod_gradr, cod_gradt, cod_gradp
   double precision :: cod_S(1:i_N), cod_bcRe(2,0:i_pH1), cod_bcIm(2,0:i_pH1)
   type(coll) :: comp_C, comp_NC
   type(phys) :: comp_gradr, comp_gradt, comp_gradp
   double precision :: comp_S(1:i_N), comp_bcRe(2,0:i_pH1), comp_bcIm(2,0:i_pH1)
   double precision :: rot_omega, rot_velTorq, rot_magTorq, rot_inertia

   character(len=512) :: fin, fout, arg
   integer :: nsteps, every, n, k, nev, uin, uout, nn, hh
   logical :: timing, icio
   integer(kind=8) :: c0, c1, crate
   double precision :: accepted_err

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   call get_command_argument(3, arg)
   read(arg, *) nsteps
   call get_command_argument(4, arg)
   read(arg, *) every
   timing = command_argument_count() >= 5
   icio = i_Nic > 1 .or. b_rot_tstep

   ! io_load_radial_txtfiles, Boussinesq without a diffusivity profile
   do n = 1, i_N
      mes_oc%eta(n,0) = 1d0
      mes_oc%eta(n,1) = 0d0
   end do
   if (i_Nic /= 1) then
      do n = 1, i_Nic
         mes_ic%eta(n,0) = 1d0
         mes_ic%eta(n,1) = 0d0
      end do
   end if
   call par_precompute()
   call mes_precompute(mes_oc, mes_ic)
   call leg_precompute()
   call var_precompute()
   call tra_precompute(1)
   call tim_precompute()
   call rot_precompute(mes_oc, rot_omega, rot_inertia, rot_velTorq, rot_magTorq)
   call vel_precompute(mes_oc, vel_Tor, vel_Pol)
   call cod_precompute(mes_oc, cod_C, cod_S, cod_bcRe, cod_bcIm)
   call comp_precompute(mes_oc, comp_C, comp_NC, comp_S, comp_bcRe, comp_bcIm)
   call mag_precompute(mes_oc, mes_ic, mag_Tor, mag_Pol, mag_icTor, mag_icPol, &
                       mag_bcRe, mag_bcIm, mag_bc_PolRe, mag_bc_PolIm)
   call non_precompute()

   ! state (replaces io_load_state): t = 0, the fixed d_timestep
   open(newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old')
   read(uin) nn, hh
   if (nn /= i_N .or. hh /= i_H) stop 'ref_step: state size does not match the build'
   call read_coll(vel_tor)
   call read_coll(vel_pol)
   call read_coll(cod_C)
   call read_coll(comp_C)
   call read_coll(mag_tor)
   call read_coll(mag_pol)
   if (icio) then
      read(uin) mag_icTor%Re(1:i_Nic, 0:i_H-1)
      read(uin) mag_icTor%Im(1:i_Nic, 0:i_H-1)
      read(uin) mag_icPol%Re(1:i_Nic, 0:i_H-1)
      read(uin) mag_icPol%Im(1:i_Nic, 0:i_H-1)
      read(uin) rot_omega
   end if
   close(uin)
   ! with d_timestep < 0 (adaptive step) io_load_state takes the initial dt
   ! from the state file (program/io.F90:429); here its magnitude stands in
   tim_t = 0d0
   tim_dt = abs(d_timestep)

   call vel_angmomclr(mes_oc, vel_tor)
   call vel_transform(mes_oc, vel_tor, vel_pol, vel_r, vel_t, vel_p, &
                      vel_curlr, vel_curlt, vel_curlp, vel_dur, vel_qv1, vel_dutr, vel_dupr, vel_qv2)
   call cod_transform(mes_oc, cod_C, cod_gradr, cod_gradt, cod_gradp)
   call comp_transform(mes_oc, comp_C, comp_gradr, comp_gradt, comp_gradp)
   call mag_transform(mag_tor, mag_pol, mag_r, mag_t, mag_p, mag_curlr, mag_curlt, mag_curlp)
   call vel_matrices(mes_oc)
   call cod_matrices(mes_oc)
   call comp_matrices(mes_oc)
   call mag_matrices(mes_oc, mes_ic)

   open(newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
   write(uout) 0, 0, tim_t, 0d0
   call write_fields()
   write(uout) mes_oc%r(1:i_N,1)
   do k = 1, 3
      write(uout) mes_oc%dr(k)%M
   end do
   write(uout) mes_oc%radLap%M, mes_oc%r1dr%M, mes_oc%intr2dr
   if (icio) write(uout) mes_ic%r(1:i_Nic,1)
   write(uout) leg_gp, leg_gw, leg_P, leg_gwP

   ! program/main.F90:100-164 (no output files)
   nev = 0
   accepted_err = 0d0
   call system_clock(c0, crate)
   do while (tim_step < nsteps)
      if (timing .and. tim_step == 1 .and. tim_it == 0) call system_clock(c0, crate)
      if (b_vel_tstep) call vel_angmomclr(mes_oc, vel_tor)
      if (b_vel_tstep) call vel_transform(mes_oc, vel_tor, vel_pol, vel_r, vel_t, vel_p, &
                                          vel_curlr, vel_curlt, vel_curlp, &
                                          vel_dur, vel_qv1, vel_dutr, vel_dupr, vel_qv2)
      if (b_cod_tstep) call cod_transform(mes_oc, cod_C, cod_gradr, cod_gradt, cod_gradp)
      if (b_comp_tstep) call comp_transform(mes_oc, comp_C, comp_gradr, comp_gradt, comp_gradp)
      if (b_mag_tstep) call mag_transform(mag_tor, mag_pol, mag_r, mag_t, mag_p, &
                                          mag_curlr, mag_curlt, mag_curlp)
      if (b_rot_tstep) call non_rotation(mes_oc, mes_ic, vel_p, vel_curlt, mag_r, mag_p, &
                                         rot_velTorq, rot_magTorq)
      if (b_vel_tstep) call non_velocity(mes_oc, vel_r, vel_t, vel_p, &
                                         vel_curlr, vel_curlt, vel_curlp, mag_r, mag_t, mag_p, &
                                         mag_curlr, mag_curlt, mag_curlp, cod_C, comp_C, vel_Ntor, vel_npol)
      if (b_cod_tstep) call non_codensity(mes_oc, vel_r, vel_t, vel_p, &
                                          vel_curlr, vel_curlt, vel_curlp, &
                                          vel_dur, vel_qv1, vel_dutr, vel_dupr, vel_qv2, &
                                          mag_curlr, mag_curlt, mag_curlp, cod_gradr, cod_gradt, cod_gradp, &
                                          cod_S, cod_NC)
      if (b_comp_tstep) call non_composition(mes_oc, vel_r, vel_t, vel_p, &
                                             vel_curlr, vel_curlt, vel_curlp, &
                                             vel_dur, vel_qv1, vel_dutr, vel_dupr, vel_qv2, &
                                             mag_curlr, mag_curlt, mag_curlp, &
                                             comp_gradr, comp_gradt, comp_gradp, comp_S, comp_NC)
      if (b_mag_tstep) call non_magnetic(mes_oc, mes_ic, vel_r, vel_t, vel_p, mag_icTor, mag_icPol, &
                                         mag_r, mag_t, mag_p, rot_omega, mag_bcRe, mag_bcIm, &
                                         mag_NTor, mag_NPol, mag_icNTor, mag_icNPol)
      nev = nev + 1
      if (tim_it == 0) then
         if (d_timestep < 0d0) then
            call non_maxtstep(mes_oc, vel_r, vel_t, vel_p, mag_r, mag_t, mag_p)
            call tim_new_tstep()
            if (tim_new_dt) then
               call vel_matrices(mes_oc)
               call cod_matrices(mes_oc)
               call comp_matrices(mes_oc)
               call mag_matrices(mes_oc, mes_ic)
            end if
         end if
         if (b_rot_tstep) call rot_predictor(rot_inertia, rot_velTorq, rot_magTorq, rot_omega)
         if (b_vel_tstep) call vel_predictor(mes_oc, vel_ntor, vel_npol, rot_omega, vel_tor, vel_pol)
         if (b_cod_tstep) call cod_predictor(cod_Nc, cod_bcRe, cod_bcIm, cod_C)
         if (b_comp_tstep) call comp_predictor(comp_NC, comp_bcRe, comp_bcIm, comp_C)
         if (b_mag_tstep) call mag_predictor(mes_oc, mes_ic, mag_ntor, mag_npol, mag_icNTor, mag_icNpol, &
                                             mag_bcRe, mag_bcIm, mag_bc_PolRe, mag_bc_PolIm, &
                                             mag_tor, mag_pol, mag_icTor, mag_icPol)
         tim_it = 1
      else
         if (b_rot_tstep) call rot_corrector(rot_inertia, rot_velTorq, rot_magTorq, rot_omega)
         if (b_vel_tstep) call vel_corrector(mes_oc, vel_ntor, vel_npol, rot_omega, vel_tor, vel_pol)
         if (b_cod_tstep) call cod_corrector(cod_NC, cod_C)
         if (b_comp_tstep) call comp_corrector(comp_NC, comp_C)
         if (b_mag_tstep) call mag_corrector(mes_oc, mes_ic, mag_ntor, mag_npol, mag_icNTor, mag_icNPol, &
                                             mag_bcRe, mag_bcIm, mag_tor, mag_pol, mag_icTor, mag_icPol)
         accepted_err = tim_dterr
         call tim_check_cgce()
      end if
      if (tim_it == 0) then
         tim_t = tim_t + tim_dt
         tim_step = tim_step + 1
         if (.not. timing .and. (mod(tim_step, every) == 0 .or. tim_step == nsteps)) then
            write(uout) tim_step, nev, tim_t, accepted_err
            call write_fields()
         end if
         nev = 0
      end if
   end do
   call system_clock(c1, crate)
   if (timing) then
      write(uout) tim_step, nev, tim_t, accepted_err
      call write_fields()
      if (nsteps > 1) print '(a,es16.8)', 'sec/step ', dble(c1 - c0) / dble(crate) / dble(nsteps - 1)
   end if
   close(uout)

 contains

   subroutine read_coll(a)
      type(coll), intent(inout) :: a
      read(uin) a%Re(1:i_N, 0:i_H-1)
      read(uin) a%Im(1:i_N, 0:i_H-1)
   end subroutine read_coll

   subroutine write_fields()
      write(uout) vel_tor%Re(1:i_N,0:i_H-1), vel_tor%Im(1:i_N,0:i_H-1)
      write(uout) vel_pol%Re(1:i_N,0:i_H-1), vel_pol%Im(1:i_N,0:i_H-1)
      write(uout) cod_C%Re(1:i_N,0:i_H-1), cod_C%Im(1:i_N,0:i_H-1)
      write(uout) comp_C%Re(1:i_N,0:i_H-1), comp_C%Im(1:i_N,0:i_H-1)
      write(uout) mag_tor%Re(1:i_N,0:i_H-1), mag_tor%Im(1:i_N,0:i_H-1)
      write(uout) mag_pol%Re(1:i_N,0:i_H-1), mag_pol%Im(1:i_N,0:i_H-1)
      if (icio) then
         write(uout) mag_icTor%Re(1:i_Nic,0:i_H-1), mag_icTor%Im(1:i_Nic,0:i_H-1)
         write(uout) mag_icPol%Re(1:i_Nic,0:i_H-1), mag_icPol%Im(1:i_Nic,0:i_H-1)
         write(uout) rot_omega, rot_velTorq, rot_magTorq
      end if
   end subroutine write_fields

 end program ref_step
