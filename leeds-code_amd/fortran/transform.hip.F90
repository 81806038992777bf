!*************************************************************************
!> Drop-in replacement of modules/transform.fftw3.F90 (selected as
!! TRANSFORM = hip in the reference's Makefile, Scripts/Makefile_template:9)
!! for the serial build (_Nth = _Nr = 1): the same public procedures with
!! the same dummy arguments, the transforms running on the MI355X through
!! the C-ABI (module lsd_c).  The reference's modules parameters and
!! variables provide the sizes and the spec/phys types; the library context
!! is the one shared with the drop-in legendre module (module lsd_dropin).
!*************************************************************************
 module transform
   use parameters, only : i_Th, i_Ph
   use lsd_c
   implicit none
   save
   type(c_ptr), private :: tra_ctx = c_null_ptr
   double precision, private :: tra_gw(i_Th)

 contains

!------------------------------------------------------------------------
!> modules/transform.fftw3.F90:65-98.  priority (FFTW planning effort) has
!! no counterpart: the GPU transforms need no plans.  The context is the
!! one the drop-in modules share (created by leg_precompute if that came
!! first).
!------------------------------------------------------------------------
   subroutine tra_precompute(priority)
      use lsd_dropin, only : lsd_dropin_context
      integer, intent(in) :: priority
      if (c_associated(tra_ctx)) return
      tra_ctx = lsd_dropin_context()
      call lsd_check(lsd_get_table(tra_ctx, LSD_T_GW, tra_gw), 'tra_precompute')
   end subroutine tra_precompute

!------------------------------------------------------------------------
!> modules/transform.fftw3.F90:472-520
!------------------------------------------------------------------------
   subroutine tra_spec2phys(s, p)
      use variables, only : spec, phys
      type(spec), intent(in)  :: s
      type(phys), intent(out) :: p
      call lsd_check(lsd_tra_spec2phys(tra_ctx, s%Re, s%Im, p%Re), 'tra_spec2phys')
   end subroutine tra_spec2phys

!------------------------------------------------------------------------
!> modules/transform.fftw3.F90:527-583
!------------------------------------------------------------------------
   subroutine tra_phys2spec(p, s)
      use variables, only : spec, phys
      type(phys), intent(in)  :: p
      type(spec), intent(out) :: s
      call lsd_check(lsd_tra_phys2spec(tra_ctx, p%Re, s%Re, s%Im), 'tra_phys2spec')
   end subroutine tra_phys2spec

!------------------------------------------------------------------------
!> modules/transform.fftw3.F90:719-818
!------------------------------------------------------------------------
   subroutine tra_qst2rtp(q, s, t, r, th, ph)
      use variables, only : spec, phys
      type(spec), intent(in)  :: q, s, t
      type(phys), intent(out) :: r, th, ph
      call lsd_check(lsd_tra_qst2rtp(tra_ctx, q%Re, q%Im, s%Re, s%Im, t%Re, t%Im, r%Re, th%Re, ph%Re), &
                     'tra_qst2rtp')
   end subroutine tra_qst2rtp

!------------------------------------------------------------------------
!> modules/transform.fftw3.F90:611-717, the anelastic branch of
!! vel_transform (modules/velocity.F90:336-339), which this drop-in does
!! not provide; present so that the reference's velocity module compiles
!! against it unchanged.  The Boussinesq branch calls tra_qst2rtp.
!------------------------------------------------------------------------
   subroutine tra_qst2rtp_vel(mes_oc, q, s, t, qr, sr, tr, r, th, ph, dur, qv1, dutr, dupr, qv2)
      use meshs, only : rdom
      use variables, only : spec, phys
      type(rdom), intent(in)  :: mes_oc
      type(spec), intent(in)  :: q, s, t, qr, sr, tr
      type(phys), intent(out) :: r, th, ph, dur, qv1, dutr, dupr, qv2
      stop 'tra_qst2rtp_vel: the anelastic branch is not part of the MI355X drop-in'
   end subroutine tra_qst2rtp_vel

!------------------------------------------------------------------------
!> modules/transform.fftw3.F90:825-949
!------------------------------------------------------------------------
   subroutine tra_rtp2qst(r, th, ph, q, s, t)
      use variables, only : spec, phys
      type(phys), intent(in)  :: r, th, ph
      type(spec), intent(out) :: q, s, t
      call lsd_check(lsd_tra_rtp2qst(tra_ctx, r%Re, th%Re, ph%Re, q%Re, q%Im, s%Re, s%Im, t%Re, t%Im), &
                     'tra_rtp2qst')
   end subroutine tra_rtp2qst

!------------------------------------------------------------------------
!> modules/transform.fftw3.F90:955-1000
!------------------------------------------------------------------------
   subroutine tra_grad(s, t, p)
      use variables, only : spec, phys
      type(spec), intent(in)  :: s
      type(phys), intent(out) :: t, p
      call lsd_check(lsd_tra_grad(tra_ctx, s%Re, s%Im, t%Re, p%Re), 'tra_grad')
   end subroutine tra_grad

!------------------------------------------------------------------------
!> modules/transform.fftw3.F90:1039-1062 (host side: a diagnostic, not on
!! the timestep path): 4 pi / Ph sum_j gwP0_j sum_m p(j,m), gwP0 = w_j / 2
!! (modules/legendre.F90:103-104).
!------------------------------------------------------------------------
   subroutine tra_surfintegral(p, S)
      double precision, intent(in)  :: p(i_Th, 0:i_Ph-1)
      double precision, intent(out) :: S
      double precision :: Y(i_Th)
      integer :: m
      Y = 0d0
      do m = 0, i_Ph-1
         Y = Y + p(:, m)
      end do
      S = dot_product(Y, 0.5d0 * tra_gw)
      S = S * 4d0 * acos(-1d0) / dble(i_Ph)
   end subroutine tra_surfintegral

 end module transform
