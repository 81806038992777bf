!*************************************************************************
!> Drop-in replacement of modules/transform.fftw3.F90 (selected as
!! TRANSFORM = hip in the reference's Makefile, Scripts/Makefile_template:9)
!! for the serial build (_Nth = _Nr = 1): the same public procedures with
!! the same dummy arguments, the transforms running on the MI355X through
!! the C-ABI (module lsd_c).  The reference's modules parameters and
!! variables provide the sizes and the spec/phys types; the context is
!! created by tra_precompute and kept for the life of the program.
!*************************************************************************
 module transform
   use parameters, only : i_L, i_M, i_Mp, i_N, i_Th, i_Ph, d_rratio, d_Ra, d_Ra_comp, d_Pr, d_Sc, &
                          d_Pm, d_E, d_timestep, d_implicit, d_dterr, d_source, d_source_comp,      &
                          b_cod_tstep, b_comp_tstep, b_mag_tstep, i_vel_bc, i_cod_bc, i_comp_bc, &
                          i_time_step_control
   use lsd_c
   implicit none
   save
   type(c_ptr), private :: tra_ctx = c_null_ptr
   double precision, private :: tra_gw(i_Th)

 contains

!------------------------------------------------------------------------
!> modules/transform.fftw3.F90:65-98.  priority (FFTW planning effort) has
!! no counterpart: the GPU transforms need no plans.
!------------------------------------------------------------------------
   subroutine tra_precompute(priority)
      integer, intent(in) :: priority
      type(lsd_config) :: cfg
      if (c_associated(tra_ctx)) return
      cfg%L = i_L; cfg%M = i_M; cfg%Mp = i_Mp; cfg%N = i_N; cfg%Th = i_Th; cfg%Ph = i_Ph
      cfg%vel = 1
      cfg%cod = merge(1, 0, b_cod_tstep)
      cfg%comp = merge(1, 0, b_comp_tstep)
      cfg%mag = merge(1, 0, b_mag_tstep)
      cfg%vel_bc = i_vel_bc; cfg%cod_bc = i_cod_bc; cfg%comp_bc = i_comp_bc
      cfg%precision = LSD_PLAIN_F64
      cfg%rratio = d_rratio; cfg%Ra = d_Ra; cfg%Ra_comp = d_Ra_comp; cfg%Pr = d_Pr; cfg%Sc = d_Sc
      cfg%Pm = d_Pm; cfg%E = d_E; cfg%dt = abs(d_timestep); cfg%implicit = d_implicit
      cfg%dterr = d_dterr; cfg%source = d_source; cfg%source_comp = d_source_comp
      cfg%step_control = i_time_step_control
      call lsd_check(lsd_create(cfg, 0_c_int, tra_ctx), 'tra_precompute')
      call lsd_check(lsd_get_table(tra_ctx, LSD_T_GW, tra_gw), 'tra_precompute')
   end subroutine tra_precompute

!> The library context, for the drop-in nonlinear and timestep modules.
   function tra_context() result(ctx)
      type(c_ptr) :: ctx
      ctx = tra_ctx
   end function tra_context

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
