!*************************************************************************
!> The library context shared by the drop-in modules (legendre, transform)
!! of the serial build (_Nth = _Nr = 1).  The first drop-in to need it --
!! leg_precompute, which program/main.F90:261-264 calls before
!! tra_precompute -- creates it from the reference's parameters module;
!! it is kept for the life of the program, as the reference keeps its
!! module arrays.
!*************************************************************************
 module lsd_dropin
   use lsd_c
   implicit none
   save
   type(c_ptr), private :: ctx = c_null_ptr

 contains

   function lsd_dropin_context() result(c)
      use parameters, only : i_L, i_M, i_Mp, i_N, i_Th, i_Ph, d_rratio, d_Ra, d_Ra_comp, d_Pr, d_Sc, &
                             d_Pm, d_E, d_timestep, d_implicit, d_dterr, d_source, d_source_comp,      &
                             b_cod_tstep, b_comp_tstep, b_mag_tstep, i_vel_bc, i_cod_bc, i_comp_bc, &
                             i_time_step_control
      type(c_ptr) :: c
      type(lsd_config) :: cfg
      if (.not. c_associated(ctx)) then
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
         cfg%Nic = 1; cfg%rot = 0
         call lsd_check(lsd_create(cfg, 0_c_int, ctx), 'lsd_dropin_context')
      end if
      c = ctx
   end function lsd_dropin_context

 end module lsd_dropin
