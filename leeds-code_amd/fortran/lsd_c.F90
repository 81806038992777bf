!*************************************************************************
!> iso_c_binding interfaces to the MI355X per-timestep library
!! (include/lsd.h).  Fortran host code -- the drop-in modules beside this
!! file and the driver program -- calls the HIP kernels only through these
!! bind(C) declarations; arrays are passed as they lie in the reference's
!! types (modules/variables.F90:65-86): coll Re(N,0:H-1), spec Re(0:H-1,N),
!! phys Re(Th,0:Ph-1,N).
!*************************************************************************
 module lsd_c
   use iso_c_binding
   implicit none

   integer(c_int), parameter :: LSD_OK = 0, LSD_E_CONFIG = -1, LSD_E_HIP = -2, &
                                LSD_E_SINGULAR = -3, LSD_E_ARG = -4, LSD_E_DIVERGED = -5
   integer(c_int), parameter :: LSD_PLAIN_F64 = 0, LSD_MIXED_F32 = 1, LSD_MIXED_BF16X3 = 2
   integer(c_int), parameter :: LSD_VEL_TOR = 0, LSD_VEL_POL = 1, LSD_COD = 2, LSD_COMP = 3, &
                                LSD_MAG_TOR = 4, LSD_MAG_POL = 5, LSD_N_VEL_TOR = 6,          &
                                LSD_N_VEL_POL = 7, LSD_N_COD = 8, LSD_N_COMP = 9,             &
                                LSD_N_MAG_TOR = 10, LSD_N_MAG_POL = 11
   integer(c_int), parameter :: LSD_T_GP = 0, LSD_T_GW = 1, LSD_T_P = 2, LSD_T_P_ = 3, LSD_T_SIN1MP = 4, &
                                LSD_T_LLP_ = 5, LSD_T_LLSIN1MP = 6, LSD_T_GWP = 7, LSD_T_GWLLP_ = 8,   &
                                LSD_T_GWLLSIN1MP = 9
   integer(c_int), parameter :: LSD_M_R = 0, LSD_M_INTR2DR = 5

   !> lsd_config: run parameters (Scripts/parameters.F90_template names)
   type, bind(C) :: lsd_config
      integer(c_int) :: L, M, Mp, N, Th, Ph
      integer(c_int) :: vel, cod, comp, mag
      integer(c_int) :: vel_bc, cod_bc, comp_bc
      integer(c_int) :: precision
      real(c_double) :: rratio, Ra, Ra_comp, Pr, Sc, Pm, E, dt, implicit, dterr, source, source_comp
      integer(c_int) :: step_control
      integer(c_int) :: Nic, rot
   end type lsd_config

   type, bind(C) :: lsd_step_stats
      integer(c_int) :: steps, evaluations, last_its, max_its
      real(c_double) :: last_dterr, t
      integer(c_int) :: status
   end type lsd_step_stats

   interface
      function lsd_last_error() bind(C, name="lsd_last_error") result(p)
         import :: c_ptr
         type(c_ptr) :: p
      end function
      function lsd_config_sizes(cfg, H, Th, Ph) bind(C, name="lsd_config_sizes") result(rc)
         import :: c_int, lsd_config
         type(lsd_config), intent(in) :: cfg
         integer(c_int), intent(out) :: H, Th, Ph
         integer(c_int) :: rc
      end function
      function lsd_create(cfg, device, ctx) bind(C, name="lsd_create") result(rc)
         import :: c_int, c_ptr, lsd_config
         type(lsd_config), intent(in) :: cfg
         integer(c_int), value :: device
         type(c_ptr), intent(out) :: ctx
         integer(c_int) :: rc
      end function
      subroutine lsd_destroy(ctx) bind(C, name="lsd_destroy")
         import :: c_ptr
         type(c_ptr), value :: ctx
      end subroutine
      function lsd_set_coll(ctx, field, re, im) bind(C, name="lsd_set_coll") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: field
         real(c_double), intent(in) :: re(*), im(*)
         integer(c_int) :: rc
      end function
      function lsd_get_coll(ctx, field, re, im) bind(C, name="lsd_get_coll") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: field
         real(c_double), intent(out) :: re(*), im(*)
         integer(c_int) :: rc
      end function
      function lsd_get_phys(ctx, field, p) bind(C, name="lsd_get_phys") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: field
         real(c_double), intent(out) :: p(*)
         integer(c_int) :: rc
      end function
      function lsd_get_table(ctx, table, out) bind(C, name="lsd_get_table") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: table
         real(c_double), intent(out) :: out(*)
         integer(c_int) :: rc
      end function
      function lsd_get_mesh(ctx, which, out) bind(C, name="lsd_get_mesh") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: which
         real(c_double), intent(out) :: out(*)
         integer(c_int) :: rc
      end function
      function lsd_transform(ctx) bind(C, name="lsd_transform") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int) :: rc
      end function
      function lsd_nonlinear(ctx) bind(C, name="lsd_nonlinear") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int) :: rc
      end function
      function lsd_predictor(ctx) bind(C, name="lsd_predictor") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int) :: rc
      end function
      function lsd_corrector(ctx, dterr) bind(C, name="lsd_corrector") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: dterr
         integer(c_int) :: rc
      end function
      function lsd_step(ctx, nsteps, stats) bind(C, name="lsd_step") result(rc)
         import :: c_int, c_ptr, lsd_step_stats
         type(c_ptr), value :: ctx
         integer(c_int), value :: nsteps
         type(lsd_step_stats), intent(out) :: stats
         integer(c_int) :: rc
      end function
      function lsd_set_graph(ctx, on) bind(C, name="lsd_set_graph") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: on
         integer(c_int) :: rc
      end function
      function lsd_set_dt(ctx, dt) bind(C, name="lsd_set_dt") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: dt
         integer(c_int) :: rc
      end function
      function lsd_tra_spec2phys(ctx, s_re, s_im, p) bind(C, name="lsd_tra_spec2phys") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: s_re(*), s_im(*)
         real(c_double), intent(out) :: p(*)
         integer(c_int) :: rc
      end function
      function lsd_tra_phys2spec(ctx, p, s_re, s_im) bind(C, name="lsd_tra_phys2spec") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: p(*)
         real(c_double), intent(out) :: s_re(*), s_im(*)
         integer(c_int) :: rc
      end function
      function lsd_tra_grad(ctx, s_re, s_im, t, p) bind(C, name="lsd_tra_grad") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: s_re(*), s_im(*)
         real(c_double), intent(out) :: t(*), p(*)
         integer(c_int) :: rc
      end function
      function lsd_tra_qst2rtp(ctx, q_re, q_im, s_re, s_im, t_re, t_im, r, th, ph) &
            bind(C, name="lsd_tra_qst2rtp") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: q_re(*), q_im(*), s_re(*), s_im(*), t_re(*), t_im(*)
         real(c_double), intent(out) :: r(*), th(*), ph(*)
         integer(c_int) :: rc
      end function
      function lsd_tra_rtp2qst(ctx, r, th, ph, q_re, q_im, s_re, s_im, t_re, t_im) &
            bind(C, name="lsd_tra_rtp2qst") result(rc)
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: r(*), th(*), ph(*)
         real(c_double), intent(out) :: q_re(*), q_im(*), s_re(*), s_im(*), t_re(*), t_im(*)
         integer(c_int) :: rc
      end function
      function lsd_sync(ctx) bind(C, name="lsd_sync") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int) :: rc
      end function
      function c_strlen(s) bind(C, name="strlen") result(n)
         import :: c_ptr, c_size_t
         type(c_ptr), value :: s
         integer(c_size_t) :: n
      end function
   end interface

 contains

   !> Message of the last failed call, as a Fortran string.
   function lsd_error_text() result(msg)
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: cs(:)
      integer :: n, i
      p = lsd_last_error()
      n = int(c_strlen(p))
      call c_f_pointer(p, cs, [n])
      allocate(character(len=n) :: msg)
      do i = 1, n
         msg(i:i) = cs(i)
      end do
   end function lsd_error_text

   !> Stops with the library's message on a non-zero code, as the reference
   !! stops on its own errors (e.g. modules/meshs.F90:454).
   subroutine lsd_check(rc, where)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      if (rc /= LSD_OK) then
         print *, where, ': ', lsd_error_text()
         stop 'lsd'
      end if
   end subroutine lsd_check

 end module lsd_c
