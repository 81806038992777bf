!*************************************************************************
!> Drop-in replacement of modules/legendre.F90 for the serial build: the
!! same public arrays (modules/legendre.F90:41-56) and leg_precompute
!! (:63-128), with the Gauss nodes and weights and the Schmidt tables
!! computed on the MI355X by the library's FP64 recursion (the same
!! GAUWTS Newton iteration and SCHNLA recurrences, equal to the
!! reference's values to rounding) and copied into the host arrays that
!! the reference's other modules read (nonlinear.F90 reads leg_sin and
!! leg_gp).  leg_qva and leg_qvb belong to the anelastic branch, which the
!! drop-in transform does not provide; they are left zero.
!*************************************************************************
 module legendre
   use parameters, only : i_Th, i_Th2, i_H1
   use lsd_c
   implicit none
   save

   double precision :: leg_th  (i_Th)
   double precision :: leg_sin (i_Th)
   double precision :: leg_sin1(i_Th)
   double precision :: leg_gp  (i_Th)
   double precision :: leg_gw  (i_Th)
   double precision :: leg_gwP0(i_Th)
   double precision :: leg_P         (i_Th2, 0:i_H1)
   double precision :: leg_P_        (i_Th2, 0:i_H1)
   double precision :: leg_sin1mP    (i_Th2, 0:i_H1)
   double precision :: leg_llP_      (i_Th2, 0:i_H1)
   double precision :: leg_llsin1mP  (i_Th2, 0:i_H1)
   double precision :: leg_qva       (i_Th2, 0:i_H1)
   double precision :: leg_qvb       (i_Th2, 0:i_H1)
   double precision :: leg_gwP       (i_Th2, 0:i_H1)
   double precision :: leg_gwllP_    (i_Th2, 0:i_H1)
   double precision :: leg_gwllsin1mP(i_Th2, 0:i_H1)

 contains

   subroutine leg_precompute()
      use lsd_dropin, only : lsd_dropin_context
      type(c_ptr) :: ctx
      integer :: j
      ctx = lsd_dropin_context()
      call lsd_check(lsd_get_table(ctx, LSD_T_GP, leg_gp), 'leg_precompute')
      call lsd_check(lsd_get_table(ctx, LSD_T_GW, leg_gw), 'leg_precompute')
      call lsd_check(lsd_get_table(ctx, LSD_T_P, leg_P), 'leg_precompute')
      call lsd_check(lsd_get_table(ctx, LSD_T_P_, leg_P_), 'leg_precompute')
      call lsd_check(lsd_get_table(ctx, LSD_T_SIN1MP, leg_sin1mP), 'leg_precompute')
      call lsd_check(lsd_get_table(ctx, LSD_T_LLP_, leg_llP_), 'leg_precompute')
      call lsd_check(lsd_get_table(ctx, LSD_T_LLSIN1MP, leg_llsin1mP), 'leg_precompute')
      call lsd_check(lsd_get_table(ctx, LSD_T_GWP, leg_gwP), 'leg_precompute')
      call lsd_check(lsd_get_table(ctx, LSD_T_GWLLP_, leg_gwllP_), 'leg_precompute')
      call lsd_check(lsd_get_table(ctx, LSD_T_GWLLSIN1MP, leg_gwllsin1mP), 'leg_precompute')
      leg_th = acos(leg_gp)
      leg_sin = sin(leg_th)
      leg_sin1 = 1d0 / leg_sin
      ! P_0^0 = 1 (Schmidt normalisation), so gwP0 = w / 2, mirrored about
      ! the equator
      do j = 1, i_Th2
         leg_gwP0(j) = leg_gw(j) * 0.5d0
         leg_gwP0(i_Th + 1 - j) = leg_gwP0(j)
      end do
      leg_qva = 0d0
      leg_qvb = 0d0
   end subroutine leg_precompute

 end module legendre
