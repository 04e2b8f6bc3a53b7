! msi2d5_hip.f90 -- drop-in Fortran module for the 5-diagonal MSI.
!
! Build this file in place of the reference's src/solvers/msi2d5-v03.f90 and
! link libmach2d_hip.so: the module name (msi2d5), routine names and dummy
! arguments are the reference's, so mod_solver_msi.f90 (class_solver_msi5d)
! and every caller compile unchanged.  Each routine forwards to the C ABI of
! include/mach2d_hip.h (all arguments by reference), which runs the gfx950
! kernels; the status of the last call is read with m2d_get_status from
! module mach2d_hip_status (msi2d9_hip.f90).

! Replaces module msi2d5 (src/solvers/msi2d5-v03.f90).
module msi2d5
   use iso_c_binding
   implicit none

   interface
      ! include/mach2d_hip.h: m2d_lu2d5 <- lu2d5, msi2d5-v03.f90:247-330
      subroutine m2d_lu2d5(b, dl, du, nij, ni, nj) bind(C, name="m2d_lu2d5")
         import :: c_int, c_double
         integer(c_int), intent(in)  :: nij, ni, nj
         real(c_double), intent(in)  :: b(nij, 5)
         real(c_double), intent(out) :: dl(nij, 4), du(nij, 3)
      end subroutine

      ! include/mach2d_hip.h: m2d_fb2d5 <- fb2d5, msi2d5-v03.f90:42-233
      subroutine m2d_fb2d5(b, dl, du, nj, ni, nij, var, gamma, nitm, rhs) &
            bind(C, name="m2d_fb2d5")
         import :: c_int, c_double
         integer(c_int), intent(in)    :: nj, ni, nij, nitm
         real(c_double), intent(in)    :: gamma
         real(c_double), intent(in)    :: b(nij, 5), dl(nij, 4), du(nij, 3), rhs(nij)
         real(c_double), intent(inout) :: var(nij)
      end subroutine
   end interface

contains

   subroutine fb2d5(b, dl, du, nj, ni, nij, var, gamma, nitm, rhs)
      implicit none
      integer, intent(in)    :: ni, nj, nij, nitm
      real(8), intent(in)    :: gamma
      real(8), intent(in)    :: b(nij, 5)
      real(8), intent(in)    :: dl(nij, 4)
      real(8), intent(in)    :: du(nij, 3)
      real(8), intent(in)    :: rhs(nij)
      real(8), intent(inout) :: var(nij)
      call m2d_fb2d5(b, dl, du, nj, ni, nij, var, gamma, nitm, rhs)
   end subroutine fb2d5

   subroutine lu2d5(b, dl, du, nij, ni, nj)
      implicit none
      integer, intent(in)  :: ni, nj, nij
      real(8), intent(in)  :: b(nij, 5)
      real(8), intent(out) :: dl(nij, 4)
      real(8), intent(out) :: du(nij, 3)
      call m2d_lu2d5(b, dl, du, nij, ni, nj)
   end subroutine lu2d5

end module msi2d5
