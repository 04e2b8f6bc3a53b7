! ref_shim.f90 -- C-callable wrappers around the REFERENCE Fortran modules.
!
! TEST INFRASTRUCTURE ONLY.  The Makefile in this directory compiles the
! reference's own sources in place (they are never copied here) together with
! this file into _ref/libm2dref.so.  Tests use it to pin the C oracle and to
! write golden vectors; the CPU-baseline leg of the benchmark may time it.
!
! Each wrapper forwards its arguments unchanged; scalars come by value.
module m2dref_shim
   use iso_c_binding
   use msi2d9
   use msi2d5
   use mtdma2d9
   use mod_class_solver_msi
   use mod_solvers, only: norm_l1_9d, norm_l1_9d_relative, norm_l1_b
   use coefficients
   implicit none
contains

   ! src/solvers/msi2d9-v03.f90:277
   subroutine ref_lu2d9(b, dl, du, nij, ni, nj) bind(C, name="ref_lu2d9")
      integer(c_int), value :: nij, ni, nj
      real(c_double), intent(in)  :: b(nij,9)
      real(c_double), intent(out) :: dl(nij,5), du(nij,4)
      call lu2d9(b, dl, du, nij, ni, nj)
   end subroutine

   ! src/solvers/msi2d9-v03.f90:42
   subroutine ref_fb2d9(b, dl, du, nj, ni, nij, var, gamma, nitm, rhs) bind(C, name="ref_fb2d9")
      integer(c_int), value :: nj, ni, nij, nitm
      real(c_double), value :: gamma
      real(c_double), intent(in)    :: b(nij,9), dl(nij,5), du(nij,4), rhs(nij)
      real(c_double), intent(inout) :: var(nij)
      call fb2d9(b, dl, du, nj, ni, nij, var, gamma, nitm, rhs)
   end subroutine

   ! src/solvers/msi2d5-v03.f90:247
   subroutine ref_lu2d5(b, dl, du, nij, ni, nj) bind(C, name="ref_lu2d5")
      integer(c_int), value :: nij, ni, nj
      real(c_double), intent(in)  :: b(nij,5)
      real(c_double), intent(out) :: dl(nij,4), du(nij,3)
      call lu2d5(b, dl, du, nij, ni, nj)
   end subroutine

   ! src/solvers/msi2d5-v03.f90:42
   subroutine ref_fb2d5(b, dl, du, nj, ni, nij, var, gamma, nitm, rhs) bind(C, name="ref_fb2d5")
      integer(c_int), value :: nj, ni, nij, nitm
      real(c_double), value :: gamma
      real(c_double), intent(in)    :: b(nij,5), dl(nij,4), du(nij,3), rhs(nij)
      real(c_double), intent(inout) :: var(nij)
      call fb2d5(b, dl, du, nj, ni, nij, var, gamma, nitm, rhs)
   end subroutine

   ! src/solvers/tdma2d9.f90:38 (built without OpenMP: the line split does
   ! not change results)
   subroutine ref_tdma2d9(nx, ny, nitm, a, b, x) bind(C, name="ref_tdma2d9")
      integer(c_int), value :: nx, ny, nitm
      real(c_double), intent(in)    :: a(nx*ny,9), b(nx*ny)
      real(c_double), intent(inout) :: x(nx*ny)
      call tdma2d9(nx, ny, nitm, a, b, x)
   end subroutine

   ! src/solvers/mod_solver_msi.f90:129-166: the plugin object, init + solve
   subroutine ref_solver_msi9d(nx, ny, nitm, tol, a, b, sol) bind(C, name="ref_solver_msi9d")
      integer(c_int), value :: nx, ny, nitm
      real(c_double), value :: tol
      real(c_double), intent(in)    :: a(nx*ny,9), b(nx*ny)
      real(c_double), intent(inout) :: sol(nx*ny)
      type(class_solver_msi9d) :: s
      call s%init(nx, ny, nitm, tol)
      call s%solve(a, b, sol)
      deallocate(s%dl, s%du)
   end subroutine

   ! src/solvers/mod_solvers.f90:128, 176, 250
   subroutine ref_norm_l1_9d(nx, ny, var, b, a, norm) bind(C, name="ref_norm_l1_9d")
      integer(c_int), value :: nx, ny
      real(c_double), intent(in) :: var(nx*ny), b(nx*ny), a(nx*ny,9)
      real(c_double), intent(inout) :: norm
      call norm_l1_9d(nx, ny, var, b, a, norm)
   end subroutine

   subroutine ref_norm_l1_9d_relative(nx, ny, var, b, a, norm) bind(C, name="ref_norm_l1_9d_relative")
      integer(c_int), value :: nx, ny
      real(c_double), intent(in) :: var(nx*ny), b(nx*ny), a(nx*ny,9)
      real(c_double), intent(inout) :: norm
      call norm_l1_9d_relative(nx, ny, var, b, a, norm)
   end subroutine

   real(c_double) function ref_norm_l1_b(nx, ny, b) bind(C, name="ref_norm_l1_b")
      integer(c_int), value :: nx, ny
      real(c_double), intent(in) :: b(nx*ny)
      ref_norm_l1_b = norm_l1_b(nx, ny, b)
   end function

end module m2dref_shim
