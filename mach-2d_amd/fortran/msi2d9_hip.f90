! msi2d9_hip.f90 -- drop-in Fortran modules for the MI355X hot path.
!
! Build this file in place of the reference's src/solvers/msi2d9-v03.f90 and
! src/solvers/tdma2d9.f90 and link libmach2d_hip.so: the module names
! (msi2d9, mtdma2d9), routine names and dummy arguments are the reference's,
! so mod_solver_msi.f90, mod_solver_tdma.f90 and every caller compile
! unchanged.  Each routine forwards to the C ABI of include/mach2d_hip.h
! (all arguments by reference), which runs the gfx950 kernels.
!
! The reference routines return nothing; the status of the last call
! (iterations, residual norms, error code) is available through
! m2d_get_status in module mach2d_hip_status.

module mach2d_hip_status
   use iso_c_binding
   implicit none

   type, bind(C) :: m2d_status
      integer(c_int) :: status
      integer(c_int) :: nit
      integer(c_int) :: nitm_reached
      integer(c_int) :: outer
      real(c_double) :: r0
      real(c_double) :: rn
      real(c_double) :: gammarp
      real(c_double) :: device_ms
   end type

   interface
      integer(c_int) function m2d_last_status(st) bind(C, name="m2d_last_status")
         import :: c_int, m2d_status
         type(m2d_status), intent(out) :: st
      end function
   end interface

contains

   !> Status of the last hot-path call of this thread.
   subroutine m2d_get_status(st)
      type(m2d_status), intent(out) :: st
      integer(c_int) :: rc
      rc = m2d_last_status(st)
   end subroutine

end module mach2d_hip_status


! Replaces module msi2d9 (src/solvers/msi2d9-v03.f90).
module msi2d9
   use iso_c_binding
   implicit none

   interface
      ! include/mach2d_hip.h: m2d_lu2d9 <- lu2d9, msi2d9-v03.f90:277-282
      subroutine m2d_lu2d9(b, dl, du, nij, ni, nj) bind(C, name="m2d_lu2d9")
         import :: c_int, c_double
         integer(c_int), intent(in)  :: nij, ni, nj
         real(c_double), intent(in)  :: b(nij, 9)
         real(c_double), intent(out) :: dl(nij, 5), du(nij, 4)
      end subroutine

      ! include/mach2d_hip.h: m2d_fb2d9 <- fb2d9, msi2d9-v03.f90:42-50
      subroutine m2d_fb2d9(b, dl, du, nj, ni, nij, var, gamma, nitm, rhs) &
            bind(C, name="m2d_fb2d9")
         import :: c_int, c_double
         integer(c_int), intent(in)    :: nj, ni, nij, nitm
         real(c_double), intent(in)    :: gamma
         real(c_double), intent(in)    :: b(nij, 9), dl(nij, 5), du(nij, 4), rhs(nij)
         real(c_double), intent(inout) :: var(nij)
      end subroutine
   end interface

contains

   subroutine fb2d9(b, dl, du, nj, ni, nij, var, gamma, nitm, rhs)
      implicit none
      integer, intent(in)    :: ni, nj, nij, nitm
      real(8), intent(in)    :: gamma
      real(8), intent(in)    :: b(nij, 9)
      real(8), intent(in)    :: dl(nij, 5)
      real(8), intent(in)    :: du(nij, 4)
      real(8), intent(in)    :: rhs(nij)
      real(8), intent(inout) :: var(nij)
      call m2d_fb2d9(b, dl, du, nj, ni, nij, var, gamma, nitm, rhs)
   end subroutine fb2d9

   subroutine lu2d9(b, dl, du, nij, ni, nj)
      implicit none
      integer, intent(in)  :: ni, nj, nij
      real(8), intent(in)  :: b(nij, 9)
      real(8), intent(out) :: dl(nij, 5)
      real(8), intent(out) :: du(nij, 4)
      call m2d_lu2d9(b, dl, du, nij, ni, nj)
   end subroutine lu2d9

end module msi2d9


! Replaces module mtdma2d9 (src/solvers/tdma2d9.f90).
module mtdma2d9
   use iso_c_binding
   implicit none

   ! Kept for interface compatibility (tdma2d9.f90:17-35); the GPU solves
   ! every line of a pass in parallel regardless of this value.
   integer, private :: tdma2d9_num_threads = 2

   interface
      ! include/mach2d_hip.h: m2d_tdma2d9 <- tdma2d9, tdma2d9.f90:38-45
      subroutine m2d_tdma2d9(nx, ny, nitm, a, b, x) bind(C, name="m2d_tdma2d9")
         import :: c_int, c_double
         integer(c_int), intent(in)    :: nx, ny, nitm
         real(c_double), intent(in)    :: a(nx*ny, 9), b(nx*ny)
         real(c_double), intent(inout) :: x(nx*ny)
      end subroutine
   end interface

contains

   subroutine set_tdma2d9_num_threads(num)
      implicit none
      integer, intent(in) :: num
      tdma2d9_num_threads = num
   end subroutine set_tdma2d9_num_threads

   integer function get_tdma2d9_num_threads()
      implicit none
      get_tdma2d9_num_threads = tdma2d9_num_threads
   end function get_tdma2d9_num_threads

   subroutine tdma2d9(nx, ny, nitm, a, b, x)
      implicit none
      integer, intent(in) :: nx
      integer, intent(in) :: ny
      integer, intent(in) :: nitm
      real(8), dimension(nx*ny, 9), intent(in)    :: a
      real(8), dimension(nx*ny),    intent(in)    :: b
      real(8), dimension(nx*ny),    intent(inout) :: x
      call m2d_tdma2d9(nx, ny, nitm, a, b, x)
   end subroutine tdma2d9

end module mtdma2d9
