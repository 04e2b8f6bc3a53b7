! tstepper_scarborough_hip.f90 -- drop-in replacement of module
! mod_tstepper_scarborough (src/tstepper/mod_tstepper_scarborough.f90): the
! same module name and public procedures; the statistics run on the GPU
! through the C ABI of libmach2d_hip.so (include/mach2d_hip.h,
! m2d_tstepper_dt0_scarborough / m2d_tstepper_dt_scarborough).  The module
! variables h0, mincc and maxcc are read by tstepper_scarborough_init from the
! parameter file as in the reference.
module mod_tstepper_scarborough
   use iso_c_binding, only: c_int, c_double
   use mod_class_ifile
   implicit none
   private
   public :: tstepper_scarborough_init, tstepper_dt0_scarborough, tstepper_dt_scarborough

   real(8), private :: h0, mincc, maxcc

   interface
      real(c_double) function c_dt0(nx, ny, Jp, u, v) bind(C, name="m2d_tstepper_dt0_scarborough")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: Jp(*), u(*), v(*)
      end function
      real(c_double) function c_dt(nx, ny, dt, au, av, at, ap, h0, mincc, maxcc) &
            bind(C, name="m2d_tstepper_dt_scarborough")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: dt, h0, mincc, maxcc
         real(c_double) :: au(*), av(*), at(*), ap(*)
      end function
   end interface

contains

   ! mod_tstepper_scarborough.f90:28-36
   subroutine tstepper_scarborough_init(ifile)
      class(class_ifile), intent(in) :: ifile
      call ifile%get_value(h0, "h0")
      call ifile%get_value(mincc, "mincc")
      call ifile%get_value(maxcc, "maxcc")
   end subroutine

   ! :40-76
   real(8) function tstepper_dt0_scarborough(nx, ny, Jp, u, v)
      integer, intent(in) :: nx, ny
      real(8), dimension(nx*ny), intent(in) :: Jp, u, v
      tstepper_dt0_scarborough = c_dt0(nx, ny, Jp, u, v)
   end function

   ! :80-123
   real(8) function tstepper_dt_scarborough(nx, ny, dt, au, av, at, ap)
      integer, intent(in) :: nx, ny
      real(8), intent(in) :: dt
      real(8), dimension(nx*ny,9), intent(in) :: au, av, at
      real(8), dimension(nx*ny,5), intent(in) :: ap
      tstepper_dt_scarborough = c_dt(nx, ny, dt, au, av, at, ap, h0, mincc, maxcc)
   end function

end module mod_tstepper_scarborough
