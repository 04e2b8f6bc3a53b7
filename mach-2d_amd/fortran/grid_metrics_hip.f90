! grid_metrics_hip.f90 -- the last three steps of grid_create
! (src/grid/mod_grid.f90:174-187) on the GPU: get_real_centroids_xy
! (:237-300), get_metrics (:304-558) and get_radius (:562-723), with the
! reference's argument lists, through the C ABI of libmach2d_hip.so
! (include/mach2d_hip.h).  mod_grid's grid_create calls them unchanged after
! `use grid_metrics_hip` in place of its own module procedures.  The status
! of the last call is read with m2d_get_status from module mach2d_hip_status
! (msi2d9_hip.f90).
module grid_metrics_hip
   use iso_c_binding, only: c_int, c_double
   implicit none
   private
   public :: get_real_centroids_xy, get_metrics, get_radius

   interface
      subroutine c_centroids(opt, nx, ny, x, y, xp, yp) bind(C, name="m2d_get_real_centroids_xy")
         import :: c_int, c_double
         integer(c_int) :: opt, nx, ny
         real(c_double) :: x(*), y(*), xp(*), yp(*)
      end subroutine
      subroutine c_metrics(nx, ny, x, y, xp, yp, xe, ye, xen, yen, xk, yk, xke, yke, Jp, &
            Je, Jn, alphae, gamman, betae, betan) bind(C, name="m2d_get_metrics")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: x(*), y(*), xp(*), yp(*), xe(*), ye(*), xen(*), yen(*), xk(*), &
            yk(*), xke(*), yke(*), Jp(*), Je(*), Jn(*), alphae(*), gamman(*), betae(*), betan(*)
      end subroutine
      subroutine c_radius(coord, nx, ny, y, yp, radius, re, rn, rp) &
            bind(C, name="m2d_get_radius")
         import :: c_int, c_double
         integer(c_int) :: coord, nx, ny
         real(c_double) :: y(*), yp(*), radius(*), re(*), rn(*), rp(*)
      end subroutine
   end interface

contains

   ! mod_grid.f90:237-300 (opt: 1 simple mean, 2 weighted mean)
   subroutine get_real_centroids_xy(opt, nx, ny, x, y, xp, yp)
      integer, intent(in) :: opt, nx, ny
      real(8), intent(in) :: x(nx*ny), y(nx*ny)
      real(8), intent(out) :: xp(nx*ny), yp(nx*ny)
      call c_centroids(opt, nx, ny, x, y, xp, yp)
   end subroutine

   ! mod_grid.f90:304-558
   subroutine get_metrics(nx, ny, x, y, xp, yp, xe, ye, xen, yen, xk, yk, xke, yke, Jp, &
         Je, Jn, alphae, gamman, betae, betan)
      integer, intent(in) :: nx, ny
      real(8), dimension(nx*ny), intent(in) :: x, y, xp, yp
      real(8), dimension(nx*ny), intent(out) :: xe, ye, xen, yen, xk, yk, xke, yke, Jp, Je, Jn
      real(8), dimension(nx*ny), intent(out) :: alphae, gamman, betae, betan
      call c_metrics(nx, ny, x, y, xp, yp, xe, ye, xen, yen, xk, yk, xke, yke, Jp, Je, Jn, &
         alphae, gamman, betae, betan)
   end subroutine

   ! mod_grid.f90:562-723 (coord: 1 axisymmetric, 0 planar)
   subroutine get_radius(coord, nx, ny, y, yp, radius, re, rn, rp)
      integer, intent(in) :: coord, nx, ny
      real(8), dimension(nx*ny), intent(in) :: y, yp
      real(8), dimension(nx*ny), intent(out) :: radius, re, rn, rp
      call c_radius(coord, nx, ny, y, yp, radius, re, rn, rp)
   end subroutine

end module grid_metrics_hip
