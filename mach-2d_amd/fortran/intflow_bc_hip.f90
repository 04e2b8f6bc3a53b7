! intflow_bc_hip.f90 -- the fictitious-row writers of the internal-flow case
! (src/user/internal_flow/mod_intflow.f90: intflow_set_bcu, intflow_set_bcv,
! intflow_set_bcpl, intflow_set_bcT), the boundary-face velocities
! (intflow_velocities_at_boundary_faces), the boundary SIMPLEC coefficients,
! the extrapolation of u, v to fictitious volumes, the boundary temperature
! from total-enthalpy conservation and the time-cycle update of the inflow,
! on the GPU, through the C ABI of
! libmach2d_hip.so (include/mach2d_hip.h, m2d_intflow_set_*).  Each routine
! keeps the reference's dummy arguments, except that the derived-type
! argument `iflow` is replaced by the fields the routine reads (plin; Tin and
! Twall), so mod_intflow's routines become one-line calls such as
!    call intflow_set_bcpl_hip(nx, ny, x, xp, pl, iflow%plin, ap, bp)
! Arrays are host arrays here (the library copies them to the GPU and back);
! status of the last call: m2d_last_status.
module intflow_bc_hip
   use iso_c_binding, only: c_int, c_double
   implicit none
   private
   public :: intflow_set_bcu_hip, intflow_set_bcv_hip, intflow_set_bcpl_hip, &
      intflow_set_bcT_hip, intflow_velocities_at_boundary_faces_hip, &
      intflow_boundary_simplec_hip, intflow_extrapolate_u_v_to_fictitious_hip, &
      intflow_boundary_T_from_H_conservation_hip, intflow_time_cycle_update_hip

   interface
      subroutine c_intflow_set_bcu(nx, ny, modvis, x, xp, xk, yk, u, v, au, b) &
            bind(C, name="m2d_intflow_set_bcu")
         import :: c_int, c_double
         integer(c_int) :: nx, ny, modvis
         real(c_double) :: x(*), xp(*), xk(*), yk(*), u(*), v(*), au(*), b(*)
      end subroutine
      subroutine c_intflow_set_bcv(nx, ny, modvis, x, xp, xk, yk, u, v, av, b) &
            bind(C, name="m2d_intflow_set_bcv")
         import :: c_int, c_double
         integer(c_int) :: nx, ny, modvis
         real(c_double) :: x(*), xp(*), xk(*), yk(*), u(*), v(*), av(*), b(*)
      end subroutine
      subroutine c_intflow_set_bcpl(nx, ny, x, xp, pl, plin, ap, bp) &
            bind(C, name="m2d_intflow_set_bcpl")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: x(*), xp(*), pl(*), plin(*), ap(*), bp(*)
      end subroutine
      subroutine c_intflow_set_bcT(nx, ny, Twall, x, xp, T, Tin, at, bt) &
            bind(C, name="m2d_intflow_set_bcT")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: Twall
         real(c_double) :: x(*), xp(*), T(*), Tin(*), at(*), bt(*)
      end subroutine
      subroutine c_intflow_velocities_at_boundary_faces(nx, ny, modvis, x, xp, xe, ye, xk, yk, &
            u, v, ue, ve, un, vn, Uce, Vcn) bind(C, name="m2d_intflow_velocities_at_boundary_faces")
         import :: c_int, c_double
         integer(c_int) :: nx, ny, modvis
         real(c_double) :: x(*), xp(*), xe(*), ye(*), xk(*), yk(*), u(*), v(*), ue(*), ve(*), &
            un(*), vn(*), Uce(*), Vcn(*)
      end subroutine
      subroutine c_intflow_boundary_simplec(nx, ny, modvis, due, dve, dun, dvn, de, dn) &
            bind(C, name="m2d_intflow_boundary_simplec")
         import :: c_int, c_double
         integer(c_int) :: nx, ny, modvis
         real(c_double) :: due(*), dve(*), dun(*), dvn(*), de(*), dn(*)
      end subroutine
      subroutine c_intflow_extrapolate_u_v_to_fictitious(nx, ny, modvis, x, xp, xk, yk, u, v) &
            bind(C, name="m2d_intflow_extrapolate_u_v_to_fictitious")
         import :: c_int, c_double
         integer(c_int) :: nx, ny, modvis
         real(c_double) :: x(*), xp(*), xk(*), yk(*), u(*), v(*)
      end subroutine
      subroutine c_intflow_boundary_T_from_H_conservation(nx, ny, Cpref, Href, Tin, u, ue, un, &
            v, ve, vn, T, Tbe, Tbw, Tbn, Tbs) bind(C, name="m2d_intflow_boundary_T_from_H_conservation")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: Cpref, Href
         real(c_double) :: Tin(*), u(*), ue(*), un(*), v(*), ve(*), vn(*), T(*), Tbe(*), &
            Tbw(*), Tbn(*), Tbs(*)
      end subroutine
      subroutine c_intflow_time_cycle_update(nx, ny, gamma, Rg, P0, T0, x, xp, u, uin, vin, &
            pin, Tin, Mw, plin, pina, p) bind(C, name="m2d_intflow_time_cycle_update")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: gamma, Rg, P0, T0
         real(c_double) :: x(*), xp(*), u(*), uin(*), vin(*), pin(*), Tin(*), Mw(*), plin(*), &
            pina(*), p(*)
      end subroutine
   end interface

contains

   ! mod_intflow.f90:214
   subroutine intflow_set_bcu_hip(nx, ny, modvis, x, xp, xk, yk, u, v, au, b)
      integer, intent(in) :: nx, ny, modvis
      real(8), intent(in) :: x(nx*ny), xp(nx*ny), xk(nx*ny), yk(nx*ny), u(nx*ny), v(nx*ny)
      real(8), intent(inout) :: au(nx*ny,9), b(nx*ny)
      call c_intflow_set_bcu(nx, ny, modvis, x, xp, xk, yk, u, v, au, b)
   end subroutine

   ! mod_intflow.f90:305
   subroutine intflow_set_bcv_hip(nx, ny, modvis, x, xp, xk, yk, u, v, av, b)
      integer, intent(in) :: nx, ny, modvis
      real(8), intent(in) :: x(nx*ny), xp(nx*ny), xk(nx*ny), yk(nx*ny), u(nx*ny), v(nx*ny)
      real(8), intent(inout) :: av(nx*ny,9), b(nx*ny)
      call c_intflow_set_bcv(nx, ny, modvis, x, xp, xk, yk, u, v, av, b)
   end subroutine

   ! mod_intflow.f90:396 (iflow%plin)
   subroutine intflow_set_bcpl_hip(nx, ny, x, xp, pl, plin, ap, bp)
      integer, intent(in) :: nx, ny
      real(8), intent(in) :: x(nx*ny), xp(nx*ny), pl(nx*ny), plin(ny)
      real(8), intent(inout) :: ap(nx*ny,5), bp(nx*ny)
      call c_intflow_set_bcpl(nx, ny, x, xp, pl, plin, ap, bp)
   end subroutine

   ! mod_intflow.f90:464 (iflow%Twall, iflow%Tin)
   subroutine intflow_set_bcT_hip(nx, ny, Twall, x, xp, T, Tin, at, bt)
      integer, intent(in) :: nx, ny
      real(8), intent(in) :: Twall
      real(8), intent(in) :: x(nx*ny), xp(nx*ny), T(nx*ny), Tin(ny)
      real(8), intent(inout) :: at(nx*ny,9), bt(nx*ny)
      call c_intflow_set_bcT(nx, ny, Twall, x, xp, T, Tin, at, bt)
   end subroutine

   ! mod_intflow.f90:692
   subroutine intflow_velocities_at_boundary_faces_hip(nx, ny, modvis, x, xp, xe, ye, xk, yk, &
         u, v, ue, ve, un, vn, Uce, Vcn)
      integer, intent(in) :: nx, ny, modvis
      real(8), intent(in) :: x(nx*ny), xp(nx*ny), xe(nx*ny), ye(nx*ny), xk(nx*ny), yk(nx*ny), &
         u(nx*ny), v(nx*ny)
      real(8), intent(inout) :: ue(nx*ny), ve(nx*ny), un(nx*ny), vn(nx*ny), Uce(nx*ny), &
         Vcn(nx*ny)
      call c_intflow_velocities_at_boundary_faces(nx, ny, modvis, x, xp, xe, ye, xk, yk, u, v, &
         ue, ve, un, vn, Uce, Vcn)
   end subroutine

   ! mod_intflow.f90:791
   subroutine intflow_boundary_simplec_hip(nx, ny, modvis, due, dve, dun, dvn, de, dn)
      integer, intent(in) :: nx, ny, modvis
      real(8), dimension(nx*ny), intent(inout) :: due, dve, dun, dvn, de, dn
      call c_intflow_boundary_simplec(nx, ny, modvis, due, dve, dun, dvn, de, dn)
   end subroutine

   ! mod_intflow.f90:881
   subroutine intflow_extrapolate_u_v_to_fictitious_hip(nx, ny, modvis, x, xp, xk, yk, u, v)
      integer, intent(in) :: nx, ny, modvis
      real(8), dimension(nx*ny), intent(in) :: x, xp, xk, yk
      real(8), dimension(nx*ny), intent(inout) :: u, v
      call c_intflow_extrapolate_u_v_to_fictitious(nx, ny, modvis, x, xp, xk, yk, u, v)
   end subroutine

   ! mod_intflow.f90:557 (iflow -> iflow%Tin)
   subroutine intflow_boundary_T_from_H_conservation_hip(nx, ny, Cpref, Href, Tin, u, ue, un, &
         v, ve, vn, T, Tbe, Tbw, Tbn, Tbs)
      integer, intent(in) :: nx, ny
      real(8), intent(in) :: Cpref, Href, Tin(ny)
      real(8), dimension(nx*ny), intent(in) :: u, ue, un, v, ve, vn
      real(8), intent(inout) :: T(nx*ny)
      real(8), intent(inout) :: Tbe(ny), Tbw(ny), Tbn(nx), Tbs(nx)
      call c_intflow_boundary_T_from_H_conservation(nx, ny, Cpref, Href, Tin, u, ue, un, v, &
         ve, vn, T, Tbe, Tbw, Tbn, Tbs)
   end subroutine

   ! mod_intflow.f90:1106 (thermomodel -> gamma = thermomodel%gm(iflow%T0),
   ! Rg = thermomodel%Rg; iflow -> P0, T0 and its inflow vectors)
   subroutine intflow_time_cycle_update_hip(nx, ny, gamma, Rg, P0, T0, x, xp, u, uin, vin, &
         pin, Tin, Mw, plin, pina, p)
      integer, intent(in) :: nx, ny
      real(8), intent(in) :: gamma, Rg, P0, T0
      real(8), dimension(nx*ny), intent(in) :: x, xp, u
      real(8), dimension(ny), intent(inout) :: uin, vin, pin, Tin, Mw, plin, pina
      real(8), intent(inout) :: p(nx*ny)
      call c_intflow_time_cycle_update(nx, ny, gamma, Rg, P0, T0, x, xp, u, uin, vin, pin, &
         Tin, Mw, plin, pina, p)
   end subroutine

end module intflow_bc_hip
