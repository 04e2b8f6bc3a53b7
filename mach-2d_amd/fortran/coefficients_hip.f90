! coefficients_hip.f90 -- drop-in replacement of the reference module
! `coefficients` (src/equations/mach2d_coef.f90): the same module name,
! subroutine names and dummy arguments, each forwarding to the gfx950
! assembly kernels through the C ABI of libmach2d_hip.so
! (include/mach2d_hip.h, m2d_get_*).  Arrays are host arrays here: the
! library copies them to the GPU and back.  Status of the last call:
! m2d_get_status (module mach2d_hip_status).
module coefficients
   use iso_c_binding, only: c_int, c_double
   implicit none
   private
   public :: get_density_at_nodes, get_density_at_faces, get_u_coefficients, &
      get_v_coefficients, get_u_source, get_v_source, get_T_coefficients_and_source, &
      get_p_coefficients, get_p_source, get_velocities_at_internal_faces, &
      get_internal_simplec_coefficients, get_pressure_correction_with_pl, &
      get_density_correction_with_pl, get_u_v_at_real_nodes_with_pl, &
      get_velocity_correction_at_faces_with_pl, initialize_boundary_faces_flags, &
      check_mass_conservation

   interface
      subroutine c_get_density_at_nodes(nx, ny, Rg, p, T, ro) &
            bind(C, name="m2d_get_density_at_nodes")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: Rg
         real(c_double) :: p(*), T(*), ro(*)
      end subroutine
      subroutine c_get_density_at_faces(nx, ny, beta, ro, Uce, Vcn, roe, ron) &
            bind(C, name="m2d_get_density_at_faces")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: beta
         real(c_double) :: ro(*), Uce(*), Vcn(*), roe(*), ron(*)
      end subroutine
      subroutine c_get_u_coefficients(nx, ny, modvis, dt, rp, re, rn, Jp, Je, Jn, ye, yk, &
         alphae, betae, betan, gamman, vle, vln, roe, ron, roa, Uce, Vcn, a) &
            bind(C, name="m2d_get_u_coefficients")
         import :: c_int, c_double
         integer(c_int) :: nx, ny, modvis
         real(c_double) :: dt
         real(c_double) :: rp(*), re(*), rn(*), Jp(*), Je(*), Jn(*), ye(*), yk(*), alphae(*), &
               betae(*), betan(*), gamman(*), vle(*), vln(*), roe(*), ron(*), roa(*), Uce(*), &
               Vcn(*), a(*)
      end subroutine
      subroutine c_get_v_coefficients(nx, ny, coord, modvis, dt, rp, re, rn, Jp, Je, Jn, xe, &
         xk, alphae, betae, betan, gamman, vle, vln, vlp, roe, ron, roa, Uce, Vcn, a) &
            bind(C, name="m2d_get_v_coefficients")
         import :: c_int, c_double
         integer(c_int) :: nx, ny, coord, modvis
         real(c_double) :: dt
         real(c_double) :: rp(*), re(*), rn(*), Jp(*), Je(*), Jn(*), xe(*), xk(*), alphae(*), &
               betae(*), betan(*), gamman(*), vle(*), vln(*), vlp(*), roe(*), ron(*), roa(*), &
               Uce(*), Vcn(*), a(*)
      end subroutine
      subroutine c_get_u_source(nx, ny, modvis, beta, dt, rp, re, rn, xe, ye, xk, yk, xke, &
         yke, xen, yen, Jp, Je, Jn, roe, ron, roa, p, vle, vln, Uce, Vcn, ua, u, v, cup, sup, &
         b) &
            bind(C, name="m2d_get_u_source")
         import :: c_int, c_double
         integer(c_int) :: nx, ny, modvis
         real(c_double) :: beta, dt
         real(c_double) :: rp(*), re(*), rn(*), xe(*), ye(*), xk(*), yk(*), xke(*), yke(*), &
               xen(*), yen(*), Jp(*), Je(*), Jn(*), roe(*), ron(*), roa(*), p(*), vle(*), &
               vln(*), Uce(*), Vcn(*), ua(*), u(*), v(*), cup(*), sup(*), b(*)
      end subroutine
      subroutine c_get_v_source(nx, ny, coord, modvis, beta, dt, rp, re, rn, xe, ye, xk, yk, &
         xke, yke, xen, yen, Jp, Je, Jn, roe, ron, roa, p, vle, vln, Uce, Vcn, va, u, v, cvp, &
         svp, b) &
            bind(C, name="m2d_get_v_source")
         import :: c_int, c_double
         integer(c_int) :: nx, ny, coord, modvis
         real(c_double) :: beta, dt
         real(c_double) :: rp(*), re(*), rn(*), xe(*), ye(*), xk(*), yk(*), xke(*), yke(*), &
               xen(*), yen(*), Jp(*), Je(*), Jn(*), roe(*), ron(*), roa(*), p(*), vle(*), &
               vln(*), Uce(*), Vcn(*), va(*), u(*), v(*), cvp(*), svp(*), b(*)
      end subroutine
      subroutine c_get_T_coefficients_and_source(nx, ny, modvis, beta, dt, rp, re, rn, xe, ye, &
         xk, yk, xke, yke, xen, yen, alphae, betae, betan, gamman, Jp, Je, Jn, roe, ron, ro, &
         roa, p, pa, cp, vle, vln, ke, kn, Uce, Vcn, u, v, ue, ve, un, vn, ua, va, T, Ta, fbe, &
         fbn, a, b) &
            bind(C, name="m2d_get_T_coefficients_and_source")
         import :: c_int, c_double
         integer(c_int) :: nx, ny, modvis
         real(c_double) :: beta, dt
         real(c_double) :: rp(*), re(*), rn(*), xe(*), ye(*), xk(*), yk(*), xke(*), yke(*), &
               xen(*), yen(*), alphae(*), betae(*), betan(*), gamman(*), Jp(*), Je(*), Jn(*), &
               roe(*), ron(*), ro(*), roa(*), p(*), pa(*), cp(*), vle(*), vln(*), ke(*), &
               kn(*), Uce(*), Vcn(*), u(*), v(*), ue(*), ve(*), un(*), vn(*), ua(*), va(*), &
               T(*), Ta(*), fbe(*), fbn(*), a(*), b(*)
      end subroutine
      subroutine c_get_p_coefficients(nx, ny, dt, rp, re, rn, Jp, Uce, Vcn, roe, ron, g, de, &
         dn, a) &
            bind(C, name="m2d_get_p_coefficients")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: dt
         real(c_double) :: rp(*), re(*), rn(*), Jp(*), Uce(*), Vcn(*), roe(*), ron(*), g(*), &
               de(*), dn(*), a(*)
      end subroutine
      subroutine c_get_p_source(nx, ny, dt, rp, re, rn, Jp, roe, ron, ro, roa, Uce, Vcn, de, &
         dn, g, pl, b) &
            bind(C, name="m2d_get_p_source")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: dt
         real(c_double) :: rp(*), re(*), rn(*), Jp(*), roe(*), ron(*), ro(*), roa(*), Uce(*), &
               Vcn(*), de(*), dn(*), g(*), pl(*), b(*)
      end subroutine
      subroutine c_get_velocities_at_internal_faces(nx, ny, dt, rp, re, rn, xe, ye, xk, yk, &
         xen, yen, xke, yke, Jp, cup, cvp, sup, svp, au, av, roa, p, u, v, uea, vea, una, vna, &
         ue, ve, un, vn, Uce, Vcn) &
            bind(C, name="m2d_get_velocities_at_internal_faces")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: dt
         real(c_double) :: rp(*), re(*), rn(*), xe(*), ye(*), xk(*), yk(*), xen(*), yen(*), &
               xke(*), yke(*), Jp(*), cup(*), cvp(*), sup(*), svp(*), au(*), av(*), roa(*), &
               p(*), u(*), v(*), uea(*), vea(*), una(*), vna(*), ue(*), ve(*), un(*), vn(*), &
               Uce(*), Vcn(*)
      end subroutine
      subroutine c_get_internal_simplec_coefficients(nx, ny, re, rn, xe, ye, xk, yk, au, av, &
         due, dve, dun, dvn, de, dn) &
            bind(C, name="m2d_get_internal_simplec_coefficients")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: re(*), rn(*), xe(*), ye(*), xk(*), yk(*), au(*), av(*), due(*), &
               dve(*), dun(*), dvn(*), de(*), dn(*)
      end subroutine
      subroutine c_get_pressure_correction_with_pl(nx, ny, pl, p) &
            bind(C, name="m2d_get_pressure_correction_with_pl")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: pl(*), p(*)
      end subroutine
      subroutine c_get_density_correction_with_pl(nx, ny, pl, g, ro) &
            bind(C, name="m2d_get_density_correction_with_pl")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: pl(*), g(*), ro(*)
      end subroutine
      subroutine c_get_u_v_at_real_nodes_with_pl(nx, ny, xe, ye, xk, yk, rp, pl, au, av, u, v) &
            bind(C, name="m2d_get_u_v_at_real_nodes_with_pl")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: xe(*), ye(*), xk(*), yk(*), rp(*), pl(*), au(*), av(*), u(*), v(*)
      end subroutine
      subroutine c_get_velocity_correction_at_faces_with_pl(nx, ny, xe, ye, xk, yk, due, dve, &
         dun, dvn, de, dn, pl, ue, ve, un, vn, Uce, Vcn) &
            bind(C, name="m2d_get_velocity_correction_at_faces_with_pl")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: xe(*), ye(*), xk(*), yk(*), due(*), dve(*), dun(*), dvn(*), de(*), &
               dn(*), pl(*), ue(*), ve(*), un(*), vn(*), Uce(*), Vcn(*)
      end subroutine
      subroutine c_initialize_boundary_faces_flags(nx, ny, fbe, fbn) &
            bind(C, name="m2d_initialize_boundary_faces_flags")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: fbe(*), fbn(*)
      end subroutine
      subroutine c_check_mass_conservation(nx, ny, dt, rp, re, rn, Jp, ro, roa, ron, roe, Uce, &
         Vcn, res) &
            bind(C, name="m2d_check_mass_conservation")
         import :: c_int, c_double
         integer(c_int) :: nx, ny
         real(c_double) :: dt
         real(c_double) :: rp(*), re(*), rn(*), Jp(*), ro(*), roa(*), ron(*), roe(*), Uce(*), &
               Vcn(*), res
      end subroutine
   end interface

contains

   ! mach2d_coef.f90:6
   subroutine get_density_at_nodes(nx, ny, Rg, p, T, ro)
      integer, intent(in) :: nx, ny
      real(8), intent(in) :: Rg
      real(8), intent(in) :: p(nx*ny), T(nx*ny)
      real(8), intent(inout) :: ro(nx*ny)
      call c_get_density_at_nodes(nx, ny, Rg, p, T, ro)
   end subroutine

   ! mach2d_coef.f90:20
   subroutine get_density_at_faces(nx, ny, beta, ro, Uce, Vcn, roe, ron)
      integer, intent(in) :: nx, ny
      real(8), intent(in) :: beta
      real(8), intent(in) :: ro(nx*ny), Uce(nx*ny), Vcn(nx*ny)
      real(8), intent(inout) :: roe(nx*ny), ron(nx*ny)
      call c_get_density_at_faces(nx, ny, beta, ro, Uce, Vcn, roe, ron)
   end subroutine

   ! mach2d_coef.f90:63
   subroutine get_u_coefficients(nx, ny, modvis, dt, rp, re, rn, Jp, Je, Jn, ye, yk, alphae, &
      betae, betan, gamman, vle, vln, roe, ron, roa, Uce, Vcn, a)
      integer, intent(in) :: nx, ny, modvis
      real(8), intent(in) :: dt
      real(8), intent(in) :: rp(nx*ny), re(nx*ny), rn(nx*ny), Jp(nx*ny), Je(nx*ny), Jn(nx*ny), &
            ye(nx*ny), yk(nx*ny), alphae(nx*ny), betae(nx*ny), betan(nx*ny), gamman(nx*ny), &
            vle(nx*ny), vln(nx*ny), roe(nx*ny), ron(nx*ny), roa(nx*ny), Uce(nx*ny), Vcn(nx*ny)
      real(8), intent(inout) :: a(nx*ny,9)
      call c_get_u_coefficients(nx, ny, modvis, dt, rp, re, rn, Jp, Je, Jn, ye, yk, alphae, &
         betae, betan, gamman, vle, vln, roe, ron, roa, Uce, Vcn, a)
   end subroutine

   ! mach2d_coef.f90:454
   subroutine get_v_coefficients(nx, ny, coord, modvis, dt, rp, re, rn, Jp, Je, Jn, xe, xk, &
      alphae, betae, betan, gamman, vle, vln, vlp, roe, ron, roa, Uce, Vcn, a)
      integer, intent(in) :: nx, ny, coord, modvis
      real(8), intent(in) :: dt
      real(8), intent(in) :: rp(nx*ny), re(nx*ny), rn(nx*ny), Jp(nx*ny), Je(nx*ny), Jn(nx*ny), &
            xe(nx*ny), xk(nx*ny), alphae(nx*ny), betae(nx*ny), betan(nx*ny), gamman(nx*ny), &
            vle(nx*ny), vln(nx*ny), vlp(nx*ny), roe(nx*ny), ron(nx*ny), roa(nx*ny), &
            Uce(nx*ny), Vcn(nx*ny)
      real(8), intent(inout) :: a(nx*ny,9)
      call c_get_v_coefficients(nx, ny, coord, modvis, dt, rp, re, rn, Jp, Je, Jn, xe, xk, &
         alphae, betae, betan, gamman, vle, vln, vlp, roe, ron, roa, Uce, Vcn, a)
   end subroutine

   ! mach2d_coef.f90:221
   subroutine get_u_source(nx, ny, modvis, beta, dt, rp, re, rn, xe, ye, xk, yk, xke, yke, &
      xen, yen, Jp, Je, Jn, roe, ron, roa, p, vle, vln, Uce, Vcn, ua, u, v, cup, sup, b)
      integer, intent(in) :: nx, ny, modvis
      real(8), intent(in) :: beta, dt
      real(8), intent(in) :: rp(nx*ny), re(nx*ny), rn(nx*ny), xe(nx*ny), ye(nx*ny), xk(nx*ny), &
            yk(nx*ny), xke(nx*ny), yke(nx*ny), xen(nx*ny), yen(nx*ny), Jp(nx*ny), Je(nx*ny), &
            Jn(nx*ny), roe(nx*ny), ron(nx*ny), roa(nx*ny), p(nx*ny), vle(nx*ny), vln(nx*ny), &
            Uce(nx*ny), Vcn(nx*ny), ua(nx*ny), u(nx*ny), v(nx*ny)
      real(8), intent(inout) :: cup(nx*ny), sup(nx*ny), b(nx*ny)
      call c_get_u_source(nx, ny, modvis, beta, dt, rp, re, rn, xe, ye, xk, yk, xke, yke, xen, &
         yen, Jp, Je, Jn, roe, ron, roa, p, vle, vln, Uce, Vcn, ua, u, v, cup, sup, b)
   end subroutine

   ! mach2d_coef.f90:616
   subroutine get_v_source(nx, ny, coord, modvis, beta, dt, rp, re, rn, xe, ye, xk, yk, xke, &
      yke, xen, yen, Jp, Je, Jn, roe, ron, roa, p, vle, vln, Uce, Vcn, va, u, v, cvp, svp, b)
      integer, intent(in) :: nx, ny, coord, modvis
      real(8), intent(in) :: beta, dt
      real(8), intent(in) :: rp(nx*ny), re(nx*ny), rn(nx*ny), xe(nx*ny), ye(nx*ny), xk(nx*ny), &
            yk(nx*ny), xke(nx*ny), yke(nx*ny), xen(nx*ny), yen(nx*ny), Jp(nx*ny), Je(nx*ny), &
            Jn(nx*ny), roe(nx*ny), ron(nx*ny), roa(nx*ny), p(nx*ny), vle(nx*ny), vln(nx*ny), &
            Uce(nx*ny), Vcn(nx*ny), va(nx*ny), u(nx*ny), v(nx*ny)
      real(8), intent(inout) :: cvp(nx*ny), svp(nx*ny), b(nx*ny)
      call c_get_v_source(nx, ny, coord, modvis, beta, dt, rp, re, rn, xe, ye, xk, yk, xke, &
         yke, xen, yen, Jp, Je, Jn, roe, ron, roa, p, vle, vln, Uce, Vcn, va, u, v, cvp, svp, &
         b)
   end subroutine

   ! mach2d_coef.f90:787
   subroutine get_T_coefficients_and_source(nx, ny, modvis, beta, dt, rp, re, rn, xe, ye, xk, &
      yk, xke, yke, xen, yen, alphae, betae, betan, gamman, Jp, Je, Jn, roe, ron, ro, roa, p, &
      pa, cp, vle, vln, ke, kn, Uce, Vcn, u, v, ue, ve, un, vn, ua, va, T, Ta, fbe, fbn, a, b)
      integer, intent(in) :: nx, ny, modvis
      real(8), intent(in) :: beta, dt
      real(8), intent(in) :: rp(nx*ny), re(nx*ny), rn(nx*ny), xe(nx*ny), ye(nx*ny), xk(nx*ny), &
            yk(nx*ny), xke(nx*ny), yke(nx*ny), xen(nx*ny), yen(nx*ny), alphae(nx*ny), &
            betae(nx*ny), betan(nx*ny), gamman(nx*ny), Jp(nx*ny), Je(nx*ny), Jn(nx*ny), &
            roe(nx*ny), ron(nx*ny), ro(nx*ny), roa(nx*ny), p(nx*ny), pa(nx*ny), cp(nx*ny), &
            vle(nx*ny), vln(nx*ny), ke(nx*ny), kn(nx*ny), Uce(nx*ny), Vcn(nx*ny), u(nx*ny), &
            v(nx*ny), ue(nx*ny), ve(nx*ny), un(nx*ny), vn(nx*ny), ua(nx*ny), va(nx*ny), &
            T(nx*ny), Ta(nx*ny), fbe(nx*ny), fbn(nx*ny)
      real(8), intent(inout) :: a(nx*ny,9), b(nx*ny)
      call c_get_T_coefficients_and_source(nx, ny, modvis, beta, dt, rp, re, rn, xe, ye, xk, &
         yk, xke, yke, xen, yen, alphae, betae, betan, gamman, Jp, Je, Jn, roe, ron, ro, roa, &
         p, pa, cp, vle, vln, ke, kn, Uce, Vcn, u, v, ue, ve, un, vn, ua, va, T, Ta, fbe, fbn, &
         a, b)
   end subroutine

   ! mach2d_coef.f90:1207
   subroutine get_p_coefficients(nx, ny, dt, rp, re, rn, Jp, Uce, Vcn, roe, ron, g, de, dn, a)
      integer, intent(in) :: nx, ny
      real(8), intent(in) :: dt
      real(8), intent(in) :: rp(nx*ny), re(nx*ny), rn(nx*ny), Jp(nx*ny), Uce(nx*ny), &
            Vcn(nx*ny), roe(nx*ny), ron(nx*ny), g(nx*ny), de(nx*ny), dn(nx*ny)
      real(8), intent(inout) :: a(nx*ny,5)
      call c_get_p_coefficients(nx, ny, dt, rp, re, rn, Jp, Uce, Vcn, roe, ron, g, de, dn, a)
   end subroutine

   ! mach2d_coef.f90:1281
   subroutine get_p_source(nx, ny, dt, rp, re, rn, Jp, roe, ron, ro, roa, Uce, Vcn, de, dn, g, &
      pl, b)
      integer, intent(in) :: nx, ny
      real(8), intent(in) :: dt
      real(8), intent(in) :: rp(nx*ny), re(nx*ny), rn(nx*ny), Jp(nx*ny), roe(nx*ny), &
            ron(nx*ny), ro(nx*ny), roa(nx*ny), Uce(nx*ny), Vcn(nx*ny), de(nx*ny), dn(nx*ny), &
            g(nx*ny), pl(nx*ny)
      real(8), intent(inout) :: b(nx*ny)
      call c_get_p_source(nx, ny, dt, rp, re, rn, Jp, roe, ron, ro, roa, Uce, Vcn, de, dn, g, &
         pl, b)
   end subroutine

   ! mach2d_coef.f90:1350
   subroutine get_velocities_at_internal_faces(nx, ny, dt, rp, re, rn, xe, ye, xk, yk, xen, &
      yen, xke, yke, Jp, cup, cvp, sup, svp, au, av, roa, p, u, v, uea, vea, una, vna, ue, ve, &
      un, vn, Uce, Vcn)
      integer, intent(in) :: nx, ny
      real(8), intent(in) :: dt
      real(8), intent(in) :: rp(nx*ny), re(nx*ny), rn(nx*ny), xe(nx*ny), ye(nx*ny), xk(nx*ny), &
            yk(nx*ny), xen(nx*ny), yen(nx*ny), xke(nx*ny), yke(nx*ny), Jp(nx*ny), cup(nx*ny), &
            cvp(nx*ny), sup(nx*ny), svp(nx*ny), au(nx*ny,9), av(nx*ny,9), roa(nx*ny), &
            p(nx*ny), u(nx*ny), v(nx*ny), uea(nx*ny), vea(nx*ny), una(nx*ny), vna(nx*ny)
      real(8), intent(inout) :: ue(nx*ny), ve(nx*ny), un(nx*ny), vn(nx*ny), Uce(nx*ny), &
            Vcn(nx*ny)
      call c_get_velocities_at_internal_faces(nx, ny, dt, rp, re, rn, xe, ye, xk, yk, xen, &
         yen, xke, yke, Jp, cup, cvp, sup, svp, au, av, roa, p, u, v, uea, vea, una, vna, ue, &
         ve, un, vn, Uce, Vcn)
   end subroutine

   ! mach2d_coef.f90:1504
   subroutine get_internal_simplec_coefficients(nx, ny, re, rn, xe, ye, xk, yk, au, av, due, &
      dve, dun, dvn, de, dn)
      integer, intent(in) :: nx, ny
      real(8), intent(in) :: re(nx*ny), rn(nx*ny), xe(nx*ny), ye(nx*ny), xk(nx*ny), yk(nx*ny), &
            au(nx*ny,9), av(nx*ny,9)
      real(8), intent(inout) :: due(nx*ny), dve(nx*ny), dun(nx*ny), dvn(nx*ny), de(nx*ny), &
            dn(nx*ny)
      call c_get_internal_simplec_coefficients(nx, ny, re, rn, xe, ye, xk, yk, au, av, due, &
         dve, dun, dvn, de, dn)
   end subroutine

   ! mach2d_coef.f90:1602
   subroutine get_pressure_correction_with_pl(nx, ny, pl, p)
      integer, intent(in) :: nx, ny
      real(8), intent(in) :: pl(nx*ny)
      real(8), intent(inout) :: p(nx*ny)
      call c_get_pressure_correction_with_pl(nx, ny, pl, p)
   end subroutine

   ! mach2d_coef.f90:1614
   subroutine get_density_correction_with_pl(nx, ny, pl, g, ro)
      integer, intent(in) :: nx, ny
      real(8), intent(in) :: pl(nx*ny), g(nx*ny)
      real(8), intent(inout) :: ro(nx*ny)
      call c_get_density_correction_with_pl(nx, ny, pl, g, ro)
   end subroutine

   ! mach2d_coef.f90:1626
   subroutine get_u_v_at_real_nodes_with_pl(nx, ny, xe, ye, xk, yk, rp, pl, au, av, u, v)
      integer, intent(in) :: nx, ny
      real(8), intent(in) :: xe(nx*ny), ye(nx*ny), xk(nx*ny), yk(nx*ny), rp(nx*ny), pl(nx*ny), &
            au(nx*ny,9), av(nx*ny,9)
      real(8), intent(inout) :: u(nx*ny), v(nx*ny)
      call c_get_u_v_at_real_nodes_with_pl(nx, ny, xe, ye, xk, yk, rp, pl, au, av, u, v)
   end subroutine

   ! mach2d_coef.f90:1679
   subroutine get_velocity_correction_at_faces_with_pl(nx, ny, xe, ye, xk, yk, due, dve, dun, &
      dvn, de, dn, pl, ue, ve, un, vn, Uce, Vcn)
      integer, intent(in) :: nx, ny
      real(8), intent(in) :: xe(nx*ny), ye(nx*ny), xk(nx*ny), yk(nx*ny), due(nx*ny), &
            dve(nx*ny), dun(nx*ny), dvn(nx*ny), de(nx*ny), dn(nx*ny), pl(nx*ny)
      real(8), intent(inout) :: ue(nx*ny), ve(nx*ny), un(nx*ny), vn(nx*ny), Uce(nx*ny), &
            Vcn(nx*ny)
      call c_get_velocity_correction_at_faces_with_pl(nx, ny, xe, ye, xk, yk, due, dve, dun, &
         dvn, de, dn, pl, ue, ve, un, vn, Uce, Vcn)
   end subroutine

   ! mach2d_coef.f90:1764
   subroutine initialize_boundary_faces_flags(nx, ny, fbe, fbn)
      integer, intent(in) :: nx, ny
      real(8), intent(inout) :: fbe(nx*ny), fbn(nx*ny)
      call c_initialize_boundary_faces_flags(nx, ny, fbe, fbn)
   end subroutine

   ! mach2d_coef.f90:1833
   subroutine check_mass_conservation(nx, ny, dt, rp, re, rn, Jp, ro, roa, ron, roe, Uce, Vcn, &
      res)
      integer, intent(in) :: nx, ny
      real(8), intent(in) :: dt
      real(8), intent(in) :: rp(nx*ny), re(nx*ny), rn(nx*ny), Jp(nx*ny), ro(nx*ny), &
            roa(nx*ny), ron(nx*ny), roe(nx*ny), Uce(nx*ny), Vcn(nx*ny)
      real(8), intent(inout) :: res
      call c_check_mass_conservation(nx, ny, dt, rp, re, rn, Jp, ro, roa, ron, roe, Uce, Vcn, &
         res)
   end subroutine

end module coefficients
