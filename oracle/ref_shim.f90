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
   use mod_class_ifile, only: class_ifile
   use mod_tstepper_scarborough, only: tstepper_scarborough_init, tstepper_dt_scarborough, &
      tstepper_dt0_scarborough
   use msi2d9
   use msi2d5
   use mtdma2d9
   use mod_class_solver_msi
   use mod_solvers, only: norm_l1_9d, norm_l1_9d_relative, norm_l1_b, norm_l1_5d
   use mod_field_operations, only: get_average
   use coefficients
   use mod_intflow, only: type_intflow, intflow_set_bcu, intflow_set_bcv, &
      intflow_set_bcpl, intflow_set_bcT, intflow_velocities_at_boundary_faces, &
      intflow_boundary_simplec, intflow_extrapolate_u_v_to_fictitious, &
      intflow_boundary_T_from_H_conservation
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

   ! src/solvers/mod_solver_msi.f90:96-133: the plugin object, init + solve
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


   ! ---- coefficient assembly, src/equations/mach2d_coef.f90 -----------------

   ! :6
   subroutine ref_density_at_nodes(nx, ny, Rg, p, T, ro) bind(C, name="ref_density_at_nodes")
      integer(c_int), value :: nx, ny
      real(c_double), value :: Rg
      real(c_double), intent(in)    :: p(nx*ny), T(nx*ny)
      real(c_double), intent(inout) :: ro(nx*ny)
      call get_density_at_nodes(nx, ny, Rg, p, T, ro)
   end subroutine

   ! :20
   subroutine ref_density_at_faces(nx, ny, beta, ro, Uce, Vcn, roe, ron) &
         bind(C, name="ref_density_at_faces")
      integer(c_int), value :: nx, ny
      real(c_double), value :: beta
      real(c_double), intent(in)    :: ro(nx*ny), Uce(nx*ny), Vcn(nx*ny)
      real(c_double), intent(inout) :: roe(nx*ny), ron(nx*ny)
      call get_density_at_faces(nx, ny, beta, ro, Uce, Vcn, roe, ron)
   end subroutine

   ! :63
   subroutine ref_u_coefficients(nx, ny, modvis, dt, rp, re, rn, Jp, Je, Jn, ye, yk, alphae, &
         betae, betan, gamman, vle, vln, roe, ron, roa, Uce, Vcn, a) &
         bind(C, name="ref_u_coefficients")
      integer(c_int), value :: nx, ny, modvis
      real(c_double), value :: dt
      real(c_double), dimension(nx*ny), intent(in) :: rp, re, rn, Jp, Je, Jn, ye, yk, alphae, &
         betae, betan, gamman, vle, vln, roe, ron, roa, Uce, Vcn
      real(c_double), intent(inout) :: a(nx*ny,9)
      call get_u_coefficients(nx, ny, modvis, dt, rp, re, rn, Jp, Je, Jn, ye, yk, alphae, &
         betae, betan, gamman, vle, vln, roe, ron, roa, Uce, Vcn, a)
   end subroutine

   ! :454
   subroutine ref_v_coefficients(nx, ny, coord, modvis, dt, rp, re, rn, Jp, Je, Jn, xe, xk, &
         alphae, betae, betan, gamman, vle, vln, vlp, roe, ron, roa, Uce, Vcn, a) &
         bind(C, name="ref_v_coefficients")
      integer(c_int), value :: nx, ny, coord, modvis
      real(c_double), value :: dt
      real(c_double), dimension(nx*ny), intent(in) :: rp, re, rn, Jp, Je, Jn, xe, xk, alphae, &
         betae, betan, gamman, vle, vln, vlp, roe, ron, roa, Uce, Vcn
      real(c_double), intent(inout) :: a(nx*ny,9)
      call get_v_coefficients(nx, ny, coord, modvis, dt, rp, re, rn, Jp, Je, Jn, xe, xk, &
         alphae, betae, betan, gamman, vle, vln, vlp, roe, ron, roa, Uce, Vcn, a)
   end subroutine

   ! :1207
   subroutine ref_p_coefficients(nx, ny, dt, rp, re, rn, Jp, Uce, Vcn, roe, ron, g, de, dn, a) &
         bind(C, name="ref_p_coefficients")
      integer(c_int), value :: nx, ny
      real(c_double), value :: dt
      real(c_double), dimension(nx*ny), intent(in) :: rp, re, rn, Jp, Uce, Vcn, roe, ron, g, &
         de, dn
      real(c_double), intent(inout) :: a(nx*ny,5)
      call get_p_coefficients(nx, ny, dt, rp, re, rn, Jp, Uce, Vcn, roe, ron, g, de, dn, a)
   end subroutine

   ! :1281
   subroutine ref_p_source(nx, ny, dt, rp, re, rn, Jp, roe, ron, ro, roa, Uce, Vcn, de, dn, &
         g, pl, b) bind(C, name="ref_p_source")
      integer(c_int), value :: nx, ny
      real(c_double), value :: dt
      real(c_double), dimension(nx*ny), intent(in) :: rp, re, rn, Jp, roe, ron, ro, roa, Uce, &
         Vcn, de, dn, g, pl
      real(c_double), intent(inout) :: b(nx*ny)
      call get_p_source(nx, ny, dt, rp, re, rn, Jp, roe, ron, ro, roa, Uce, Vcn, de, dn, g, &
         pl, b)
   end subroutine

   ! :221
   subroutine ref_u_source(nx, ny, modvis, beta, dt, rp, re, rn, xe, ye, xk, yk, xke, yke, &
         xen, yen, Jp, Je, Jn, roe, ron, roa, p, vle, vln, Uce, Vcn, ua, u, v, cup, sup, b) &
         bind(C, name="ref_u_source")
      integer(c_int), value :: nx, ny, modvis
      real(c_double), value :: beta, dt
      real(c_double), dimension(nx*ny), intent(in) :: rp, re, rn, xe, ye, xk, yk, xke, yke, &
         xen, yen, Jp, Je, Jn, roe, ron, roa, p, vle, vln, Uce, Vcn, ua, u, v
      real(c_double), dimension(nx*ny), intent(inout) :: cup, sup, b
      call get_u_source(nx, ny, modvis, beta, dt, rp, re, rn, xe, ye, xk, yk, xke, yke, &
         xen, yen, Jp, Je, Jn, roe, ron, roa, p, vle, vln, Uce, Vcn, ua, u, v, cup, sup, b)
   end subroutine

   ! :616
   subroutine ref_v_source(nx, ny, coord, modvis, beta, dt, rp, re, rn, xe, ye, xk, yk, xke, &
         yke, xen, yen, Jp, Je, Jn, roe, ron, roa, p, vle, vln, Uce, Vcn, va, u, v, cvp, svp, &
         b) bind(C, name="ref_v_source")
      integer(c_int), value :: nx, ny, coord, modvis
      real(c_double), value :: beta, dt
      real(c_double), dimension(nx*ny), intent(in) :: rp, re, rn, xe, ye, xk, yk, xke, yke, &
         xen, yen, Jp, Je, Jn, roe, ron, roa, p, vle, vln, Uce, Vcn, va, u, v
      real(c_double), dimension(nx*ny), intent(inout) :: cvp, svp, b
      call get_v_source(nx, ny, coord, modvis, beta, dt, rp, re, rn, xe, ye, xk, yk, xke, &
         yke, xen, yen, Jp, Je, Jn, roe, ron, roa, p, vle, vln, Uce, Vcn, va, u, v, cvp, svp, b)
   end subroutine

   ! :787
   subroutine ref_T_coefficients_and_source(nx, ny, modvis, beta, dt, rp, re, rn, xe, ye, &
         xk, yk, xke, yke, xen, yen, alphae, betae, betan, gamman, Jp, Je, Jn, roe, ron, ro, &
         roa, p, pa, cp, vle, vln, ke, kn, Uce, Vcn, u, v, ue, ve, un, vn, ua, va, T, Ta, fbe, &
         fbn, a, b) bind(C, name="ref_T_coefficients_and_source")
      integer(c_int), value :: nx, ny, modvis
      real(c_double), value :: beta, dt
      real(c_double), dimension(nx*ny), intent(in) :: rp, re, rn, xe, ye, xk, yk, xke, yke, &
         xen, yen, alphae, betae, betan, gamman, Jp, Je, Jn, roe, ron, ro, roa, p, pa, cp, &
         vle, vln, ke, kn, Uce, Vcn, u, v, ue, ve, un, vn, ua, va, T, Ta, fbe, fbn
      real(c_double), intent(inout) :: a(nx*ny,9), b(nx*ny)
      call get_T_coefficients_and_source(nx, ny, modvis, beta, dt, rp, re, rn, xe, ye, xk, &
         yk, xke, yke, xen, yen, alphae, betae, betan, gamman, Jp, Je, Jn, roe, ron, ro, roa, &
         p, pa, cp, vle, vln, ke, kn, Uce, Vcn, u, v, ue, ve, un, vn, ua, va, T, Ta, fbe, fbn, &
         a, b)
   end subroutine

   ! :1350
   subroutine ref_velocities_at_internal_faces(nx, ny, dt, rp, re, rn, xe, ye, xk, yk, xen, &
         yen, xke, yke, Jp, cup, cvp, sup, svp, au, av, roa, p, u, v, uea, vea, una, vna, ue, &
         ve, un, vn, Uce, Vcn) bind(C, name="ref_velocities_at_internal_faces")
      integer(c_int), value :: nx, ny
      real(c_double), value :: dt
      real(c_double), dimension(nx*ny), intent(in) :: rp, re, rn, xe, ye, xk, yk, xen, yen, &
         xke, yke, Jp, cup, cvp, sup, svp, roa, p, u, v, uea, vea, una, vna
      real(c_double), intent(in) :: au(nx*ny,9), av(nx*ny,9)
      real(c_double), dimension(nx*ny), intent(inout) :: ue, ve, un, vn, Uce, Vcn
      call get_velocities_at_internal_faces(nx, ny, dt, rp, re, rn, xe, ye, xk, yk, xen, yen, &
         xke, yke, Jp, cup, cvp, sup, svp, au, av, roa, p, u, v, uea, vea, una, vna, ue, ve, &
         un, vn, Uce, Vcn)
   end subroutine

   ! :1504
   subroutine ref_internal_simplec_coefficients(nx, ny, re, rn, xe, ye, xk, yk, au, av, due, &
         dve, dun, dvn, de, dn) bind(C, name="ref_internal_simplec_coefficients")
      integer(c_int), value :: nx, ny
      real(c_double), dimension(nx*ny), intent(in) :: re, rn, xe, ye, xk, yk
      real(c_double), intent(in) :: au(nx*ny,9), av(nx*ny,9)
      real(c_double), dimension(nx*ny), intent(inout) :: due, dve, dun, dvn, de, dn
      call get_internal_simplec_coefficients(nx, ny, re, rn, xe, ye, xk, yk, au, av, due, dve, &
         dun, dvn, de, dn)
   end subroutine

   ! :1602
   subroutine ref_pressure_correction_with_pl(nx, ny, pl, p) &
         bind(C, name="ref_pressure_correction_with_pl")
      integer(c_int), value :: nx, ny
      real(c_double), intent(in) :: pl(nx*ny)
      real(c_double), intent(inout) :: p(nx*ny)
      call get_pressure_correction_with_pl(nx, ny, pl, p)
   end subroutine

   ! :1614
   subroutine ref_density_correction_with_pl(nx, ny, pl, g, ro) &
         bind(C, name="ref_density_correction_with_pl")
      integer(c_int), value :: nx, ny
      real(c_double), intent(in) :: pl(nx*ny), g(nx*ny)
      real(c_double), intent(inout) :: ro(nx*ny)
      call get_density_correction_with_pl(nx, ny, pl, g, ro)
   end subroutine

   ! :1626
   subroutine ref_u_v_at_real_nodes_with_pl(nx, ny, xe, ye, xk, yk, rp, pl, au, av, u, v) &
         bind(C, name="ref_u_v_at_real_nodes_with_pl")
      integer(c_int), value :: nx, ny
      real(c_double), dimension(nx*ny), intent(in) :: xe, ye, xk, yk, rp, pl
      real(c_double), intent(in) :: au(nx*ny,9), av(nx*ny,9)
      real(c_double), dimension(nx*ny), intent(inout) :: u, v
      call get_u_v_at_real_nodes_with_pl(nx, ny, xe, ye, xk, yk, rp, pl, au, av, u, v)
   end subroutine

   ! :1679
   subroutine ref_velocity_correction_at_faces_with_pl(nx, ny, xe, ye, xk, yk, due, dve, dun, &
         dvn, de, dn, pl, ue, ve, un, vn, Uce, Vcn) &
         bind(C, name="ref_velocity_correction_at_faces_with_pl")
      integer(c_int), value :: nx, ny
      real(c_double), dimension(nx*ny), intent(in) :: xe, ye, xk, yk, due, dve, dun, dvn, de, &
         dn, pl
      real(c_double), dimension(nx*ny), intent(inout) :: ue, ve, un, vn, Uce, Vcn
      call get_velocity_correction_at_faces_with_pl(nx, ny, xe, ye, xk, yk, due, dve, dun, dvn, &
         de, dn, pl, ue, ve, un, vn, Uce, Vcn)
   end subroutine

   ! :1764
   subroutine ref_boundary_faces_flags(nx, ny, fbe, fbn) bind(C, name="ref_boundary_faces_flags")
      integer(c_int), value :: nx, ny
      real(c_double), dimension(nx*ny), intent(inout) :: fbe, fbn
      call initialize_boundary_faces_flags(nx, ny, fbe, fbn)
   end subroutine

   ! :1833
   subroutine ref_mass_conservation(nx, ny, dt, rp, re, rn, Jp, ro, roa, ron, roe, Uce, Vcn, &
         res) bind(C, name="ref_mass_conservation")
      integer(c_int), value :: nx, ny
      real(c_double), value :: dt
      real(c_double), dimension(nx*ny), intent(in) :: rp, re, rn, Jp, ro, roa, ron, roe, Uce, &
         Vcn
      real(c_double), intent(inout) :: res
      call check_mass_conservation(nx, ny, dt, rp, re, rn, Jp, ro, roa, ron, roe, Uce, Vcn, res)
   end subroutine

   ! src/user/internal_flow/mod_intflow.f90:214
   subroutine ref_intflow_set_bcu(nx, ny, modvis, x, xp, xk, yk, u, v, au, b) &
         bind(C, name="ref_intflow_set_bcu")
      integer(c_int), value :: nx, ny, modvis
      real(c_double), dimension(nx*ny), intent(in) :: x, xp, xk, yk, u, v
      real(c_double), intent(inout) :: au(nx*ny,9), b(nx*ny)
      call intflow_set_bcu(nx, ny, modvis, x, xp, xk, yk, u, v, au, b)
   end subroutine

   ! mod_intflow.f90:305
   subroutine ref_intflow_set_bcv(nx, ny, modvis, x, xp, xk, yk, u, v, av, b) &
         bind(C, name="ref_intflow_set_bcv")
      integer(c_int), value :: nx, ny, modvis
      real(c_double), dimension(nx*ny), intent(in) :: x, xp, xk, yk, u, v
      real(c_double), intent(inout) :: av(nx*ny,9), b(nx*ny)
      call intflow_set_bcv(nx, ny, modvis, x, xp, xk, yk, u, v, av, b)
   end subroutine

   ! mod_intflow.f90:396; iflow carries only plin here (the routine reads nothing else)
   subroutine ref_intflow_set_bcpl(nx, ny, x, xp, pl, plin, ap, bp) &
         bind(C, name="ref_intflow_set_bcpl")
      integer(c_int), value :: nx, ny
      real(c_double), dimension(nx*ny), intent(in) :: x, xp, pl
      real(c_double), intent(in) :: plin(ny)
      real(c_double), intent(inout) :: ap(nx*ny,5), bp(nx*ny)
      type(type_intflow) :: iflow
      allocate(iflow%plin(ny))
      iflow%plin = plin
      call intflow_set_bcpl(nx, ny, x, xp, pl, iflow, ap, bp)
   end subroutine

   ! mod_intflow.f90:464; iflow carries Tin and Twall (the fields the routine reads)
   subroutine ref_intflow_set_bcT(nx, ny, Twall, x, xp, T, Tin, at, bt) &
         bind(C, name="ref_intflow_set_bcT")
      integer(c_int), value :: nx, ny
      real(c_double), value :: Twall
      real(c_double), dimension(nx*ny), intent(in) :: x, xp, T
      real(c_double), intent(in) :: Tin(ny)
      real(c_double), intent(inout) :: at(nx*ny,9), bt(nx*ny)
      type(type_intflow) :: iflow
      allocate(iflow%Tin(ny))
      iflow%Tin = Tin
      iflow%Twall = Twall
      call intflow_set_bcT(nx, ny, x, xp, T, iflow, at, bt)
   end subroutine

   ! mod_intflow.f90:692
   subroutine ref_intflow_velocities_at_boundary_faces(nx, ny, modvis, x, xp, xe, ye, xk, yk, &
         u, v, ue, ve, un, vn, Uce, Vcn) bind(C, name="ref_intflow_velocities_at_boundary_faces")
      integer(c_int), value :: nx, ny, modvis
      real(c_double), dimension(nx*ny), intent(in) :: x, xp, xe, ye, xk, yk, u, v
      real(c_double), dimension(nx*ny), intent(inout) :: ue, ve, un, vn, Uce, Vcn
      call intflow_velocities_at_boundary_faces(nx, ny, modvis, x, xp, xe, ye, xk, yk, u, v, &
         ue, ve, un, vn, Uce, Vcn)
   end subroutine

   ! src/util/class_ifile.f90: init + load a parameter file, then get_value of
   ! one variable as an integer (kind 0), a real(8) (1) or a string (2)
   subroutine ref_ifile_get(fname, fs, varname, kind, ival, rval, sval, slen) &
         bind(C, name="ref_ifile_get")
      character(kind=c_char), intent(in) :: fname(*), fs(*), varname(*)
      integer(c_int), value :: kind, slen
      integer(c_int), intent(out) :: ival
      real(c_double), intent(out) :: rval
      character(kind=c_char), intent(out) :: sval(slen)
      type(class_ifile) :: f
      character(len=1000) :: fn, fsp, vn, sv
      integer :: k, iv
      real(8) :: rv
      call c_str(fname, fn)
      call c_str(fs, fsp)
      call c_str(varname, vn)
      call f%init(trim(fn), trim(fsp))
      call f%load()
      ival = 0
      rval = 0.d0
      sval = ' '
      if (kind == 0) then
         call f%get_value(iv, trim(vn))
         ival = iv
      else if (kind == 1) then
         call f%get_value(rv, trim(vn))
         rval = rv
      else
         call f%get_value(sv, trim(vn))
         do k = 1, min(slen, len_trim(sv))
            sval(k) = sv(k:k)
         end do
      end if
   contains
      subroutine c_str(c, s)
         character(kind=c_char), intent(in) :: c(*)
         character(len=*), intent(out) :: s
         integer :: i
         s = ' '
         do i = 1, len(s)
            if (c(i) == c_null_char) exit
            s(i:i) = c(i)
         end do
      end subroutine
   end subroutine

   ! the residual row and header formats of src/mach2d_main.f90:495-517
   subroutine ref_format_residual(it, dt, norm, normpl, rmass, row, header, slen) &
         bind(C, name="ref_format_residual")
      integer(c_int), value :: it, slen
      real(c_double), value :: dt, norm, normpl, rmass
      character(kind=c_char), intent(out) :: row(slen), header(slen)
      character(len=200) :: r, h
      integer :: k
      write(r, "(I10, 4(1X,ES14.7), A)") it, dt, norm, normpl, rmass, ""
      write(h, "(A10, 4(1X, A14), A)") 'it', 'dt', 'norm', 'max(|pl|)/p_avg', 'rmass', ""
      row = ' '
      header = ' '
      do k = 1, min(slen, 200)
         row(k) = r(k:k)
         header(k) = h(k:k)
      end do
   end subroutine

   ! src/tstepper/mod_tstepper_scarborough.f90: init from a parameter file
   ! (h0, mincc, maxcc), then tstepper_dt_scarborough (:80) / dt0 (:40)
   real(c_double) function ref_tstepper_dt_scarborough(fname, nx, ny, dt, au, av, at, ap) &
         bind(C, name="ref_tstepper_dt_scarborough")
      character(kind=c_char), intent(in) :: fname(*)
      integer(c_int), value :: nx, ny
      real(c_double), value :: dt
      real(c_double), intent(in) :: au(nx*ny,9), av(nx*ny,9), at(nx*ny,9), ap(nx*ny,5)
      type(class_ifile) :: f
      character(len=1000) :: fn
      integer :: i
      fn = ' '
      do i = 1, len(fn)
         if (fname(i) == c_null_char) exit
         fn(i:i) = fname(i)
      end do
      call f%init(trim(fn), "&")
      call f%load()
      call tstepper_scarborough_init(f)
      ref_tstepper_dt_scarborough = tstepper_dt_scarborough(nx, ny, dt, au, av, at, ap)
   end function

   real(c_double) function ref_tstepper_dt0_scarborough(nx, ny, Jp, u, v) &
         bind(C, name="ref_tstepper_dt0_scarborough")
      integer(c_int), value :: nx, ny
      real(c_double), dimension(nx*ny), intent(in) :: Jp, u, v
      ref_tstepper_dt0_scarborough = tstepper_dt0_scarborough(nx, ny, Jp, u, v)
   end function

   ! mod_solvers.f90:91
   subroutine ref_norm_l1_5d(nx, ny, var, b, a, norm) bind(C, name="ref_norm_l1_5d")
      integer(c_int), value :: nx, ny
      real(c_double), intent(in) :: var(nx*ny), b(nx*ny), a(nx*ny,5)
      real(c_double), intent(inout) :: norm
      call norm_l1_5d(nx, ny, var, b, a, norm)
   end subroutine

   ! mod_field_operations.f90:115
   real(c_double) function ref_get_average(nx, ny, F) bind(C, name="ref_get_average")
      integer(c_int), value :: nx, ny
      real(c_double), intent(in) :: F(nx*ny)
      real(8) :: avg
      call get_average(nx, ny, F, avg)
      ref_get_average = avg
   end function

   ! mod_intflow.f90:791
   subroutine ref_intflow_boundary_simplec(nx, ny, modvis, due, dve, dun, dvn, de, dn) &
         bind(C, name="ref_intflow_boundary_simplec")
      integer(c_int), value :: nx, ny, modvis
      real(c_double), dimension(nx*ny), intent(inout) :: due, dve, dun, dvn, de, dn
      call intflow_boundary_simplec(nx, ny, modvis, due, dve, dun, dvn, de, dn)
   end subroutine

   ! mod_intflow.f90:881
   subroutine ref_intflow_extrapolate_u_v_to_fictitious(nx, ny, modvis, x, xp, xk, yk, u, v) &
         bind(C, name="ref_intflow_extrapolate_u_v_to_fictitious")
      integer(c_int), value :: nx, ny, modvis
      real(c_double), dimension(nx*ny), intent(in) :: x, xp, xk, yk
      real(c_double), dimension(nx*ny), intent(inout) :: u, v
      call intflow_extrapolate_u_v_to_fictitious(nx, ny, modvis, x, xp, xk, yk, u, v)
   end subroutine

   ! mod_intflow.f90:557; iflow carries Tin (the field the routine reads)
   subroutine ref_intflow_boundary_T_from_H_conservation(nx, ny, Cpref, Href, Tin, u, ue, un, &
         v, ve, vn, T, Tbe, Tbw, Tbn, Tbs) bind(C, name="ref_intflow_boundary_T_from_H_conservation")
      integer(c_int), value :: nx, ny
      real(c_double), value :: Cpref, Href
      real(c_double), intent(in) :: Tin(ny)
      real(c_double), dimension(nx*ny), intent(in) :: u, ue, un, v, ve, vn
      real(c_double), intent(inout) :: T(nx*ny), Tbe(ny), Tbw(ny), Tbn(nx), Tbs(nx)
      type(type_intflow) :: iflow
      allocate(iflow%Tin(ny))
      iflow%Tin = Tin
      call intflow_boundary_T_from_H_conservation(nx, ny, Cpref, Href, iflow, u, ue, un, v, ve, &
         vn, T, Tbe, Tbw, Tbn, Tbs)
   end subroutine

end module m2dref_shim
