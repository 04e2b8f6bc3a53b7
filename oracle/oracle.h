/*
 * oracle.h -- CPU restatement of the MACH-2D SIMPLEC linear-algebra hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP product path
 * is compared against.  Only tests, the smoke check and the CPU-baseline leg
 * of the benchmark may load it; the product path never links or calls it.
 *
 * Every routine restates one reference subroutine, keeping its argument
 * meaning, its special-case rows and its floating-point operation order, so
 * that a build with -ffp-contract=off reproduces the reference bit for bit.
 * Arrays are the reference's column-major Fortran arrays seen from C:
 * a(nij,9) is `double a[9][nij]`, diagonal k (1-based) at a + (k-1)*nij.
 * Point index np = nx*(j-1)+i (1-based, i fastest).
 *
 * 9-diagonal slots: 1=SW 2=S 3=SE 4=W 5=P 6=E 7=NW 8=N 9=NE.
 * 5-diagonal slots: 1=S 2=W 3=P 4=E 5=N.
 */
#ifndef MACH2D_ORACLE_H
#define MACH2D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* src/solvers/msi2d9-v03.f90:277-432 (lu2d9).  alpha0 is 0.5 in the
 * reference (:287); it is a parameter here so tests can force alpha = 0. */
void orc_lu2d9(const double *b, double *dl, double *du, int nij, int ni,
               int nj, double alpha0);

/* src/solvers/msi2d9-v03.f90:42-262 (fb2d9).  Extra outputs, which the
 * reference does not return: *nit_out (iterations run), *r0_out (||r0||_2),
 * hist[nit-1] = ||r_nit||_2 (hist may be NULL; hist_cap bounds it).
 * max_iter_cap guards nitm <= 0, which never exits on count in the
 * reference (:259); pass 0 for no guard. */
void orc_fb2d9(const double *b, const double *dl, const double *du, int nj,
               int ni, int nij, double *var, double gamma, int nitm,
               const double *rhs, int *nit_out, double *r0_out, double *hist,
               int hist_cap, int max_iter_cap);

/* src/solvers/msi2d5-v03.f90:247-330 (lu2d5) and :42-233 (fb2d5). */
void orc_lu2d5(const double *b, double *dl, double *du, int nij, int ni,
               int nj, double alpha0);
void orc_fb2d5(const double *b, const double *dl, const double *du, int nj,
               int ni, int nij, double *var, double gamma, int nitm,
               const double *rhs, int *nit_out, double *r0_out, double *hist,
               int hist_cap, int max_iter_cap);

/* src/solvers/tdma2d9.f90:38-322 (tdma2d9); line-Jacobi, serial here
 * (the reference's OpenMP split over lines does not change results). */
void orc_tdma2d9(int nx, int ny, int nitm, const double *a, const double *b,
                 double *x);

/* src/solvers/mod_solvers.f90:128-170, 176-245, 250-276. */
void orc_norm_l1_9d(int nx, int ny, const double *var, const double *b,
                    const double *a, double *norm);
void orc_norm_l1_9d_relative(int nx, int ny, const double *var,
                             const double *b, const double *a, double *norm);
double orc_norm_l1_b(int nx, int ny, const double *b);

/* r = A.var - rhs over all nij rows, the residual of fb2d9 (:59-131). */
void orc_residual9(const double *b, int ni, int nj, const double *var,
                   const double *rhs, double *r);

#ifdef __cplusplus
}
#endif

/* coef.c: coefficient assembly of src/equations/mach2d_coef.f90 */
void orc_density_at_nodes(int nx, int ny, double Rg, const double *p, const double *T,
                          double *ro);
void orc_density_at_faces(int nx, int ny, double beta, const double *ro, const double *Uce,
                          const double *Vcn, double *roe, double *ron);
void orc_u_coefficients(int nx, int ny, int modvis, double dt, const double *rp, const double *re,
                        const double *rn, const double *Jp, const double *Je, const double *Jn,
                        const double *ye, const double *yk, const double *alphae,
                        const double *betae, const double *betan, const double *gamman,
                        const double *vle, const double *vln, const double *roe,
                        const double *ron, const double *roa, const double *Uce,
                        const double *Vcn, double *a);
void orc_v_coefficients(int nx, int ny, int coord, int modvis, double dt, const double *rp,
                        const double *re, const double *rn, const double *Jp, const double *Je,
                        const double *Jn, const double *xe, const double *xk,
                        const double *alphae, const double *betae, const double *betan,
                        const double *gamman, const double *vle, const double *vln,
                        const double *vlp, const double *roe, const double *ron,
                        const double *roa, const double *Uce, const double *Vcn, double *a);
void orc_p_coefficients(int nx, int ny, double dt, const double *rp, const double *re,
                        const double *rn, const double *Jp, const double *Uce, const double *Vcn,
                        const double *roe, const double *ron, const double *g, const double *de,
                        const double *dn, double *a);
void orc_p_source(int nx, int ny, double dt, const double *rp, const double *re,
                  const double *rn, const double *Jp, const double *roe, const double *ron,
                  const double *ro, const double *roa, const double *Uce, const double *Vcn,
                  const double *de, const double *dn, const double *g, const double *pl,
                  double *b);

/* bc.c: fictitious-row writers of src/user/internal_flow/mod_intflow.f90 */
void orc_u_v_north_slip(int nx, int ny, const double *x, const double *xp, const double *xk,
                        const double *yk, const double *u, const double *v, double *ubn,
                        double *vbn);
void orc_intflow_set_bcu(int nx, int ny, int modvis, const double *x, const double *xp,
                         const double *xk, const double *yk, const double *u, const double *v,
                         double *au, double *b);
void orc_intflow_set_bcv(int nx, int ny, int modvis, const double *x, const double *xp,
                         const double *xk, const double *yk, const double *u, const double *v,
                         double *av, double *b);
void orc_intflow_set_bcpl(int nx, int ny, const double *x, const double *xp, const double *pl,
                          const double *plin, double *a, double *bp);
void orc_intflow_set_bcT(int nx, int ny, double Twall, const double *x, const double *xp,
                         const double *T, const double *Tin, double *a, double *bt);
void orc_intflow_velocities_at_boundary_faces(int nx, int ny, int modvis, const double *x,
                                              const double *xp, const double *xe,
                                              const double *ye, const double *xk,
                                              const double *yk, const double *u, const double *v,
                                              double *ue, double *ve, double *un, double *vn,
                                              double *Uce, double *Vcn);

/* tstep.c: Scarborough time-step statistics, src/tstepper/mod_tstepper_scarborough.f90 */
double orc_get_mean_cc_9d(int nx, int ny, const double *a);
double orc_get_mean_cc_5d(int nx, int ny, const double *a);
double orc_tstepper_dt0_scarborough(int nx, int ny, const double *Jp, const double *u,
                                    const double *v);
double orc_tstepper_dt_scarborough(int nx, int ny, double dt, const double *au,
                                   const double *av, const double *at, const double *ap,
                                   double h0, double mincc, double maxcc);


/* grid.c: centroids, metrics and radii, src/grid/mod_grid.f90:237-723 */
void orc_get_real_centroids_xy(int opt, int nx, int ny, const double *x, const double *y,
                               double *xp, double *yp);
void orc_get_metrics(int nx, int ny, const double *x, const double *y, const double *xp,
                     const double *yp, double *xe, double *ye, double *xen, double *yen,
                     double *xk, double *yk, double *xke, double *yke, double *Jp, double *Je,
                     double *Jn, double *alphae, double *gamman, double *betae, double *betan);
void orc_get_radius(int coord, int nx, int ny, const double *y, const double *yp,
                    double *radius, double *re, double *rn, double *rp);


/* intflow.c: the rest of one internal-flow time step, mod_intflow.f90:557-1134,
 * mod_solvers.f90:91-125, mod_field_operations.f90:115-142 */
void orc_intflow_boundary_T_from_H_conservation(int nx, int ny, double Cpref, double Href,
                                                const double *Tin, const double *u,
                                                const double *ue, const double *un,
                                                const double *v, const double *ve,
                                                const double *vn, double *T, double *Tbe,
                                                double *Tbw, double *Tbn, double *Tbs);
void orc_intflow_boundary_simplec(int nx, int ny, int modvis, double *due, double *dve,
                                  double *dun, double *dvn, double *de, double *dn);
void orc_intflow_extrapolate_u_v_to_fictitious(int nx, int ny, int modvis, const double *x,
                                               const double *xp, const double *xk,
                                               const double *yk, double *u, double *v);
void orc_get_uin_vin_pin_Tin_Mw(int nx, int ny, double gamma, double Rg, double po, double T0,
                                const double *u, double *uin, double *vin, double *pin,
                                double *Tin, double *Mw);
void orc_get_plin_and_p_fictitious(int nx, int ny, const double *x, const double *xp,
                                   const double *pina, const double *pin, double *plin,
                                   double *p);
void orc_intflow_time_cycle_update(int nx, int ny, double gamma, double Rg, double P0,
                                   double T0, const double *x, const double *xp,
                                   const double *u, double *uin, double *vin, double *pin,
                                   double *Tin, double *Mw, double *plin, double *pina,
                                   double *p);
void orc_norm_l1_5d(int nx, int ny, const double *var, const double *b, const double *a,
                    double *norm);
double orc_get_average(int nx, int ny, const double *F);

#endif
