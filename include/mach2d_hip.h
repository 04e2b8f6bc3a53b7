/*
 * mach2d_hip.h -- C ABI of the MI355X (gfx950) SIMPLEC linear-algebra hot
 * path: 9-diagonal MSI (modified strongly implicit) factorization and solve,
 * alternating-direction line TDMA, and L1 residual norms.
 *
 * Layout everywhere: the reference's column-major Fortran arrays seen from C.
 * A 9-diagonal coefficient array a(nij,9) is `double a[9][nij]` (diagonal k,
 * 1-based, at a + (k-1)*nij), slots 1=SW 2=S 3=SE 4=W 5=P 6=E 7=NW 8=N 9=NE;
 * the point index is np = ni*(j-1)+i (1-based, i fastest).
 *
 * Two families of entry points:
 *
 *  1. Reference-signature entry points (host arrays, synchronous, every
 *     argument by reference, so a Fortran `bind(C)` interface with no VALUE
 *     attributes binds them directly).  They replace the reference routines
 *     named in each comment and keep their argument order and meaning.
 *     Status, iteration count and residual norms of the last call, which the
 *     reference signatures cannot return, are read with m2d_last_status().
 *
 *  2. A handle API for device-resident data (device pointers; no host
 *     copies), used for batching, graph capture and the mixed-precision and
 *     learned-guess extensions.
 *
 * Every entry point requires the invariant the reference states at
 * src/solvers/msi2d9-v03.f90:16-20: coefficients that couple a point to a
 * neighbour outside the grid are exactly zero.  Inputs that break it are
 * rejected with M2D_ERR_OUTSIDE_COUPLING (the reference would silently couple
 * across row ends through its flat index).
 */
#ifndef MACH2D_HIP_H
#define MACH2D_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (m2d_status.status and the int return values). */
enum {
    M2D_OK = 0,
    M2D_ERR_NONFINITE = 1,          /* zero/non-finite pivot or residual     */
    M2D_ERR_HIP = 3,                /* a HIP runtime call failed             */
    M2D_ERR_ARG = 4,                /* invalid argument (sizes, pointers)    */
    M2D_ERR_OUTSIDE_COUPLING = 5,   /* a coupling to an outside neighbour    */
    M2D_ERR_SYNC_TIMEOUT = 6,       /* inter-workgroup wait gave up          */
    M2D_ERR_ITER_CAP = 7            /* nitm <= 0 and the safety cap reached  */
};

/* Precision of a handle. */
enum {
    M2D_FP64 = 0,       /* FP64 factorization and sweeps (parity mode)        */
    M2D_FP32_IR = 1     /* FP32 factorization and sweeps inside FP64
                           iterative refinement (FP64 residual and update)    */
};

typedef struct m2d_status {
    int status;         /* M2D_OK or an error code above                      */
    int nit;            /* MSI iterations run (fb2d9's nit)                   */
    int nitm_reached;   /* 1 when the loop ended on nit == nitm               */
    int outer;          /* refinement steps (M2D_FP32_IR), else 0             */
    double r0;          /* ||A.var_in - rhs||_2                               */
    double rn;          /* ||A.var_out - rhs||_2                              */
    double gammarp;     /* gamma * r0, the stopping threshold                 */
    double device_ms;   /* device time of the call's kernels                  */
} m2d_status;

/* ---- 1. Reference-signature entry points (host arrays) ----------------- */

/* Replaces lu2d9(b,dl,du,nij,ni,nj), src/solvers/msi2d9-v03.f90:277-282.
 * Note the argument order (nij, ni, nj) of the reference. */
void m2d_lu2d9(const double *b, double *dl, double *du, const int *nij,
               const int *ni, const int *nj);

/* Replaces fb2d9(b,dl,du,nj,ni,nij,var,gamma,nitm,rhs),
 * src/solvers/msi2d9-v03.f90:42-50.  Note the order (nj, ni, nij).
 * Stops when ||r||_2 < gamma*||r0||_2 or after nitm iterations; at least one
 * iteration always runs; nitm <= 0 means no count limit (a safety cap of
 * 1,000,000 iterations then applies, status M2D_ERR_ITER_CAP). */
void m2d_fb2d9(const double *b, const double *dl, const double *du,
               const int *nj, const int *ni, const int *nij, double *var,
               const double *gamma, const int *nitm, const double *rhs);

/* Replaces tdma2d9(nx,ny,nitm,a,b,x), src/solvers/tdma2d9.f90:38-45. */
void m2d_tdma2d9(const int *nx, const int *ny, const int *nitm,
                 const double *a, const double *b, double *x);

/* Replaces class_solver_msi9d%solve(a,b,sol) after init(nx,ny,nitm,tol),
 * src/solvers/mod_solver_msi.f90:96-133: lu2d9 then fb2d9. */
void m2d_solve_msi9d(const int *nx, const int *ny, const int *nitm,
                     const double *tol, const double *a, const double *b,
                     double *sol);

/* Replaces lu2d5(b,dl,du,nij,ni,nj), src/solvers/msi2d5-v03.f90:247-330:
 * the 5-diagonal MSI factorization.  b(nij,5) holds the slots S, W, P, E, N
 * (src/solvers/tdma2d5.f90:47-51); out: dl(nij,4) (P, W, SE, S) and
 * du(nij,3) (E, NW, N).  Computed as the 9-diagonal factorization of the
 * zero-corner embedding of b, whose terms reduce to lu2d5's. */
void m2d_lu2d5(const double *b, double *dl, double *du, const int *nij,
               const int *ni, const int *nj);

/* Replaces fb2d5(b,dl,du,nj,ni,nij,var,gamma,nitm,rhs),
 * src/solvers/msi2d5-v03.f90:42-233 (argument order (nj, ni, nij)); the
 * stopping rule and limits of m2d_fb2d9, with the 5-diagonal residual and
 * sweeps. */
void m2d_fb2d5(const double *b, const double *dl, const double *du,
               const int *nj, const int *ni, const int *nij, double *var,
               const double *gamma, const int *nitm, const double *rhs);

/* Replaces class_solver_msi5d%solve(a,b,sol) after init(nx,ny,nitm,tol),
 * src/solvers/mod_solver_msi.f90:56-93: lu2d5 then fb2d5. */
void m2d_solve_msi5d(const int *nx, const int *ny, const int *nitm,
                     const double *tol, const double *a, const double *b,
                     double *sol);

/* No counterpart in the reference.  m2d_solve_msi9d on DEVICE arrays a, b,
 * sol (same meaning, FP64, the same results bit for bit) that allocates
 * nothing: for callers that cannot call hipMalloc (tight serving loops,
 * harnesses that forbid allocation).  Every buffer beyond a, b and sol
 * comes from a 48 MiB device array the library holds on each GPU it runs
 * on until the process ends (grids up to about 550x270 fit; larger ones
 * return M2D_ERR_ARG with a message).  Calls that overlap on one GPU race
 * on that array.  Runs on `stream` (hipStream_t as void*, NULL = default)
 * and returns when sol is final; returns the status (also in
 * m2d_last_status). */
int m2d_solve_msi9d_noalloc(const int *nx, const int *ny, const int *nitm,
                            const double *tol, const double *a, const double *b,
                            double *sol, void *stream);

/* Replace norm_l1_9d, norm_l1_9d_relative (accumulate into *norm) and
 * norm_l1_b, src/solvers/mod_solvers.f90:128-170, 176-245, 250-276. */
void m2d_norm_l1_9d(const int *nx, const int *ny, const double *var,
                    const double *b, const double *a, double *norm);
void m2d_norm_l1_9d_relative(const int *nx, const int *ny, const double *var,
                             const double *b, const double *a, double *norm);
double m2d_norm_l1_b(const int *nx, const int *ny, const double *b);
/* norm_l1_5d, mod_solvers.f90:91-125 (a: 5 diagonals S, W, P, E, N; host
 * or device arrays, norm a host scalar) */
void m2d_norm_l1_5d(const int *nx, const int *ny, const double *var, const double *b,
                    const double *a, double *norm);
/* get_average, src/util/mod_field_operations.f90:115-142: the mean over
 * real volumes, added in the reference's order (bit for bit); host or
 * device array */
double m2d_get_average(const int *nx, const int *ny, const double *F);

/* Status of the last call made by this host thread (no reference
 * counterpart: the reference routines return nothing and fail silently). */
int m2d_last_status(m2d_status *out);

/* ---- 2. Handle API (device pointers) ----------------------------------- */

typedef struct m2d_system m2d_system;

/* Create a solver for one nx*ny grid; precision is M2D_FP64 or M2D_FP32_IR.
 * Allocates the factor workspace (dl, du: 9*nx*ny values) on the current
 * device.  Mirrors class_solver_msi9d%init, mod_solver_msi.f90:96-115. */
int m2d_create(int nx, int ny, int precision, m2d_system **out);

/* As m2d_create for A with ndiag = 9 or 5 diagonals.  ndiag = 5: A holds the
 * slots S, W, P, E, N (5*nx*ny doubles) of msi2d5 (class_solver_msi5d,
 * mod_solver_msi.f90:56-93); factor, solve, residual norm and guess scaling
 * then follow lu2d5 / fb2d5, and m2d_get_factor returns dl 4*nx*ny and
 * du 3*nx*ny values. */
int m2d_create_stencil(int nx, int ny, int precision, int ndiag, m2d_system **out);
int m2d_destroy(m2d_system *s);

/* Run on a HIP stream (hipStream_t passed as void*; NULL = default). */
int m2d_set_stream(m2d_system *s, void *stream);

/* Factor the device array d_a (9*nx*ny doubles, layout above).  In FP32_IR
 * mode the factor and an FP32 copy of A are kept in FP32. */
int m2d_factor(m2d_system *s, const double *d_a);

/* Solve A.var = rhs (device pointers) from the var passed in, with the
 * stopping rule of fb2d9: ||r||_2 < gamma*||r0||_2 or nitm iterations.
 * d_hist, if not NULL, receives ||r_k||_2 for k = 1..nit (capacity hist_cap).
 * In FP32_IR mode, gamma and nitm bound the FP64 outer residual; inner FP32
 * sweeps run until their own residual drops by `inner_gamma` (set with
 * m2d_set_inner) or `inner_nitm` iterations (defaults 1e-4 and 25: short
 * inner runs between FP64 residuals converge in about as many iterations as
 * plain FP64 MSI). */
int m2d_solve(m2d_system *s, const double *d_a, const double *d_rhs,
              double *d_var, double gamma, int nitm, double *d_hist,
              int hist_cap, m2d_status *st);

int m2d_set_inner(m2d_system *s, double inner_gamma, int inner_nitm);

/* ||A.var - rhs||_2 in FP64 (device pointers), the norm fb2d9 starts from
 * (msi2d9-v03.f90:59-133); used to judge an initial guess against the zero
 * guess before a solve. */
int m2d_residual_norm(m2d_system *s, const double *d_a, const double *d_rhs,
                      const double *d_var, double *out);

/* Scale an initial guess g in place by the alpha that minimises
 * ||A (alpha g) - rhs||_2, alpha = (A g . rhs) / (A g . A g) (0 when A g = 0),
 * so the scaled guess is never worse than the zero guess.  No reference
 * counterpart. */
int m2d_guess_scale(m2d_system *s, const double *d_a, const double *d_rhs, double *d_g,
                    double *alpha_out);

/* As m2d_guess_scale with alpha = (g . rhs) / (g . A g) (0 when g.A g <= 0):
 * the amplitude that minimises the energy functional of A (for a symmetric
 * positive definite A, the A-norm of the error), which weighs the smooth
 * error components MSI removes slowest.  No reference counterpart. */
int m2d_guess_scale_energy(m2d_system *s, const double *d_a, const double *d_rhs, double *d_g,
                           double *alpha_out);

/* m2d_guess_scale(_energy) (energy = 0 / 1) for `count` systems of the
 * handle's grid (device arrays d_a[q], d_rhs[q], d_g[q]) with two host reads
 * in all: alpha_out[q] and rnorm_out[q] = ||rhs_q - A_q g_q|| of the scaled
 * guess.  No reference counterpart (batched learned guesses). */
int m2d_guess_prepare_many(m2d_system *s, int count, const double *const *d_a,
                           const double *const *d_rhs, double *const *d_g, int energy,
                           double *alpha_out, double *rnorm_out);

/* Device copies of the factor (dl 5*nx*ny, du 4*nx*ny doubles, reference
 * layout); FP64 handles only. */
int m2d_get_factor(m2d_system *s, double *d_dl, double *d_du);

/* Launch-shape tuning of the wavefront kernels (autotuner hook):
 * publish_every = columns between progress publications of a band;
 * xcd_stride = 1 (bands on consecutive XCDs) or 8 (bands of one system on
 * one XCD).  Default: 8 for systems of five or more 64-row bands (ny > 256),
 * else 1. */
int m2d_set_tuning(m2d_system *s, int publish_every, int xcd_stride);

/* Sweep operand path: 1 = the LDS-staged sweeps (a second wave loads each
 * chunk's operands into an LDS ring ahead of use), 0 (default) = operands
 * loaded into registers by the computing wave.  Same results either way. */
int m2d_set_staging(m2d_system *s, int staged);

/* ---- 3. Batched API (device pointers) ---------------------------------- */
/* Many independent systems of uneven grid size solved together: one launch
 * per phase for the whole batch, iterations replayed from a captured
 * hipGraph, launch placement picked by an autotuner.  No reference
 * counterpart (the reference solves one system per call); each system
 * follows fb2d9's semantics exactly as m2d_solve does.  Arrays of device
 * pointers are host arrays with `count` entries, indexed like nx/ny. */
typedef struct m2d_batch m2d_batch;

int m2d_batch_create(int count, const int *nx, const int *ny, int precision,
                     m2d_batch **out);
int m2d_batch_destroy(m2d_batch *b);

/* A non-NULL stream for the batch's work; NULL restores the batch's own
 * stream, which first waits for the legacy default stream.  Calls return
 * after the work finished. */
int m2d_batch_set_stream(m2d_batch *b, void *stream);

/* Factor every system (lu2d9 on each d_a[q]). */
int m2d_batch_factor(m2d_batch *b, const double *const *d_a);

/* Solve every system from its var: stopping rule of fb2d9 per system
 * (gamma, nitm as in m2d_solve); st (count entries, may be NULL) receives
 * each system's status; device_ms is the batch's time.  Returns the first
 * non-zero status. */
int m2d_batch_solve(m2d_batch *b, const double *const *d_a, const double *const *d_rhs,
                    double *const *d_var, double gamma, int nitm, m2d_status *st);

/* As m2d_batch_solve with one gamma per system (host array of count
 * values): lets every system stop at its own threshold, e.g. a fixed
 * residual level when systems start from different initial guesses. */
int m2d_batch_solve_each(m2d_batch *b, const double *const *d_a, const double *const *d_rhs,
                         double *const *d_var, const double *gammas, int nitm,
                         m2d_status *st);

/* FP32_IR: the inner stopping rule (as m2d_set_inner). */
int m2d_batch_set_inner(m2d_batch *b, double inner_gamma, int inner_nitm);

/* Launch placement: layout 0 = systems in order, 1 = each system's bands on
 * one XCD (longest-processing-time assignment), 2 = largest systems first;
 * layouts 3..5 = the same placements with the LDS-staged sweeps (operands
 * loaded into an LDS ring ahead of use by a second wave; more LDS per
 * workgroup, deeper prefetch); chunk_iters = iterations per graph replay
 * (1..64). */
int m2d_batch_set_tuning(m2d_batch *b, int layout, int chunk_iters);
int m2d_batch_get_tuning(m2d_batch *b, int *layout, int *chunk_iters);

/* Autotuner: time `iters` iterations of each layout on these systems (from
 * a zero iterate in scratch memory; the caller's var is not touched) and
 * keep the fastest.  ms_out (6 entries, may be NULL): time per iteration. */
int m2d_batch_tune(m2d_batch *b, const double *const *d_a, const double *const *d_rhs,
                   int iters, int *layout_out, double *ms_out);

/* ---- 4. Coefficient assembly (src/equations/mach2d_coef.f90) ----------- */
/* Each entry mirrors a routine of module `coefficients` with the same
 * argument list and order, every argument by pointer (bind(C) from Fortran
 * unchanged).  Arrays: N = nx*ny doubles, reference layout; coefficient
 * arrays a(N,9) / a(N,5) column-major.  Arrays may be host memory (copied,
 * call returns with the results) or device memory (all of one kind;
 * processed on the assembly stream, call returns once enqueued).  Only the
 * points the reference writes are written.  Status: m2d_last_status. */
void m2d_assembly_set_stream(void *stream);

/* get_density_at_nodes, mach2d_coef.f90:6 (every point) */
void m2d_get_density_at_nodes(const int *nx, const int *ny, const double *Rg, const double *p,
                              const double *T, double *ro);
/* get_density_at_faces, mach2d_coef.f90:20 */
void m2d_get_density_at_faces(const int *nx, const int *ny, const double *beta,
                              const double *ro, const double *Uce, const double *Vcn,
                              double *roe, double *ron);
/* get_u_coefficients, mach2d_coef.f90:63 */
void m2d_get_u_coefficients(const int *nx, const int *ny, const int *modvis, const double *dt,
                            const double *rp, const double *re, const double *rn,
                            const double *Jp, const double *Je, const double *Jn,
                            const double *ye, const double *yk, const double *alphae,
                            const double *betae, const double *betan, const double *gamman,
                            const double *vle, const double *vln, const double *roe,
                            const double *ron, const double *roa, const double *Uce,
                            const double *Vcn, double *a);
/* get_u_source, mach2d_coef.f90:221 */
void m2d_get_u_source(const int *nx, const int *ny, const int *modvis, const double *beta,
                      const double *dt, const double *rp, const double *re, const double *rn,
                      const double *xe, const double *ye, const double *xk, const double *yk,
                      const double *xke, const double *yke, const double *xen,
                      const double *yen, const double *Jp, const double *Je, const double *Jn,
                      const double *roe, const double *ron, const double *roa, const double *p,
                      const double *vle, const double *vln, const double *Uce,
                      const double *Vcn, const double *ua, const double *u, const double *v,
                      double *cup, double *sup, double *b);
/* get_v_coefficients, mach2d_coef.f90:454 */
void m2d_get_v_coefficients(const int *nx, const int *ny, const int *coord, const int *modvis,
                            const double *dt, const double *rp, const double *re,
                            const double *rn, const double *Jp, const double *Je,
                            const double *Jn, const double *xe, const double *xk,
                            const double *alphae, const double *betae, const double *betan,
                            const double *gamman, const double *vle, const double *vln,
                            const double *vlp, const double *roe, const double *ron,
                            const double *roa, const double *Uce, const double *Vcn, double *a);
/* get_v_source, mach2d_coef.f90:616 */
void m2d_get_v_source(const int *nx, const int *ny, const int *coord, const int *modvis,
                      const double *beta, const double *dt, const double *rp, const double *re,
                      const double *rn, const double *xe, const double *ye, const double *xk,
                      const double *yk, const double *xke, const double *yke,
                      const double *xen, const double *yen, const double *Jp, const double *Je,
                      const double *Jn, const double *roe, const double *ron, const double *roa,
                      const double *p, const double *vle, const double *vln, const double *Uce,
                      const double *Vcn, const double *va, const double *u, const double *v,
                      double *cvp, double *svp, double *b);
/* get_T_coefficients_and_source, mach2d_coef.f90:787 */
void m2d_get_T_coefficients_and_source(
    const int *nx, const int *ny, const int *modvis, const double *beta, const double *dt,
    const double *rp, const double *re, const double *rn, const double *xe, const double *ye,
    const double *xk, const double *yk, const double *xke, const double *yke, const double *xen,
    const double *yen, const double *alphae, const double *betae, const double *betan,
    const double *gamman, const double *Jp, const double *Je, const double *Jn,
    const double *roe, const double *ron, const double *ro, const double *roa, const double *p,
    const double *pa, const double *cp, const double *vle, const double *vln, const double *ke,
    const double *kn, const double *Uce, const double *Vcn, const double *u, const double *v,
    const double *ue, const double *ve, const double *un, const double *vn, const double *ua,
    const double *va, const double *T, const double *Ta, const double *fbe, const double *fbn,
    double *a, double *b);
/* get_p_coefficients, mach2d_coef.f90:1207 (a: 5 diagonals) */
void m2d_get_p_coefficients(const int *nx, const int *ny, const double *dt, const double *rp,
                            const double *re, const double *rn, const double *Jp,
                            const double *Uce, const double *Vcn, const double *roe,
                            const double *ron, const double *g, const double *de,
                            const double *dn, double *a);
/* get_p_source, mach2d_coef.f90:1281 */
void m2d_get_p_source(const int *nx, const int *ny, const double *dt, const double *rp,
                      const double *re, const double *rn, const double *Jp, const double *roe,
                      const double *ron, const double *ro, const double *roa, const double *Uce,
                      const double *Vcn, const double *de, const double *dn, const double *g,
                      const double *pl, double *b);

/* SIMPLEC glue between the solves (mach2d_coef.f90:1350-1880); au, av are
 * the 9-diagonal u / v coefficient arrays.  In/out arrays as in the
 * reference (p, ro, u, v, ue, ve, un, vn, Uce, Vcn where marked inout). */
/* get_velocities_at_internal_faces, :1350 */
void m2d_get_velocities_at_internal_faces(
    const int *nx, const int *ny, const double *dt, const double *rp, const double *re,
    const double *rn, const double *xe, const double *ye, const double *xk, const double *yk,
    const double *xen, const double *yen, const double *xke, const double *yke, const double *Jp,
    const double *cup, const double *cvp, const double *sup, const double *svp,
    const double *au, const double *av, const double *roa, const double *p, const double *u,
    const double *v, const double *uea, const double *vea, const double *una, const double *vna,
    double *ue, double *ve, double *un, double *vn, double *Uce, double *Vcn);
/* get_internal_simplec_coefficients, :1504 */
void m2d_get_internal_simplec_coefficients(
    const int *nx, const int *ny, const double *re, const double *rn, const double *xe,
    const double *ye, const double *xk, const double *yk, const double *au, const double *av,
    double *due, double *dve, double *dun, double *dvn, double *de, double *dn);
/* get_pressure_correction_with_pl, :1602 (p inout, every point) */
void m2d_get_pressure_correction_with_pl(const int *nx, const int *ny, const double *pl,
                                         double *p);
/* get_density_correction_with_pl, :1614 (ro inout, every point) */
void m2d_get_density_correction_with_pl(const int *nx, const int *ny, const double *pl,
                                        const double *g, double *ro);
/* get_u_v_at_real_nodes_with_pl, :1626 (u, v inout) */
void m2d_get_u_v_at_real_nodes_with_pl(const int *nx, const int *ny, const double *xe,
                                       const double *ye, const double *xk, const double *yk,
                                       const double *rp, const double *pl, const double *au,
                                       const double *av, double *u, double *v);
/* get_velocity_correction_at_faces_with_pl, :1679 (last six inout) */
void m2d_get_velocity_correction_at_faces_with_pl(
    const int *nx, const int *ny, const double *xe, const double *ye, const double *xk,
    const double *yk, const double *due, const double *dve, const double *dun, const double *dvn,
    const double *de, const double *dn, const double *pl, double *ue, double *ve, double *un,
    double *vn, double *Uce, double *Vcn);
/* initialize_boundary_faces_flags, :1764 */
void m2d_initialize_boundary_faces_flags(const int *nx, const int *ny, double *fbe,
                                         double *fbn);
/* check_mass_conservation, :1833 (res: L1 norm, summed in the reference's
 * loop order, so bit-identical) */
void m2d_check_mass_conservation(const int *nx, const int *ny, const double *dt,
                                 const double *rp, const double *re, const double *rn,
                                 const double *Jp, const double *ro, const double *roa,
                                 const double *ron, const double *roe, const double *Uce,
                                 const double *Vcn, double *res);

/* Fictitious-row writers of the internal-flow case,
 * src/user/internal_flow/mod_intflow.f90.  Each overwrites the coefficient
 * rows (all diagonals) and the source entries of the boundary it owns and
 * leaves every other row as the caller passed it.  The reference passes
 * `type(type_intflow) :: iflow`; the fields these routines read cross the C
 * ABI as plain arguments: plin / Tin (iflow%plin / iflow%Tin, ny values)
 * and Twall (iflow%Twall, a scalar, listed first like the other scalars).
 * Host or device arrays, as for the assembly entries above. */
/* intflow_set_bcu, mod_intflow.f90:214 (Euler north rows use the slip
 * velocity of get_u_v_north_slip, :142, which needs nx >= 4) */
void m2d_intflow_set_bcu(const int *nx, const int *ny, const int *modvis, const double *x,
                         const double *xp, const double *xk, const double *yk, const double *u,
                         const double *v, double *au, double *b);
/* intflow_set_bcv, :305 */
void m2d_intflow_set_bcv(const int *nx, const int *ny, const int *modvis, const double *x,
                         const double *xp, const double *xk, const double *yk, const double *u,
                         const double *v, double *av, double *b);
/* intflow_set_bcpl, :396 (5 diagonals S, W, P, E, N; corner rows untouched) */
void m2d_intflow_set_bcpl(const int *nx, const int *ny, const double *x, const double *xp,
                          const double *pl, const double *plin, double *ap, double *bp);
/* intflow_set_bcT, :464 (Twall > 0: prescribed north wall temperature) */
void m2d_intflow_set_bcT(const int *nx, const int *ny, const double *Twall, const double *x,
                         const double *xp, const double *T, const double *Tin, double *at,
                         double *bt);
/* intflow_velocities_at_boundary_faces, :692-787 (in/out: only the boundary
 * faces are written; modvis = 0 gives the north face the Euler slip
 * velocity) */
void m2d_intflow_velocities_at_boundary_faces(const int *nx, const int *ny, const int *modvis,
                                              const double *x, const double *xp,
                                              const double *xe, const double *ye,
                                              const double *xk, const double *yk,
                                              const double *u, const double *v, double *ue,
                                              double *ve, double *un, double *vn, double *Uce,
                                              double *Vcn);

/* intflow_boundary_simplec, :791-877 (in/out; modvis unused, kept for the
 * reference's argument list) */
void m2d_intflow_boundary_simplec(const int *nx, const int *ny, const int *modvis, double *due,
                                  double *dve, double *dun, double *dvn, double *de,
                                  double *dn);
/* intflow_extrapolate_u_v_to_fictitious, :881-964 (u, v in/out; nx >= 4,
 * the north slip extrapolation reads the third column) */
void m2d_intflow_extrapolate_u_v_to_fictitious(const int *nx, const int *ny, const int *modvis,
                                               const double *x, const double *xp,
                                               const double *xk, const double *yk, double *u,
                                               double *v);
/* intflow_boundary_T_from_H_conservation, :557-688 (iflow contributes Tin,
 * ny values; T in/out; Tbe, Tbw: ny values, Tbn, Tbs: nx values) */
void m2d_intflow_boundary_T_from_H_conservation(const int *nx, const int *ny,
                                                const double *Cpref, const double *Href,
                                                const double *Tin, const double *u,
                                                const double *ue, const double *un,
                                                const double *v, const double *ve,
                                                const double *vn, double *T, double *Tbe,
                                                double *Tbw, double *Tbn, double *Tbs);
/* intflow_time_cycle_update, :1106-1134: gamma = thermomodel%gm(T0),
 * Rg = thermomodel%Rg, P0 and T0 of iflow; the inflow vectors (ny values)
 * are the iflow fields of the same names; p in/out.  The inflow pressure is
 * a real power: the device pow may differ from the host libm in the last
 * bit, so pin, plin and the fictitious pressures agree to 2 ulp. */
void m2d_intflow_time_cycle_update(const int *nx, const int *ny, const double *gamma,
                                   const double *Rg, const double *P0, const double *T0,
                                   const double *x, const double *xp, const double *u,
                                   double *uin, double *vin, double *pin, double *Tin,
                                   double *Mw, double *plin, double *pina, double *p);

/* Device arrays of the internal-flow pressure-correction loop: N-vectors in
 * the reference layout (N = nx*ny), plin of ny values. */
typedef struct m2d_pl_loop_fields {
    /* grid and metrics */
    const double *x, *xp, *xe, *ye, *xk, *yk, *rp, *re, *rn, *Jp;
    /* fields frozen during the loop: SIMPLEC coefficients, g = 1/(R T),
     * density of the previous time step, nodal velocities, inlet pl */
    const double *due, *dve, *dun, *dvn, *de, *dn, *g, *roa, *u, *v, *plin;
    /* updated in place */
    double *ue, *ve, *un, *vn, *Uce, *Vcn, *roe, *ron, *ro, *p, *pl;
    /* the p' system, persistent as in the reference: ap (5 N, slots S, W, P,
     * E, N) and bp (N); the corner rows keep their initial values (the
     * reference allocates ap = 0 with ap(:,3) = 1, mod_mach2d_data.f90:
     * 401-403) */
    double *ap, *bp;
} m2d_pl_loop_fields;

/* The pressure-correction loop of one SIMPLEC iteration of the internal-flow
 * case, src/mach2d_main.f90:257-309: pl = 0, then for itp = 1..itpmax:
 * get_p_coefficients, intflow_set_bcpl, get_p_source, the 5-diagonal MSI
 * solve (lu2d5 + fb2d5 from the current pl, tol_p, nitm_p),
 * get_velocity_correction_at_faces_with_pl,
 * intflow_velocities_at_boundary_faces, get_pressure_correction_with_pl,
 * get_density_correction_with_pl, get_density_at_faces, and the exit test
 * max|pl| / p_avg < tolm.  Everything runs on the handle's stream on device
 * arrays.  The exit test runs on the device: once it fires (or the solve
 * reports an error) every kernel of the iterations enqueued after it
 * returns at once.  The iterations go out in batches, replayed from HIP
 * graphs captured on first use (m2d_set_pl_graph), each batch followed by
 * one host read; the first batch is as long as the previous call's loop,
 * later ones one iteration each, so a steady run reads the host once per
 * call.  s5: a 5-diagonal FP64 handle (m2d_create_stencil) of the grid,
 * not shared with another thread while the call runs.  itp_out: itp
 * iterations run; normpl_out: the last max|pl| / p_avg. */
int m2d_intflow_pl_loop(m2d_system *s5, const m2d_pl_loop_fields *f, int modvis, double dt,
                        double beta, double p_avg, double tolm, int itpmax, int nitm_p,
                        double tol_p, int *itp_out, double *normpl_out);

/* No reference counterpart (tuning and introspection of the loop above).
 * m2d_set_pl_graph: 1 (default) replays the batches from captured HIP
 * graphs, 0 enqueues the same kernels directly; the results are the same
 * bit for bit.  m2d_pl_loop_info: the host reads and batches of the last
 * m2d_intflow_pl_loop call on this handle. */
int m2d_set_pl_graph(m2d_system *s5, int use_graph);
int m2d_pl_loop_info(const m2d_system *s5, int *host_reads, int *batches);
/* No reference counterpart.  MSI iterations enqueued per p' solve of the
 * loop above (the solve's own cap stays nitm_p): 0 (default) the previous
 * call's longest solve plus 2, rounded up to a multiple of 4 (all nitm_p on
 * a handle's first call); cap > 0 that many; cap < 0 all nitm_p.  A solve
 * still iterating after its enqueued iterations ends the loop early, and
 * the call runs again from the updated fields as they were at its start
 * with all nitm_p enqueued: the results are the same bit for bit, the
 * iterations past a solve's exit no longer cost launches.
 * m2d_pl_loop_cap_info: the last call's enqueued iterations per solve and
 * whether it ran again (0 or 1). */
int m2d_set_pl_solve_cap(m2d_system *s5, int cap);
int m2d_pl_loop_cap_info(const m2d_system *s5, int *cap, int *reruns);

/* Device arrays of one internal-flow time step (src/mach2d_main.f90:94-
 * 404): N-vectors in the reference layout unless stated. */
typedef struct m2d_time_step_fields {
    /* grid and metrics */
    const double *x, *xp, *xe, *ye, *xk, *yk, *xke, *yke, *xen, *yen, *rp, *re, *rn, *Jp, *Je,
        *Jn, *alphae, *betae, *betan, *gamman;
    /* properties: laminar viscosity at faces and nodes (zeros for Euler);
     * for the T equation (modvis = 1) cp, ke, kn and the boundary-face
     * flags fbe, fbn (may be NULL for Euler) */
    const double *vle, *vln, *vlp, *cp, *ke, *kn, *fbe, *fbn;
    /* state, updated in place */
    double *u, *v, *ue, *ve, *un, *vn, *Uce, *Vcn, *p, *T, *ro, *roe, *ron, *pl;
    /* previous time level (written by the step) */
    double *ua, *va, *uea, *vea, *una, *vna, *pa, *Ta, *roa;
    /* linear systems (au, av, at: 9 N; ap: 5 N, its corner rows keep their
     * initial values as in the reference) and work arrays */
    double *au, *av, *at, *ap, *bu, *bv, *bt, *bp, *cup, *sup, *cvp, *svp, *due, *dve, *dun,
        *dvn, *de, *dn, *g;
    /* iflow: inflow vectors of ny values; boundary temperatures Tbe, Tbw (ny
     * values), Tbn, Tbs (nx values) */
    double *uin, *vin, *pin, *Tin, *Mw, *plin, *pina, *Tbe, *Tbw, *Tbn, *Tbs;
} m2d_time_step_fields;

typedef struct m2d_time_step_params {
    int modvis;      /* 0 Euler (T from total enthalpy), 1 Navier-Stokes (T equation) */
    int coord;       /* 1 axisymmetric, 0 planar */
    int solver9d;    /* u, v, T solver: 0 MSI (lu2d9 + fb2d9), 1 TDMA (tdma2d9) */
    int itmmax, itpmax, nitm_u, nitm_p;
    double tol_u, tol_p, tolm, beta;
    double Rg, gamma, P0, T0, Cpref, Href, Twall;  /* gamma = thermomodel%gm(T0) */
    double h0, mincc, maxcc;                        /* Scarborough time stepper */
} m2d_time_step_params;

typedef struct m2d_time_step_result {
    double norm;     /* sum of the relative L1 residual norms of the last mass cycle */
    double rmass;    /* check_mass_conservation */
    double normpl;   /* max|pl| / p_avg of the last pressure-correction iteration */
    int itp;         /* pressure-correction iterations of the last mass cycle */
    int stage;       /* 0, or the part of the step that failed: 1 time levels,
                        2 inflow update, 3 p average and g, 4 u, v assembly,
                        5 u solve, 6 v solve, 7 face velocities, 8 p' loop,
                        9 node velocities, 10 mass check, 11 temperature,
                        12 densities, 13 next dt */
} m2d_time_step_result;

/* One time step of the internal-flow case with constant thermophysical
 * properties, src/mach2d_main.f90:94-404: time levels,
 * intflow_time_cycle_update, get_average(p), g; itmmax mass cycles (u, v
 * coefficients, sources and fictitious rows, SIMPLEC coefficients, the u
 * and v solves, face velocities, the pressure-correction loop
 * (m2d_intflow_pl_loop), norm_l1_5d, the node velocity correction and the
 * extrapolation to fictitious volumes); check_mass_conservation; T from
 * total enthalpy (Euler) or the T equation (NS); the state equation and
 * face densities; the Scarborough time step (*dt in: this step's, out: the
 * next).  All on the handles' stream (s5) with device arrays; the host reads
 * a few scalars (the exit tests, p_avg, the norms, rmass, dt).  s9: a
 * 9-diagonal FP64 handle (solver9d = 0; may be NULL for TDMA); s5: a
 * 5-diagonal FP64 handle of the same grid. */
int m2d_intflow_time_step(m2d_system *s9, m2d_system *s5, const m2d_time_step_fields *f,
                          const m2d_time_step_params *p, double *dt,
                          m2d_time_step_result *res);

/* ---- Time-step statistics (src/tstepper/mod_tstepper_scarborough.f90) ---
 * Host or device arrays (all of one kind; device arrays on the assembly
 * stream); each call returns its value and synchronises.  Status of the last
 * call: m2d_last_status. */

/* get_mean_cc_9d / get_mean_cc_5d, :260-320: the mean over real volumes of
 * (sum |a_nb|) / |a_P| (9 slots, or 5 slots S, W, P, E, N), bit for bit. */
double m2d_get_mean_cc_9d(const int *nx, const int *ny, const double *a);
double m2d_get_mean_cc_5d(const int *nx, const int *ny, const double *a);

/* tstepper_dt0_scarborough, :40-76 */
double m2d_tstepper_dt0_scarborough(const int *nx, const int *ny, const double *Jp,
                                    const double *u, const double *v);

/* tstepper_dt_scarborough, :80-123; h0, mincc, maxcc are the values
 * tstepper_scarborough_init (:28-36) reads from the parameter file. */
double m2d_tstepper_dt_scarborough(const int *nx, const int *ny, const double *dt,
                                   const double *au, const double *av, const double *at,
                                   const double *ap, const double *h0, const double *mincc,
                                   const double *maxcc);

/* ---- Grid centroids, metrics and radii (src/grid/mod_grid.f90) ---------
 * The steps of grid_create after the corner coordinates (:174-187).  Every
 * point of every output is written, bit for bit the reference's values.
 * Host or device arrays (all of one kind): device arrays are processed on
 * the assembly stream (enqueue only), host arrays are copied in and out.
 * Status: m2d_last_status. */

/* get_real_centroids_xy, :237-300 (opt 1 simple mean, else weighted mean) */
void m2d_get_real_centroids_xy(const int *opt, const int *nx, const int *ny, const double *x,
                               const double *y, double *xp, double *yp);

/* get_metrics, :304-558 */
void m2d_get_metrics(const int *nx, const int *ny, const double *x, const double *y,
                     const double *xp, const double *yp, double *xe, double *ye, double *xen,
                     double *yen, double *xk, double *yk, double *xke, double *yke, double *Jp,
                     double *Je, double *Jn, double *alphae, double *gamman, double *betae,
                     double *betan);

/* get_radius, :562-723 (coord 1 axisymmetric, else every radius 1) */
void m2d_get_radius(const int *coord, const int *nx, const int *ny, const double *y,
                    const double *yp, double *radius, double *re, double *rn, double *rp);

/* ---- 5. Learned initial guess: U-Net kernels (device pointers) ---------- */
/* No reference counterpart (the reference starts MSI from the caller's var).
 * BF16 NHWC tensors with channel counts that are multiples of 8 and 16-byte
 * aligned; FP32 accumulation on the matrix cores.  Weights are bf16
 * [cout rounded up to 16][kpad], k = tap*cin + ci (tap = 3*ky + kx), kpad =
 * 9*cin rounded up to 32, zero beyond.  mode1: 0 direct, 1 the 2x2 average
 * of a 2h x 2w tensor, 2 the nearest 2x upsampling of an h/2 x w/2 tensor;
 * in2 (c2 channels, may be 0) is concatenated after in1.  stream: a
 * hipStream_t or NULL.  All return M2D_OK or an error code. */
int m2d_unet_conv3x3(const void *in1, int c1, int mode1, const void *in2, int c2,
                     const void *w, const float *bias, int cout, int cout_stride, int kpad,
                     int n, int h, int w_, int relu, void *out, float *out32, int accumulate,
                     void *stream);
/* dw[co][k] += sum over pixels of g[p][co] * im2col(p, k); db[co] += sum g. */
int m2d_unet_conv3x3_wgrad(const void *in1, int c1, int mode1, const void *in2, int c2,
                           const void *g, int cout, int kpad, int n, int h, int w_, float *dw,
                           float *db, void *stream);
/* g *= (y > 0) */
int m2d_unet_relu_bwd(void *g, const void *y, long count, void *stream);
/* dst32 += gradient of channels c0..c0+c-1 of gi (cs channels, h x w) carried
 * back through input mode `mode` to the input tensor's own resolution. */
int m2d_unet_mode_bwd(const float *gi, int cs, int c0, int c, int mode, int n, int h, int w,
                      float *dst32, void *stream);
int m2d_unet_to_bf16(const float *src, void *dst, long count, void *stream);
/* Adam (bias-corrected, step >= 1) on FP32 parameters. */
int m2d_unet_adam(float *w, float *m, float *v, const float *g, long count, float lr, float b1,
                  float b2, float eps, int step, void *stream);
/* bf16 forward weights wb and input-gradient weights wr [cin_pad][kpad_r]
 * (taps rotated by 180 degrees, channels transposed) from FP32 w. */
int m2d_unet_pack_weights(const float *w, int cout, int cin, int kpad, void *wb, void *wr,
                          int kpad_r, int cin_pad, void *stream);
/* 0.5*sum(mask*(pred - target)^2)*inv_count: per-256-pixel partial sums and
 * the bf16 gradient (gs channels, value in channel 0). */
int m2d_unet_mse(const float *pred, int ps, const float *target, const float *mask, long npix,
                 float inv_count, void *grad, int gs, float *partial, void *stream);
/* Network input of nb systems (a: 9-diagonal FP64 per system, b: rhs) on an
 * h x w >= ny x nx tensor: b/(a_P s), a_W, a_E, a_S, a_N and the corner sum
 * over a_P; inv_s = 1/s. */
int m2d_unet_pack_input(const double *a, const double *b, int nx, int ny, int h, int w,
                        double inv_s, void *x, int nb, void *stream);
/* var = s * channel 0 of the FP32 network output (os channels, width w). */
int m2d_unet_unpack_output(const float *out, int os, int nx, int ny, int w, double s,
                           double *var, void *stream);

/* tdma2d9(nx,ny,nitm,a,b,x), src/solvers/tdma2d9.f90:38-322, on device
 * pointers (same results bit for bit; 3 <= nx, ny <= 2048, else
 * M2D_ERR_ARG).  Runs on `stream`
 * and returns when x is final. */
int m2d_tdma2d9_device(int nx, int ny, int nitm, const double *d_a,
                       const double *d_b, double *d_x, void *stream);
/* Line solver of the device TDMA (this host thread's later calls of
 * m2d_tdma2d9, m2d_tdma2d9_device, m2d_tdma2d9_device_many and the time
 * step):
 * 0 (default) the reference's Thomas recurrence, bit for bit; 1 parallel
 * cyclic reduction of each line (grids up to 400 x 400, larger ones keep the
 * recurrence), the same line-Jacobi iteration up to rounding (not the
 * reference's rounding).  No reference counterpart (an alternative solver
 * of the line systems of tdma2d9.f90:79-316). */
int m2d_set_tdma_mode(int mode);
/* tdma2d9 on `count` independent systems of one grid size (d_a[q], d_b[q],
 * d_x[q] device pointers) in the same launches: the u and v systems of a
 * SIMPLEC iteration, which the reference solves one after the other with
 * the same solver9d (src/mach2d_main.f90:226,233).  Each x[q] is what
 * m2d_tdma2d9_device gives for that system alone. */
int m2d_tdma2d9_device_many(int count, int nx, int ny, int nitm, const double *const *d_a,
                            const double *const *d_b, double *const *d_x, void *stream);

/* Per-kernel device time of this handle's launches (HIP events on the
 * handle's stream around each launch; launches after the stopping rule
 * fired are not counted).  Kernel ids: 0 factor, 1 both sweeps of an
 * iteration (one launch; with staged sweeps the forward sweep alone),
 * 2 backward sweep (staged sweeps only), 3 update + residual and its norm;
 * +4 for the FP32 kernels of M2D_FP32_IR.  Enabling resets the totals.  No reference counterpart (the
 * reference times only the whole run with cpu_time). */
int m2d_enable_kernel_timing(m2d_system *s, int on);
int m2d_kernel_timing(m2d_system *s, int kernel, double *total_ms, long *launches);

/* Library version string. */
const char *m2d_version(void);

#ifdef __cplusplus
}
#endif
#endif
