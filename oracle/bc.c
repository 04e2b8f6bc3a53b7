/* bc.c -- ORACLE (test infrastructure only): C restatement of the
 * fictitious-row writers of the internal-flow case,
 * src/user/internal_flow/mod_intflow.f90 (intflow_set_bcu / bcv / bcpl /
 * bcT and get_u_v_north_slip), in the reference's operation order (left to
 * right, no contraction).  Never linked into the product.
 *
 * Layout as in coef.c: flat index np = nx*(j-1) + i, diagonal k (1-based) of
 * a(N, 9) / a(N, 5) at a + (k-1)*N.  Each writer overwrites whole rows of the
 * boundary it owns (all diagonals zeroed first) and leaves every other row
 * untouched, as the reference does. */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

#define P(i, j) ((long)nx * ((j) - 1) + (i) - 1)
#define A(k, p) a[(long)((k) - 1) * N + (p)]

static void clear_row(double *a, long N, int nd, long p)
{
    for (int k = 1; k <= nd; ++k) A(k, p) = 0.0;
}

/* get_u_v_north_slip: mod_intflow.f90:142-210.  ubn/vbn have nx entries. */
void orc_u_v_north_slip(int nx, int ny, const double *x, const double *xp, const double *xk,
                        const double *yk, const double *u, const double *v, double *ubn,
                        double *vbn)
{
    const int j = ny - 1;
    for (int i = 2; i <= nx - 1; ++i) {
        const long np = P(i, j);
        const double Up = sqrt(u[np] * u[np] + v[np] * v[np]);
        const double signal = copysign(1.0, xk[np] * u[np] + yk[np] * v[np]);
        const double t = sqrt(xk[np] * xk[np] + yk[np] * yk[np]);
        const double tx = xk[np] / t, ty = yk[np] / t;
        ubn[i - 1] = signal * Up * tx;
        vbn[i - 1] = signal * Up * ty;
    }
    {  /* west fictitious: :196-203 */
        const long np = P(1, j), npe = np + 1, npee = npe + 1;
        ubn[0] = ubn[1] + (ubn[1] - ubn[2]) * (xp[npe] - x[np]) / (xp[npee] - xp[npe]);
        vbn[0] = vbn[1] + (vbn[1] - vbn[2]) * (xp[npe] - x[np]) / (xp[npee] - xp[npe]);
    }
    {  /* east fictitious: :200-208 */
        const long np = P(nx, j), npw = np - 1, npww = npw - 1;
        ubn[nx - 1] = ubn[nx - 2] + (ubn[nx - 2] - ubn[nx - 3]) * (x[npw] - xp[npw]) /
                                        (xp[npw] - xp[npww]);
        vbn[nx - 1] = vbn[nx - 2] + (vbn[nx - 2] - vbn[nx - 3]) * (x[npw] - xp[npw]) /
                                        (xp[npw] - xp[npww]);
    }
}

/* intflow_set_bcu (:214-301) and intflow_set_bcv (:305-392): `comp` 0 = u, 1 = v */
static void set_bc_uv(int comp, int nx, int ny, int modvis, const double *x, const double *xp,
                      const double *xk, const double *yk, const double *u, const double *v,
                      double *a, double *b)
{
    const long N = (long)nx * ny;
    for (int j = 2; j <= ny - 1; ++j) {  /* west, corners excluded */
        const long np = P(1, j), npe = np + 1, npee = npe + 1;
        clear_row(a, N, 9, np);
        if (comp == 0) {
            A(6, np) = -1.0;
            A(5, np) = 1.0;
            b[np] = (2.0 * xp[npe]) / (xp[npee] - xp[npe]) * (u[npe] - u[npee]);
        } else {
            A(6, np) = 1.0;
            A(5, np) = 1.0;
            b[np] = 0.0;
        }
    }
    for (int j = 2; j <= ny - 1; ++j) {  /* east, corners excluded */
        const long np = P(nx, j), npw = np - 1, npww = npw - 1;
        clear_row(a, N, 9, np);
        A(4, np) = -1.0;
        A(5, np) = 1.0;
        if (comp == 0)
            b[np] = 2.0 * (x[npw] - xp[npw]) / (xp[npw] - xp[npww]) * (u[npw] - u[npww]);
        else
            b[np] = (v[npw] - v[npww]) * 2.0 * (x[npw] - xp[npw]) / (xp[npw] - xp[npww]);
    }
    for (int i = 1; i <= nx; ++i) {  /* south, corners included */
        const long np = P(i, 1);
        clear_row(a, N, 9, np);
        A(8, np) = comp == 0 ? -1.0 : 1.0;
        A(5, np) = 1.0;
        b[np] = 0.0;
    }
    double *ubn = NULL, *vbn = NULL;
    if (modvis != 1) {
        ubn = malloc(sizeof(double) * nx);
        vbn = malloc(sizeof(double) * nx);
        orc_u_v_north_slip(nx, ny, x, xp, xk, yk, u, v, ubn, vbn);
    }
    for (int i = 1; i <= nx; ++i) {  /* north, corners included */
        const long np = P(i, ny);
        clear_row(a, N, 9, np);
        A(2, np) = 1.0;
        A(5, np) = 1.0;
        b[np] = modvis == 1 ? 0.0 : 2.0 * (comp == 0 ? ubn : vbn)[i - 1];
    }
    free(ubn);
    free(vbn);
}

void orc_intflow_set_bcu(int nx, int ny, int modvis, const double *x, const double *xp,
                         const double *xk, const double *yk, const double *u, const double *v,
                         double *au, double *b)
{
    set_bc_uv(0, nx, ny, modvis, x, xp, xk, yk, u, v, au, b);
}

void orc_intflow_set_bcv(int nx, int ny, int modvis, const double *x, const double *xp,
                         const double *xk, const double *yk, const double *u, const double *v,
                         double *av, double *b)
{
    set_bc_uv(1, nx, ny, modvis, x, xp, xk, yk, u, v, av, b);
}

/* intflow_set_bcpl: mod_intflow.f90:396-459.  5-diagonal ap (S, W, P, E,
 * N); plin has ny entries (iflow%plin); corners are not written. */
void orc_intflow_set_bcpl(int nx, int ny, const double *x, const double *xp, const double *pl,
                          const double *plin, double *a, double *bp)
{
    const long N = (long)nx * ny;
    for (int j = 2; j <= ny - 1; ++j) {
        const long np = P(1, j);
        clear_row(a, N, 5, np);
        A(4, np) = 1.0;
        A(3, np) = 1.0;
        bp[np] = 2.0 * plin[j - 1];
    }
    for (int j = 2; j <= ny - 1; ++j) {
        const long np = P(nx, j), npw = np - 1, npww = npw - 1;
        clear_row(a, N, 5, np);
        A(2, np) = -1.0;
        A(3, np) = 1.0;
        bp[np] = 2.0 * (x[npw] - xp[npw]) / (xp[npw] - xp[npww]) * (pl[npw] - pl[npww]);
    }
    for (int i = 2; i <= nx - 1; ++i) {
        const long np = P(i, 1);
        clear_row(a, N, 5, np);
        A(5, np) = -1.0;
        A(3, np) = 1.0;
        bp[np] = 0.0;
    }
    for (int i = 2; i <= nx - 1; ++i) {
        const long np = P(i, ny);
        clear_row(a, N, 5, np);
        A(1, np) = -1.0;
        A(3, np) = 1.0;
        bp[np] = 0.0;
    }
}

/* intflow_set_bcT: mod_intflow.f90:464-550.  Tin has ny entries
 * (iflow%Tin); Twall > 0 prescribes the north wall temperature, otherwise
 * the north wall is adiabatic. */
void orc_intflow_set_bcT(int nx, int ny, double Twall, const double *x, const double *xp,
                         const double *T, const double *Tin, double *a, double *bt)
{
    const long N = (long)nx * ny;
    const int ccTw = Twall > 0.0;
    for (int j = 2; j <= ny - 1; ++j) {
        const long np = P(1, j);
        clear_row(a, N, 9, np);
        A(6, np) = 1.0;
        A(5, np) = 1.0;
        bt[np] = 2.0 * Tin[j - 1];
    }
    for (int j = 2; j <= ny - 1; ++j) {
        const long np = P(nx, j), npw = np - 1, npww = npw - 1;
        clear_row(a, N, 9, np);
        A(4, np) = -1.0;
        A(5, np) = 1.0;
        bt[np] = 2.0 * (x[npw] - xp[npw]) / (xp[npw] - xp[npww]) * (T[npw] - T[npww]);
    }
    for (int i = 1; i <= nx; ++i) {
        const long np = P(i, 1);
        clear_row(a, N, 9, np);
        A(8, np) = -1.0;
        A(5, np) = 1.0;
        bt[np] = 0.0;
    }
    for (int i = 1; i <= nx; ++i) {
        const long np = P(i, ny);
        clear_row(a, N, 9, np);
        A(2, np) = ccTw ? 1.0 : -1.0;
        A(5, np) = 1.0;
        bt[np] = ccTw ? 2.0 * Twall : 0.0;
    }
}

/* intflow_velocities_at_boundary_faces: mod_intflow.f90:692-787.  Velocities
 * on the boundary faces: west (i = 1) and east (i = nx-1) faces of rows
 * 2..ny-1, south (j = 1) and north (j = ny-1) faces of columns 2..nx-1; the
 * north face takes the Euler slip velocity (modvis = 0) or zero (NS). */
void orc_intflow_velocities_at_boundary_faces(int nx, int ny, int modvis, const double *x,
                                              const double *xp, const double *xe,
                                              const double *ye, const double *xk,
                                              const double *yk, const double *u, const double *v,
                                              double *ue, double *ve, double *un, double *vn,
                                              double *Uce, double *Vcn)
{
    for (int j = 2; j <= ny - 1; ++j) {  /* west */
        const long np = P(1, j), npe = np + 1;
        ue[np] = (u[np] + u[npe]) / 2.0;
        ve[np] = 0.0;
        Uce[np] = ue[np] * ye[np] - ve[np] * xe[np];
    }
    for (int j = 2; j <= ny - 1; ++j) {  /* east */
        const long np = P(nx - 1, j), npe = np + 1;
        ue[np] = (u[np] + u[npe]) / 2.0;
        ve[np] = (v[np] + v[npe]) / 2.0;
        Uce[np] = ue[np] * ye[np] - ve[np] * xe[np];
    }
    for (int i = 2; i <= nx - 1; ++i) {  /* south */
        const long np = P(i, 1), npn = np + nx;
        un[np] = (u[np] + u[npn]) / 2.0;
        vn[np] = 0.0;
        Vcn[np] = 0.0;
    }
    double *ubn = (double *)calloc((size_t)nx, sizeof(double));
    double *vbn = (double *)calloc((size_t)nx, sizeof(double));
    if (modvis == 0) orc_u_v_north_slip(nx, ny, x, xp, xk, yk, u, v, ubn, vbn);
    for (int i = 2; i <= nx - 1; ++i) {  /* north */
        const long np = P(i, ny - 1);
        un[np] = ubn[i - 1];
        vn[np] = vbn[i - 1];
        Vcn[np] = 0.0;
    }
    free(ubn);
    free(vbn);
}
