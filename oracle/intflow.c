/* intflow.c -- ORACLE (test infrastructure only): C restatement of the
 * internal-flow routines of one time step that are not coefficient writers,
 * loop by loop in the reference's order (no contraction):
 * src/user/internal_flow/mod_intflow.f90 intflow_boundary_T_from_H_conservation
 * (:557-688), intflow_boundary_simplec (:791-877),
 * intflow_extrapolate_u_v_to_fictitious (:881-964), get_uin_vin_pin_Tin_Mw
 * (:968-1006), get_plin_and_p_fictitious (:1010-1101) and
 * intflow_time_cycle_update (:1106-1134); norm_l1_5d
 * (src/solvers/mod_solvers.f90:91-125); get_average
 * (src/util/mod_field_operations.f90:115-142).  Never linked into the
 * product.  Flat index np = nx*(j-1) + i, 1-based; vectors of ny (nx)
 * values are indexed by j (i), 1-based. */
#include <math.h>

#include "oracle.h"

#define P(i, j) ((long)nx * ((j) - 1) + (i) - 1)

void orc_intflow_boundary_T_from_H_conservation(int nx, int ny, double Cpref, double Href,
                                                const double *Tin, const double *u,
                                                const double *ue, const double *un,
                                                const double *v, const double *ve,
                                                const double *vn, double *T, double *Tbe,
                                                double *Tbw, double *Tbn, double *Tbs)
{
    (void)u;
    (void)v;
    for (int j = 2; j <= ny - 1; ++j) {  /* east */
        const long np = P(nx - 1, j);
        Tbe[j - 1] = (Href - (ue[np] * ue[np] + ve[np] * ve[np]) / 2.0) / Cpref;
        T[np + 1] = 2.0 * Tbe[j - 1] - T[np];
    }
    for (int j = 2; j <= ny - 1; ++j) {  /* west */
        const long np = P(1, j);
        Tbw[j - 1] = Tin[j - 1];
        T[np] = 2.0 * Tbw[j - 1] - T[np + 1];
    }
    for (int i = 2; i <= nx - 1; ++i) {  /* north */
        const long np = P(i, ny - 1);
        Tbn[i - 1] = (Href - (un[np] * un[np] + vn[np] * vn[np]) / 2.0) / Cpref;
        T[np + nx] = 2.0 * Tbn[i - 1] - T[np];
    }
    for (int i = 2; i <= nx - 1; ++i) {  /* south */
        const long np = P(i, 1);
        Tbs[i - 1] = (Href - (un[np] * un[np] + vn[np] * vn[np]) / 2.0) / Cpref;
        T[np] = 2.0 * Tbs[i - 1] - T[np + nx];
    }
    T[P(1, 1)] = T[P(1, 2)];
    T[P(nx, 1)] = T[P(nx, 2)];
    T[P(1, ny)] = T[P(1, ny - 1)];
    T[P(nx, ny)] = T[P(nx, ny - 1)];
}

void orc_intflow_boundary_simplec(int nx, int ny, int modvis, double *due, double *dve,
                                  double *dun, double *dvn, double *de, double *dn)
{
    (void)modvis;
    for (int j = 2; j <= ny - 1; ++j) {  /* west */
        const long np = P(1, j), npe = np + 1;
        due[np] = due[npe];
        dve[np] = 0.0;
        de[np] = de[npe];
    }
    for (int j = 2; j <= ny - 1; ++j) {  /* east face of the last real volume */
        const long np = P(nx - 1, j), npw = np - 1;
        due[np] = due[npw];
        dve[np] = dve[npw];
        de[np] = de[npw];
    }
    for (int i = 2; i <= nx - 1; ++i) {  /* south */
        const long np = P(i, 1), npn = np + nx;
        dun[np] = dun[npn];
        dvn[np] = 0.0;
        dn[np] = 0.0;
    }
    for (int i = 2; i <= nx - 1; ++i) {  /* north face of the last real volume */
        const long np = P(i, ny - 1);
        dun[np] = 0.0;
        dvn[np] = 0.0;
        dn[np] = 0.0;
    }
}

/* get_u_v_north_slip, mod_intflow.f90:142-210 */
static void north_slip(int nx, int ny, const double *x, const double *xp, const double *xk,
                       const double *yk, const double *u, const double *v, double *ubn,
                       double *vbn)
{
    for (int i = 2; i <= nx - 1; ++i) {
        const long np = P(i, ny - 1);
        const double Up = sqrt(u[np] * u[np] + v[np] * v[np]);
        const double signal = copysign(1.0, xk[np] * u[np] + yk[np] * v[np]);
        const double t = sqrt(xk[np] * xk[np] + yk[np] * yk[np]);
        const double tx = xk[np] / t, ty = yk[np] / t;
        ubn[i - 1] = signal * Up * tx;
        vbn[i - 1] = signal * Up * ty;
    }
    long np = P(1, ny - 1), npe = np + 1, npee = npe + 1;
    ubn[0] = ubn[1] + (ubn[1] - ubn[2]) * (xp[npe] - x[np]) / (xp[npee] - xp[npe]);
    vbn[0] = vbn[1] + (vbn[1] - vbn[2]) * (xp[npe] - x[np]) / (xp[npee] - xp[npe]);
    np = P(nx, ny - 1);
    const long npw = np - 1, npww = npw - 1;
    ubn[nx - 1] = ubn[nx - 2] + (ubn[nx - 2] - ubn[nx - 3]) * (x[npw] - xp[npw]) /
                                    (xp[npw] - xp[npww]);
    vbn[nx - 1] = vbn[nx - 2] + (vbn[nx - 2] - vbn[nx - 3]) * (x[npw] - xp[npw]) /
                                    (xp[npw] - xp[npww]);
}

void orc_intflow_extrapolate_u_v_to_fictitious(int nx, int ny, int modvis, const double *x,
                                               const double *xp, const double *xk,
                                               const double *yk, double *u, double *v)
{
    for (int j = 2; j <= ny - 1; ++j) {  /* west */
        const long np = P(1, j), npe = np + 1, npee = np + 2;
        u[np] = u[npe] + 2.0 * xp[npe] / (xp[npee] - xp[npe]) * (u[npe] - u[npee]);
        v[np] = -v[npe];
    }
    for (int j = 2; j <= ny - 1; ++j) {  /* east */
        const long np = P(nx, j), npw = np - 1, npww = np - 2;
        u[np] = u[npw] + 2.0 * (x[npw] - xp[npw]) / (xp[npw] - xp[npww]) * (u[npw] - u[npww]);
        v[np] = v[npw] + 2.0 * (x[npw] - xp[npw]) / (xp[npw] - xp[npww]) * (v[npw] - v[npww]);
    }
    for (int i = 1; i <= nx; ++i) {  /* south */
        const long np = P(i, 1), npn = np + nx;
        u[np] = u[npn];
        v[np] = -v[npn];
    }
    if (modvis == 1) {
        for (int i = 1; i <= nx; ++i) {
            const long np = P(i, ny), nps = np - nx;
            u[np] = -u[nps];
            v[np] = -v[nps];
        }
    } else {
        double ubn[nx], vbn[nx];
        north_slip(nx, ny, x, xp, xk, yk, u, v, ubn, vbn);
        for (int i = 1; i <= nx; ++i) {
            const long np = P(i, ny), nps = np - nx;
            u[np] = -u[nps] + 2.0 * ubn[i - 1];
            v[np] = -v[nps] + 2.0 * vbn[i - 1];
        }
    }
}

void orc_get_uin_vin_pin_Tin_Mw(int nx, int ny, double gamma, double Rg, double po, double T0,
                                const double *u, double *uin, double *vin, double *pin,
                                double *Tin, double *Mw)
{
    for (int j = 2; j <= ny - 1; ++j) {
        const long np = P(1, j), npe = np + 1;
        const int k = j - 1;
        uin[k] = (u[np] + u[npe]) / 2.0;
        vin[k] = 0.0;
        Tin[k] = T0 - (gamma - 1.0) / (2.0 * gamma * Rg) * (uin[k] * uin[k] + vin[k] * vin[k]);
        Mw[k] = sqrt((uin[k] * uin[k] + vin[k] * vin[k]) / (gamma * Rg * Tin[k]));
        pin[k] = po * pow(1.0 + 0.5 * (gamma - 1.0) * (Mw[k] * Mw[k]), -gamma / (gamma - 1.0));
    }
}

void orc_get_plin_and_p_fictitious(int nx, int ny, const double *x, const double *xp,
                                   const double *pina, const double *pin, double *plin,
                                   double *p)
{
    for (int j = 2; j <= ny - 1; ++j) {  /* west */
        const long np = P(1, j), npe = np + 1;
        p[np] = 2.0 * pin[j - 1] - p[npe];
        plin[j - 1] = pin[j - 1] - pina[j - 1];
    }
    for (int j = 2; j <= ny - 1; ++j) {  /* east */
        const long np = P(nx, j), npw = np - 1, npww = npw - 1;
        p[np] = p[npw] + 2.0 * (x[npw] - xp[npw]) / (xp[npw] - xp[npww]) * (p[npw] - p[npww]);
    }
    for (int i = 2; i <= nx - 1; ++i) p[P(i, 1)] = p[P(i, 2)];
    for (int i = 2; i <= nx - 1; ++i) p[P(i, ny)] = p[P(i, ny - 1)];
    p[P(1, 1)] = p[P(1, 2)];
    p[P(nx, 1)] = p[P(nx, 2)];
    p[P(1, ny)] = p[P(1, ny - 1)];
    p[P(nx, ny)] = p[P(nx, ny - 1)];
}

void orc_intflow_time_cycle_update(int nx, int ny, double gamma, double Rg, double P0,
                                   double T0, const double *x, const double *xp,
                                   const double *u, double *uin, double *vin, double *pin,
                                   double *Tin, double *Mw, double *plin, double *pina,
                                   double *p)
{
    for (int k = 0; k < ny; ++k) pina[k] = pin[k];
    orc_get_uin_vin_pin_Tin_Mw(nx, ny, gamma, Rg, P0, T0, u, uin, vin, pin, Tin, Mw);
    orc_get_plin_and_p_fictitious(nx, ny, x, xp, pina, pin, plin, p);
}

void orc_norm_l1_5d(int nx, int ny, const double *var, const double *b, const double *a,
                    double *norm)
{
    const long N = (long)nx * ny;
    double s = *norm;
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j);
            s = s + fabs(+a[np] * var[np - nx] + a[N + np] * var[np - 1] + a[2 * N + np] * var[np] +
                         a[3 * N + np] * var[np + 1] + a[4 * N + np] * var[np + nx] - b[np]);
        }
    *norm = s;
}

double orc_get_average(int nx, int ny, const double *F)
{
    double s = 0.0;
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) s = s + F[P(i, j)];
    return s / ((double)(nx - 2) * (double)(ny - 2));
}
