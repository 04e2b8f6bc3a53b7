/* tstep.c -- ORACLE (test infrastructure only): C restatement of the
 * Scarborough time-step statistics of src/tstepper/mod_tstepper_scarborough.f90
 * (tstepper_dt0_scarborough :40-76, tstepper_dt_scarborough :80-123,
 * get_mean_cc_9d :260-288, get_mean_cc_5d :292-320), in the reference's
 * loop and operation order (i outer, j inner; no contraction).  Never linked
 * into the product.  Arrays as in coef.c: a(N, k) diagonal-major. */
#include <math.h>

#include "oracle.h"

#define P(i, j) ((long)nx * ((j) - 1) + (i) - 1)

double orc_get_mean_cc_9d(int nx, int ny, const double *a)
{
    const long N = (long)nx * ny;
    double m = 0.0;
    for (int i = 2; i <= nx - 1; ++i)
        for (int j = 2; j <= ny - 1; ++j) {
            const long np = P(i, j);
            const double s1 = fabs(a[np]) + fabs(a[N + np]) + fabs(a[2 * N + np]) +
                              fabs(a[3 * N + np]);
            const double s2 = fabs(a[5 * N + np]) + fabs(a[6 * N + np]) + fabs(a[7 * N + np]) +
                              fabs(a[8 * N + np]);
            m = (s1 + s2) / fabs(a[4 * N + np]) + m;
        }
    return m / (double)((nx - 2) * (ny - 2));
}

double orc_get_mean_cc_5d(int nx, int ny, const double *a)
{
    const long N = (long)nx * ny;
    double m = 0.0;
    for (int i = 2; i <= nx - 1; ++i)
        for (int j = 2; j <= ny - 1; ++j) {
            const long np = P(i, j);
            const double s1 = fabs(a[np]) + fabs(a[N + np]);
            const double s2 = fabs(a[3 * N + np]) + fabs(a[4 * N + np]);
            m = (s1 + s2) / fabs(a[2 * N + np]) + m;
        }
    return m / (double)((nx - 2) * (ny - 2));
}

double orc_tstepper_dt0_scarborough(int nx, int ny, const double *Jp, const double *u,
                                    const double *v)
{
    double dt = 1.0e10;
    for (int i = 2; i <= nx - 1; ++i)
        for (int j = 2; j <= ny - 1; ++j) {
            const long np = P(i, j);
            const double aux = 1.0 / sqrt(Jp[np] * (u[np] * u[np] + v[np] * v[np]));
            if (dt > aux) dt = aux;
        }
    return dt * 0.25;
}

double orc_tstepper_dt_scarborough(int nx, int ny, double dt, const double *au,
                                   const double *av, const double *at, const double *ap,
                                   double h0, double mincc, double maxcc)
{
    const double cu = orc_get_mean_cc_9d(nx, ny, au), cv = orc_get_mean_cc_9d(nx, ny, av);
    const double ct = orc_get_mean_cc_9d(nx, ny, at), cp = orc_get_mean_cc_5d(nx, ny, ap);
    double c = cu;
    if (cv > c) c = cv;
    if (ct > c) c = ct;
    if (cp > c) c = cp;
    double h;
    if (c < mincc)
        h = h0 * sqrt(1.0 - c / mincc);
    else if (maxcc < c)
        h = -h0 * sqrt(c / maxcc - 1.0);
    else
        h = 0.0;
    return dt * (1.0 + h);
}
