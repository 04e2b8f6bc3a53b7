/* grid.c -- ORACLE (test infrastructure only): C restatement of the grid
 * centroids, metrics and radii of src/grid/mod_grid.f90
 * (get_real_centroids_xy :237-300, get_metrics :304-558, get_radius
 * :562-723), loop by loop in the reference's order (no contraction).  Never
 * linked into the product.  Flat index np = nx*(j-1) + i, 1-based (i, j);
 * x, y: coordinates of the north-east corner of each volume.
 *
 * Parity unpinned against the compiled reference: mod_grid.f90 uses
 * mod_extflow, whose geometry module needs the absent depp_interface, so
 * the module does not build here.  The restatement is pinned instead by
 * the mathematics (tests/test_oracle_golden.py): closed forms on uniform
 * and stretched Cartesian grids (xe = dy, xke = dx, Je = 1/(dx dy),
 * beta = 0), exactness of the one-sided boundary differences on affine
 * grids, rotation invariance of alpha, gamma, beta and J, the polygon area
 * of each volume for 1/Jp, and the centroid of the two triangles against
 * the shoelace centroid of the quadrilateral. */
#include <string.h>
#include <math.h>

#include "oracle.h"

#define P(i, j) ((long)nx * ((j) - 1) + (i) - 1)

void orc_get_real_centroids_xy(int opt, int nx, int ny, const double *x, const double *y,
                               double *xp, double *yp)
{
    const long N = (long)nx * ny;
    memset(xp, 0, sizeof(double) * N);
    memset(yp, 0, sizeof(double) * N);
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j), nps = np - nx, npw = np - 1, npsw = nps - 1;
            if (opt == 1) {
                xp[np] = (x[np] + x[npw] + x[npsw] + x[nps]) / 4.0;
                yp[np] = (y[np] + y[npw] + y[npsw] + y[nps]) / 4.0;
            } else {
                const double A1 = 0.5 * (x[npsw] * (y[nps] - y[np]) + x[nps] * (y[np] - y[npsw]) +
                                         x[np] * (y[npsw] - y[nps]));
                const double A2 = 0.5 * (x[npsw] * (y[np] - y[npw]) + x[np] * (y[npw] - y[npsw]) +
                                         x[npw] * (y[npsw] - y[np]));
                xp[np] = ((x[npsw] + x[nps] + x[np]) * A1 + (x[npsw] + x[np] + x[npw]) * A2) /
                         (3.0 * (A1 + A2));
                yp[np] = ((y[npsw] + y[nps] + y[np]) * A1 + (y[npsw] + y[np] + y[npw]) * A2) /
                         (3.0 * (A1 + A2));
            }
        }
}

void orc_get_metrics(int nx, int ny, const double *x, const double *y, const double *xp,
                     const double *yp, double *xe, double *ye, double *xen, double *yen,
                     double *xk, double *yk, double *xke, double *yke, double *Jp, double *Je,
                     double *Jn, double *alphae, double *gamman, double *betae, double *betan)
{
    const long N = (long)nx * ny;
    double *outs[15] = {xe, ye, xen, yen, xk, yk, xke, yke, Jp, Je, Jn, alphae, gamman, betae,
                        betan};
    for (int k = 0; k < 15; ++k) memset(outs[k], 0, sizeof(double) * N);
    for (int j = 2; j <= ny - 1; ++j)  /* eta-derivatives at east faces */
        for (int i = 1; i <= nx - 1; ++i) {
            const long np = P(i, j), nps = np - nx;
            xe[np] = x[np] - x[nps];
            ye[np] = y[np] - y[nps];
            alphae[np] = xe[np] * xe[np] + ye[np] * ye[np];
        }
    for (int j = 1; j <= ny - 1; ++j)  /* csi-derivatives at north faces */
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j), npw = np - 1;
            xk[np] = x[np] - x[npw];
            yk[np] = y[np] - y[npw];
            gamman[np] = xk[np] * xk[np] + yk[np] * yk[np];
        }
    for (int j = 2; j <= ny - 1; ++j)  /* csi-derivatives at inner east faces */
        for (int i = 2; i <= nx - 2; ++i) {
            const long np = P(i, j), npe = np + 1;
            xke[np] = xp[npe] - xp[np];
            yke[np] = yp[npe] - yp[np];
        }
    for (int j = 2; j <= ny - 1; ++j) {  /* west boundary face */
        const long np = P(2, j), nps = np - nx, npw = np - 1, npsw = nps - 1;
        double fw = (x[npw] + x[npsw]) / 2.0, fe = (x[np] + x[nps]) / 2.0;
        xke[npw] = -3.0 * fw + 4.0 * xp[np] - fe;
        fw = (y[npw] + y[npsw]) / 2.0;
        fe = (y[np] + y[nps]) / 2.0;
        yke[npw] = -3.0 * fw + 4.0 * yp[np] - fe;
    }
    for (int j = 2; j <= ny - 1; ++j) {  /* east boundary face */
        const long np = P(nx - 1, j), nps = np - nx, npw = np - 1, npsw = nps - 1;
        double fw = (x[npw] + x[npsw]) / 2.0, fe = (x[np] + x[nps]) / 2.0;
        xke[np] = fw - 4.0 * xp[np] + 3.0 * fe;
        fw = (y[npw] + y[npsw]) / 2.0;
        fe = (y[np] + y[nps]) / 2.0;
        yke[np] = fw - 4.0 * yp[np] + 3.0 * fe;
    }
    for (int j = 2; j <= ny - 2; ++j)  /* eta-derivatives at inner north faces */
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j), npn = np + nx;
            xen[np] = xp[npn] - xp[np];
            yen[np] = yp[npn] - yp[np];
        }
    for (int i = 2; i <= nx - 1; ++i) {  /* south boundary face */
        const long np = P(i, 2), nps = np - nx, npw = np - 1, npsw = nps - 1;
        double fs = (x[nps] + x[npsw]) / 2.0, fn = (x[np] + x[npw]) / 2.0;
        xen[nps] = -3.0 * fs + 4.0 * xp[np] - fn;
        fs = (y[nps] + y[npsw]) / 2.0;
        fn = (y[np] + y[npw]) / 2.0;
        yen[nps] = -3.0 * fs + 4.0 * yp[np] - fn;
    }
    for (int i = 2; i <= nx - 1; ++i) {  /* north boundary face */
        const long np = P(i, ny - 1), nps = np - nx, npw = np - 1, npsw = nps - 1;
        double fs = (x[nps] + x[npsw]) / 2.0, fn = (x[np] + x[npw]) / 2.0;
        xen[np] = fs - 4.0 * xp[np] + 3.0 * fn;
        fs = (y[nps] + y[npsw]) / 2.0;
        fn = (y[np] + y[npw]) / 2.0;
        yen[np] = fs - 4.0 * yp[np] + 3.0 * fn;
    }
    for (int j = 2; j <= ny - 1; ++j)  /* beta, J at east faces */
        for (int i = 1; i <= nx - 1; ++i) {
            const long np = P(i, j);
            betae[np] = xke[np] * xe[np] + yke[np] * ye[np];
            Je[np] = 1.0 / (xke[np] * ye[np] - xe[np] * yke[np]);
        }
    for (int j = 1; j <= ny - 1; ++j)  /* beta, J at north faces */
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j);
            betan[np] = xk[np] * xen[np] + yk[np] * yen[np];
            Jn[np] = 1.0 / (xk[np] * yen[np] - xen[np] * yk[np]);
        }
    for (int j = 2; j <= ny - 1; ++j)  /* J at volume centres */
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j), nps = np - nx, npw = np - 1, npsw = nps - 1;
            double fw = (x[npw] + x[npsw]) / 2.0, fe = (x[np] + x[nps]) / 2.0;
            double fs = (x[nps] + x[npsw]) / 2.0, fn = (x[np] + x[npw]) / 2.0;
            const double xkp = fe - fw, xep = fn - fs;
            fw = (y[npw] + y[npsw]) / 2.0;
            fe = (y[np] + y[nps]) / 2.0;
            fs = (y[nps] + y[npsw]) / 2.0;
            fn = (y[np] + y[npw]) / 2.0;
            const double ykp = fe - fw, yep = fn - fs;
            Jp[np] = 1.0 / (xkp * yep - xep * ykp);
        }
    for (long q = 0; q < N; ++q) {
        Je[q] = fabs(Je[q]);
        Jn[q] = fabs(Jn[q]);
        Jp[q] = fabs(Jp[q]);
    }
}

void orc_get_radius(int coord, int nx, int ny, const double *y, const double *yp,
                    double *radius, double *re, double *rn, double *rp)
{
    const long N = (long)nx * ny;
    for (long q = 0; q < N; ++q) radius[q] = re[q] = rn[q] = rp[q] = 1.0;
    if (coord != 1) return;
    for (long q = 0; q < N; ++q) radius[q] = y[q];
    for (int i = 1; i <= nx - 1; ++i)
        for (int j = 2; j <= ny - 1; ++j) {
            const long np = P(i, j), nps = np - nx;
            re[np] = (y[np] + y[nps]) / 2.0;
        }
    for (int i = 2; i <= nx - 1; ++i)
        for (int j = 1; j <= ny - 1; ++j) {
            const long np = P(i, j), npw = np - 1;
            rn[np] = (y[np] + y[npw]) / 2.0;
        }
    for (long q = 0; q < N; ++q) rp[q] = yp[q];
    for (int i = 2; i <= nx - 1; ++i) {  /* south */
        const long np = P(i, 1), npn = np + nx;
        rp[np] = rn[np] - rp[npn] + rn[np];
    }
    for (int i = 2; i <= nx - 1; ++i) {  /* north */
        const long np = P(i, ny), nps = np - nx;
        rp[np] = rn[nps] - rp[nps] + rn[nps];
    }
    for (int j = 2; j <= ny - 1; ++j) {  /* west */
        const long np = P(1, j), npe = np + 1;
        rp[np] = re[np] - rp[npe] + re[np];
    }
    for (int j = 2; j <= ny - 1; ++j) {  /* east */
        const long np = P(nx, j), npw = np - 1;
        rp[np] = re[npw] - rp[npw] + re[npw];
    }
    long np = P(1, 1);
    rp[np] = 4.0 * y[np] - rp[np + 1] - rp[np + nx] - rp[np + nx + 1];
    np = P(nx, 1);
    rp[np] = 4.0 * y[np - 1] - rp[np - 1] - rp[np + nx] - rp[np + nx - 1];
    np = P(1, ny);
    rp[np] = 4.0 * y[np - nx] - rp[np + 1] - rp[np - nx] - rp[np - nx + 1];
    np = P(nx, ny);
    rp[np] = 4.0 * y[np - nx - 1] - rp[np - 1] - rp[np - nx] - rp[np - nx - 1];
}
