/* coef.c -- ORACLE (test infrastructure only): C restatement of the
 * coefficient assembly of src/equations/mach2d_coef.f90 (module
 * `coefficients`), in the reference's operation order (left to right, no
 * contraction), for the parity tests of the HIP assembly kernels.  Never
 * linked into the product.
 *
 * Arrays are the reference's: N = nx*ny values per field, flat index
 * np = nx*(j-1) + i (1-based i, j; here P(i, j) gives the 0-based offset);
 * coefficient arrays a(N, 9) / a(N, 5) are column-major, i.e. diagonal k
 * (1-based) starts at a + (k-1)*N.  Only interior points (2..nx-1,
 * 2..ny-1) are written, as in the reference. */
#include <math.h>

#include "oracle.h"

#define P(i, j) ((long)nx * ((j) - 1) + (i) - 1)
#define A(k, p) a[(long)((k) - 1) * N + (p)]

/* get_density_at_nodes: mach2d_coef.f90:6-17 (every point) */
void orc_density_at_nodes(int nx, int ny, double Rg, const double *p, const double *T,
                          double *ro)
{
    const long N = (long)nx * ny;
    for (long q = 0; q < N; ++q) ro[q] = p[q] / (Rg * T[q]);
}

/* get_density_at_faces: mach2d_coef.f90:20-60 */
void orc_density_at_faces(int nx, int ny, double beta, const double *ro, const double *Uce,
                          const double *Vcn, double *roe, double *ron)
{
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            const long np = P(i, j), npe = np + 1;
            const double ae = copysign(0.5, Uce[np]);
            roe[np] = (0.5 + ae) * ro[np] + (0.5 - ae) * ro[npe] + beta * ae * (ro[npe] - ro[np]);
        }
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j), npn = np + nx;
            const double an = copysign(0.5, Vcn[np]);
            ron[np] = (0.5 + an) * ro[np] + (0.5 - an) * ro[npn] + beta * an * (ro[npn] - ro[np]);
        }
}

/* Advection and diffusion part shared by get_u_coefficients (:63-218) and
 * get_v_coefficients (:454-613); the reference spells it out twice. */
static void uv_advection_diffusion(int nx, int ny, int modvis, double dt, const double *rp,
                                   const double *re, const double *rn, const double *Jp,
                                   const double *Je, const double *Jn, const double *alphae,
                                   const double *betae, const double *betan,
                                   const double *gamman, const double *vle, const double *vln,
                                   const double *roe, const double *ron, const double *roa,
                                   const double *Uce, const double *Vcn, double *a, int i, int j)
{
    const long N = (long)nx * ny;
    const long np = P(i, j), nps = np - nx, npw = np - 1;
    const double fme = roe[np] * re[np] * Uce[np];
    const double fmw = roe[npw] * re[npw] * Uce[npw];
    const double fmn = ron[np] * rn[np] * Vcn[np];
    const double fms = ron[nps] * rn[nps] * Vcn[nps];
    const double mpa = roa[np] * rp[np] / Jp[np];
    const double as = copysign(0.5, Vcn[nps]);
    const double an = copysign(0.5, Vcn[np]);
    const double aw = copysign(0.5, Uce[npw]);
    const double ae = copysign(0.5, Uce[np]);
    if (modvis == 0) {
        A(1, np) = 0.0;
        A(3, np) = 0.0;
        A(7, np) = 0.0;
        A(9, np) = 0.0;
        A(2, np) = -fms * (0.5 + as);
        A(4, np) = -fmw * (0.5 + aw);
        A(6, np) = fme * (0.5 - ae);
        A(8, np) = fmn * (0.5 - an);
        A(5, np) = mpa / dt - (A(8, np) + A(2, np) + A(4, np) + A(6, np));
        return;
    }
    const double ew = vle[npw] * re[npw] * Je[npw] * betae[npw];
    const double ep = vle[np] * re[np] * Je[np] * betae[np];
    const double ns = vln[nps] * rn[nps] * Jn[nps] * betan[nps];
    const double nq = vln[np] * rn[np] * Jn[np] * betan[np];
    A(1, np) = (ew + ns) / 4.0;
    A(2, np) = -(fms * (0.5 + as)) - vln[nps] * rn[nps] * Jn[nps] * gamman[nps] + (ew - ep) / 4.0;
    A(3, np) = -(ep + ns) / 4.0;
    A(4, np) = -(fmw * (0.5 + aw)) - vle[npw] * re[npw] * Je[npw] * alphae[npw] + (ns - nq) / 4.0;
    A(6, np) = fme * (0.5 - ae) - vle[np] * re[np] * Je[np] * alphae[np] + (-ns + nq) / 4.0;
    A(7, np) = (-ew - nq) / 4.0;
    A(8, np) = fmn * (0.5 - an) - vln[np] * rn[np] * Jn[np] * gamman[np] - (ew - ep) / 4.0;
    A(9, np) = (ep + nq) / 4.0;
    A(5, np) = mpa / dt - (A(8, np) + A(2, np) + A(4, np) + A(6, np));
}

/* get_u_coefficients: mach2d_coef.f90:63-218 */
void orc_u_coefficients(int nx, int ny, int modvis, double dt, const double *rp, const double *re,
                        const double *rn, const double *Jp, const double *Je, const double *Jn,
                        const double *ye, const double *yk, const double *alphae,
                        const double *betae, const double *betan, const double *gamman,
                        const double *vle, const double *vln, const double *roe,
                        const double *ron, const double *roa, const double *Uce,
                        const double *Vcn, double *a)
{
    const long N = (long)nx * ny;
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            uv_advection_diffusion(nx, ny, modvis, dt, rp, re, rn, Jp, Je, Jn, alphae, betae,
                                   betan, gamman, vle, vln, roe, ron, roa, Uce, Vcn, a, i, j);
            if (modvis == 0) continue;
            const long np = P(i, j), nps = np - nx, npw = np - 1;
            const double S1p = rp[np] * Je[np] * vle[np] * (ye[np] * ye[np]) / 3.0 +
                               rp[np] * Je[npw] * vle[npw] * (ye[npw] * ye[npw]) / 3.0;
            const double S2p = rp[np] * Jn[np] * vln[np] * (yk[np] * yk[np]) / 3.0 +
                               rp[np] * Jn[nps] * vln[nps] * (yk[nps] * yk[nps]) / 3.0;
            A(5, np) = A(5, np) + S1p + S2p;
        }
}

/* get_v_coefficients: mach2d_coef.f90:454-613 */
void orc_v_coefficients(int nx, int ny, int coord, int modvis, double dt, const double *rp,
                        const double *re, const double *rn, const double *Jp, const double *Je,
                        const double *Jn, const double *xe, const double *xk,
                        const double *alphae, const double *betae, const double *betan,
                        const double *gamman, const double *vle, const double *vln,
                        const double *vlp, const double *roe, const double *ron,
                        const double *roa, const double *Uce, const double *Vcn, double *a)
{
    const long N = (long)nx * ny;
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            uv_advection_diffusion(nx, ny, modvis, dt, rp, re, rn, Jp, Je, Jn, alphae, betae,
                                   betan, gamman, vle, vln, roe, ron, roa, Uce, Vcn, a, i, j);
            if (modvis == 0) continue;
            const long np = P(i, j), nps = np - nx, npw = np - 1;
            const double S1p = vle[np] * re[np] * Je[np] * (xe[np] * xe[np]) / 3.0 +
                               vle[npw] * re[npw] * Je[npw] * (xe[npw] * xe[npw]) / 3.0;
            const double S2p = vln[np] * rn[np] * Jn[np] * (xk[np] * xk[np]) / 3.0 +
                               vln[nps] * rn[nps] * Jn[nps] * (xk[nps] * xk[nps]) / 3.0;
            const double S7p = (double)coord * vlp[np] * 4.0 / (3.0 * rp[np] * Jp[np]);
            A(5, np) = A(5, np) + S1p + S2p + S7p;
        }
}

/* get_p_coefficients: mach2d_coef.f90:1207-1278 (a: 5 diagonals S W P E N) */
void orc_p_coefficients(int nx, int ny, double dt, const double *rp, const double *re,
                        const double *rn, const double *Jp, const double *Uce, const double *Vcn,
                        const double *roe, const double *ron, const double *g, const double *de,
                        const double *dn, double *a)
{
    const long N = (long)nx * ny;
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j), nps = np - nx, npn = np + nx, npw = np - 1, npe = np + 1;
            const double as = copysign(0.5, Vcn[nps]);
            const double an = copysign(0.5, Vcn[np]);
            const double aw = copysign(0.5, Uce[npw]);
            const double ae = copysign(0.5, Uce[np]);
            const double roem = roe[np], rowm = roe[npw], ronm = ron[np], rosm = ron[nps];
            A(1, np) = -rn[nps] * ((0.5 + as) * Vcn[nps] * g[nps] + rosm * dn[nps]);
            A(2, np) = -re[npw] * ((0.5 + aw) * Uce[npw] * g[npw] + rowm * de[npw]);
            A(3, np) = g[np] * (rp[np] / (Jp[np] * dt) + (0.5 + ae) * re[np] * Uce[np] -
                                (0.5 - aw) * re[npw] * Uce[npw] + (0.5 + an) * rn[np] * Vcn[np] -
                                (0.5 - as) * rn[nps] * Vcn[nps]) +
                       roem * re[np] * de[np] + rowm * re[npw] * de[npw] +
                       ronm * rn[np] * dn[np] + rosm * rn[nps] * dn[nps];
            A(4, np) = re[np] * ((0.5 - ae) * Uce[np] * g[npe] - roem * de[np]);
            A(5, np) = rn[np] * ((0.5 - an) * Vcn[np] * g[npn] - ronm * dn[np]);
        }
}

/* get_p_source: mach2d_coef.f90:1281-1348 */
void orc_p_source(int nx, int ny, double dt, const double *rp, const double *re,
                  const double *rn, const double *Jp, const double *roe, const double *ron,
                  const double *ro, const double *roa, const double *Uce, const double *Vcn,
                  const double *de, const double *dn, const double *g, const double *pl,
                  double *b)
{
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j), nps = np - nx, npn = np + nx, npw = np - 1, npe = np + 1;
            const double as = copysign(0.5, Vcn[nps]);
            const double an = copysign(0.5, Vcn[np]);
            const double aw = copysign(0.5, Uce[npw]);
            const double ae = copysign(0.5, Uce[np]);
            const double roel = g[np] * pl[np] * (0.5 + ae) + g[npe] * pl[npe] * (0.5 - ae);
            const double rowl = g[npw] * pl[npw] * (0.5 + aw) + g[np] * pl[np] * (0.5 - aw);
            const double ronl = g[np] * pl[np] * (0.5 + an) + g[npn] * pl[npn] * (0.5 - an);
            const double rosl = g[nps] * pl[nps] * (0.5 + as) + g[np] * pl[np] * (0.5 - as);
            const double Ucel = de[np] * (pl[np] - pl[npe]);
            const double Ucwl = de[npw] * (pl[npw] - pl[np]);
            const double Vcnl = dn[np] * (pl[np] - pl[npn]);
            const double Vcsl = dn[nps] * (pl[nps] - pl[np]);
            b[np] = -rp[np] * (ro[np] - roa[np]) / (Jp[np] * dt) - roe[np] * re[np] * Uce[np] +
                    roe[npw] * re[npw] * Uce[npw] - ron[np] * rn[np] * Vcn[np] +
                    ron[nps] * rn[nps] * Vcn[nps] - roel * re[np] * Ucel +
                    rowl * re[npw] * Ucwl - ronl * rn[np] * Vcnl + rosl * rn[nps] * Vcsl;
        }
}

/* get_u_source: mach2d_coef.f90:221-451 (outputs cup, sup, b; sup is only
 * written for modvis = 1) */
void orc_u_source(int nx, int ny, int modvis, double beta, double dt, const double *rp,
                  const double *re, const double *rn, const double *xe, const double *ye,
                  const double *xk, const double *yk, const double *xke, const double *yke,
                  const double *xen, const double *yen, const double *Jp, const double *Je,
                  const double *Jn, const double *roe, const double *ron, const double *roa,
                  const double *p, const double *vle, const double *vln, const double *Uce,
                  const double *Vcn, const double *ua, const double *u, const double *v,
                  double *cup, double *sup, double *b)
{
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j), nps = np - nx, npn = np + nx, npw = np - 1, npe = np + 1;
            const double fme = roe[np] * re[np] * Uce[np];
            const double fmw = roe[npw] * re[npw] * Uce[npw];
            const double fmn = ron[np] * rn[np] * Vcn[np];
            const double fms = ron[nps] * rn[nps] * Vcn[nps];
            const double mpa = roa[np] * rp[np] / Jp[np];
            const double as = copysign(0.5, Vcn[nps]);
            const double an = copysign(0.5, Vcn[np]);
            const double aw = copysign(0.5, Uce[npw]);
            const double ae = copysign(0.5, Uce[np]);
            b[np] = mpa * ua[np] / dt;
            cup[np] = -(beta * (fme * ae * (u[npe] - u[np]) - fmw * aw * (u[np] - u[npw]) +
                                fmn * an * (u[npn] - u[np]) - fms * as * (u[np] - u[nps])));
            b[np] = b[np] + cup[np];
            b[np] = b[np] + 0.5 * rp[np] *
                                (yk[np] * (p[npn] + p[np]) - yk[nps] * (p[nps] + p[np]) -
                                 ye[np] * (p[npe] + p[np]) + ye[npw] * (p[npw] + p[np]));
        }
    if (modvis != 1) return;
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j), nps = np - nx, npn = np + nx, npw = np - 1, npe = np + 1;
            const long npsw = nps - 1, npse = nps + 1, npnw = npn - 1, npne = npn + 1;
            const double S1 =
                rp[np] / 3.0 *
                (Je[np] * vle[np] *
                     (ye[np] * ye[np] * u[npe] -
                      yke[np] * ye[np] * (u[npn] + u[npne] - u[nps] - u[npse]) / 4.0) -
                 Je[npw] * vle[npw] *
                     (-(ye[npw] * ye[npw] * u[npw]) -
                      yke[npw] * ye[npw] * (u[npn] + u[npnw] - u[nps] - u[npsw]) / 4.0));
            const double S2 =
                rp[np] / 3.0 *
                (Jn[np] * vln[np] *
                     (yk[np] * yk[np] * u[npn] -
                      yk[np] * yen[np] * (u[npe] + u[npne] - u[npw] - u[npnw]) / 4.0) -
                 Jn[nps] * vln[nps] *
                     (-(yk[nps] * yk[nps] * u[nps]) -
                      yk[nps] * yen[nps] * (u[npe] + u[npse] - u[npw] - u[npsw]) / 4.0));
            const double S3 =
                Je[np] * re[np] * vle[np] * xe[np] *
                    (yke[np] * (v[npn] + v[npne] - v[nps] - v[npse]) / 4.0 -
                     ye[np] * (v[npe] - v[np])) -
                Je[npw] * re[npw] * vle[npw] * xe[npw] *
                    (yke[npw] * (v[npn] + v[npnw] - v[nps] - v[npsw]) / 4.0 -
                     ye[npw] * (v[np] - v[npw]));
            const double S4 =
                Jn[np] * rn[np] * vln[np] * xk[np] *
                    (yen[np] * (v[npne] + v[npe] - v[npnw] - v[npw]) / 4.0 -
                     yk[np] * (v[npn] - v[np])) -
                Jn[nps] * rn[nps] * vln[nps] * xk[nps] *
                    (yen[nps] * (v[npse] + v[npe] - v[npsw] - v[npw]) / 4.0 -
                     yk[nps] * (v[np] - v[nps]));
            /* the two face terms of S5 / S6; a face on the axis (r = 0) drops its term */
            const double X1 =
                2.0 * rp[np] / 3.0 * Je[np] * vle[np] / re[np] * ye[np] *
                (xke[np] * (rp[npn] * v[npn] + rp[npne] * v[npne] - rp[nps] * v[nps] -
                            rp[npse] * v[npse]) / 4.0 -
                 xe[np] * (rp[npe] * v[npe] - rp[np] * v[np]));
            const double X2 =
                2.0 * rp[np] / 3.0 * Je[npw] * vle[npw] / re[npw] * ye[npw] *
                (xke[npw] * (rp[npn] * v[npn] + rp[npnw] * v[npnw] - rp[nps] * v[nps] -
                             rp[npsw] * v[npsw]) / 4.0 -
                 xe[npw] * (rp[np] * v[np] - rp[npw] * v[npw]));
            double S5;
            if (fabs(re[npw]) < 1e-15)
                S5 = -X1;
            else if (fabs(re[np]) < 1e-15)
                S5 = X2;
            else
                S5 = -X1 + X2;
            const double Y1 =
                2.0 * rp[np] / 3.0 * Jn[np] * vln[np] / rn[np] * yk[np] *
                (xen[np] * (rp[npne] * v[npne] + rp[npe] * v[npe] - rp[npnw] * v[npnw] -
                            rp[npw] * v[npw]) / 4.0 -
                 xk[np] * (rp[npn] * v[npn] - rp[np] * v[np]));
            const double Y2 =
                2.0 * rp[np] / 3.0 * Jn[nps] * vln[nps] / rn[nps] * yk[nps] *
                (xen[nps] * (rp[npe] * v[npe] + rp[npse] * v[npse] - rp[npw] * v[npw] -
                             rp[npsw] * v[npsw]) / 4.0 -
                 xk[nps] * (rp[np] * v[np] - rp[nps] * v[nps]));
            double S6;
            if (fabs(rn[nps]) < 1e-15)
                S6 = -Y1;
            else if (fabs(rn[np]) < 1e-15)
                S6 = Y2;
            else
                S6 = -Y1 + Y2;
            sup[np] = S1 + S2 + S3 + S4 + S5 + S6;
            b[np] = b[np] + sup[np];
        }
}

/* get_v_source: mach2d_coef.f90:616-783 */
void orc_v_source(int nx, int ny, int coord, int modvis, double beta, double dt,
                  const double *rp, const double *re, const double *rn, const double *xe,
                  const double *ye, const double *xk, const double *yk, const double *xke,
                  const double *yke, const double *xen, const double *yen, const double *Jp,
                  const double *Je, const double *Jn, const double *roe, const double *ron,
                  const double *roa, const double *p, const double *vle, const double *vln,
                  const double *Uce, const double *Vcn, const double *va, const double *u,
                  const double *v, double *cvp, double *svp, double *b)
{
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j), nps = np - nx, npn = np + nx, npw = np - 1, npe = np + 1;
            const double fme = roe[np] * re[np] * Uce[np];
            const double fmw = roe[npw] * re[npw] * Uce[npw];
            const double fmn = ron[np] * rn[np] * Vcn[np];
            const double fms = ron[nps] * rn[nps] * Vcn[nps];
            const double mpa = roa[np] * rp[np] / Jp[np];
            const double as = copysign(0.5, Vcn[nps]);
            const double an = copysign(0.5, Vcn[np]);
            const double aw = copysign(0.5, Uce[npw]);
            const double ae = copysign(0.5, Uce[np]);
            b[np] = mpa * va[np] / dt;
            cvp[np] = -(beta * (fme * ae * (v[npe] - v[np]) - fmw * aw * (v[np] - v[npw]) +
                                fmn * an * (v[npn] - v[np]) - fms * as * (v[np] - v[nps])));
            b[np] = b[np] + cvp[np];
            b[np] = b[np] + 0.5 * rp[np] *
                                (xe[np] * (p[np] + p[npe]) - xe[npw] * (p[np] + p[npw]) -
                                 xk[np] * (p[np] + p[npn]) + xk[nps] * (p[np] + p[nps]));
        }
    if (modvis != 1) return;
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j), nps = np - nx, npn = np + nx, npw = np - 1, npe = np + 1;
            const long npsw = nps - 1, npse = nps + 1, npnw = npn - 1, npne = npn + 1;
            const double S1 =
                vle[np] * re[np] * Je[np] / 3.0 *
                    (xe[np] * xe[np] * v[npe] -
                     xke[np] * xe[np] * (v[npn] + v[npne] - v[nps] - v[npse]) / 4.0) -
                vle[npw] * re[npw] * Je[npw] / 3.0 *
                    (-(xe[npw] * xe[npw] * v[npw]) -
                     xke[npw] * xe[npw] * (v[npn] + v[npnw] - v[nps] - v[npsw]) / 4.0);
            const double S2 =
                vln[np] * rn[np] * Jn[np] / 3.0 *
                    (xk[np] * xk[np] * v[npn] -
                     xk[np] * xen[np] * (v[npe] + v[npne] - v[npnw] - v[npw]) / 4.0) -
                vln[nps] * rn[nps] * Jn[nps] / 3.0 *
                    (-(xk[nps] * xk[nps] * v[nps]) -
                     xk[nps] * xen[nps] * (v[npe] + v[npse] - v[npsw] - v[npw]) / 4.0);
            const double S3 =
                rp[np] * vle[np] * Je[np] * ye[np] *
                    (xke[np] * (u[npn] + u[npne] - u[nps] - u[npse]) / 4.0 -
                     xe[np] * (u[npe] - u[np])) -
                rp[np] * vle[npw] * Je[npw] * ye[npw] *
                    (xke[npw] * (u[npn] + u[npnw] - u[nps] - u[npsw]) / 4.0 -
                     xe[npw] * (u[np] - u[npw]));
            const double S4 =
                rp[np] * vln[np] * Jn[np] * yk[np] *
                    (xen[np] * (u[npne] + u[npe] - u[npnw] - u[npw]) / 4.0 -
                     xk[np] * (u[npn] - u[np])) -
                rp[np] * vln[nps] * Jn[nps] * yk[nps] *
                    (xen[nps] * (u[npse] + u[npe] - u[npsw] - u[npw]) / 4.0 -
                     xk[nps] * (u[np] - u[nps]));
            const double S5 =
                -(2.0 / 3.0 * rp[np] * vle[np] * Je[np] * xe[np] *
                  (yke[np] * (u[npn] + u[npne] - u[nps] - u[npse]) / 4.0 -
                   ye[np] * (u[npe] - u[np]))) +
                2.0 / 3.0 * rp[np] * vle[npw] * Je[npw] * xe[npw] *
                    (yke[npw] * (u[npn] + u[npnw] - u[nps] - u[npsw]) / 4.0 -
                     ye[npw] * (u[np] - u[npw]));
            const double S6 =
                -(2.0 / 3.0 * rp[np] * vln[np] * Jn[np] * xk[np] *
                  (yen[np] * (u[npe] + u[npne] - u[npw] - u[npnw]) / 4.0 -
                   yk[np] * (u[npn] - u[np]))) +
                2.0 / 3.0 * rp[np] * vln[nps] * Jn[nps] * xk[nps] *
                    (yen[nps] * (u[npe] + u[npse] - u[npw] - u[npsw]) / 4.0 -
                     yk[nps] * (u[np] - u[nps]));
            const double S8 = 2.0 / 3.0 * (double)coord * v[np] *
                              (-(vln[np] * xk[np]) + vln[nps] * xk[nps] + vle[np] * xe[np] -
                               vle[npw] * xe[npw]);
            svp[np] = S1 + S2 + S3 + S4 + S5 + S6 + S8;
            b[np] = b[np] + svp[np];
        }
}

/* get_T_coefficients_and_source: mach2d_coef.f90:787-1203 */
void orc_T_coefficients_and_source(
    int nx, int ny, int modvis, double beta, double dt, const double *rp, const double *re,
    const double *rn, const double *xe, const double *ye, const double *xk, const double *yk,
    const double *xke, const double *yke, const double *xen, const double *yen,
    const double *alphae, const double *betae, const double *betan, const double *gamman,
    const double *Jp, const double *Je, const double *Jn, const double *roe, const double *ron,
    const double *ro, const double *roa, const double *p, const double *pa, const double *cp,
    const double *vle, const double *vln, const double *ke, const double *kn, const double *Uce,
    const double *Vcn, const double *u, const double *v, const double *ue, const double *ve,
    const double *un, const double *vn, const double *ua, const double *va, const double *T,
    const double *Ta, const double *fbe, const double *fbn, double *a, double *b)
{
    const long N = (long)nx * ny;
    const double eps = 2.220446049250313080847e-16; /* epsilon(1.d0) */
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j), nps = np - nx, npn = np + nx, npw = np - 1, npe = np + 1;
            const long npsw = nps - 1, npse = nps + 1, npnw = npn - 1, npne = npn + 1;
            const double fme = roe[np] * re[np] * Uce[np];
            const double fmw = roe[npw] * re[npw] * Uce[npw];
            const double fmn = ron[np] * rn[np] * Vcn[np];
            const double fms = ron[nps] * rn[nps] * Vcn[nps];
            const double mpa = roa[np] * rp[np] / Jp[np];
            const double as = copysign(0.5, Vcn[nps]);
            const double an = copysign(0.5, Vcn[np]);
            const double aw = copysign(0.5, Uce[npw]);
            const double ae = copysign(0.5, Uce[np]);
            if (modvis == 0) {
                A(1, np) = 0.0;
                A(3, np) = 0.0;
                A(7, np) = 0.0;
                A(9, np) = 0.0;
                A(2, np) = -fms * (0.5 + as) * cp[np];
                A(4, np) = -fmw * (0.5 + aw) * cp[np];
                A(6, np) = fme * (0.5 - ae) * cp[np];
                A(8, np) = fmn * (0.5 - an) * cp[np];
            } else {
                const double ew = ke[npw] * re[npw] * Je[npw] * betae[npw];
                const double ep = ke[np] * re[np] * Je[np] * betae[np];
                const double ns = kn[nps] * rn[nps] * Jn[nps] * betan[nps];
                const double nq = kn[np] * rn[np] * Jn[np] * betan[np];
                A(1, np) = (ew + ns) / 4.0;
                A(2, np) = -(fms * (0.5 + as) * cp[np]) - kn[nps] * rn[nps] * Jn[nps] * gamman[nps] +
                           (ew - ep) / 4.0;
                A(3, np) = -(ep + ns) / 4.0;
                A(4, np) = -(fmw * (0.5 + aw) * cp[np]) - ke[npw] * re[npw] * Je[npw] * alphae[npw] +
                           (ns - nq) / 4.0;
                A(6, np) = fme * (0.5 - ae) * cp[np] - ke[np] * re[np] * Je[np] * alphae[np] +
                           (-ns + nq) / 4.0;
                A(7, np) = (-ew - nq) / 4.0;
                A(8, np) = fmn * (0.5 - an) * cp[np] - kn[np] * rn[np] * Jn[np] * gamman[np] -
                           (ew - ep) / 4.0;
                A(9, np) = (ep + nq) / 4.0;
            }
            A(5, np) = mpa * cp[np] / dt - (A(8, np) + A(2, np) + A(4, np) + A(6, np));
            b[np] = mpa * Ta[np] * cp[np] / dt;
            b[np] = b[np] - beta * cp[np] *
                                (fme * ae * (T[npe] - T[np]) - fmw * aw * (T[np] - T[npw]) +
                                 fmn * an * (T[npn] - T[np]) - fms * as * (T[np] - T[nps]));
            const double KpE = (u[npe] * u[npe] + v[npe] * v[npe]) / 2.0;
            const double KpW = (u[npw] * u[npw] + v[npw] * v[npw]) / 2.0;
            const double KpN = (u[npn] * u[npn] + v[npn] * v[npn]) / 2.0;
            const double KpS = (u[nps] * u[nps] + v[nps] * v[nps]) / 2.0;
            const double KpP = (u[np] * u[np] + v[np] * v[np]) / 2.0;
            const double KpPa = (ua[np] * ua[np] + va[np] * va[np]) / 2.0;
            double Kfe = (0.5 + ae) * KpP + (0.5 - ae) * KpE + beta * ae * (KpE - KpP);
            double Kfw = (0.5 + aw) * KpW + (0.5 - aw) * KpP + beta * aw * (KpP - KpW);
            double Kfn = (0.5 + an) * KpP + (0.5 - an) * KpN + beta * an * (KpN - KpP);
            double Kfs = (0.5 + as) * KpS + (0.5 - as) * KpP + beta * as * (KpP - KpS);
            const double Kfbe = (ue[np] * ue[np] + ve[np] * ve[np]) / 2.0;
            const double Kfbw = (ue[npw] * ue[npw] + ve[npw] * ve[npw]) / 2.0;
            const double Kfbn = (un[np] * un[np] + vn[np] * vn[np]) / 2.0;
            const double Kfbs = (un[nps] * un[nps] + vn[nps] * vn[nps]) / 2.0;
            Kfe = Kfe * (1.0 - fbe[np]) + Kfbe * fbe[np];
            Kfw = Kfw * (1.0 - fbe[npw]) + Kfbw * fbe[npw];
            Kfn = Kfn * (1.0 - fbn[np]) + Kfbn * fbn[np];
            Kfs = Kfs * (1.0 - fbn[nps]) + Kfbs * fbn[nps];
            double raux = rp[np] / Jp[np] * (ro[np] * KpP - roa[np] * KpPa) / dt + fme * Kfe -
                          fmw * Kfw + fmn * Kfn - fms * Kfs;
            b[np] = b[np] + rp[np] / Jp[np] * (p[np] - pa[np]) / dt - raux;
            if (modvis == 0) continue;
            const double dudk_e = u[npe] - u[np];
            const double dudk_w = u[np] - u[npw];
            const double dudk_n = (u[npne] + u[npe] - u[npnw] - u[npw]) / 4.0;
            const double dudk_s = (u[npse] + u[npe] - u[npsw] - u[npw]) / 4.0;
            const double dude_e = (u[npne] + u[npn] - u[npse] - u[nps]) / 4.0;
            const double dude_w = (u[npnw] + u[npn] - u[npsw] - u[nps]) / 4.0;
            const double dude_n = u[npn] - u[np];
            const double dude_s = u[np] - u[nps];
            const double dvdk_e = v[npe] - v[np];
            const double dvdk_w = v[np] - v[npw];
            const double dvdk_n = (v[npne] + v[npe] - v[npnw] - v[npw]) / 4.0;
            const double dvdk_s = (v[npse] + v[npe] - v[npsw] - v[npw]) / 4.0;
            const double dvde_e = (v[npne] + v[npn] - v[npse] - v[nps]) / 4.0;
            const double dvde_w = (v[npnw] + v[npn] - v[npsw] - v[nps]) / 4.0;
            const double dvde_n = v[npn] - v[np];
            const double dvde_s = v[np] - v[nps];
            const double dudx_e = Je[np] * (ye[np] * dudk_e - yke[np] * dude_e);
            const double dudx_w = Je[npw] * (ye[npw] * dudk_w - yke[npw] * dude_w);
            const double dudx_n = Jn[np] * (yen[np] * dudk_n - yk[np] * dude_n);
            const double dudx_s = Jn[nps] * (yen[nps] * dudk_s - yk[nps] * dude_s);
            const double dudy_e = Je[np] * (xke[np] * dude_e - xe[np] * dudk_e);
            const double dudy_w = Je[npw] * (xke[npw] * dude_w - xe[npw] * dudk_w);
            const double dudy_n = Jn[np] * (xk[np] * dude_n - xen[np] * dudk_n);
            const double dudy_s = Jn[nps] * (xk[nps] * dude_s - xen[nps] * dudk_s);
            const double dvdx_e = Je[np] * (ye[np] * dvdk_e - yke[np] * dvde_e);
            const double dvdx_w = Je[npw] * (ye[npw] * dvdk_w - yke[npw] * dvde_w);
            const double dvdx_n = Jn[np] * (yen[np] * dvdk_n - yk[np] * dvde_n);
            const double dvdx_s = Jn[nps] * (yen[nps] * dvdk_s - yk[nps] * dvde_s);
            const double dvdy_e = Je[np] * (xke[np] * dvde_e - xe[np] * dvdk_e);
            const double dvdy_w = Je[npw] * (xke[npw] * dvde_w - xe[npw] * dvdk_w);
            const double dvdy_n = Jn[np] * (xk[np] * dvde_n - xen[np] * dvdk_n);
            const double dvdy_s = Jn[nps] * (xk[nps] * dvde_s - xen[nps] * dvdk_s);
            const double divU_e = dudx_e + dvdy_e + ve[np] / (fabs(re[np]) + eps);
            const double divU_w = dudx_w + dvdy_w + ve[npw] / (fabs(re[npw]) + eps);
            const double divU_n = dudx_n + dvdy_n + vn[np] / (fabs(rn[np]) + eps);
            const double divU_s = dudx_s + dvdy_s + vn[nps] / (fabs(rn[nps]) + eps);
            const double Txx_e = -2.0 * vle[np] * dudx_e + 2.0 / 3.0 * vle[np] * divU_e;
            const double Txx_w = -2.0 * vle[npw] * dudx_w + 2.0 / 3.0 * vle[npw] * divU_w;
            const double Txx_n = -2.0 * vln[np] * dudx_n + 2.0 / 3.0 * vln[np] * divU_n;
            const double Txx_s = -2.0 * vln[nps] * dudx_s + 2.0 / 3.0 * vln[nps] * divU_s;
            const double Txy_e = -vle[np] * (dvdx_e + dudy_e);
            const double Txy_w = -vle[npw] * (dvdx_w + dudy_w);
            const double Txy_n = -vln[np] * (dvdx_n + dudy_n);
            const double Txy_s = -vln[nps] * (dvdx_s + dudy_s);
            const double Tyy_e = -2.0 * vle[np] * dvdy_e + 2.0 / 3.0 * vle[np] * divU_e;
            const double Tyy_w = -2.0 * vle[npw] * dvdy_w + 2.0 / 3.0 * vle[npw] * divU_w;
            const double Tyy_n = -2.0 * vln[np] * dvdy_n + 2.0 / 3.0 * vln[np] * divU_n;
            const double Tyy_s = -2.0 * vln[nps] * dvdy_s + 2.0 / 3.0 * vln[nps] * divU_s;
            const double Lx_e = ve[np] * Txy_e + ue[np] * Txx_e;
            const double Lx_w = ve[npw] * Txy_w + ue[npw] * Txx_w;
            const double Lx_n = vn[np] * Txy_n + un[np] * Txx_n;
            const double Lx_s = vn[nps] * Txy_s + un[nps] * Txx_s;
            const double Ly_e = ve[np] * Tyy_e + ue[np] * Txy_e;
            const double Ly_w = ve[npw] * Tyy_w + ue[npw] * Txy_w;
            const double Ly_n = vn[np] * Tyy_n + un[np] * Txy_n;
            const double Ly_s = vn[nps] * Tyy_s + un[nps] * Txy_s;
            const double Lk_e = Lx_e * ye[np] - Ly_e * xe[np];
            const double Lk_w = Lx_w * ye[npw] - Ly_w * xe[npw];
            const double Le_n = Ly_n * xk[np] - Lx_n * yk[np];
            const double Le_s = Ly_s * xk[nps] - Lx_s * yk[nps];
            raux = re[np] * Lk_e - re[npw] * Lk_w + rn[np] * Le_n - rn[nps] * Le_s;
            b[np] = b[np] - raux;
        }
}

/* ---- SIMPLEC glue ---------------------------------------------------------- */

/* sum(a(np,:)) of a 9-diagonal array, left to right */
static double sum9(const double *a, long N, long np)
{
    double s = a[np];
    for (int k = 1; k < 9; ++k) s = s + a[k * N + np];
    return s;
}

/* get_velocities_at_internal_faces: mach2d_coef.f90:1350-1501 */
void orc_velocities_at_internal_faces(
    int nx, int ny, double dt, const double *rp, const double *re, const double *rn,
    const double *xe, const double *ye, const double *xk, const double *yk, const double *xen,
    const double *yen, const double *xke, const double *yke, const double *Jp,
    const double *cup, const double *cvp, const double *sup, const double *svp,
    const double *au, const double *av, const double *roa, const double *p, const double *u,
    const double *v, const double *uea, const double *vea, const double *una,
    const double *vna, double *ue, double *ve, double *un, double *vn, double *Uce,
    double *Vcn)
{
    const long N = (long)nx * ny;
#define AU(k, q) au[(long)((k) - 1) * N + (q)]
#define AV(k, q) av[(long)((k) - 1) * N + (q)]
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 2; ++i) {
            const long np = P(i, j), nps = np - nx, npn = np + nx, npw = np - 1, npe = np + 1;
            const long npsw = nps - 1, npse = nps + 1, npnw = npn - 1, npne = npn + 1;
            const long npee = npe + 1, npsee = npse + 1, npnee = npne + 1;
            const double mpa = roa[np] * rp[np] / Jp[np];
            const double mea = roa[npe] * rp[npe] / Jp[npe];
            const double sumup = AU(1, np) * u[npsw] + AU(2, np) * u[nps] + AU(3, np) * u[npse] +
                                 AU(4, np) * u[npw] + AU(6, np) * u[npe] + AU(7, np) * u[npnw] +
                                 AU(8, np) * u[npn] + AU(9, np) * u[npne];
            const double sumue = AU(1, npe) * u[nps] + AU(2, npe) * u[npse] +
                                 AU(3, npe) * u[npsee] + AU(4, npe) * u[np] +
                                 AU(6, npe) * u[npee] + AU(7, npe) * u[npn] +
                                 AU(8, npe) * u[npne] + AU(9, npe) * u[npnee];
            const double sumvp = AV(1, np) * v[npsw] + AV(2, np) * v[nps] + AV(3, np) * v[npse] +
                                 AV(4, np) * v[npw] + AV(6, np) * v[npe] + AV(7, np) * v[npnw] +
                                 AV(8, np) * v[npn] + AV(9, np) * v[npne];
            const double sumve = AV(1, npe) * v[nps] + AV(2, npe) * v[npse] +
                                 AV(3, npe) * v[npsee] + AV(4, npe) * v[np] +
                                 AV(6, npe) * v[npee] + AV(7, npe) * v[npn] +
                                 AV(8, npe) * v[npne] + AV(9, npe) * v[npnee];
            const double aux = (p[npn] + p[npne] - p[nps] - p[npse]) / 4.0;
            const double pue = re[np] * (yke[np] * aux + ye[np] * (p[np] - p[npe]));
            const double pve = re[np] * (-xke[np] * aux + xe[np] * (p[npe] - p[np]));
            ue[np] = ((mpa + mea) * uea[np] / dt + cup[np] + cup[npe] + sup[np] + sup[npe] -
                      sumup - sumue + 2.0 * pue) /
                     (AU(5, np) + AU(5, npe));
            ve[np] = ((mpa + mea) * vea[np] / dt + cvp[np] + cvp[npe] + svp[np] + svp[npe] -
                      sumvp - sumve + 2.0 * pve) /
                     (AV(5, np) + AV(5, npe));
            Uce[np] = ue[np] * ye[np] - ve[np] * xe[np];
        }
    for (int j = 2; j <= ny - 2; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j), nps = np - nx, npn = np + nx, npw = np - 1, npe = np + 1;
            const long npsw = nps - 1, npse = nps + 1, npnw = npn - 1, npne = npn + 1;
            const long npnn = npn + nx, npnnw = npnn - 1, npnne = npnn + 1;
            const double mpa = roa[np] * rp[np] / Jp[np];
            const double mna = roa[npn] * rp[npn] / Jp[npn];
            const double sumup = AU(1, np) * u[npsw] + AU(2, np) * u[nps] + AU(3, np) * u[npse] +
                                 AU(4, np) * u[npw] + AU(6, np) * u[npe] + AU(7, np) * u[npnw] +
                                 AU(8, np) * u[npn] + AU(9, np) * u[npne];
            const double sumun = AU(1, npn) * u[npw] + AU(2, npn) * u[np] + AU(3, npn) * u[npe] +
                                 AU(4, npn) * u[npnw] + AU(6, npn) * u[npne] +
                                 AU(7, npn) * u[npnnw] + AU(8, npn) * u[npnn] +
                                 AU(9, npn) * u[npnne];
            const double sumvp = AV(1, np) * v[npsw] + AV(2, np) * v[nps] + AV(3, np) * v[npse] +
                                 AV(4, np) * v[npw] + AV(6, np) * v[npe] + AV(7, np) * v[npnw] +
                                 AV(8, np) * v[npn] + AV(9, np) * v[npne];
            const double sumvn = AV(1, npn) * v[npw] + AV(2, npn) * v[np] + AV(3, npn) * v[npe] +
                                 AV(4, npn) * v[npnw] + AV(6, npn) * v[npne] +
                                 AV(7, npn) * v[npnnw] + AV(8, npn) * v[npnn] +
                                 AV(9, npn) * v[npnne];
            const double aux = (p[npne] + p[npe] - p[npnw] - p[npw]) / 4.0;
            const double pun = rn[np] * (yk[np] * (p[npn] - p[np]) - yen[np] * aux);
            const double pvn = rn[np] * (xk[np] * (p[np] - p[npn]) + xen[np] * aux);
            un[np] = ((mpa + mna) * una[np] / dt + cup[np] + cup[npn] + sup[np] + sup[npn] -
                      sumup - sumun + 2.0 * pun) /
                     (AU(5, np) + AU(5, npn));
            vn[np] = ((mpa + mna) * vna[np] / dt + cvp[np] + cvp[npn] + svp[np] + svp[npn] -
                      sumvp - sumvn + 2.0 * pvn) /
                     (AV(5, np) + AV(5, npn));
            Vcn[np] = vn[np] * xk[np] - un[np] * yk[np];
        }
#undef AU
#undef AV
}

/* get_internal_simplec_coefficients: mach2d_coef.f90:1504-1600 */
void orc_internal_simplec_coefficients(int nx, int ny, const double *re, const double *rn,
                                       const double *xe, const double *ye, const double *xk,
                                       const double *yk, const double *au, const double *av,
                                       double *due, double *dve, double *dun, double *dvn,
                                       double *de, double *dn)
{
    const long N = (long)nx * ny;
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 2; ++i) {
            const long np = P(i, j), npe = np + 1;
            const double Lue = 2.0 * re[np] / (sum9(au, N, np) + sum9(au, N, npe));
            const double Lve = 2.0 * re[np] / (sum9(av, N, np) + sum9(av, N, npe));
            due[np] = Lue * ye[np];
            dve[np] = Lve * xe[np];
            de[np] = ye[np] * due[np] + xe[np] * dve[np];
        }
    for (int j = 2; j <= ny - 2; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j), npn = np + nx;
            const double Lun = 2.0 * rn[np] / (sum9(au, N, np) + sum9(au, N, npn));
            const double Lvn = 2.0 * rn[np] / (sum9(av, N, np) + sum9(av, N, npn));
            dun[np] = Lun * yk[np];
            dvn[np] = Lvn * xk[np];
            dn[np] = xk[np] * dvn[np] + yk[np] * dun[np];
        }
}

/* get_pressure_correction_with_pl: mach2d_coef.f90:1602-1611 (every point) */
void orc_pressure_correction_with_pl(int nx, int ny, const double *pl, double *p)
{
    const long N = (long)nx * ny;
    for (long q = 0; q < N; ++q) p[q] = p[q] + pl[q];
}

/* get_density_correction_with_pl: mach2d_coef.f90:1614-1624 (every point) */
void orc_density_correction_with_pl(int nx, int ny, const double *pl, const double *g,
                                    double *ro)
{
    const long N = (long)nx * ny;
    for (long q = 0; q < N; ++q) ro[q] = ro[q] + pl[q] * g[q];
}

/* get_u_v_at_real_nodes_with_pl: mach2d_coef.f90:1626-1675 */
void orc_u_v_at_real_nodes_with_pl(int nx, int ny, const double *xe, const double *ye,
                                   const double *xk, const double *yk, const double *rp,
                                   const double *pl, const double *au, const double *av,
                                   double *u, double *v)
{
    const long N = (long)nx * ny;
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j), nps = np - nx, npn = np + nx, npw = np - 1, npe = np + 1;
            const double plup = 0.5 * rp[np] *
                                (yk[np] * (pl[npn] + pl[np]) - yk[nps] * (pl[nps] + pl[np]) -
                                 ye[np] * (pl[npe] + pl[np]) + ye[npw] * (pl[npw] + pl[np]));
            const double plvp = 0.5 * rp[np] *
                                (xe[np] * (pl[np] + pl[npe]) - xe[npw] * (pl[np] + pl[npw]) -
                                 xk[np] * (pl[np] + pl[npn]) + xk[nps] * (pl[np] + pl[nps]));
            u[np] = u[np] + plup / sum9(au, N, np);
            v[np] = v[np] + plvp / sum9(av, N, np);
        }
}

/* get_velocity_correction_at_faces_with_pl: mach2d_coef.f90:1679-1758 */
void orc_velocity_correction_at_faces_with_pl(int nx, int ny, const double *xe,
                                              const double *ye, const double *xk,
                                              const double *yk, const double *due,
                                              const double *dve, const double *dun,
                                              const double *dvn, const double *de,
                                              const double *dn, const double *pl, double *ue,
                                              double *ve, double *un, double *vn, double *Uce,
                                              double *Vcn)
{
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            const long np = P(i, j), npe = np + 1;
            Uce[np] = ue[np] * ye[np] - ve[np] * xe[np] + de[np] * (pl[np] - pl[npe]);
            ue[np] = ue[np] + due[np] * (pl[np] - pl[npe]);
            ve[np] = ve[np] + dve[np] * (pl[npe] - pl[np]);
        }
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j), npn = np + nx;
            Vcn[np] = vn[np] * xk[np] - un[np] * yk[np] + dn[np] * (pl[np] - pl[npn]);
            un[np] = un[np] + dun[np] * (pl[npn] - pl[np]);
            vn[np] = vn[np] + dvn[np] * (pl[np] - pl[npn]);
        }
}

/* initialize_boundary_faces_flags: mach2d_coef.f90:1764-1828 */
void orc_boundary_faces_flags(int nx, int ny, double *fbe, double *fbn)
{
    const long N = (long)nx * ny;
    for (long q = 0; q < N; ++q) fbe[q] = 0.0;
    for (int j = 2; j <= ny - 1; ++j) fbe[P(1, j)] = 1.0;
    for (int j = 2; j <= ny - 1; ++j) fbe[P(nx - 1, j)] = 1.0;
    for (long q = 0; q < N; ++q) fbn[q] = 0.0;
    for (int i = 2; i <= nx - 1; ++i) fbn[P(i, ny - 1)] = 1.0;
    for (int i = 2; i <= nx - 1; ++i) fbn[P(i, 1)] = 1.0;
}

/* check_mass_conservation: mach2d_coef.f90:1833-1880 (L1 sum in loop order) */
void orc_mass_conservation(int nx, int ny, double dt, const double *rp, const double *re,
                           const double *rn, const double *Jp, const double *ro,
                           const double *roa, const double *ron, const double *roe,
                           const double *Uce, const double *Vcn, double *res)
{
    double r = 0.0;
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            const long np = P(i, j), nps = np - nx, npw = np - 1;
            const double fme = roe[np] * re[np] * Uce[np];
            const double fmw = roe[npw] * re[npw] * Uce[npw];
            const double fmn = ron[np] * rn[np] * Vcn[np];
            const double fms = ron[nps] * rn[nps] * Vcn[nps];
            r = r + fabs((ro[np] - roa[np]) * rp[np] / (Jp[np] * dt) + fme - fmw + fmn - fms);
        }
    *res = r;
}
