/*
 * tdma2d9.c -- oracle restatement of the alternating-direction line TDMA for
 * 9-diagonal systems (reference: src/solvers/tdma2d9.f90:38-322), and of the
 * L1 residual norms (src/solvers/mod_solvers.f90:128-276).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 *
 * Each tdma2d9 iteration is two line-Jacobi passes: every x-line (fixed j) is
 * solved with W/E implicit and the other six neighbours taken from x as it
 * was before the pass; then every y-line (fixed i) with S/N implicit.  The
 * reference splits the lines over OpenMP threads; all forward eliminations
 * of a pass finish before any back-substitution writes x, so the result does
 * not depend on the split and a serial loop reproduces it.
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

#define A(np, k) a[(size_t)((k) - 1) * nxy + (np) - 1]
#define BB(np) b[(np) - 1]
#define X(np) x[(np) - 1]
#define PP(np) P[(np) - 1]
#define QQ(np) Q[(np) - 1]

void orc_tdma2d9(int nx, int ny, int nitm, const double *a, const double *b,
                 double *x)
{
    const int nxy = nx * ny;
    int i, j, it, np, nps, npn, npw, npe, npsw, npse, npnw, npne;
    double D;
    double *P = (double *)malloc(sizeof(double) * (size_t)nxy);
    double *Q = (double *)malloc(sizeof(double) * (size_t)nxy);

    for (it = 1; it <= nitm; ++it) {
        /* x-lines, forward elimination: tdma2d9.f90:79-173 */
        for (j = 1; j <= ny; ++j) {
            i = 1;
            np = nx * (j - 1) + i;
            nps = np - nx;
            npn = np + nx;
            npse = nps + 1;
            npne = npn + 1;
            if (j == 1)
                D = BB(np) - A(np, 8) * X(npn) - A(np, 9) * X(npne);
            else if (j == ny)
                D = BB(np) - A(np, 2) * X(nps) - A(np, 3) * X(npse);
            else
                D = BB(np) - A(np, 2) * X(nps) - A(np, 3) * X(npse)
                  - A(np, 8) * X(npn) - A(np, 9) * X(npne);
            PP(np) = -A(np, 6) / A(np, 5);
            QQ(np) = D / A(np, 5);

            for (i = 2; i <= nx - 1; ++i) {
                np = nx * (j - 1) + i;
                nps = np - nx;
                npn = np + nx;
                npsw = nps - 1;
                npse = nps + 1;
                npnw = npn - 1;
                npne = npn + 1;
                if (j == 1)
                    D = BB(np) - A(np, 7) * X(npnw) - A(np, 8) * X(npn) - A(np, 9) * X(npne);
                else if (j == ny)
                    D = BB(np) - A(np, 1) * X(npsw) - A(np, 2) * X(nps) - A(np, 3) * X(npse);
                else
                    D = BB(np) - A(np, 1) * X(npsw) - A(np, 2) * X(nps) - A(np, 3) * X(npse)
                      - A(np, 7) * X(npnw) - A(np, 8) * X(npn) - A(np, 9) * X(npne);
                PP(np) = -A(np, 6) / (A(np, 5) + A(np, 4) * PP(np - 1));
                QQ(np) = (D - A(np, 4) * QQ(np - 1)) / (A(np, 5) + A(np, 4) * PP(np - 1));
            }

            i = nx;
            np = nx * (j - 1) + i;
            nps = np - nx;
            npn = np + nx;
            npsw = nps - 1;
            npnw = npn - 1;
            if (j == 1)
                D = BB(np) - A(np, 7) * X(npnw) - A(np, 8) * X(npn);
            else if (j == ny)
                D = BB(np) - A(np, 1) * X(npsw) - A(np, 2) * X(nps);
            else
                D = BB(np) - A(np, 1) * X(npsw) - A(np, 2) * X(nps)
                  - A(np, 7) * X(npnw) - A(np, 8) * X(npn);
            PP(np) = -A(np, 6) / (A(np, 5) + A(np, 4) * PP(np - 1));
            QQ(np) = (D - A(np, 4) * QQ(np - 1)) / (A(np, 5) + A(np, 4) * PP(np - 1));
        }
        /* x-lines, back substitution: :175-193 */
        for (j = 1; j <= ny; ++j) {
            np = nx * (j - 1) + nx;
            X(np) = QQ(np);
            for (i = nx - 1; i >= 1; --i) {
                np = nx * (j - 1) + i;
                X(np) = PP(np) * X(np + 1) + QQ(np);
            }
        }

        /* y-lines, forward elimination: :200-296 */
        for (i = 1; i <= nx; ++i) {
            j = 1;
            np = nx * (j - 1) + i;
            npn = np + nx;
            npw = np - 1;
            npe = np + 1;
            npnw = npn - 1;
            npne = npn + 1;
            if (i == 1)
                D = BB(np) - A(np, 6) * X(npe) - A(np, 9) * X(npne);
            else if (i == nx)
                D = BB(np) - A(np, 4) * X(npw) - A(np, 7) * X(npnw);
            else
                D = BB(np) - A(np, 4) * X(npw) - A(np, 7) * X(npnw)
                  - A(np, 6) * X(npe) - A(np, 9) * X(npne);
            PP(np) = -A(np, 8) / A(np, 5);
            QQ(np) = D / A(np, 5);

            for (j = 2; j <= ny - 1; ++j) {
                np = nx * (j - 1) + i;
                nps = np - nx;
                npn = np + nx;
                npw = np - 1;
                npe = np + 1;
                npsw = nps - 1;
                npse = nps + 1;
                npnw = npn - 1;
                npne = npn + 1;
                if (i == 1)
                    D = BB(np) - A(np, 3) * X(npse) - A(np, 6) * X(npe) - A(np, 9) * X(npne);
                else if (i == nx)
                    D = BB(np) - A(np, 1) * X(npsw) - A(np, 4) * X(npw) - A(np, 7) * X(npnw);
                else
                    D = BB(np) - A(np, 1) * X(npsw) - A(np, 4) * X(npw) - A(np, 7) * X(npnw)
                      - A(np, 3) * X(npse) - A(np, 6) * X(npe) - A(np, 9) * X(npne);
                PP(np) = -A(np, 8) / (A(np, 5) + A(np, 2) * PP(np - nx));
                QQ(np) = (D - A(np, 2) * QQ(np - nx)) / (A(np, 5) + A(np, 2) * PP(np - nx));
            }

            j = ny;
            np = nx * (j - 1) + i;
            nps = np - nx;
            npw = np - 1;
            npe = np + 1;
            npsw = nps - 1;
            npse = nps + 1;
            if (i == 1)
                D = BB(np) - A(np, 3) * X(npse) - A(np, 6) * X(npe);
            else if (i == nx)
                D = BB(np) - A(np, 1) * X(npsw) - A(np, 4) * X(npw);
            else
                D = BB(np) - A(np, 1) * X(npsw) - A(np, 4) * X(npw)
                  - A(np, 3) * X(npse) - A(np, 6) * X(npe);
            PP(np) = -A(np, 8) / (A(np, 5) + A(np, 2) * PP(np - nx));
            QQ(np) = (D - A(np, 2) * QQ(np - nx)) / (A(np, 5) + A(np, 2) * PP(np - nx));
        }
        /* y-lines, back substitution: :298-316 */
        for (i = 1; i <= nx; ++i) {
            np = nx * (ny - 1) + i;
            X(np) = QQ(np);
            for (j = ny - 1; j >= 1; --j) {
                np = nx * (j - 1) + i;
                X(np) = X(np + nx) * PP(np) + QQ(np);
            }
        }
    }
    free(P);
    free(Q);
}

/* --- L1 norms over real volumes, src/solvers/mod_solvers.f90 --------------- */

#define VAR(np) var[(np) - 1]

static double row_l1(int nx, int nxy, const double *a, const double *var,
                     const double *b, int np)
{
    int nps = np - nx, npn = np + nx, npw = np - 1, npe = np + 1;
    int npsw = nps - 1, npse = nps + 1, npnw = npn - 1, npne = npn + 1;
    /* the leading unary '+' of the reference changes nothing */
    return fabs(A(np, 1) * VAR(npsw) + A(np, 2) * VAR(nps) + A(np, 3) * VAR(npse)
              + A(np, 4) * VAR(npw) + A(np, 5) * VAR(np) + A(np, 6) * VAR(npe)
              + A(np, 7) * VAR(npnw) + A(np, 8) * VAR(npn) + A(np, 9) * VAR(npne)
              - BB(np));
}

/* mod_solvers.f90:128-170 */
void orc_norm_l1_9d(int nx, int ny, const double *var, const double *b,
                    const double *a, double *norm)
{
    const int nxy = nx * ny;
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i)
            *norm = *norm + row_l1(nx, nxy, a, var, b, nx * (j - 1) + i);
}

/* mod_solvers.f90:176-245 */
void orc_norm_l1_9d_relative(int nx, int ny, const double *var,
                             const double *b, const double *a, double *norm)
{
    const int nxy = nx * ny;
    const double eps = 10.0 * 2.220446049250313080847e-16; /* 10*epsilon(1d0) */
    double norma = 0.0, normb = 0.0;
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) {
            int np = nx * (j - 1) + i;
            norma = norma + row_l1(nx, nxy, a, var, b, np);
            normb = normb + fabs(BB(np));
        }
    if (normb > eps)
        *norm = *norm + norma / normb;
    else
        *norm = *norm + norma;
}

/* mod_solvers.f90:250-276 */
double orc_norm_l1_b(int nx, int ny, const double *b)
{
    double norm = 0.0;
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx - 1; ++i) norm = norm + fabs(BB(nx * (j - 1) + i));
    return norm;
}
