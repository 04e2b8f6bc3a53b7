/*
 * msi2d9.c -- oracle restatement of the 9-diagonal MSI solver
 * (reference: src/solvers/msi2d9-v03.f90).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 *
 * The flat-index special rows and the operation order of every expression
 * follow the reference so that, built with -ffp-contract=off, the results
 * equal the reference's bit for bit (the reference is compiled without FMA
 * contraction on x86-64).  1-based macros keep the index arithmetic the
 * reference's.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define B(ip, k) b[(size_t)((k) - 1) * nij + (ip) - 1]
#define DL(ip, k) dl[(size_t)((k) - 1) * nij + (ip) - 1]
#define DU(ip, k) du[(size_t)((k) - 1) * nij + (ip) - 1]
#define V(ip) var[(ip) - 1]
#define RHS(ip) rhs[(ip) - 1]
#define R(ip) r[(ip) - 1]
#define W(ip) w[(ip) - 1]
#define Z(ip) z[(ip) - 1]

/* Residual r = A.var - rhs, src/solvers/msi2d9-v03.f90:59-131 (repeated
 * verbatim in the iteration at :183-255).  Seven row classes: the four
 * corners, the first/last grid rows and the interior with i=1 / i=ni
 * branches (mod(ip,ni)). */
static void residual9(const double *b, int ni, int nij, const double *var,
                      const double *rhs, double *r)
{
    int ip, iw, ie, is, in, isw, ise, inw, ine;

    ip = 1;
    ie = ip + 1;
    in = ip + ni;
    ine = in + 1;
    R(ip) = -RHS(ip) + B(ip, 5) * V(ip) + B(ip, 6) * V(ie)
          + B(ip, 8) * V(in) + B(ip, 9) * V(ine);

    for (ip = 2; ip <= ni - 1; ++ip) {
        iw = ip - 1;
        ie = ip + 1;
        in = ip + ni;
        inw = in - 1;
        ine = in + 1;
        R(ip) = -RHS(ip) + B(ip, 4) * V(iw) + B(ip, 5) * V(ip) + B(ip, 6) * V(ie)
              + B(ip, 7) * V(inw) + B(ip, 8) * V(in) + B(ip, 9) * V(ine);
    }

    ip = ni;
    iw = ip - 1;
    in = ip + ni;
    inw = in - 1;
    R(ip) = -RHS(ip) + B(ip, 4) * V(iw) + B(ip, 5) * V(ip)
          + B(ip, 7) * V(inw) + B(ip, 8) * V(in);

    for (ip = ni + 1; ip <= nij - ni; ++ip) {
        iw = ip - 1;
        ie = ip + 1;
        is = ip - ni;
        in = ip + ni;
        isw = is - 1;
        ise = is + 1;
        inw = in - 1;
        ine = in + 1;
        if (ip % ni == 1) {
            R(ip) = -RHS(ip) + B(ip, 2) * V(is) + B(ip, 3) * V(ise)
                  + B(ip, 5) * V(ip) + B(ip, 6) * V(ie)
                  + B(ip, 8) * V(in) + B(ip, 9) * V(ine);
        } else if (ip % ni == 0) {
            R(ip) = -RHS(ip) + B(ip, 1) * V(isw) + B(ip, 2) * V(is)
                  + B(ip, 4) * V(iw) + B(ip, 5) * V(ip)
                  + B(ip, 7) * V(inw) + B(ip, 8) * V(in);
        } else {
            R(ip) = -RHS(ip) + B(ip, 1) * V(isw) + B(ip, 2) * V(is) + B(ip, 3) * V(ise)
                  + B(ip, 4) * V(iw) + B(ip, 5) * V(ip) + B(ip, 6) * V(ie)
                  + B(ip, 7) * V(inw) + B(ip, 8) * V(in) + B(ip, 9) * V(ine);
        }
    }

    ip = nij - ni + 1;
    ie = ip + 1;
    is = ip - ni;
    ise = is + 1;
    R(ip) = -RHS(ip) + B(ip, 2) * V(is) + B(ip, 3) * V(ise)
          + B(ip, 5) * V(ip) + B(ip, 6) * V(ie);

    for (ip = nij - ni + 2; ip <= nij - 1; ++ip) {
        iw = ip - 1;
        ie = ip + 1;
        is = ip - ni;
        isw = is - 1;
        ise = is + 1;
        R(ip) = -RHS(ip) + B(ip, 1) * V(isw) + B(ip, 2) * V(is) + B(ip, 3) * V(ise)
              + B(ip, 4) * V(iw) + B(ip, 5) * V(ip) + B(ip, 6) * V(ie);
    }

    ip = nij;
    iw = ip - 1;
    is = ip - ni;
    isw = is - 1;
    R(ip) = -RHS(ip) + B(ip, 1) * V(isw) + B(ip, 2) * V(is)
          + B(ip, 4) * V(iw) + B(ip, 5) * V(ip);
}

void orc_residual9(const double *b, int ni, int nj, const double *var,
                   const double *rhs, double *r)
{
    residual9(b, ni, ni * nj, var, rhs, r);
}

/* Fortran dot_product(r,r): a left-to-right sum. */
static double l2(const double *r, int n)
{
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += r[k] * r[k];
    return sqrt(s);
}

/* src/solvers/msi2d9-v03.f90:42-262 */
void orc_fb2d9(const double *b, const double *dl, const double *du, int nj,
               int ni, int nij, double *var, double gamma, int nitm,
               const double *rhs, int *nit_out, double *r0_out, double *hist,
               int hist_cap, int max_iter_cap)
{
    int ip, nit;
    double gammarp, rn;
    double *r = (double *)malloc(sizeof(double) * (size_t)nij);
    double *w = (double *)malloc(sizeof(double) * (size_t)nij);
    double *z = (double *)malloc(sizeof(double) * (size_t)nij);
    (void)nj;

    residual9(b, ni, nij, var, rhs, r);
    rn = l2(r, nij);
    gammarp = gamma * rn;                                   /* :133 */
    if (r0_out) *r0_out = rn;

    nit = 0;
    for (;;) {
        nit = nit + 1;
        /* Lw = -r, :139-156 */
        ip = 1;
        W(ip) = -R(ip) / DL(ip, 1);
        for (ip = 2; ip <= ni - 1; ++ip)
            W(ip) = -(R(ip) + DL(ip, 2) * W(ip - 1)) / DL(ip, 1);
        ip = ni;
        W(ip) = -(R(ip) + DL(ip, 2) * W(ip - 1) + DL(ip, 3) * W(ip - ni + 1)) / DL(ip, 1);
        ip = ni + 1;
        W(ip) = -(R(ip) + DL(ip, 2) * W(ip - 1) + DL(ip, 3) * W(ip - ni + 1)
                  + DL(ip, 4) * W(ip - ni)) / DL(ip, 1);
        for (ip = ni + 2; ip <= nij; ++ip)
            W(ip) = -(R(ip) + DL(ip, 2) * W(ip - 1) + DL(ip, 3) * W(ip - ni + 1)
                      + DL(ip, 4) * W(ip - ni) + DL(ip, 5) * W(ip - ni - 1)) / DL(ip, 1);

        /* Uz = w, :158-176 */
        ip = nij;
        Z(ip) = W(ip);
        for (ip = nij - 1; ip >= nij - ni + 2; --ip)
            Z(ip) = W(ip) - DU(ip, 1) * Z(ip + 1);
        ip = nij - ni + 1;
        Z(ip) = W(ip) - DU(ip, 1) * Z(ip + 1) - DU(ip, 2) * Z(ip + ni - 1);
        ip = nij - ni;
        Z(ip) = W(ip) - DU(ip, 1) * Z(ip + 1) - DU(ip, 2) * Z(ip + ni - 1)
              - DU(ip, 3) * Z(ip + ni);
        for (ip = nij - ni - 1; ip >= 1; --ip)
            Z(ip) = W(ip) - DU(ip, 1) * Z(ip + 1) - DU(ip, 2) * Z(ip + ni - 1)
                  - DU(ip, 3) * Z(ip + ni) - DU(ip, 4) * Z(ip + ni + 1);

        /* var = z + var, :181 */
        for (ip = 1; ip <= nij; ++ip) V(ip) = Z(ip) + V(ip);

        residual9(b, ni, nij, var, rhs, r);                 /* :183-255 */
        rn = l2(r, nij);
        if (hist && nit <= hist_cap) hist[nit - 1] = rn;
        if (rn < gammarp) break;                            /* :258 */
        if (nit == nitm) break;                             /* :259 */
        if (max_iter_cap > 0 && nit >= max_iter_cap) break;
    }
    if (nit_out) *nit_out = nit;
    free(r);
    free(w);
    free(z);
}

/* src/solvers/msi2d9-v03.f90:277-432 */
void orc_lu2d9(const double *b, double *dl, double *du, int nij, int ni,
               int nj, double alpha0)
{
    int isw, is, ise, iw, ip;
    double alpha, phi1, phi2, phi3, phi4;
    (void)nj;

    memset(du, 0, sizeof(double) * 4 * (size_t)nij);       /* :290-291 */
    memset(dl, 0, sizeof(double) * 5 * (size_t)nij);

    /* first point, :295-301 */
    ip = 1;
    DL(ip, 1) = B(ip, 5);
    DU(ip, 1) = B(ip, 6) / DL(ip, 1);
    DU(ip, 2) = B(ip, 7) / DL(ip, 1);
    DU(ip, 3) = B(ip, 8) / DL(ip, 1);
    DU(ip, 4) = B(ip, 9) / DL(ip, 1);

    /* rest of the first grid row, :304-312 */
    for (ip = 2; ip <= ni - 1; ++ip) {
        iw = ip - 1;
        DL(ip, 2) = B(ip, 4);
        DL(ip, 1) = B(ip, 5) - DL(ip, 2) * DU(iw, 1);
        DU(ip, 1) = B(ip, 6) / DL(ip, 1);
        DU(ip, 2) = (B(ip, 7) - DL(ip, 2) * DU(iw, 3)) / DL(ip, 1);
        DU(ip, 3) = (B(ip, 8) - DL(ip, 2) * DU(iw, 4)) / DL(ip, 1);
        DU(ip, 4) = B(ip, 9) / DL(ip, 1);
    }

    /* ip = ni, :315-326 */
    ip = ni;
    iw = ip - 1;
    is = ip - ni;
    ise = is + 1;
    DL(ip, 3) = B(ip, 3);
    DL(ip, 2) = B(ip, 4);
    DL(ip, 1) = B(ip, 5) - DL(ip, 2) * DU(iw, 1) - DL(ip, 3) * DU(ise, 2);
    DU(ip, 1) = (B(ip, 6) - DL(ip, 3) * DU(ise, 3)) / DL(ip, 1);
    DU(ip, 2) = (B(ip, 7) - DL(ip, 2) * DU(iw, 3)) / DL(ip, 1);
    DU(ip, 3) = (B(ip, 8) - DL(ip, 2) * DU(iw, 4)) / DL(ip, 1);
    DU(ip, 4) = B(ip, 9) / DL(ip, 1);

    /* ip = ni+1, :329-341 */
    ip = ni + 1;
    iw = ip - 1;
    is = ip - ni;
    ise = is + 1;
    DL(ip, 4) = B(ip, 2);
    DL(ip, 3) = B(ip, 3) - DL(ip, 4) * DU(is, 1);
    DL(ip, 2) = B(ip, 4) - DL(ip, 4) * DU(is, 2);
    DL(ip, 1) = B(ip, 5) - DL(ip, 2) * DU(iw, 1) - DL(ip, 3) * DU(ise, 2)
              - DL(ip, 4) * DU(is, 3);
    DU(ip, 1) = (B(ip, 6) - DL(ip, 3) * DU(ise, 3) - DL(ip, 4) * DU(is, 4)) / DL(ip, 1);
    DU(ip, 2) = (B(ip, 7) - DL(ip, 2) * DU(iw, 3)) / DL(ip, 1);
    DU(ip, 3) = (B(ip, 8) - DL(ip, 2) * DU(iw, 4)) / DL(ip, 1);
    DU(ip, 4) = B(ip, 9) / DL(ip, 1);

    /* interior with partial cancellation, :344-370 */
    for (ip = ni + 2; ip <= nij - ni - 1; ++ip) {
        if (ip % ni == 0 || ip % ni == 1)
            alpha = 0.0;
        else
            alpha = alpha0;
        iw = ip - 1;
        is = ip - ni;
        isw = is - 1;
        ise = is + 1;
        DL(ip, 5) = B(ip, 1);
        DL(ip, 4) = (B(ip, 2) - DL(ip, 5) * DU(isw, 1) - alpha * B(ip, 3) * DU(ise, 1))
                  / (1.0 - alpha * DU(is, 1) * DU(ise, 1));
        DL(ip, 3) = B(ip, 3) - DL(ip, 4) * DU(is, 1);
        DL(ip, 2) = (B(ip, 4) - DL(ip, 4) * DU(is, 2) - DL(ip, 5) * DU(isw, 3)
                     - 2.0 * alpha * DL(ip, 5) * DU(isw, 2))
                  / (1 + 2.0 * alpha * DU(iw, 2));
        phi1 = DL(ip, 3) * DU(ise, 1);
        phi2 = DL(ip, 5) * DU(isw, 2);
        phi3 = DL(ip, 3) * DU(ise, 4);
        phi4 = DL(ip, 2) * DU(iw, 2);
        DL(ip, 1) = B(ip, 5) - DL(ip, 2) * DU(iw, 1) - DL(ip, 3) * DU(ise, 2)
                  - DL(ip, 4) * DU(is, 3) - DL(ip, 5) * DU(isw, 4)
                  + alpha * (2.0 * phi1 + phi2 + phi3 + 2.0 * phi4);
        DU(ip, 1) = (B(ip, 6) - DL(ip, 3) * DU(ise, 3) - DL(ip, 4) * DU(is, 4)
                     - 2.0 * alpha * (phi1 + phi3)) / DL(ip, 1);
        DU(ip, 2) = (B(ip, 7) - DL(ip, 2) * DU(iw, 3)) / DL(ip, 1);
        DU(ip, 3) = (B(ip, 8) - DL(ip, 2) * DU(iw, 4) - alpha * phi4) / DL(ip, 1);
        DU(ip, 4) = B(ip, 9) / DL(ip, 1);
    }

    /* ip = nij-ni, :373-386 */
    ip = nij - ni;
    iw = ip - 1;
    is = ip - ni;
    isw = is - 1;
    ise = is + 1;
    DL(ip, 5) = B(ip, 1);
    DL(ip, 4) = B(ip, 2) - DL(ip, 5) * DU(isw, 1);
    DL(ip, 3) = B(ip, 3) - DL(ip, 4) * DU(is, 1);
    DL(ip, 2) = B(ip, 4) - DL(ip, 4) * DU(is, 2) - DL(ip, 5) * DU(isw, 3);
    DL(ip, 1) = B(ip, 5) - DL(ip, 2) * DU(iw, 1) - DL(ip, 3) * DU(ise, 2)
              - DL(ip, 4) * DU(is, 3) - DL(ip, 5) * DU(isw, 4);
    DU(ip, 1) = (B(ip, 6) - DL(ip, 3) * DU(ise, 3) - DL(ip, 4) * DU(is, 4)) / DL(ip, 1);
    DU(ip, 2) = (B(ip, 7) - DL(ip, 2) * DU(iw, 3)) / DL(ip, 1);
    DU(ip, 3) = (B(ip, 8) - DL(ip, 2) * DU(iw, 4)) / DL(ip, 1);

    /* ip = nij-ni+1, :389-401 */
    ip = nij - ni + 1;
    iw = ip - 1;
    is = ip - ni;
    isw = is - 1;
    ise = is + 1;
    DL(ip, 5) = B(ip, 1);
    DL(ip, 4) = B(ip, 2) - DL(ip, 5) * DU(isw, 1);
    DL(ip, 3) = B(ip, 3) - DL(ip, 4) * DU(is, 1);
    DL(ip, 2) = B(ip, 4) - DL(ip, 4) * DU(is, 2) - DL(ip, 5) * DU(isw, 3);
    DL(ip, 1) = B(ip, 5) - DL(ip, 2) * DU(iw, 1) - DL(ip, 3) * DU(ise, 2)
              - DL(ip, 4) * DU(is, 3) - DL(ip, 5) * DU(isw, 4);
    DU(ip, 1) = (B(ip, 6) - DL(ip, 3) * DU(ise, 3) - DL(ip, 4) * DU(is, 4)) / DL(ip, 1);
    DU(ip, 2) = (B(ip, 7) - DL(ip, 2) * DU(iw, 3)) / DL(ip, 1);

    /* last grid row, :404-416 */
    for (ip = nij - ni + 2; ip <= nij - 1; ++ip) {
        iw = ip - 1;
        is = ip - ni;
        isw = is - 1;
        ise = is + 1;
        DL(ip, 5) = B(ip, 1);
        DL(ip, 4) = B(ip, 2) - DL(ip, 5) * DU(isw, 1);
        DL(ip, 3) = B(ip, 3) - DL(ip, 4) * DU(is, 1);
        DL(ip, 2) = B(ip, 4) - DL(ip, 4) * DU(is, 2) - DL(ip, 5) * DU(isw, 3);
        DL(ip, 1) = B(ip, 5) - DL(ip, 2) * DU(iw, 1) - DL(ip, 3) * DU(ise, 2)
                  - DL(ip, 4) * DU(is, 3) - DL(ip, 5) * DU(isw, 4);
        DU(ip, 1) = (B(ip, 6) - DL(ip, 3) * DU(ise, 3) - DL(ip, 4) * DU(is, 4)) / DL(ip, 1);
    }

    /* last point, :419-429 */
    ip = nij;
    iw = ip - 1;
    is = ip - ni;
    isw = is - 1;
    ise = is + 1;
    DL(ip, 5) = B(ip, 1);
    DL(ip, 4) = B(ip, 2) - DL(ip, 5) * DU(isw, 1);
    DL(ip, 3) = B(ip, 3) - DL(ip, 4) * DU(is, 1);
    DL(ip, 2) = B(ip, 4) - DL(ip, 4) * DU(is, 2) - DL(ip, 5) * DU(isw, 3);
    DL(ip, 1) = B(ip, 5) - DL(ip, 2) * DU(iw, 1) - DL(ip, 3) * DU(ise, 2)
              - DL(ip, 4) * DU(is, 3) - DL(ip, 5) * DU(isw, 4);
}
