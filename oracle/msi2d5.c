/*
 * msi2d5.c -- oracle restatement of the 5-diagonal MSI solver
 * (reference: src/solvers/msi2d5-v03.f90), the solver the reference uses for
 * its pressure-correction system.  Used to pin the zero-corner embedding:
 * lu2d9/fb2d9 on a 5-point system placed in slots 2,4,5,6,8 give the same
 * iterates as lu2d5/fb2d5.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 * Operation order follows the reference (build with -ffp-contract=off).
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

/* src/solvers/msi2d5-v03.f90:59-107 (and :158-207 in the loop) */
static void residual5(const double *b, int ni, int nij, const double *var,
                      const double *rhs, double *r)
{
    int ip, iw, ie, is, in;

    ip = 1;
    ie = ip + 1;
    in = ip + ni;
    R(ip) = -RHS(ip) + B(ip, 3) * V(ip) + B(ip, 4) * V(ie) + B(ip, 5) * V(in);

    for (ip = 2; ip <= ni - 1; ++ip) {
        iw = ip - 1;
        ie = ip + 1;
        in = ip + ni;
        R(ip) = -RHS(ip) + B(ip, 2) * V(iw) + B(ip, 3) * V(ip) + B(ip, 4) * V(ie)
              + B(ip, 5) * V(in);
    }

    ip = ni;
    iw = ip - 1;
    in = ip + ni;
    R(ip) = -RHS(ip) + B(ip, 2) * V(iw) + B(ip, 3) * V(ip) + B(ip, 5) * V(in);

    for (ip = ni + 1; ip <= nij - ni; ++ip) {
        iw = ip - 1;
        ie = ip + 1;
        is = ip - ni;
        in = ip + ni;
        if (ip % ni == 0) {
            R(ip) = -RHS(ip) + B(ip, 1) * V(is) + B(ip, 2) * V(iw) + B(ip, 3) * V(ip)
                  + B(ip, 5) * V(in);
        } else if (ip % ni == 1) {
            R(ip) = -RHS(ip) + B(ip, 1) * V(is) + B(ip, 3) * V(ip) + B(ip, 4) * V(ie)
                  + B(ip, 5) * V(in);
        } else {
            R(ip) = -RHS(ip) + B(ip, 1) * V(is) + B(ip, 2) * V(iw) + B(ip, 3) * V(ip)
                  + B(ip, 4) * V(ie) + B(ip, 5) * V(in);
        }
    }

    ip = nij - ni + 1;
    ie = ip + 1;
    is = ip - ni;
    R(ip) = -RHS(ip) + B(ip, 1) * V(is) + B(ip, 3) * V(ip) + B(ip, 4) * V(ie);

    for (ip = nij - ni + 2; ip <= nij - 1; ++ip) {
        iw = ip - 1;
        ie = ip + 1;
        is = ip - ni;
        R(ip) = -RHS(ip) + B(ip, 1) * V(is) + B(ip, 2) * V(iw) + B(ip, 3) * V(ip)
              + B(ip, 4) * V(ie);
    }

    ip = nij;
    iw = ip - 1;
    is = ip - ni;
    R(ip) = -RHS(ip) + B(ip, 1) * V(is) + B(ip, 2) * V(iw) + B(ip, 3) * V(ip);
}

static double l2(const double *r, int n)
{
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += r[k] * r[k];
    return sqrt(s);
}

/* src/solvers/msi2d5-v03.f90:42-233 */
void orc_fb2d5(const double *b, const double *dl, const double *du, int nj,
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

    residual5(b, ni, nij, var, rhs, r);
    rn = l2(r, nij);
    gammarp = gamma * rn;
    if (r0_out) *r0_out = rn;

    nit = 0;
    for (;;) {
        nit = nit + 1;
        ip = 1;
        W(ip) = -R(ip) / DL(ip, 1);
        for (ip = 2; ip <= ni - 1; ++ip)
            W(ip) = -(R(ip) + DL(ip, 2) * W(ip - 1)) / DL(ip, 1);
        ip = ni;
        W(ip) = -(R(ip) + DL(ip, 2) * W(ip - 1) + DL(ip, 3) * W(ip - ni + 1)) / DL(ip, 1);
        for (ip = ni + 1; ip <= nij; ++ip)
            W(ip) = -(R(ip) + DL(ip, 2) * W(ip - 1) + DL(ip, 3) * W(ip - ni + 1)
                      + DL(ip, 4) * W(ip - ni)) / DL(ip, 1);

        ip = nij;
        Z(ip) = W(ip);
        for (ip = nij - 1; ip >= nij - ni + 2; --ip)
            Z(ip) = W(ip) - DU(ip, 1) * Z(ip + 1);
        ip = nij - ni + 1;
        Z(ip) = W(ip) - DU(ip, 1) * Z(ip + 1) - DU(ip, 2) * Z(ip + ni - 1);
        for (ip = nij - ni; ip >= 1; --ip)
            Z(ip) = W(ip) - DU(ip, 1) * Z(ip + 1) - DU(ip, 2) * Z(ip + ni - 1)
                  - DU(ip, 3) * Z(ip + ni);

        for (ip = 1; ip <= nij; ++ip) V(ip) = Z(ip) + V(ip);

        residual5(b, ni, nij, var, rhs, r);
        rn = l2(r, nij);
        if (hist && nit <= hist_cap) hist[nit - 1] = rn;
        if (rn < gammarp) break;
        if (nit == nitm) break;
        if (max_iter_cap > 0 && nit >= max_iter_cap) break;
    }
    if (nit_out) *nit_out = nit;
    free(r);
    free(w);
    free(z);
}

/* src/solvers/msi2d5-v03.f90:247-330 */
void orc_lu2d5(const double *b, double *dl, double *du, int nij, int ni,
               int nj, double alpha0)
{
    int is, ise, iw, ip;
    double alpha, phi1, phi4;
    (void)nj;

    memset(du, 0, sizeof(double) * 3 * (size_t)nij);
    memset(dl, 0, sizeof(double) * 4 * (size_t)nij);

    ip = 1;
    DL(ip, 1) = B(ip, 3);
    DU(ip, 1) = B(ip, 4) / DL(ip, 1);
    DU(ip, 3) = B(ip, 5) / DL(ip, 1);

    for (ip = 2; ip <= ni; ++ip) {
        iw = ip - 1;
        DL(ip, 2) = B(ip, 2);
        DL(ip, 1) = B(ip, 3) - DL(ip, 2) * DU(iw, 1);
        DU(ip, 1) = B(ip, 4) / DL(ip, 1);
        DU(ip, 2) = -DL(ip, 2) * DU(iw, 3) / DL(ip, 1);
        DU(ip, 3) = B(ip, 5) / DL(ip, 1);
    }

    for (ip = ni + 1; ip <= nij - ni; ++ip) {
        iw = ip - 1;
        is = ip - ni;
        ise = is + 1;
        if (ip % ni == 0 || ip % ni == 1)
            alpha = 0.0;
        else
            alpha = alpha0;
        DL(ip, 4) = B(ip, 1) / (1.0 - alpha * DU(is, 1) * DU(ise, 1));
        DL(ip, 3) = -DL(ip, 4) * DU(is, 1);
        DL(ip, 2) = (B(ip, 2) - DL(ip, 4) * DU(is, 2)) / (1.0 + 2.0 * alpha * DU(iw, 2));
        phi1 = DL(ip, 3) * DU(ise, 1);
        phi4 = DL(ip, 2) * DU(iw, 2);
        DL(ip, 1) = B(ip, 3) - DL(ip, 2) * DU(iw, 1) - DL(ip, 3) * DU(ise, 2)
                  - DL(ip, 4) * DU(is, 3) + 2.0 * alpha * (phi1 + phi4);
        DU(ip, 1) = (B(ip, 4) - DL(ip, 3) * DU(ise, 3) - 2.0 * alpha * phi1) / DL(ip, 1);
        DU(ip, 2) = -DL(ip, 2) * DU(iw, 3) / DL(ip, 1);
        DU(ip, 3) = (B(ip, 5) - alpha * phi4) / DL(ip, 1);
    }

    ip = nij - ni + 1;
    iw = ip - 1;
    is = ip - ni;
    ise = is + 1;
    DL(ip, 4) = B(ip, 1);
    DL(ip, 3) = -DL(ip, 4) * DU(is, 1);
    DL(ip, 2) = (B(ip, 2) - DL(ip, 4) * DU(is, 2));
    DL(ip, 1) = B(ip, 3) - DL(ip, 2) * DU(iw, 1) - DL(ip, 3) * DU(ise, 2)
              - DL(ip, 4) * DU(is, 3);
    DU(ip, 1) = (B(ip, 4) - DL(ip, 3) * DU(ise, 3)) / DL(ip, 1);
    DU(ip, 2) = -DL(ip, 2) * DU(iw, 3) / DL(ip, 1);

    for (ip = nij - ni + 2; ip <= nij - 1; ++ip) {
        iw = ip - 1;
        is = ip - ni;
        ise = is + 1;
        DL(ip, 4) = B(ip, 1);
        DL(ip, 3) = -DL(ip, 4) * DU(is, 1);
        DL(ip, 2) = B(ip, 2) - DL(ip, 4) * DU(is, 2);
        DL(ip, 1) = B(ip, 3) - DL(ip, 2) * DU(iw, 1) - DL(ip, 3) * DU(ise, 2)
                  - DL(ip, 4) * DU(is, 3);
        DU(ip, 1) = (B(ip, 4) - DL(ip, 3) * DU(ise, 3)) / DL(ip, 1);
    }

    ip = nij;
    iw = ip - 1;
    is = ip - ni;
    ise = is + 1;
    DL(ip, 4) = B(ip, 1);
    DL(ip, 3) = -DL(ip, 4) * DU(is, 1);
    DL(ip, 2) = B(ip, 2) - DL(ip, 4) * DU(is, 2);
    DL(ip, 1) = B(ip, 3) - DL(ip, 2) * DU(iw, 1) - DL(ip, 3) * DU(ise, 2)
              - DL(ip, 4) * DU(is, 3);
}
