// assembly.hip -- coefficient assembly of the SIMPLEC equations on the GPU:
// the 9-point u, v, T and 5-point p' coefficients and sources of
// src/equations/mach2d_coef.f90 (module `coefficients`), plus the face
// densities.  One thread per point; every formula keeps the reference's
// operation order and is compiled without contraction, so the results equal
// the reference bit for bit.
//
// Memory: all inputs are N-vectors in the reference layout (flat index
// np = nx*(j-1) + i); a point reads its own value and the W/E/S/N (and, for
// the viscous terms, diagonal) neighbours of a few fields, so a 256-wide
// row tile reads each field once from HBM with the neighbours from the same
// or adjacent lines.  The kernels are HBM-bound; the algorithmic traffic of
// each is (inputs + outputs) x 8 bytes per point.
//
// Entry points take the reference's argument lists, every argument by
// pointer (a Fortran bind(C) interface calls them unchanged).  The arrays may
// be host or device memory (all of one kind): device arrays are processed in
// place on the assembly stream (m2d_assembly_set_stream) and the call
// returns once the kernel is enqueued; host arrays are copied to the device
// and back, and the call returns when the results are in place.  As in the
// reference, only the points the routine defines are written.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "m2d_internal.h"
#include "msi_stencil.h"
#include "seqsum.h"

namespace m2d {

constexpr int kTx = 256;  // points per workgroup along i
constexpr int kMaxIn = 48;

struct Fields {
    const double *in[kMaxIn];
    double *out[6];
    int nx, ny;
    int modvis, coord;
    double dt, beta;
    double twall;
    // set by a loop that is enqueued ahead of its exit test (the pressure-
    // correction loop): when gate is non-null and *gate != 0 the kernel
    // returns at once; when sc is non-null, dt = sc[0] and beta = sc[1] are
    // read from device memory, so one captured loop serves every dt
    const int *gate;
    const double *sc;
};

__device__ __forceinline__ bool gated(const Fields &F) { return F.gate && *F.gate; }
__device__ __forceinline__ double field_dt(const Fields &F) { return F.sc ? F.sc[0] : F.dt; }
__device__ __forceinline__ double field_beta(const Fields &F) { return F.sc ? F.sc[1] : F.beta; }

__device__ __forceinline__ double dsign5(double x) { return copysign(0.5, x); }

// Advection and diffusion part of get_u_coefficients / get_v_coefficients
// (mach2d_coef.f90:100-197, :491-598).
struct UV {
    const double *rp, *re, *rn, *Jp, *Je, *Jn, *alphae, *betae, *betan, *gamman, *vle, *vln,
        *roe, *ron, *roa, *Uce, *Vcn;
};

__device__ __forceinline__ void uv_point(const UV &f, int modvis, double dt, double *a, long N,
                                         long np, long nx)
{
    const long nps = np - nx, npw = np - 1;
    const double fme = f.roe[np] * f.re[np] * f.Uce[np];
    const double fmw = f.roe[npw] * f.re[npw] * f.Uce[npw];
    const double fmn = f.ron[np] * f.rn[np] * f.Vcn[np];
    const double fms = f.ron[nps] * f.rn[nps] * f.Vcn[nps];
    const double mpa = f.roa[np] * f.rp[np] / f.Jp[np];
    const double as = dsign5(f.Vcn[nps]), an = dsign5(f.Vcn[np]);
    const double aw = dsign5(f.Uce[npw]), ae = dsign5(f.Uce[np]);
    double c[9];
    if (modvis == 0) {
        c[0] = c[2] = c[6] = c[8] = 0.0;
        c[1] = -fms * (0.5 + as);
        c[3] = -fmw * (0.5 + aw);
        c[5] = fme * (0.5 - ae);
        c[7] = fmn * (0.5 - an);
    } else {
        const double ew = f.vle[npw] * f.re[npw] * f.Je[npw] * f.betae[npw];
        const double ep = f.vle[np] * f.re[np] * f.Je[np] * f.betae[np];
        const double ns = f.vln[nps] * f.rn[nps] * f.Jn[nps] * f.betan[nps];
        const double nq = f.vln[np] * f.rn[np] * f.Jn[np] * f.betan[np];
        c[0] = (ew + ns) / 4.0;
        c[1] = -(fms * (0.5 + as)) - f.vln[nps] * f.rn[nps] * f.Jn[nps] * f.gamman[nps] +
               (ew - ep) / 4.0;
        c[2] = -(ep + ns) / 4.0;
        c[3] = -(fmw * (0.5 + aw)) - f.vle[npw] * f.re[npw] * f.Je[npw] * f.alphae[npw] +
               (ns - nq) / 4.0;
        c[5] = fme * (0.5 - ae) - f.vle[np] * f.re[np] * f.Je[np] * f.alphae[np] +
               (-ns + nq) / 4.0;
        c[6] = (-ew - nq) / 4.0;
        c[7] = fmn * (0.5 - an) - f.vln[np] * f.rn[np] * f.Jn[np] * f.gamman[np] -
               (ew - ep) / 4.0;
        c[8] = (ep + nq) / 4.0;
    }
    c[4] = mpa / dt - (c[7] + c[1] + c[3] + c[5]);
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k * N + np] = c[k];
}

// Interior point (i, j) of this thread (1-based, 2..nx-1 x 2..ny-1); false
// outside.
__device__ __forceinline__ bool interior(int nx, int ny, long &np)
{
    const int i = blockIdx.x * kTx + threadIdx.x + 2;
    const int j = blockIdx.y + 2;
    if (i > nx - 1 || j > ny - 1) return false;
    np = (long)nx * (j - 1) + i - 1;
    return true;
}

// get_density_at_nodes: mach2d_coef.f90:6-17 (ro = p / (Rg T), every
// point; F.beta carries Rg).  Grid over all N.
__global__ __launch_bounds__(kTx) void k_density_at_nodes(Fields F)
{
    const long q = (long)(blockIdx.y * gridDim.x + blockIdx.x) * kTx + threadIdx.x;
    if (q >= (long)F.nx * F.ny) return;
    F.out[0][q] = F.in[0][q] / (F.beta * F.in[1][q]);
}

// get_density_at_faces: mach2d_coef.f90:20-60 (east faces j = 2..ny-1,
// i = 1..nx-1; north faces j = 1..ny-1, i = 2..nx-1).  Grid over
// i = 1..nx-1, j = 1..ny-1.
__global__ __launch_bounds__(kTx) void k_density_at_faces(Fields F)
{
    if (gated(F)) return;
    const int nx = F.nx, ny = F.ny;
    const int i = blockIdx.x * kTx + threadIdx.x + 1;
    const int j = blockIdx.y + 1;
    if (i > nx - 1 || j > ny - 1) return;
    const double *ro = F.in[0], *Uce = F.in[1], *Vcn = F.in[2];
    const double beta = field_beta(F);
    const long np = (long)nx * (j - 1) + i - 1;
    if (j >= 2) {
        const long npe = np + 1;
        const double ae = dsign5(Uce[np]);
        F.out[0][np] = (0.5 + ae) * ro[np] + (0.5 - ae) * ro[npe] + beta * ae * (ro[npe] - ro[np]);
    }
    if (i >= 2) {
        const long npn = np + nx;
        const double an = dsign5(Vcn[np]);
        F.out[1][np] = (0.5 + an) * ro[np] + (0.5 - an) * ro[npn] + beta * an * (ro[npn] - ro[np]);
    }
}

// get_u_coefficients: mach2d_coef.f90:63-218
__global__ __launch_bounds__(kTx) void k_u_coefficients(Fields F)
{
    long np;
    if (!interior(F.nx, F.ny, np)) return;
    const double *const *q = F.in;
    // rp re rn Jp Je Jn ye yk alphae betae betan gamman vle vln roe ron roa Uce Vcn
    const UV f = {q[0], q[1], q[2], q[3], q[4], q[5], q[8], q[9], q[10], q[11], q[12], q[13],
                  q[14], q[15], q[16], q[17], q[18]};
    const long N = (long)F.nx * F.ny, nps = np - F.nx, npw = np - 1;
    double *a = F.out[0];
    uv_point(f, F.modvis, F.dt, a, N, np, F.nx);
    if (F.modvis == 0) return;
    const double *rp = q[0], *Je = q[4], *Jn = q[5], *ye = q[6], *yk = q[7], *vle = q[12],
                 *vln = q[13];
    const double S1p = rp[np] * Je[np] * vle[np] * (ye[np] * ye[np]) / 3.0 +
                       rp[np] * Je[npw] * vle[npw] * (ye[npw] * ye[npw]) / 3.0;
    const double S2p = rp[np] * Jn[np] * vln[np] * (yk[np] * yk[np]) / 3.0 +
                       rp[np] * Jn[nps] * vln[nps] * (yk[nps] * yk[nps]) / 3.0;
    a[4 * N + np] = a[4 * N + np] + S1p + S2p;
}

// get_v_coefficients: mach2d_coef.f90:454-613
__global__ __launch_bounds__(kTx) void k_v_coefficients(Fields F)
{
    long np;
    if (!interior(F.nx, F.ny, np)) return;
    const double *const *q = F.in;
    // rp re rn Jp Je Jn xe xk alphae betae betan gamman vle vln vlp roe ron roa Uce Vcn
    const UV f = {q[0], q[1], q[2], q[3], q[4], q[5], q[8], q[9], q[10], q[11], q[12], q[13],
                  q[15], q[16], q[17], q[18], q[19]};
    const long N = (long)F.nx * F.ny, nps = np - F.nx, npw = np - 1;
    double *a = F.out[0];
    uv_point(f, F.modvis, F.dt, a, N, np, F.nx);
    if (F.modvis == 0) return;
    const double *rp = q[0], *re = q[1], *rn = q[2], *Jp = q[3], *Je = q[4], *Jn = q[5],
                 *xe = q[6], *xk = q[7], *vle = q[12], *vln = q[13], *vlp = q[14];
    const double S1p = vle[np] * re[np] * Je[np] * (xe[np] * xe[np]) / 3.0 +
                       vle[npw] * re[npw] * Je[npw] * (xe[npw] * xe[npw]) / 3.0;
    const double S2p = vln[np] * rn[np] * Jn[np] * (xk[np] * xk[np]) / 3.0 +
                       vln[nps] * rn[nps] * Jn[nps] * (xk[nps] * xk[nps]) / 3.0;
    const double S7p = (double)F.coord * vlp[np] * 4.0 / (3.0 * rp[np] * Jp[np]);
    a[4 * N + np] = a[4 * N + np] + S1p + S2p + S7p;
}

// get_p_coefficients: mach2d_coef.f90:1207-1278 (5 diagonals S W P E N)
__global__ __launch_bounds__(kTx) void k_p_coefficients(Fields F)
{
    if (gated(F)) return;
    long np;
    if (!interior(F.nx, F.ny, np)) return;
    const double *const *q = F.in;
    const double *rp = q[0], *re = q[1], *rn = q[2], *Jp = q[3], *Uce = q[4], *Vcn = q[5],
                 *roe = q[6], *ron = q[7], *g = q[8], *de = q[9], *dn = q[10];
    const long N = (long)F.nx * F.ny;
    const long nps = np - F.nx, npn = np + F.nx, npw = np - 1, npe = np + 1;
    const double dt = field_dt(F);
    const double as = dsign5(Vcn[nps]), an = dsign5(Vcn[np]);
    const double aw = dsign5(Uce[npw]), ae = dsign5(Uce[np]);
    const double roem = roe[np], rowm = roe[npw], ronm = ron[np], rosm = ron[nps];
    double *a = F.out[0];
    a[np] = -rn[nps] * ((0.5 + as) * Vcn[nps] * g[nps] + rosm * dn[nps]);
    a[N + np] = -re[npw] * ((0.5 + aw) * Uce[npw] * g[npw] + rowm * de[npw]);
    a[2 * N + np] = g[np] * (rp[np] / (Jp[np] * dt) + (0.5 + ae) * re[np] * Uce[np] -
                             (0.5 - aw) * re[npw] * Uce[npw] + (0.5 + an) * rn[np] * Vcn[np] -
                             (0.5 - as) * rn[nps] * Vcn[nps]) +
                    roem * re[np] * de[np] + rowm * re[npw] * de[npw] + ronm * rn[np] * dn[np] +
                    rosm * rn[nps] * dn[nps];
    a[3 * N + np] = re[np] * ((0.5 - ae) * Uce[np] * g[npe] - roem * de[np]);
    a[4 * N + np] = rn[np] * ((0.5 - an) * Vcn[np] * g[npn] - ronm * dn[np]);
}

// get_p_source: mach2d_coef.f90:1281-1348
__global__ __launch_bounds__(kTx) void k_p_source(Fields F)
{
    if (gated(F)) return;
    long np;
    if (!interior(F.nx, F.ny, np)) return;
    const double *const *q = F.in;
    const double *rp = q[0], *re = q[1], *rn = q[2], *Jp = q[3], *roe = q[4], *ron = q[5],
                 *ro = q[6], *roa = q[7], *Uce = q[8], *Vcn = q[9], *de = q[10], *dn = q[11],
                 *g = q[12], *pl = q[13];
    const long nps = np - F.nx, npn = np + F.nx, npw = np - 1, npe = np + 1;
    const double dt = field_dt(F);
    const double as = dsign5(Vcn[nps]), an = dsign5(Vcn[np]);
    const double aw = dsign5(Uce[npw]), ae = dsign5(Uce[np]);
    const double roel = g[np] * pl[np] * (0.5 + ae) + g[npe] * pl[npe] * (0.5 - ae);
    const double rowl = g[npw] * pl[npw] * (0.5 + aw) + g[np] * pl[np] * (0.5 - aw);
    const double ronl = g[np] * pl[np] * (0.5 + an) + g[npn] * pl[npn] * (0.5 - an);
    const double rosl = g[nps] * pl[nps] * (0.5 + as) + g[np] * pl[np] * (0.5 - as);
    const double Ucel = de[np] * (pl[np] - pl[npe]);
    const double Ucwl = de[npw] * (pl[npw] - pl[np]);
    const double Vcnl = dn[np] * (pl[np] - pl[npn]);
    const double Vcsl = dn[nps] * (pl[nps] - pl[np]);
    F.out[0][np] = -rp[np] * (ro[np] - roa[np]) / (Jp[np] * dt) - roe[np] * re[np] * Uce[np] +
                   roe[npw] * re[npw] * Uce[npw] - ron[np] * rn[np] * Vcn[np] +
                   ron[nps] * rn[nps] * Vcn[nps] - roel * re[np] * Ucel +
                   rowl * re[npw] * Ucwl - ronl * rn[np] * Vcnl + rosl * rn[nps] * Vcsl;
}

// get_u_source: mach2d_coef.f90:221-451 (outputs cup, sup, b)
__global__ __launch_bounds__(kTx) void k_u_source(Fields F)
{
    long np;
    if (!interior(F.nx, F.ny, np)) return;
    const double *const *q = F.in;
    const double *rp = q[0], *re = q[1], *rn = q[2], *xe = q[3], *ye = q[4], *xk = q[5],
                 *yk = q[6], *xke = q[7], *yke = q[8], *xen = q[9], *yen = q[10], *Jp = q[11],
                 *Je = q[12], *Jn = q[13], *roe = q[14], *ron = q[15], *roa = q[16], *p = q[17],
                 *vle = q[18], *vln = q[19], *Uce = q[20], *Vcn = q[21], *ua = q[22], *u = q[23],
                 *v = q[24];
    double *cup = F.out[0], *sup = F.out[1], *b = F.out[2];
    const long nx = F.nx;
    const long nps = np - nx, npn = np + nx, npw = np - 1, npe = np + 1;
    const double fme = roe[np] * re[np] * Uce[np];
    const double fmw = roe[npw] * re[npw] * Uce[npw];
    const double fmn = ron[np] * rn[np] * Vcn[np];
    const double fms = ron[nps] * rn[nps] * Vcn[nps];
    const double mpa = roa[np] * rp[np] / Jp[np];
    const double as = dsign5(Vcn[nps]), an = dsign5(Vcn[np]);
    const double aw = dsign5(Uce[npw]), ae = dsign5(Uce[np]);
    double bb = mpa * ua[np] / F.dt;
    const double c = -(F.beta * (fme * ae * (u[npe] - u[np]) - fmw * aw * (u[np] - u[npw]) +
                                 fmn * an * (u[npn] - u[np]) - fms * as * (u[np] - u[nps])));
    cup[np] = c;
    bb = bb + c;
    bb = bb + 0.5 * rp[np] *
                  (yk[np] * (p[npn] + p[np]) - yk[nps] * (p[nps] + p[np]) -
                   ye[np] * (p[npe] + p[np]) + ye[npw] * (p[npw] + p[np]));
    if (F.modvis == 1) {
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
        const double S3 = Je[np] * re[np] * vle[np] * xe[np] *
                              (yke[np] * (v[npn] + v[npne] - v[nps] - v[npse]) / 4.0 -
                               ye[np] * (v[npe] - v[np])) -
                          Je[npw] * re[npw] * vle[npw] * xe[npw] *
                              (yke[npw] * (v[npn] + v[npnw] - v[nps] - v[npsw]) / 4.0 -
                               ye[npw] * (v[np] - v[npw]));
        const double S4 = Jn[np] * rn[np] * vln[np] * xk[np] *
                              (yen[np] * (v[npne] + v[npe] - v[npnw] - v[npw]) / 4.0 -
                               yk[np] * (v[npn] - v[np])) -
                          Jn[nps] * rn[nps] * vln[nps] * xk[nps] *
                              (yen[nps] * (v[npse] + v[npe] - v[npsw] - v[npw]) / 4.0 -
                               yk[nps] * (v[np] - v[nps]));
        // face terms of S5 / S6; a face on the axis (r = 0) drops its term
        // (mach2d_coef.f90:378-441).  Both terms are evaluated and the
        // branch selects, which equals the reference also when a dropped
        // term would divide by zero.
        double S5, S6;
        const bool axw = fabs(re[npw]) < 1e-15, axe = fabs(re[np]) < 1e-15;
        const double X1 = 2.0 * rp[np] / 3.0 * Je[np] * vle[np] / re[np] * ye[np] *
                                    (xke[np] *
                                         (rp[npn] * v[npn] + rp[npne] * v[npne] -
                                          rp[nps] * v[nps] - rp[npse] * v[npse]) / 4.0 -
                                     xe[np] * (rp[npe] * v[npe] - rp[np] * v[np]));
        const double X2 = 2.0 * rp[np] / 3.0 * Je[npw] * vle[npw] / re[npw] * ye[npw] *
                                    (xke[npw] *
                                         (rp[npn] * v[npn] + rp[npnw] * v[npnw] -
                                          rp[nps] * v[nps] - rp[npsw] * v[npsw]) / 4.0 -
                                     xe[npw] * (rp[np] * v[np] - rp[npw] * v[npw]));
        if (axw)
            S5 = -X1;
        else if (axe)
            S5 = X2;
        else
            S5 = -X1 + X2;
        const bool axs = fabs(rn[nps]) < 1e-15, axn = fabs(rn[np]) < 1e-15;
        const double Y1 = 2.0 * rp[np] / 3.0 * Jn[np] * vln[np] / rn[np] * yk[np] *
                                    (xen[np] *
                                         (rp[npne] * v[npne] + rp[npe] * v[npe] -
                                          rp[npnw] * v[npnw] - rp[npw] * v[npw]) / 4.0 -
                                     xk[np] * (rp[npn] * v[npn] - rp[np] * v[np]));
        const double Y2 = 2.0 * rp[np] / 3.0 * Jn[nps] * vln[nps] / rn[nps] * yk[nps] *
                                    (xen[nps] *
                                         (rp[npe] * v[npe] + rp[npse] * v[npse] -
                                          rp[npw] * v[npw] - rp[npsw] * v[npsw]) / 4.0 -
                                     xk[nps] * (rp[np] * v[np] - rp[nps] * v[nps]));
        if (axs)
            S6 = -Y1;
        else if (axn)
            S6 = Y2;
        else
            S6 = -Y1 + Y2;
        const double s = S1 + S2 + S3 + S4 + S5 + S6;
        sup[np] = s;
        bb = bb + s;
    }
    b[np] = bb;
}

// get_v_source: mach2d_coef.f90:616-783 (outputs cvp, svp, b)
__global__ __launch_bounds__(kTx) void k_v_source(Fields F)
{
    long np;
    if (!interior(F.nx, F.ny, np)) return;
    const double *const *q = F.in;
    const double *rp = q[0], *re = q[1], *rn = q[2], *xe = q[3], *ye = q[4], *xk = q[5],
                 *yk = q[6], *xke = q[7], *yke = q[8], *xen = q[9], *yen = q[10], *Jp = q[11],
                 *Je = q[12], *Jn = q[13], *roe = q[14], *ron = q[15], *roa = q[16], *p = q[17],
                 *vle = q[18], *vln = q[19], *Uce = q[20], *Vcn = q[21], *va = q[22], *u = q[23],
                 *v = q[24];
    double *cvp = F.out[0], *svp = F.out[1], *b = F.out[2];
    const long nx = F.nx;
    const long nps = np - nx, npn = np + nx, npw = np - 1, npe = np + 1;
    const double fme = roe[np] * re[np] * Uce[np];
    const double fmw = roe[npw] * re[npw] * Uce[npw];
    const double fmn = ron[np] * rn[np] * Vcn[np];
    const double fms = ron[nps] * rn[nps] * Vcn[nps];
    const double mpa = roa[np] * rp[np] / Jp[np];
    const double as = dsign5(Vcn[nps]), an = dsign5(Vcn[np]);
    const double aw = dsign5(Uce[npw]), ae = dsign5(Uce[np]);
    double bb = mpa * va[np] / F.dt;
    const double c = -(F.beta * (fme * ae * (v[npe] - v[np]) - fmw * aw * (v[np] - v[npw]) +
                                 fmn * an * (v[npn] - v[np]) - fms * as * (v[np] - v[nps])));
    cvp[np] = c;
    bb = bb + c;
    bb = bb + 0.5 * rp[np] *
                  (xe[np] * (p[np] + p[npe]) - xe[npw] * (p[np] + p[npw]) -
                   xk[np] * (p[np] + p[npn]) + xk[nps] * (p[np] + p[nps]));
    if (F.modvis == 1) {
        const long npsw = nps - 1, npse = nps + 1, npnw = npn - 1, npne = npn + 1;
        const double S1 = vle[np] * re[np] * Je[np] / 3.0 *
                              (xe[np] * xe[np] * v[npe] -
                               xke[np] * xe[np] * (v[npn] + v[npne] - v[nps] - v[npse]) / 4.0) -
                          vle[npw] * re[npw] * Je[npw] / 3.0 *
                              (-(xe[npw] * xe[npw] * v[npw]) -
                               xke[npw] * xe[npw] * (v[npn] + v[npnw] - v[nps] - v[npsw]) / 4.0);
        const double S2 = vln[np] * rn[np] * Jn[np] / 3.0 *
                              (xk[np] * xk[np] * v[npn] -
                               xk[np] * xen[np] * (v[npe] + v[npne] - v[npnw] - v[npw]) / 4.0) -
                          vln[nps] * rn[nps] * Jn[nps] / 3.0 *
                              (-(xk[nps] * xk[nps] * v[nps]) -
                               xk[nps] * xen[nps] * (v[npe] + v[npse] - v[npsw] - v[npw]) / 4.0);
        const double S3 = rp[np] * vle[np] * Je[np] * ye[np] *
                              (xke[np] * (u[npn] + u[npne] - u[nps] - u[npse]) / 4.0 -
                               xe[np] * (u[npe] - u[np])) -
                          rp[np] * vle[npw] * Je[npw] * ye[npw] *
                              (xke[npw] * (u[npn] + u[npnw] - u[nps] - u[npsw]) / 4.0 -
                               xe[npw] * (u[np] - u[npw]));
        const double S4 = rp[np] * vln[np] * Jn[np] * yk[np] *
                              (xen[np] * (u[npne] + u[npe] - u[npnw] - u[npw]) / 4.0 -
                               xk[np] * (u[npn] - u[np])) -
                          rp[np] * vln[nps] * Jn[nps] * yk[nps] *
                              (xen[nps] * (u[npse] + u[npe] - u[npsw] - u[npw]) / 4.0 -
                               xk[nps] * (u[np] - u[nps]));
        const double S5 = -(2.0 / 3.0 * rp[np] * vle[np] * Je[np] * xe[np] *
                            (yke[np] * (u[npn] + u[npne] - u[nps] - u[npse]) / 4.0 -
                             ye[np] * (u[npe] - u[np]))) +
                          2.0 / 3.0 * rp[np] * vle[npw] * Je[npw] * xe[npw] *
                              (yke[npw] * (u[npn] + u[npnw] - u[nps] - u[npsw]) / 4.0 -
                               ye[npw] * (u[np] - u[npw]));
        const double S6 = -(2.0 / 3.0 * rp[np] * vln[np] * Jn[np] * xk[np] *
                            (yen[np] * (u[npe] + u[npne] - u[npw] - u[npnw]) / 4.0 -
                             yk[np] * (u[npn] - u[np]))) +
                          2.0 / 3.0 * rp[np] * vln[nps] * Jn[nps] * xk[nps] *
                              (yen[nps] * (u[npe] + u[npse] - u[npw] - u[npsw]) / 4.0 -
                               yk[nps] * (u[np] - u[nps]));
        const double S8 = 2.0 / 3.0 * (double)F.coord * v[np] *
                          (-(vln[np] * xk[np]) + vln[nps] * xk[nps] + vle[np] * xe[np] -
                           vle[npw] * xe[npw]);
        const double s = S1 + S2 + S3 + S4 + S5 + S6 + S8;
        svp[np] = s;
        bb = bb + s;
    }
    b[np] = bb;
}

// get_T_coefficients_and_source: mach2d_coef.f90:787-1203
__global__ __launch_bounds__(kTx) void k_T_coefficients_and_source(Fields F)
{
    long np;
    if (!interior(F.nx, F.ny, np)) return;
    const double *const *q = F.in;
    const double *rp = q[0], *re = q[1], *rn = q[2], *xe = q[3], *ye = q[4], *xk = q[5],
                 *yk = q[6], *xke = q[7], *yke = q[8], *xen = q[9], *yen = q[10],
                 *alphae = q[11], *betae = q[12], *betan = q[13], *gamman = q[14], *Jp = q[15],
                 *Je = q[16], *Jn = q[17], *roe = q[18], *ron = q[19], *ro = q[20],
                 *roa = q[21], *p = q[22], *pa = q[23], *cp = q[24], *vle = q[25], *vln = q[26],
                 *ke = q[27], *kn = q[28], *Uce = q[29], *Vcn = q[30], *u = q[31], *v = q[32],
                 *ue = q[33], *ve = q[34], *un = q[35], *vn = q[36], *ua = q[37], *va = q[38],
                 *T = q[39], *Ta = q[40], *fbe = q[41], *fbn = q[42];
    double *a = F.out[0], *b = F.out[1];
    const long nx = F.nx, N = (long)F.nx * F.ny;
    const double dt = F.dt, beta = F.beta;
    const double eps = 2.220446049250313080847e-16;  // epsilon(1.d0)
    const long nps = np - nx, npn = np + nx, npw = np - 1, npe = np + 1;
    const double fme = roe[np] * re[np] * Uce[np];
    const double fmw = roe[npw] * re[npw] * Uce[npw];
    const double fmn = ron[np] * rn[np] * Vcn[np];
    const double fms = ron[nps] * rn[nps] * Vcn[nps];
    const double mpa = roa[np] * rp[np] / Jp[np];
    const double as = dsign5(Vcn[nps]), an = dsign5(Vcn[np]);
    const double aw = dsign5(Uce[npw]), ae = dsign5(Uce[np]);
    double c[9];
    if (F.modvis == 0) {
        c[0] = c[2] = c[6] = c[8] = 0.0;
        c[1] = -fms * (0.5 + as) * cp[np];
        c[3] = -fmw * (0.5 + aw) * cp[np];
        c[5] = fme * (0.5 - ae) * cp[np];
        c[7] = fmn * (0.5 - an) * cp[np];
    } else {
        const double ew = ke[npw] * re[npw] * Je[npw] * betae[npw];
        const double ep = ke[np] * re[np] * Je[np] * betae[np];
        const double ns = kn[nps] * rn[nps] * Jn[nps] * betan[nps];
        const double nq = kn[np] * rn[np] * Jn[np] * betan[np];
        c[0] = (ew + ns) / 4.0;
        c[1] = -(fms * (0.5 + as) * cp[np]) - kn[nps] * rn[nps] * Jn[nps] * gamman[nps] +
               (ew - ep) / 4.0;
        c[2] = -(ep + ns) / 4.0;
        c[3] = -(fmw * (0.5 + aw) * cp[np]) - ke[npw] * re[npw] * Je[npw] * alphae[npw] +
               (ns - nq) / 4.0;
        c[5] = fme * (0.5 - ae) * cp[np] - ke[np] * re[np] * Je[np] * alphae[np] +
               (-ns + nq) / 4.0;
        c[6] = (-ew - nq) / 4.0;
        c[7] = fmn * (0.5 - an) * cp[np] - kn[np] * rn[np] * Jn[np] * gamman[np] - (ew - ep) / 4.0;
        c[8] = (ep + nq) / 4.0;
    }
    c[4] = mpa * cp[np] / dt - (c[7] + c[1] + c[3] + c[5]);
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k * N + np] = c[k];
    double bb = mpa * Ta[np] * cp[np] / dt;
    bb = bb - beta * cp[np] *
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
    bb = bb + rp[np] / Jp[np] * (p[np] - pa[np]) / dt - raux;
    if (F.modvis != 0) {
        const long npsw = nps - 1, npse = nps + 1, npnw = npn - 1, npne = npn + 1;
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
        bb = bb - raux;
    }
    b[np] = bb;
}

// ---- SIMPLEC glue --------------------------------------------------------

// sum(a(np,:)) of a 9-diagonal array, left to right (Fortran SUM)
__device__ __forceinline__ double sum9(const double *a, long N, long np)
{
    double s = a[np];
#pragma unroll
    for (int k = 1; k < 9; ++k) s = s + a[k * N + np];
    return s;
}

// get_velocities_at_internal_faces: mach2d_coef.f90:1350-1501.  Grid over
// i = 2..nx-1, j = 2..ny-1: east faces for i <= nx-2, north faces for
// j <= ny-2.
__global__ __launch_bounds__(kTx) void k_velocities_at_internal_faces(Fields F)
{
    long np;
    if (!interior(F.nx, F.ny, np)) return;
    const double *const *q = F.in;
    const double *rp = q[0], *re = q[1], *rn = q[2], *xe = q[3], *ye = q[4], *xk = q[5],
                 *yk = q[6], *xen = q[7], *yen = q[8], *xke = q[9], *yke = q[10], *Jp = q[11],
                 *cup = q[12], *cvp = q[13], *sup = q[14], *svp = q[15], *au = q[16],
                 *av = q[17], *roa = q[18], *p = q[19], *u = q[20], *v = q[21], *uea = q[22],
                 *vea = q[23], *una = q[24], *vna = q[25];
    double *ue = F.out[0], *ve = F.out[1], *un = F.out[2], *vn = F.out[3], *Uce = F.out[4],
           *Vcn = F.out[5];
    const long nx = F.nx, N = (long)F.nx * F.ny;
    const double dt = F.dt;
    const int i = (int)(np % nx) + 1, j = (int)(np / nx) + 1;
    const long nps = np - nx, npn = np + nx, npw = np - 1, npe = np + 1;
    const long npsw = nps - 1, npse = nps + 1, npnw = npn - 1, npne = npn + 1;
#define AU(k, r) au[(long)((k) - 1) * N + (r)]
#define AV(k, r) av[(long)((k) - 1) * N + (r)]
    const double mpa = roa[np] * rp[np] / Jp[np];
    const double sumup = AU(1, np) * u[npsw] + AU(2, np) * u[nps] + AU(3, np) * u[npse] +
                         AU(4, np) * u[npw] + AU(6, np) * u[npe] + AU(7, np) * u[npnw] +
                         AU(8, np) * u[npn] + AU(9, np) * u[npne];
    const double sumvp = AV(1, np) * v[npsw] + AV(2, np) * v[nps] + AV(3, np) * v[npse] +
                         AV(4, np) * v[npw] + AV(6, np) * v[npe] + AV(7, np) * v[npnw] +
                         AV(8, np) * v[npn] + AV(9, np) * v[npne];
    if (i <= F.nx - 2) {
        const long npee = npe + 1, npsee = npse + 1, npnee = npne + 1;
        const double mea = roa[npe] * rp[npe] / Jp[npe];
        const double sumue = AU(1, npe) * u[nps] + AU(2, npe) * u[npse] + AU(3, npe) * u[npsee] +
                             AU(4, npe) * u[np] + AU(6, npe) * u[npee] + AU(7, npe) * u[npn] +
                             AU(8, npe) * u[npne] + AU(9, npe) * u[npnee];
        const double sumve = AV(1, npe) * v[nps] + AV(2, npe) * v[npse] + AV(3, npe) * v[npsee] +
                             AV(4, npe) * v[np] + AV(6, npe) * v[npee] + AV(7, npe) * v[npn] +
                             AV(8, npe) * v[npne] + AV(9, npe) * v[npnee];
        const double aux = (p[npn] + p[npne] - p[nps] - p[npse]) / 4.0;
        const double pue = re[np] * (yke[np] * aux + ye[np] * (p[np] - p[npe]));
        const double pve = re[np] * (-xke[np] * aux + xe[np] * (p[npe] - p[np]));
        const double a = ((mpa + mea) * uea[np] / dt + cup[np] + cup[npe] + sup[np] + sup[npe] -
                          sumup - sumue + 2.0 * pue) /
                         (AU(5, np) + AU(5, npe));
        const double b = ((mpa + mea) * vea[np] / dt + cvp[np] + cvp[npe] + svp[np] + svp[npe] -
                          sumvp - sumve + 2.0 * pve) /
                         (AV(5, np) + AV(5, npe));
        ue[np] = a;
        ve[np] = b;
        Uce[np] = a * ye[np] - b * xe[np];
    }
    if (j <= F.ny - 2) {
        const long npnn = npn + nx, npnnw = npnn - 1, npnne = npnn + 1;
        const double mna = roa[npn] * rp[npn] / Jp[npn];
        const double sumun = AU(1, npn) * u[npw] + AU(2, npn) * u[np] + AU(3, npn) * u[npe] +
                             AU(4, npn) * u[npnw] + AU(6, npn) * u[npne] + AU(7, npn) * u[npnnw] +
                             AU(8, npn) * u[npnn] + AU(9, npn) * u[npnne];
        const double sumvn = AV(1, npn) * v[npw] + AV(2, npn) * v[np] + AV(3, npn) * v[npe] +
                             AV(4, npn) * v[npnw] + AV(6, npn) * v[npne] + AV(7, npn) * v[npnnw] +
                             AV(8, npn) * v[npnn] + AV(9, npn) * v[npnne];
        const double aux = (p[npne] + p[npe] - p[npnw] - p[npw]) / 4.0;
        const double pun = rn[np] * (yk[np] * (p[npn] - p[np]) - yen[np] * aux);
        const double pvn = rn[np] * (xk[np] * (p[np] - p[npn]) + xen[np] * aux);
        const double a = ((mpa + mna) * una[np] / dt + cup[np] + cup[npn] + sup[np] + sup[npn] -
                          sumup - sumun + 2.0 * pun) /
                         (AU(5, np) + AU(5, npn));
        const double b = ((mpa + mna) * vna[np] / dt + cvp[np] + cvp[npn] + svp[np] + svp[npn] -
                          sumvp - sumvn + 2.0 * pvn) /
                         (AV(5, np) + AV(5, npn));
        un[np] = a;
        vn[np] = b;
        Vcn[np] = b * xk[np] - a * yk[np];
    }
#undef AU
#undef AV
}

// get_internal_simplec_coefficients: mach2d_coef.f90:1504-1600
__global__ __launch_bounds__(kTx) void k_internal_simplec_coefficients(Fields F)
{
    long np;
    if (!interior(F.nx, F.ny, np)) return;
    const double *const *q = F.in;
    const double *re = q[0], *rn = q[1], *xe = q[2], *ye = q[3], *xk = q[4], *yk = q[5],
                 *au = q[6], *av = q[7];
    const long nx = F.nx, N = (long)F.nx * F.ny;
    const int i = (int)(np % nx) + 1, j = (int)(np / nx) + 1;
    const double sau = sum9(au, N, np), sav = sum9(av, N, np);
    if (i <= F.nx - 2) {
        const long npe = np + 1;
        const double Lue = 2.0 * re[np] / (sau + sum9(au, N, npe));
        const double Lve = 2.0 * re[np] / (sav + sum9(av, N, npe));
        const double due = Lue * ye[np], dve = Lve * xe[np];
        F.out[0][np] = due;
        F.out[1][np] = dve;
        F.out[4][np] = ye[np] * due + xe[np] * dve;
    }
    if (j <= F.ny - 2) {
        const long npn = np + nx;
        const double Lun = 2.0 * rn[np] / (sau + sum9(au, N, npn));
        const double Lvn = 2.0 * rn[np] / (sav + sum9(av, N, npn));
        const double dun = Lun * yk[np], dvn = Lvn * xk[np];
        F.out[2][np] = dun;
        F.out[3][np] = dvn;
        F.out[5][np] = xk[np] * dvn + yk[np] * dun;
    }
}

// Every point: get_pressure_correction_with_pl (:1602, p += pl) and
// get_density_correction_with_pl (:1614, ro += pl*g).  Grid over all N.
__global__ __launch_bounds__(kTx) void k_pressure_correction_with_pl(Fields F)
{
    if (gated(F)) return;
    const long q = (long)(blockIdx.y * gridDim.x + blockIdx.x) * kTx + threadIdx.x;
    if (q >= (long)F.nx * F.ny) return;
    F.out[0][q] = F.out[0][q] + F.in[0][q];
}
__global__ __launch_bounds__(kTx) void k_density_correction_with_pl(Fields F)
{
    if (gated(F)) return;
    const long q = (long)(blockIdx.y * gridDim.x + blockIdx.x) * kTx + threadIdx.x;
    if (q >= (long)F.nx * F.ny) return;
    F.out[0][q] = F.out[0][q] + F.in[0][q] * F.in[1][q];
}

// get_u_v_at_real_nodes_with_pl: mach2d_coef.f90:1626-1675
__global__ __launch_bounds__(kTx) void k_u_v_at_real_nodes_with_pl(Fields F)
{
    long np;
    if (!interior(F.nx, F.ny, np)) return;
    const double *const *q = F.in;
    const double *xe = q[0], *ye = q[1], *xk = q[2], *yk = q[3], *rp = q[4], *pl = q[5],
                 *au = q[6], *av = q[7];
    const long nx = F.nx, N = (long)F.nx * F.ny;
    const long nps = np - nx, npn = np + nx, npw = np - 1, npe = np + 1;
    const double plup = 0.5 * rp[np] *
                        (yk[np] * (pl[npn] + pl[np]) - yk[nps] * (pl[nps] + pl[np]) -
                         ye[np] * (pl[npe] + pl[np]) + ye[npw] * (pl[npw] + pl[np]));
    const double plvp = 0.5 * rp[np] *
                        (xe[np] * (pl[np] + pl[npe]) - xe[npw] * (pl[np] + pl[npw]) -
                         xk[np] * (pl[np] + pl[npn]) + xk[nps] * (pl[np] + pl[nps]));
    F.out[0][np] = F.out[0][np] + plup / sum9(au, N, np);
    F.out[1][np] = F.out[1][np] + plvp / sum9(av, N, np);
}

// get_velocity_correction_at_faces_with_pl: mach2d_coef.f90:1679-1758 (east
// faces j = 2..ny-1, i = 1..nx-1; north faces j = 1..ny-1, i = 2..nx-1).
// Grid over i = 1..nx-1, j = 1..ny-1.
__global__ __launch_bounds__(kTx) void k_velocity_correction_at_faces_with_pl(Fields F)
{
    if (gated(F)) return;
    const int nx = F.nx, ny = F.ny;
    const int i = blockIdx.x * kTx + threadIdx.x + 1;
    const int j = blockIdx.y + 1;
    if (i > nx - 1 || j > ny - 1) return;
    const double *const *q = F.in;
    const double *xe = q[0], *ye = q[1], *xk = q[2], *yk = q[3], *due = q[4], *dve = q[5],
                 *dun = q[6], *dvn = q[7], *de = q[8], *dn = q[9], *pl = q[10];
    double *ue = F.out[0], *ve = F.out[1], *un = F.out[2], *vn = F.out[3], *Uce = F.out[4],
           *Vcn = F.out[5];
    const long np = (long)nx * (j - 1) + i - 1;
    if (j >= 2) {
        const long npe = np + 1;
        Uce[np] = ue[np] * ye[np] - ve[np] * xe[np] + de[np] * (pl[np] - pl[npe]);
        ue[np] = ue[np] + due[np] * (pl[np] - pl[npe]);
        ve[np] = ve[np] + dve[np] * (pl[npe] - pl[np]);
    }
    if (i >= 2) {
        const long npn = np + nx;
        Vcn[np] = vn[np] * xk[np] - un[np] * yk[np] + dn[np] * (pl[np] - pl[npn]);
        un[np] = un[np] + dun[np] * (pl[npn] - pl[np]);
        vn[np] = vn[np] + dvn[np] * (pl[np] - pl[npn]);
    }
}

// initialize_boundary_faces_flags: mach2d_coef.f90:1764-1828.  Grid over all N.
__global__ __launch_bounds__(kTx) void k_boundary_faces_flags(Fields F)
{
    const int nx = F.nx, ny = F.ny;
    const long q = (long)(blockIdx.y * gridDim.x + blockIdx.x) * kTx + threadIdx.x;
    if (q >= (long)nx * ny) return;
    const int i = (int)(q % nx) + 1, j = (int)(q / nx) + 1;
    const bool e = (i == 1 || i == nx - 1) && j >= 2 && j <= ny - 1;
    const bool n = (j == 1 || j == ny - 1) && i >= 2 && i <= nx - 1;
    F.out[0][q] = e ? 1.0 : 0.0;
    F.out[1][q] = n ? 1.0 : 0.0;
}

// check_mass_conservation: mach2d_coef.f90:1833-1880.  The per-point terms
// in parallel (into the scratch output out[1], interior points in loop
// order), then one lane adds them in that order, which is the reference's
// summation order, so the L1 norm is bit-identical.
__global__ __launch_bounds__(kTx) void k_mass_terms(Fields F)
{
    long np;
    if (!interior(F.nx, F.ny, np)) return;
    const double *const *q = F.in;
    const double *rp = q[0], *re = q[1], *rn = q[2], *Jp = q[3], *ro = q[4], *roa = q[5],
                 *ron = q[6], *roe = q[7], *Uce = q[8], *Vcn = q[9];
    const long nps = np - F.nx, npw = np - 1;
    const double fme = roe[np] * re[np] * Uce[np];
    const double fmw = roe[npw] * re[npw] * Uce[npw];
    const double fmn = ron[np] * rn[np] * Vcn[np];
    const double fms = ron[nps] * rn[nps] * Vcn[nps];
    const int i = (int)(np % F.nx) + 1, j = (int)(np / F.nx) + 1;
    const long t = (long)(j - 2) * (F.nx - 2) + (i - 2);
    F.out[1][t] = fabs((ro[np] - roa[np]) * rp[np] / (Jp[np] * F.dt) + fme - fmw + fmn - fms);
}
// the terms added in the reference's order by one wave (seq_sum)
__global__ __launch_bounds__(kSeqThreads) void k_mass_sum(Fields F)
{
    const long n = (long)(F.nx - 2) * (F.ny - 2);
    const double *t = F.out[1];
    __shared__ double lds[kSeqSlots * kSeqChunk];
    const double r = seq_sum(t, n, lds);
    if (threadIdx.x == 0) F.out[0][0] = r;
}

// ---------------------------------------------------------------------------
// Fictitious-row writers of the internal-flow case,
// src/user/internal_flow/mod_intflow.f90:214-550.  Grid over the boundary:
// thread q < nx is (q+1, 1) of the south row, nx <= q < 2nx the north row,
// then the west column (1, 2..ny-1) and the east column (nx, 2..ny-1).
// Each point zeroes its whole coefficient row and writes its own entries,
// so rows the routine does not own keep the caller's values.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool boundary_point(int nx, int ny, int &i, int &j)
{
    const long q = (long)blockIdx.x * kTx + threadIdx.x;
    const long r = q - 2L * nx;
    if (q < nx) {
        i = (int)q + 1;
        j = 1;
    } else if (q < 2L * nx) {
        i = (int)(q - nx) + 1;
        j = ny;
    } else if (r < ny - 2) {
        i = 1;
        j = (int)r + 2;
    } else if (r < 2L * (ny - 2)) {
        i = nx;
        j = (int)(r - (ny - 2)) + 2;
    } else {
        return false;
    }
    return true;
}

// Tangential wall velocity component C (0: u, 1: v) at np of row ny-1:
// get_u_v_north_slip, mod_intflow.f90:158-186.
template <int C>
__device__ __forceinline__ double slip_at(const double *xk, const double *yk, const double *u,
                                          const double *v, long np)
{
    const double Up = sqrt(u[np] * u[np] + v[np] * v[np]);
    const double signal = copysign(1.0, xk[np] * u[np] + yk[np] * v[np]);
    const double t = sqrt(xk[np] * xk[np] + yk[np] * yk[np]);
    const double tc = (C == 0 ? xk[np] : yk[np]) / t;
    return signal * Up * tc;
}

// ubn(i) / vbn(i) of get_u_v_north_slip (:142-210): the slip velocity on
// i = 2..nx-1, linear extrapolation to the fictitious i = 1 and i = nx.
template <int C>
__device__ double north_slip(const double *x, const double *xp, const double *xk,
                             const double *yk, const double *u, const double *v, int nx, int ny,
                             int i)
{
    const long np = (long)nx * (ny - 2) + i - 1;
    if (i == 1) {
        const long npe = np + 1, npee = npe + 1;
        const double b2 = slip_at<C>(xk, yk, u, v, npe), b3 = slip_at<C>(xk, yk, u, v, npee);
        return b2 + (b2 - b3) * (xp[npe] - x[np]) / (xp[npee] - xp[npe]);
    }
    if (i == nx) {
        const long npw = np - 1, npww = npw - 1;
        const double b1 = slip_at<C>(xk, yk, u, v, npw), b2 = slip_at<C>(xk, yk, u, v, npww);
        return b1 + (b1 - b2) * (x[npw] - xp[npw]) / (xp[npw] - xp[npww]);
    }
    return slip_at<C>(xk, yk, u, v, np);
}

// intflow_set_bcu (C = 0, :214-301) and intflow_set_bcv (C = 1, :305-392).
// in: x, xp, xk, yk, u, v.  out: a (9 diagonals), b.
template <int C>
__global__ __launch_bounds__(kTx) void k_intflow_bc_uv(Fields F)
{
    const int nx = F.nx, ny = F.ny;
    int i, j;
    if (!boundary_point(nx, ny, i, j)) return;
    const double *x = F.in[0], *xp = F.in[1], *xk = F.in[2], *yk = F.in[3], *u = F.in[4],
                 *v = F.in[5];
    const long N = (long)nx * ny, np = (long)nx * (j - 1) + i - 1;
    double c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double b;
    if (j == 1) {  // south, corners included
        c[kN] = C == 0 ? -1.0 : 1.0;
        c[kP] = 1.0;
        b = 0.0;
    } else if (j == ny) {  // north, corners included
        c[kS] = 1.0;
        c[kP] = 1.0;
        b = F.modvis == 1 ? 0.0 : 2.0 * north_slip<C>(x, xp, xk, yk, u, v, nx, ny, i);
    } else if (i == 1) {  // west
        const long npe = np + 1, npee = npe + 1;
        c[kE] = C == 0 ? -1.0 : 1.0;
        c[kP] = 1.0;
        b = C == 0 ? (2.0 * xp[npe]) / (xp[npee] - xp[npe]) * (u[npe] - u[npee]) : 0.0;
    } else {  // east
        const long npw = np - 1, npww = npw - 1;
        c[kW] = -1.0;
        c[kP] = 1.0;
        b = C == 0 ? 2.0 * (x[npw] - xp[npw]) / (xp[npw] - xp[npww]) * (u[npw] - u[npww])
                   : (v[npw] - v[npww]) * 2.0 * (x[npw] - xp[npw]) / (xp[npw] - xp[npww]);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) F.out[0][k * N + np] = c[k];
    F.out[1][np] = b;
}

// intflow_set_bcpl: :396-459.  5 diagonals (S, W, P, E, N); corners are not
// written.  in: x, xp, pl, plin (ny values).  out: ap, bp.
__global__ __launch_bounds__(kTx) void k_intflow_bc_pl(Fields F)
{
    if (gated(F)) return;
    const int nx = F.nx, ny = F.ny;
    int i, j;
    if (!boundary_point(nx, ny, i, j)) return;
    if ((j == 1 || j == ny) && (i == 1 || i == nx)) return;
    const double *x = F.in[0], *xp = F.in[1], *pl = F.in[2], *plin = F.in[3];
    const long N = (long)nx * ny, np = (long)nx * (j - 1) + i - 1;
    double c[5] = {0, 0, 0, 0, 0};
    double b;
    if (j == 1) {
        c[4] = -1.0;
        c[2] = 1.0;
        b = 0.0;
    } else if (j == ny) {
        c[0] = -1.0;
        c[2] = 1.0;
        b = 0.0;
    } else if (i == 1) {
        c[3] = 1.0;
        c[2] = 1.0;
        b = 2.0 * plin[j - 1];
    } else {
        const long npw = np - 1, npww = npw - 1;
        c[1] = -1.0;
        c[2] = 1.0;
        b = 2.0 * (x[npw] - xp[npw]) / (xp[npw] - xp[npww]) * (pl[npw] - pl[npww]);
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) F.out[0][k * N + np] = c[k];
    F.out[1][np] = b;
}

// intflow_set_bcT: :464-550.  F.twall > 0 prescribes the north wall
// temperature, otherwise the north wall is adiabatic.  in: x, xp, T, Tin
// (ny values).  out: at, bt.
__global__ __launch_bounds__(kTx) void k_intflow_bc_T(Fields F)
{
    const int nx = F.nx, ny = F.ny;
    int i, j;
    if (!boundary_point(nx, ny, i, j)) return;
    const double *x = F.in[0], *xp = F.in[1], *T = F.in[2], *Tin = F.in[3];
    const long N = (long)nx * ny, np = (long)nx * (j - 1) + i - 1;
    const bool ccTw = F.twall > 0.0;
    double c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double b;
    if (j == 1) {
        c[kN] = -1.0;
        c[kP] = 1.0;
        b = 0.0;
    } else if (j == ny) {
        c[kS] = ccTw ? 1.0 : -1.0;
        c[kP] = 1.0;
        b = ccTw ? 2.0 * F.twall : 0.0;
    } else if (i == 1) {
        c[kE] = 1.0;
        c[kP] = 1.0;
        b = 2.0 * Tin[j - 1];
    } else {
        const long npw = np - 1, npww = npw - 1;
        c[kW] = -1.0;
        c[kP] = 1.0;
        b = 2.0 * (x[npw] - xp[npw]) / (xp[npw] - xp[npww]) * (T[npw] - T[npww]);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) F.out[0][k * N + np] = c[k];
    F.out[1][np] = b;
}

// intflow_velocities_at_boundary_faces: :692-787.  One thread per boundary
// face: west (i = 1) and east (i = nx-1) faces of rows 2..ny-1, south (j = 1)
// and north (j = ny-1) faces of columns 2..nx-1; the north face takes the
// Euler slip velocity (modvis = 0) or zero.  in: x, xp, xe, ye, xk, yk, u, v.
// in/out: ue, ve, un, vn, Uce, Vcn.
__global__ __launch_bounds__(kTx) void k_intflow_vel_bf(Fields F)
{
    if (gated(F)) return;
    const int nx = F.nx, ny = F.ny;
    const long q = (long)blockIdx.x * kTx + threadIdx.x;
    const long nr = ny - 2, nc = nx - 2;
    const double *x = F.in[0], *xp = F.in[1], *xe = F.in[2], *ye = F.in[3], *xk = F.in[4],
                 *yk = F.in[5], *u = F.in[6], *v = F.in[7];
    double *ue = F.out[0], *ve = F.out[1], *un = F.out[2], *vn = F.out[3], *Uce = F.out[4],
           *Vcn = F.out[5];
    if (q < 2 * nr) {  // west, then east
        const bool east = q >= nr;
        const int j = (int)(east ? q - nr : q) + 2, i = east ? nx - 1 : 1;
        const long np = (long)nx * (j - 1) + i - 1, npe = np + 1;
        const double uf = (u[np] + u[npe]) / 2.0;
        const double vf = east ? (v[np] + v[npe]) / 2.0 : 0.0;
        ue[np] = uf;
        ve[np] = vf;
        Uce[np] = uf * ye[np] - vf * xe[np];
    } else if (q < 2 * nr + 2 * nc) {  // south, then north
        const long r = q - 2 * nr;
        const bool north = r >= nc;
        const int i = (int)(north ? r - nc : r) + 2, j = north ? ny - 1 : 1;
        const long np = (long)nx * (j - 1) + i - 1;
        if (north) {
            un[np] = F.modvis == 0 ? slip_at<0>(xk, yk, u, v, np) : 0.0;
            vn[np] = F.modvis == 0 ? slip_at<1>(xk, yk, u, v, np) : 0.0;
        } else {
            un[np] = (u[np] + u[np + nx]) / 2.0;
            vn[np] = 0.0;
        }
        Vcn[np] = 0.0;
    }
    (void)x;
    (void)xp;
}

}  // namespace m2d

using namespace m2d;

// ---------------------------------------------------------------------------
// host side: one routine = kernel + field counts
// ---------------------------------------------------------------------------
namespace {

thread_local hipStream_t g_stream = nullptr;
thread_local const int *g_gate = nullptr;
thread_local const double *g_sc = nullptr;

enum Grid { kInterior = 0, kFaces = 1, kAll = 2, kBoundary = 3 };

struct Routine {
    void (*kernel)(Fields);
    int nin;
    int in9[2];   // inputs that are 9-diagonal arrays (-1: none)
    int nout;
    int outk[6];  // N-vectors per output; 0 = one scalar
    int grid;     // kInterior, kFaces (1..nx-1 x 1..ny-1), kAll or kBoundary
    void (*post)(Fields);  // one-workgroup follow-up over a scratch output
    int iny1;     // 1 + index of an input of ny values (0: none)
};

bool is_device_ptr(const void *p)
{
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();  // clear the error of an unknown host pointer
        return false;
    }
    return at.type == hipMemoryTypeDevice;
}

long in_len(const Routine &R, int k, long N, int ny)
{
    if (k + 1 == R.iny1) return ny;
    return (k == R.in9[0] || k == R.in9[1]) ? 9 * N : N;
}
long out_len(const Routine &R, int k, long N) { return R.outk[k] ? R.outk[k] * N : 1; }

int run(const Routine &R, Fields F)
{
    const int nx = F.nx, ny = F.ny;
    if (nx < 3 || ny < 3) return M2D_ERR_ARG;
    for (int k = 0; k < R.nin; ++k)
        if (!F.in[k]) return M2D_ERR_ARG;
    for (int k = 0; k < R.nout; ++k)
        if (!F.out[k]) return M2D_ERR_ARG;
    const long N = (long)nx * ny;
    const bool dev = is_device_ptr(R.nin ? (const void *)F.in[0] : (const void *)F.out[0]);
    for (int k = 0; k < R.nin; ++k)
        if (is_device_ptr(F.in[k]) != dev) return M2D_ERR_ARG;
    for (int k = 0; k < R.nout; ++k)
        if (is_device_ptr(F.out[k]) != dev) return M2D_ERR_ARG;
    dim3 grid, block(kTx);
    if (R.grid == kFaces)
        grid = dim3((unsigned)((nx - 1 + kTx - 1) / kTx), (unsigned)(ny - 1));
    else if (R.grid == kInterior)
        grid = dim3((unsigned)((nx - 2 + kTx - 1) / kTx), (unsigned)(ny - 2));
    else if (R.grid == kBoundary)
        grid = dim3((unsigned)((2L * nx + 2L * (ny - 2) + kTx - 1) / kTx), 1);
    else
        grid = dim3((unsigned)((N + kTx - 1) / kTx), 1);
    // device buffer: staged host arrays and the scratch output of `post`
    long tot = R.post ? N : 0;
    if (!dev) {
        for (int k = 0; k < R.nin; ++k) tot += in_len(R, k, N, ny);
        for (int k = 0; k < R.nout; ++k) tot += out_len(R, k, N);
    }
    double *d = nullptr;
    if (tot && hipMalloc((void **)&d, sizeof(double) * tot) != hipSuccess) return M2D_ERR_HIP;
    Fields G = F;
    hipStream_t s = dev ? g_stream : nullptr;
    if (dev) {
        G.gate = g_gate;
        G.sc = g_sc;
    }
    int rc = M2D_OK;
    long off = 0;
    if (!dev) {
        for (int k = 0; k < R.nin && !rc; ++k) {
            const long n = in_len(R, k, N, ny);
            if (hipMemcpy(d + off, F.in[k], sizeof(double) * n, hipMemcpyHostToDevice) !=
                hipSuccess)
                rc = M2D_ERR_HIP;
            G.in[k] = d + off;
            off += n;
        }
        for (int k = 0; k < R.nout && !rc; ++k) {
            const long n = out_len(R, k, N);
            if (hipMemcpy(d + off, F.out[k], sizeof(double) * n, hipMemcpyHostToDevice) !=
                hipSuccess)
                rc = M2D_ERR_HIP;
            G.out[k] = d + off;
            off += n;
        }
    }
    if (R.post) G.out[R.nout] = d + off;  // scratch
    if (!rc) {
        hipLaunchKernelGGL(R.kernel, grid, block, 0, s, G);
        if (R.post) hipLaunchKernelGGL(R.post, dim3(1), dim3(kSeqThreads), 0, s, G);
        if (hipGetLastError() != hipSuccess) rc = M2D_ERR_HIP;
    }
    if (!dev)
        for (int k = 0; k < R.nout && !rc; ++k)
            if (hipMemcpy(F.out[k], G.out[k], sizeof(double) * out_len(R, k, N),
                          hipMemcpyDeviceToHost) != hipSuccess)
                rc = M2D_ERR_HIP;
    if (d) {
        if (dev && hipStreamSynchronize(s) != hipSuccess) rc = M2D_ERR_HIP;  // scratch in use
        (void)hipFree(d);
    }
    return rc;
}

const Routine kNodes = {k_density_at_nodes, 2, {-1, -1}, 1, {1}, kAll, nullptr};
const Routine kDensity = {k_density_at_faces, 3, {-1, -1}, 2, {1, 1}, kFaces, nullptr};
const Routine kU = {k_u_coefficients, 19, {-1, -1}, 1, {9}, kInterior, nullptr};
const Routine kV = {k_v_coefficients, 20, {-1, -1}, 1, {9}, kInterior, nullptr};
const Routine kUs = {k_u_source, 25, {-1, -1}, 3, {1, 1, 1}, kInterior, nullptr};
const Routine kVs = {k_v_source, 25, {-1, -1}, 3, {1, 1, 1}, kInterior, nullptr};
const Routine kT = {k_T_coefficients_and_source, 43, {-1, -1}, 2, {9, 1}, kInterior, nullptr};
const Routine kPc = {k_p_coefficients, 11, {-1, -1}, 1, {5}, kInterior, nullptr};
const Routine kPs = {k_p_source, 14, {-1, -1}, 1, {1}, kInterior, nullptr};
const Routine kVel = {k_velocities_at_internal_faces, 26, {16, 17}, 6, {1, 1, 1, 1, 1, 1},
                      kInterior, nullptr};
const Routine kSimplec = {k_internal_simplec_coefficients, 8, {6, 7}, 6, {1, 1, 1, 1, 1, 1},
                          kInterior, nullptr};
const Routine kPcor = {k_pressure_correction_with_pl, 1, {-1, -1}, 1, {1}, kAll, nullptr};
const Routine kRcor = {k_density_correction_with_pl, 2, {-1, -1}, 1, {1}, kAll, nullptr};
const Routine kUVcor = {k_u_v_at_real_nodes_with_pl, 8, {6, 7}, 2, {1, 1}, kInterior, nullptr};
const Routine kFcor = {k_velocity_correction_at_faces_with_pl, 11, {-1, -1}, 6,
                       {1, 1, 1, 1, 1, 1}, kFaces, nullptr};
const Routine kFlags = {k_boundary_faces_flags, 0, {-1, -1}, 2, {1, 1}, kAll, nullptr};
const Routine kMass = {k_mass_terms, 10, {-1, -1}, 1, {0}, kInterior, k_mass_sum};
const Routine kBcU = {k_intflow_bc_uv<0>, 6, {-1, -1}, 2, {9, 1}, kBoundary, nullptr};
const Routine kBcV = {k_intflow_bc_uv<1>, 6, {-1, -1}, 2, {9, 1}, kBoundary, nullptr};
const Routine kBcPl = {k_intflow_bc_pl, 4, {-1, -1}, 2, {5, 1}, kBoundary, nullptr, 4};
const Routine kBcT = {k_intflow_bc_T, 4, {-1, -1}, 2, {9, 1}, kBoundary, nullptr, 4};
// the boundary grid (2 nx + 2 (ny-2) threads) covers its 2 (nx-2) + 2 (ny-2) faces
const Routine kVelBf = {k_intflow_vel_bf, 8, {-1, -1}, 6, {1, 1, 1, 1, 1, 1}, kBoundary, nullptr};

Fields fields(const int *nx, const int *ny)
{
    Fields F;
    memset(&F, 0, sizeof(F));
    F.nx = *nx;
    F.ny = *ny;
    return F;
}

}  // namespace

void m2d_set_last_status_code(int status);

#define IN(k, x) F.in[k] = (x)
#define FINISH(R)                               \
    do {                                        \
        const int rc_ = run(R, F);              \
        m2d_set_last_status_code(rc_);          \
    } while (0)
#define CHECK_SCALARS(...)                                      \
    do {                                                        \
        const void *s_[] = {__VA_ARGS__};                       \
        for (const void *p_ : s_)                               \
            if (!p_) {                                          \
                m2d_set_last_status_code(M2D_ERR_ARG);          \
                return;                                         \
            }                                                   \
    } while (0)

extern "C" void m2d_assembly_set_stream(void *stream) { g_stream = (hipStream_t)stream; }

void m2d::mass_terms_enqueue(int nx, int ny, double dt, const double *const *in, double *t,
                             hipStream_t s)
{
    Fields F;
    memset(&F, 0, sizeof(F));
    F.nx = nx;
    F.ny = ny;
    F.dt = dt;
    for (int k = 0; k < 10; ++k) F.in[k] = in[k];
    F.out[1] = t;
    const dim3 grid((unsigned)((nx - 2 + kTx - 1) / kTx), (unsigned)(ny - 2));
    hipLaunchKernelGGL(k_mass_terms, grid, dim3(kTx), 0, s, F);
}

// Stream of this thread's device-array assembly calls; returns the previous one.
hipStream_t m2d::assembly_stream(hipStream_t s)
{
    const hipStream_t old = g_stream;
    g_stream = s;
    return old;
}

hipStream_t m2d::assembly_stream_current() { return g_stream; }

void m2d::assembly_gate(const int *gate, const double *sc)
{
    g_gate = gate;
    g_sc = sc;
}

bool m2d::is_device_array(const void *p) { return is_device_ptr(p); }

extern "C" void m2d_get_density_at_nodes(const int *nx, const int *ny, const double *Rg,
                                         const double *p, const double *T, double *ro)
{
    CHECK_SCALARS(nx, ny, Rg);
    Fields F = fields(nx, ny);
    F.beta = *Rg;
    IN(0, p);
    IN(1, T);
    F.out[0] = ro;
    FINISH(kNodes);
}

extern "C" void m2d_get_density_at_faces(const int *nx, const int *ny, const double *beta,
                                         const double *ro, const double *Uce, const double *Vcn,
                                         double *roe, double *ron)
{
    CHECK_SCALARS(nx, ny, beta);
    Fields F = fields(nx, ny);
    F.beta = *beta;
    IN(0, ro);
    IN(1, Uce);
    IN(2, Vcn);
    F.out[0] = roe;
    F.out[1] = ron;
    FINISH(kDensity);
}

extern "C" void m2d_get_u_coefficients(const int *nx, const int *ny, const int *modvis,
                                       const double *dt, const double *rp, const double *re,
                                       const double *rn, const double *Jp, const double *Je,
                                       const double *Jn, const double *ye, const double *yk,
                                       const double *alphae, const double *betae,
                                       const double *betan, const double *gamman,
                                       const double *vle, const double *vln, const double *roe,
                                       const double *ron, const double *roa, const double *Uce,
                                       const double *Vcn, double *a)
{
    CHECK_SCALARS(nx, ny, modvis, dt);
    Fields F = fields(nx, ny);
    F.modvis = *modvis;
    F.dt = *dt;
    const double *in[] = {rp, re, rn, Jp, Je, Jn, ye, yk, alphae, betae,
                          betan, gamman, vle, vln, roe, ron, roa, Uce, Vcn};
    for (int k = 0; k < 19; ++k) IN(k, in[k]);
    F.out[0] = a;
    FINISH(kU);
}

extern "C" void m2d_get_v_coefficients(const int *nx, const int *ny, const int *coord,
                                       const int *modvis, const double *dt, const double *rp,
                                       const double *re, const double *rn, const double *Jp,
                                       const double *Je, const double *Jn, const double *xe,
                                       const double *xk, const double *alphae,
                                       const double *betae, const double *betan,
                                       const double *gamman, const double *vle,
                                       const double *vln, const double *vlp, const double *roe,
                                       const double *ron, const double *roa, const double *Uce,
                                       const double *Vcn, double *a)
{
    CHECK_SCALARS(nx, ny, coord, modvis, dt);
    Fields F = fields(nx, ny);
    F.coord = *coord;
    F.modvis = *modvis;
    F.dt = *dt;
    const double *in[] = {rp,    re,    rn,     Jp,  Je,  Jn,  xe,  xk,  alphae, betae,
                          betan, gamman, vle, vln, vlp, roe, ron, roa, Uce,    Vcn};
    for (int k = 0; k < 20; ++k) IN(k, in[k]);
    F.out[0] = a;
    FINISH(kV);
}

extern "C" void m2d_get_u_source(const int *nx, const int *ny, const int *modvis,
                                 const double *beta, const double *dt, const double *rp,
                                 const double *re, const double *rn, const double *xe,
                                 const double *ye, const double *xk, const double *yk,
                                 const double *xke, const double *yke, const double *xen,
                                 const double *yen, const double *Jp, const double *Je,
                                 const double *Jn, const double *roe, const double *ron,
                                 const double *roa, const double *p, const double *vle,
                                 const double *vln, const double *Uce, const double *Vcn,
                                 const double *ua, const double *u, const double *v, double *cup,
                                 double *sup, double *b)
{
    CHECK_SCALARS(nx, ny, modvis, beta, dt);
    Fields F = fields(nx, ny);
    F.modvis = *modvis;
    F.beta = *beta;
    F.dt = *dt;
    const double *in[] = {rp, re, rn, xe,  ye,  xk,  yk, xke, yke, xen, yen, Jp, Je,
                          Jn, roe, ron, roa, p, vle, vln, Uce, Vcn, ua, u, v};
    for (int k = 0; k < 25; ++k) IN(k, in[k]);
    F.out[0] = cup;
    F.out[1] = sup;
    F.out[2] = b;
    FINISH(kUs);
}

extern "C" void m2d_get_v_source(const int *nx, const int *ny, const int *coord,
                                 const int *modvis, const double *beta, const double *dt,
                                 const double *rp, const double *re, const double *rn,
                                 const double *xe, const double *ye, const double *xk,
                                 const double *yk, const double *xke, const double *yke,
                                 const double *xen, const double *yen, const double *Jp,
                                 const double *Je, const double *Jn, const double *roe,
                                 const double *ron, const double *roa, const double *p,
                                 const double *vle, const double *vln, const double *Uce,
                                 const double *Vcn, const double *va, const double *u,
                                 const double *v, double *cvp, double *svp, double *b)
{
    CHECK_SCALARS(nx, ny, coord, modvis, beta, dt);
    Fields F = fields(nx, ny);
    F.coord = *coord;
    F.modvis = *modvis;
    F.beta = *beta;
    F.dt = *dt;
    const double *in[] = {rp, re, rn, xe,  ye,  xk,  yk, xke, yke, xen, yen, Jp, Je,
                          Jn, roe, ron, roa, p, vle, vln, Uce, Vcn, va, u, v};
    for (int k = 0; k < 25; ++k) IN(k, in[k]);
    F.out[0] = cvp;
    F.out[1] = svp;
    F.out[2] = b;
    FINISH(kVs);
}

extern "C" void m2d_get_T_coefficients_and_source(
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
    double *a, double *b)
{
    CHECK_SCALARS(nx, ny, modvis, beta, dt);
    Fields F = fields(nx, ny);
    F.modvis = *modvis;
    F.beta = *beta;
    F.dt = *dt;
    const double *in[] = {rp,  re,  rn,  xe,  ye,  xk,  yk,  xke, yke, xen, yen,
                          alphae, betae, betan, gamman, Jp, Je, Jn, roe, ron, ro, roa,
                          p,   pa,  cp,  vle, vln, ke,  kn,  Uce, Vcn, u,   v,
                          ue,  ve,  un,  vn,  ua,  va,  T,   Ta,  fbe, fbn};
    for (int k = 0; k < 43; ++k) IN(k, in[k]);
    F.out[0] = a;
    F.out[1] = b;
    FINISH(kT);
}

extern "C" void m2d_get_p_coefficients(const int *nx, const int *ny, const double *dt,
                                       const double *rp, const double *re, const double *rn,
                                       const double *Jp, const double *Uce, const double *Vcn,
                                       const double *roe, const double *ron, const double *g,
                                       const double *de, const double *dn, double *a)
{
    CHECK_SCALARS(nx, ny, dt);
    Fields F = fields(nx, ny);
    F.dt = *dt;
    const double *in[] = {rp, re, rn, Jp, Uce, Vcn, roe, ron, g, de, dn};
    for (int k = 0; k < 11; ++k) IN(k, in[k]);
    F.out[0] = a;
    FINISH(kPc);
}

extern "C" void m2d_get_p_source(const int *nx, const int *ny, const double *dt,
                                 const double *rp, const double *re, const double *rn,
                                 const double *Jp, const double *roe, const double *ron,
                                 const double *ro, const double *roa, const double *Uce,
                                 const double *Vcn, const double *de, const double *dn,
                                 const double *g, const double *pl, double *b)
{
    CHECK_SCALARS(nx, ny, dt);
    Fields F = fields(nx, ny);
    F.dt = *dt;
    const double *in[] = {rp, re, rn, Jp, roe, ron, ro, roa, Uce, Vcn, de, dn, g, pl};
    for (int k = 0; k < 14; ++k) IN(k, in[k]);
    F.out[0] = b;
    FINISH(kPs);
}

extern "C" void m2d_get_velocities_at_internal_faces(
    const int *nx, const int *ny, const double *dt, const double *rp, const double *re,
    const double *rn, const double *xe, const double *ye, const double *xk, const double *yk,
    const double *xen, const double *yen, const double *xke, const double *yke, const double *Jp,
    const double *cup, const double *cvp, const double *sup, const double *svp,
    const double *au, const double *av, const double *roa, const double *p, const double *u,
    const double *v, const double *uea, const double *vea, const double *una, const double *vna,
    double *ue, double *ve, double *un, double *vn, double *Uce, double *Vcn)
{
    CHECK_SCALARS(nx, ny, dt);
    Fields F = fields(nx, ny);
    F.dt = *dt;
    const double *in[] = {rp,  re,  rn,  xe, ye, xk, yk, xen, yen, xke, yke, Jp, cup,
                          cvp, sup, svp, au, av, roa, p, u,  v,   uea, vea, una, vna};
    for (int k = 0; k < 26; ++k) IN(k, in[k]);
    double *out[] = {ue, ve, un, vn, Uce, Vcn};
    for (int k = 0; k < 6; ++k) F.out[k] = out[k];
    FINISH(kVel);
}

extern "C" void m2d_get_internal_simplec_coefficients(
    const int *nx, const int *ny, const double *re, const double *rn, const double *xe,
    const double *ye, const double *xk, const double *yk, const double *au, const double *av,
    double *due, double *dve, double *dun, double *dvn, double *de, double *dn)
{
    CHECK_SCALARS(nx, ny);
    Fields F = fields(nx, ny);
    const double *in[] = {re, rn, xe, ye, xk, yk, au, av};
    for (int k = 0; k < 8; ++k) IN(k, in[k]);
    double *out[] = {due, dve, dun, dvn, de, dn};
    for (int k = 0; k < 6; ++k) F.out[k] = out[k];
    FINISH(kSimplec);
}

extern "C" void m2d_get_pressure_correction_with_pl(const int *nx, const int *ny,
                                                    const double *pl, double *p)
{
    CHECK_SCALARS(nx, ny);
    Fields F = fields(nx, ny);
    IN(0, pl);
    F.out[0] = p;
    FINISH(kPcor);
}

extern "C" void m2d_get_density_correction_with_pl(const int *nx, const int *ny,
                                                   const double *pl, const double *g, double *ro)
{
    CHECK_SCALARS(nx, ny);
    Fields F = fields(nx, ny);
    IN(0, pl);
    IN(1, g);
    F.out[0] = ro;
    FINISH(kRcor);
}

extern "C" void m2d_get_u_v_at_real_nodes_with_pl(const int *nx, const int *ny,
                                                  const double *xe, const double *ye,
                                                  const double *xk, const double *yk,
                                                  const double *rp, const double *pl,
                                                  const double *au, const double *av, double *u,
                                                  double *v)
{
    CHECK_SCALARS(nx, ny);
    Fields F = fields(nx, ny);
    const double *in[] = {xe, ye, xk, yk, rp, pl, au, av};
    for (int k = 0; k < 8; ++k) IN(k, in[k]);
    F.out[0] = u;
    F.out[1] = v;
    FINISH(kUVcor);
}

extern "C" void m2d_get_velocity_correction_at_faces_with_pl(
    const int *nx, const int *ny, const double *xe, const double *ye, const double *xk,
    const double *yk, const double *due, const double *dve, const double *dun, const double *dvn,
    const double *de, const double *dn, const double *pl, double *ue, double *ve, double *un,
    double *vn, double *Uce, double *Vcn)
{
    CHECK_SCALARS(nx, ny);
    Fields F = fields(nx, ny);
    const double *in[] = {xe, ye, xk, yk, due, dve, dun, dvn, de, dn, pl};
    for (int k = 0; k < 11; ++k) IN(k, in[k]);
    double *out[] = {ue, ve, un, vn, Uce, Vcn};
    for (int k = 0; k < 6; ++k) F.out[k] = out[k];
    FINISH(kFcor);
}

extern "C" void m2d_initialize_boundary_faces_flags(const int *nx, const int *ny, double *fbe,
                                                    double *fbn)
{
    CHECK_SCALARS(nx, ny);
    Fields F = fields(nx, ny);
    F.out[0] = fbe;
    F.out[1] = fbn;
    FINISH(kFlags);
}

extern "C" void m2d_check_mass_conservation(const int *nx, const int *ny, const double *dt,
                                            const double *rp, const double *re, const double *rn,
                                            const double *Jp, const double *ro,
                                            const double *roa, const double *ron,
                                            const double *roe, const double *Uce,
                                            const double *Vcn, double *res)
{
    CHECK_SCALARS(nx, ny, dt);
    Fields F = fields(nx, ny);
    F.dt = *dt;
    const double *in[] = {rp, re, rn, Jp, ro, roa, ron, roe, Uce, Vcn};
    for (int k = 0; k < 10; ++k) IN(k, in[k]);
    F.out[0] = res;
    FINISH(kMass);
}

// --- fictitious-row writers of the internal-flow case (mod_intflow.f90) ----
// Rows the routine does not own keep the caller's values, as in the
// reference; plin / Tin (iflow%plin, iflow%Tin) hold ny values, Twall is
// iflow%Twall.  The Euler north-slip extrapolation (modvis /= 1) reads
// columns 2 and 3 (nx-1 and nx-2), so it needs nx >= 4.

extern "C" void m2d_intflow_set_bcu(const int *nx, const int *ny, const int *modvis,
                                    const double *x, const double *xp, const double *xk,
                                    const double *yk, const double *u, const double *v,
                                    double *au, double *b)
{
    CHECK_SCALARS(nx, ny, modvis);
    if (*nx < 4 && *modvis != 1) return m2d_set_last_status_code(M2D_ERR_ARG);
    Fields F = fields(nx, ny);
    F.modvis = *modvis;
    const double *in[] = {x, xp, xk, yk, u, v};
    for (int k = 0; k < 6; ++k) IN(k, in[k]);
    F.out[0] = au;
    F.out[1] = b;
    FINISH(kBcU);
}

extern "C" void m2d_intflow_set_bcv(const int *nx, const int *ny, const int *modvis,
                                    const double *x, const double *xp, const double *xk,
                                    const double *yk, const double *u, const double *v,
                                    double *av, double *b)
{
    CHECK_SCALARS(nx, ny, modvis);
    if (*nx < 4 && *modvis != 1) return m2d_set_last_status_code(M2D_ERR_ARG);
    Fields F = fields(nx, ny);
    F.modvis = *modvis;
    const double *in[] = {x, xp, xk, yk, u, v};
    for (int k = 0; k < 6; ++k) IN(k, in[k]);
    F.out[0] = av;
    F.out[1] = b;
    FINISH(kBcV);
}

extern "C" void m2d_intflow_set_bcpl(const int *nx, const int *ny, const double *x,
                                     const double *xp, const double *pl, const double *plin,
                                     double *ap, double *bp)
{
    CHECK_SCALARS(nx, ny);
    Fields F = fields(nx, ny);
    const double *in[] = {x, xp, pl, plin};
    for (int k = 0; k < 4; ++k) IN(k, in[k]);
    F.out[0] = ap;
    F.out[1] = bp;
    FINISH(kBcPl);
}

extern "C" void m2d_intflow_set_bcT(const int *nx, const int *ny, const double *Twall,
                                    const double *x, const double *xp, const double *T,
                                    const double *Tin, double *at, double *bt)
{
    CHECK_SCALARS(nx, ny, Twall);
    Fields F = fields(nx, ny);
    F.twall = *Twall;
    const double *in[] = {x, xp, T, Tin};
    for (int k = 0; k < 4; ++k) IN(k, in[k]);
    F.out[0] = at;
    F.out[1] = bt;
    FINISH(kBcT);
}

extern "C" void m2d_intflow_velocities_at_boundary_faces(
    const int *nx, const int *ny, const int *modvis, const double *x, const double *xp,
    const double *xe, const double *ye, const double *xk, const double *yk, const double *u,
    const double *v, double *ue, double *ve, double *un, double *vn, double *Uce, double *Vcn)
{
    CHECK_SCALARS(nx, ny, modvis);
    Fields F = fields(nx, ny);
    F.modvis = *modvis;
    const double *in[] = {x, xp, xe, ye, xk, yk, u, v};
    for (int k = 0; k < 8; ++k) IN(k, in[k]);
    double *out[] = {ue, ve, un, vn, Uce, Vcn};
    for (int k = 0; k < 6; ++k) F.out[k] = out[k];
    FINISH(kVelBf);
}
