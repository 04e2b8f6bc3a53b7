// tdma9.hip -- gfx950 kernels of the alternating-direction line TDMA for
// 9-diagonal systems (reference: src/solvers/tdma2d9.f90:38-322) and of the
// L1 residual norms (src/solvers/mod_solvers.f90:128-276).
//
// One TDMA iteration is two line-Jacobi passes.  x-pass: every grid row is a
// tridiagonal system in its W/E couplings with the other six neighbours
// taken from x as it was before the pass; y-pass: every column, S/N
// implicit, explicit neighbours from the x-pass result.  All forward
// eliminations of a pass read only the pass's input, so each pass reads one
// buffer and writes the other (x -> tmp -> x) and needs no grid-wide barrier
// between elimination and back substitution.
//
// One lane per line runs the Thomas recurrences in the reference's order
// (bit-identical results).  y-lines: lane i walks column i, so a wave reads
// 64 consecutive points per step (coalesced).  x-lines: a workgroup owns 64
// rows and stages 64x64 tiles of the nine diagonals, b and the three input
// rows through LDS with coalesced loads, then each lane walks its row in LDS.
// P and Q spill to global memory (they are re-read once by the back
// substitution, from L2).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>

#include "m2d_internal.h"
#include "msi_stencil.h"

namespace m2d {

constexpr int kTile = 64;
// threads of an x-pass workgroup: all stage the tiles, the first kTile run
// the rows
constexpr int kLoad = 4 * kTile;

// x-pass: rows j = blockIdx.x*64 + lane.  xin is the pass input, xout its
// output (P, Q scratch of N each).  Tiles of 64 rows x 16 columns are staged
// through LDS (97 KB) with lanes mapped to (row, column) pairs so that every
// wave-wide load reads runs of 16 consecutive points.
constexpr int kTX = 16;

template <typename T>
__global__ __launch_bounds__(kLoad) void k_tdma_x(const T *__restrict__ a, const T *__restrict__ b,
                                                  const T *__restrict__ xin, T *__restrict__ xout,
                                                  T *__restrict__ Pw, T *__restrict__ Qw, int nx,
                                                  int ny)
{
    const long nn = (long)nx * ny;
    const int lane = threadIdx.x;
    const int j0 = blockIdx.x * kTile;
    const int rows = min(kTile, ny - j0);
    __shared__ T ta[9][kTile][kTX + 1];
    __shared__ T tb[kTile][kTX + 1];
    __shared__ T tx[kTile + 2][kTX + 3];  // rows j0-1 .. j0+64, columns c0-1 .. c0+16

    const int j = j0 + lane;
    const bool row_ok = lane < rows;
    T Pm = 0, Qm = 0;  // P, Q of the previous column in this row
    for (int c0 = 0; c0 < nx; c0 += kTX) {
        const int cols = min(kTX, nx - c0);
        __syncthreads();
        for (int idx = lane; idx < rows * cols; idx += kLoad) {
            const int rr = idx / cols, cc = idx - rr * cols;
            const long p = (long)(j0 + rr) * nx + c0 + cc;
#pragma unroll
            for (int k = 0; k < 9; ++k) ta[k][rr][cc] = a[k * nn + p];
            tb[rr][cc] = b[p];
        }
        const int xr = rows + 2, xc = cols + 2;
        for (int idx = lane; idx < xr * xc; idx += kLoad) {
            const int rr = idx / xc, cc = idx - rr * xc;
            const int jj = j0 + rr - 1, col = c0 + cc - 1;
            if (jj >= 0 && jj < ny && col >= 0 && col < nx) tx[rr][cc] = xin[(long)jj * nx + col];
        }
        __syncthreads();
        if (row_ok) {
            for (int ci = 0; ci < cols; ++ci) {
                const int i = c0 + ci;
                const T a1 = ta[kSW][lane][ci], a2 = ta[kS][lane][ci], a3 = ta[kSE][lane][ci];
                const T a4 = ta[kW][lane][ci], a5 = ta[kP][lane][ci], a6 = ta[kE][lane][ci];
                const T a7 = ta[kNW][lane][ci], a8 = ta[kN][lane][ci], a9 = ta[kNE][lane][ci];
                const T bp = tb[lane][ci];
                // neighbour x: tx[row - j0 + 1][col - c0 + 1]; entries outside the
                // grid are never staged and never used by the branch taken
                const T xs = tx[lane][ci + 1], xse = tx[lane][ci + 2], xsw = tx[lane][ci];
                const T xn = tx[lane + 2][ci + 1], xne = tx[lane + 2][ci + 2],
                        xnw = tx[lane + 2][ci];
                T D;
                if (i == 0) {
                    if (j == 0)
                        D = bp - a8 * xn - a9 * xne;
                    else if (j == ny - 1)
                        D = bp - a2 * xs - a3 * xse;
                    else
                        D = bp - a2 * xs - a3 * xse - a8 * xn - a9 * xne;
                    Pm = -a6 / a5;
                    Qm = D / a5;
                } else {
                    if (i < nx - 1) {
                        if (j == 0)
                            D = bp - a7 * xnw - a8 * xn - a9 * xne;
                        else if (j == ny - 1)
                            D = bp - a1 * xsw - a2 * xs - a3 * xse;
                        else
                            D = bp - a1 * xsw - a2 * xs - a3 * xse - a7 * xnw - a8 * xn -
                                a9 * xne;
                    } else {
                        if (j == 0)
                            D = bp - a7 * xnw - a8 * xn;
                        else if (j == ny - 1)
                            D = bp - a1 * xsw - a2 * xs;
                        else
                            D = bp - a1 * xsw - a2 * xs - a7 * xnw - a8 * xn;
                    }
                    const T Pn = -a6 / (a5 + a4 * Pm);
                    const T Qn = (D - a4 * Qm) / (a5 + a4 * Pm);
                    Pm = Pn;
                    Qm = Qn;
                }
                // column-major P, Q: the lanes (rows) of one column are
                // adjacent, so these stores and the loads below coalesce
                Pw[(long)i * ny + j] = Pm;
                Qw[(long)i * ny + j] = Qm;
            }
        }
    }
    // back substitution: tdma2d9.f90:175-193 (the loads do not depend on the
    // recurrence; unrolling lets them run ahead of it)
    if (row_ok) {
        const long rb = (long)j * nx;
        T xv = Qw[(long)(nx - 1) * ny + j];
        xout[rb + nx - 1] = xv;
#pragma unroll 8
        for (int i = nx - 2; i >= 0; --i) {
            xv = Pw[(long)i * ny + j] * xv + Qw[(long)i * ny + j];
            xout[rb + i] = xv;
        }
    }
}

// y-pass: column i = blockIdx.x*kTile + lane (tdma2d9.f90:200-316).  Chunks
// of kRY rows of the 64 columns (coefficients, b and the x halo) are staged
// through LDS by all four waves, so the recurrence of each column reads LDS
// instead of waiting on HBM row by row.
constexpr int kRY = 16;

template <typename T>
__global__ __launch_bounds__(kLoad) void k_tdma_y(const T *__restrict__ a, const T *__restrict__ b,
                                                  const T *__restrict__ xin, T *__restrict__ xout,
                                                  T *__restrict__ Pw, T *__restrict__ Qw, int nx,
                                                  int ny)
{
    const long nn = (long)nx * ny;
    const int lane = threadIdx.x;
    const int i0 = blockIdx.x * kTile;
    const int cols = min(kTile, nx - i0);
    const int i = i0 + lane;
    const bool col_ok = lane < cols;
    const bool west = i == 0, east = i == nx - 1;
    __shared__ T ta[9][kRY][kTile];
    __shared__ T tb[kRY][kTile];
    __shared__ T tx[kRY + 2][kTile + 2];  // rows r0-1 .. r0+kRY, columns i0-1 .. i0+64
    T Pm = 0, Qm = 0;
    for (int r0 = 0; r0 < ny; r0 += kRY) {
        const int rows = min(kRY, ny - r0);
        __syncthreads();
        for (int idx = lane; idx < rows * cols; idx += kLoad) {
            const int rr = idx / cols, cc = idx - rr * cols;
            const long p = (long)(r0 + rr) * nx + i0 + cc;
#pragma unroll
            for (int k = 0; k < 9; ++k) ta[k][rr][cc] = a[k * nn + p];
            tb[rr][cc] = b[p];
        }
        const int xr = rows + 2, xc = cols + 2;
        for (int idx = lane; idx < xr * xc; idx += kLoad) {
            const int rr = idx / xc, cc = idx - rr * xc;
            const int jj = r0 + rr - 1, col = i0 + cc - 1;
            if (jj >= 0 && jj < ny && col >= 0 && col < nx) tx[rr][cc] = xin[(long)jj * nx + col];
        }
        __syncthreads();
        if (!col_ok) continue;
        for (int rj = 0; rj < rows; ++rj) {
            const int j = r0 + rj;
            const long p = (long)j * nx + i;
            const T a1 = ta[kSW][rj][lane], a2 = ta[kS][rj][lane], a3 = ta[kSE][rj][lane];
            const T a4 = ta[kW][rj][lane], a5 = ta[kP][rj][lane], a6 = ta[kE][rj][lane];
            const T a7 = ta[kNW][rj][lane], a8 = ta[kN][rj][lane], a9 = ta[kNE][rj][lane];
            const T bp = tb[rj][lane];
            // neighbour x: tx[row - r0 + 1][col - i0 + 1]; entries outside the
            // grid are never staged and never used by the branch taken
            const T xw = west ? T(0) : tx[rj + 1][lane];
            const T xe = east ? T(0) : tx[rj + 1][lane + 2];
            const T xnw = (west || j == ny - 1) ? T(0) : tx[rj + 2][lane];
            const T xne = (east || j == ny - 1) ? T(0) : tx[rj + 2][lane + 2];
            const T xsw = (west || j == 0) ? T(0) : tx[rj][lane];
            const T xse = (east || j == 0) ? T(0) : tx[rj][lane + 2];
            T D;
            if (j == 0) {
                if (west)
                    D = bp - a6 * xe - a9 * xne;
                else if (east)
                    D = bp - a4 * xw - a7 * xnw;
                else
                    D = bp - a4 * xw - a7 * xnw - a6 * xe - a9 * xne;
                Pm = -a8 / a5;
                Qm = D / a5;
            } else {
                if (j < ny - 1) {
                    if (west)
                        D = bp - a3 * xse - a6 * xe - a9 * xne;
                    else if (east)
                        D = bp - a1 * xsw - a4 * xw - a7 * xnw;
                    else
                        D = bp - a1 * xsw - a4 * xw - a7 * xnw - a3 * xse - a6 * xe - a9 * xne;
                } else {
                    if (west)
                        D = bp - a3 * xse - a6 * xe;
                    else if (east)
                        D = bp - a1 * xsw - a4 * xw;
                    else
                        D = bp - a1 * xsw - a4 * xw - a3 * xse - a6 * xe;
                }
                const T Pn = -a8 / (a5 + a2 * Pm);
                const T Qn = (D - a2 * Qm) / (a5 + a2 * Pm);
                Pm = Pn;
                Qm = Qn;
            }
            Pw[p] = Pm;
            Qw[p] = Qm;
        }
    }
    if (!col_ok) return;
    // back substitution: tdma2d9.f90:298-316
    long p = (long)(ny - 1) * nx + i;
    T xv = Qw[p];
    xout[p] = xv;
#pragma unroll 8
    for (int j = ny - 2; j >= 0; --j) {
        p = (long)j * nx + i;
        xv = xv * Pw[p] + Qw[p];
        xout[p] = xv;
    }
}

template <typename T>
void launch_tdma_iteration(const T *a, const T *b, T *x, T *tmp, T *Pw, T *Qw, int nx, int ny,
                           hipStream_t s)
{
    hipLaunchKernelGGL(k_tdma_x<T>, dim3((ny + kTile - 1) / kTile), dim3(kLoad), 0, s, a, b, x,
                       tmp, Pw, Qw, nx, ny);
    hipLaunchKernelGGL(k_tdma_y<T>, dim3((nx + kTile - 1) / kTile), dim3(kLoad), 0, s, a, b, tmp,
                       x, Pw, Qw, nx, ny);
}
template void launch_tdma_iteration<double>(const double *, const double *, double *, double *,
                                            double *, double *, int, int, hipStream_t);

// ---------------------------------------------------------------------------
// L1 norms over real volumes (mod_solvers.f90:128-276): per-block partial
// sums in a fixed order, then one block finishes.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_norm_l1_partial(const double *__restrict__ a,
                                                         const double *__restrict__ b,
                                                         const double *__restrict__ var, int nx,
                                                         int ny, double *__restrict__ pa,
                                                         double *__restrict__ pb)
{
    const long nn = (long)nx * ny;
    const int i = blockIdx.x * 256 + threadIdx.x + 1;  // real columns 1..nx-2 (0-based)
    const int j = blockIdx.y + 1;                      // real rows 1..ny-2
    double ra = 0.0, rbv = 0.0;
    if (i <= nx - 2) {
        const long p = (long)j * nx + i;
        const double s = a[kSW * nn + p] * var[p - nx - 1] + a[kS * nn + p] * var[p - nx] +
                         a[kSE * nn + p] * var[p - nx + 1] + a[kW * nn + p] * var[p - 1] +
                         a[kP * nn + p] * var[p] + a[kE * nn + p] * var[p + 1] +
                         a[kNW * nn + p] * var[p + nx - 1] + a[kN * nn + p] * var[p + nx] +
                         a[kNE * nn + p] * var[p + nx + 1] - b[p];
        ra = fabs(s);
        rbv = fabs(b[p]);
    }
    __shared__ double sa[256], sb[256];
    sa[threadIdx.x] = ra;
    sb[threadIdx.x] = rbv;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sa[threadIdx.x] += sa[threadIdx.x + o];
            sb[threadIdx.x] += sb[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const long q = (long)blockIdx.y * gridDim.x + blockIdx.x;
        pa[q] = sa[0];
        pb[q] = sb[0];
    }
}

__global__ __launch_bounds__(256) void k_sum2(const double *__restrict__ pa,
                                              const double *__restrict__ pb, int np,
                                              double *__restrict__ out)
{
    __shared__ double sa[256], sb[256];
    double x = 0.0, y = 0.0;
    for (int q = threadIdx.x; q < np; q += 256) {
        x += pa[q];
        y += pb[q];
    }
    sa[threadIdx.x] = x;
    sb[threadIdx.x] = y;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sa[threadIdx.x] += sa[threadIdx.x + o];
            sb[threadIdx.x] += sb[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sa[0];
        out[1] = sb[0];
    }
}

// out[0] = sum |A.var - b|, out[1] = sum |b| over real volumes.
int norm_l1_device(const double *a, const double *b, const double *var, int nx, int ny,
                   double *d_work, double *h_out, hipStream_t s)
{
    const int gx = (nx - 2 + 255) / 256, gy = ny - 2;
    const int np = gx * gy;
    double *pa = d_work, *pb = d_work + np, *res = d_work + 2 * np;
    hipLaunchKernelGGL(k_norm_l1_partial, dim3(gx, gy), dim3(256), 0, s, a, b, var, nx, ny, pa,
                       pb);
    hipLaunchKernelGGL(k_sum2, dim3(1), dim3(256), 0, s, pa, pb, np, res);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        fprintf(stderr, "mach2d_hip: norm kernels failed to launch: %s\n", hipGetErrorString(e));
        return M2D_ERR_HIP;
    }
    if (hipMemcpyAsync(h_out, res, 2 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
        return M2D_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return M2D_ERR_HIP;
    return M2D_OK;
}

int norm_l1_work_size(int nx, int ny) { return 2 * ((nx - 2 + 255) / 256) * (ny - 2) + 2; }

}  // namespace m2d

using namespace m2d;

// ---------------------------------------------------------------------------
// C ABI: TDMA and norms
// ---------------------------------------------------------------------------
void m2d_set_last_status_code(int status);

extern "C" int m2d_tdma2d9_device(int nx, int ny, int nitm, const double *d_a, const double *d_b,
                                  double *d_x, void *stream)
{
    if (nx < 3 || ny < 3 || !d_a || !d_b || !d_x) return M2D_ERR_ARG;
    const long nn = (long)nx * ny;
    double *work = nullptr;
    if (hipMalloc((void **)&work, sizeof(double) * 3 * nn) != hipSuccess) return M2D_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    for (int it = 0; it < nitm; ++it)
        launch_tdma_iteration<double>(d_a, d_b, d_x, work, work + nn, work + 2 * nn, nx, ny, s);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(work);
    return e == hipSuccess ? M2D_OK : M2D_ERR_HIP;
}

extern "C" void m2d_tdma2d9(const int *nx, const int *ny, const int *nitm, const double *a,
                            const double *b, double *x)
{
    if (!nx || !ny || !nitm || !a || !b || !x || *nx < 3 || *ny < 3) {
        m2d_set_last_status_code(M2D_ERR_ARG);
        return;
    }
    const long N = (long)*nx * *ny;
    double *d = nullptr;
    if (hipMalloc((void **)&d, sizeof(double) * 11 * N) != hipSuccess) {
        m2d_set_last_status_code(M2D_ERR_HIP);
        return;
    }
    double *da = d, *db = d + 9 * N, *dx = d + 10 * N;
    int rc = M2D_OK;
    if (hipMemcpy(da, a, sizeof(double) * 9 * N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db, b, sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dx, x, sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess)
        rc = M2D_ERR_HIP;
    if (!rc) rc = m2d_tdma2d9_device(*nx, *ny, *nitm, da, db, dx, nullptr);
    if (!rc && hipMemcpy(x, dx, sizeof(double) * N, hipMemcpyDeviceToHost) != hipSuccess)
        rc = M2D_ERR_HIP;
    (void)hipFree(d);
    m2d_set_last_status_code(rc);
}

static int host_norm(int nx, int ny, const double *var, const double *b, const double *a,
                     double out[2])
{
    const long nn = (long)nx * ny;
    const int nw = norm_l1_work_size(nx, ny);
    double *d = nullptr;
    if (hipMalloc((void **)&d, sizeof(double) * (11 * nn + nw)) != hipSuccess) return M2D_ERR_HIP;
    double *da = d, *db = d + 9 * nn, *dv = d + 10 * nn, *dw = d + 11 * nn;
    int rc = M2D_OK;
    if (hipMemcpy(da, a, sizeof(double) * 9 * nn, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db, b, sizeof(double) * nn, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dv, var, sizeof(double) * nn, hipMemcpyHostToDevice) != hipSuccess)
        rc = M2D_ERR_HIP;
    if (!rc) rc = norm_l1_device(da, db, dv, nx, ny, dw, out, nullptr);
    (void)hipFree(d);
    return rc;
}

extern "C" void m2d_norm_l1_9d(const int *nx, const int *ny, const double *var, const double *b,
                               const double *a, double *norm)
{
    double out[2] = {0.0, 0.0};
    if (!nx || !ny || !var || !b || !a || !norm || *nx < 3 || *ny < 3) {
        m2d_set_last_status_code(M2D_ERR_ARG);
        return;
    }
    int rc = host_norm(*nx, *ny, var, b, a, out);
    if (!rc) *norm = *norm + out[0];
    m2d_set_last_status_code(rc);
}

extern "C" void m2d_norm_l1_9d_relative(const int *nx, const int *ny, const double *var,
                                        const double *b, const double *a, double *norm)
{
    double out[2] = {0.0, 0.0};
    if (!nx || !ny || !var || !b || !a || !norm || *nx < 3 || *ny < 3) {
        m2d_set_last_status_code(M2D_ERR_ARG);
        return;
    }
    int rc = host_norm(*nx, *ny, var, b, a, out);
    const double eps = 10.0 * 2.220446049250313080847e-16;  // 10*epsilon(1d0)
    if (!rc) *norm = (out[1] > eps) ? *norm + out[0] / out[1] : *norm + out[0];
    m2d_set_last_status_code(rc);
}

extern "C" double m2d_norm_l1_b(const int *nx, const int *ny, const double *b)
{
    if (!nx || !ny || !b || *nx < 3 || *ny < 3) {
        m2d_set_last_status_code(M2D_ERR_ARG);
        return 0.0;
    }
    const long N = (long)*nx * *ny;
    // |A.0 - b| summed is the L1 norm of b: reuse the kernel with a zero var
    double *z = (double *)calloc((size_t)N, sizeof(double));
    double *za = (double *)calloc((size_t)(9 * N), sizeof(double));
    double out[2] = {0.0, 0.0};
    int rc = (z && za) ? host_norm(*nx, *ny, z, b, za, out) : M2D_ERR_ARG;
    free(z);
    free(za);
    m2d_set_last_status_code(rc);
    return out[1];
}
