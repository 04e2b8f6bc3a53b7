// grid.hip -- the grid centroids, metric terms and radii of
// src/grid/mod_grid.f90 on the GPU: get_real_centroids_xy (:237-300),
// get_metrics (:304-558) and get_radius (:562-723), the three steps of
// grid_create after the corner coordinates are known (:174-187).
//
// The reference runs each output over its own loop nest (the one-sided
// boundary differences and the corner radii after the rest); every value it
// writes depends only on the inputs and on values of the same call that
// are themselves closed forms of the inputs, so one thread per point
// computes every output of that point, with the reference's operation
// order and no contraction (bit for bit).  Every point of every output is
// written, as the reference zero- (or one-) initialises its outputs.
//
// Memory: HBM-bound; per point the metrics read x, y, xp, yp (neighbours
// from the same or adjacent lines of a 256-wide row tile) and write 15
// N-vectors: 19 x 8 bytes; centroids 4 x 8, radii 6 x 8.
//
// Arrays are host or device memory (all of one kind).  Device arrays are
// processed on the assembly stream and the call returns once the kernel is
// enqueued; host arrays are copied to the device and back and the call
// returns with the results in place.  Status: m2d_last_status.
#include <math.h>

#include "m2d_internal.h"

using namespace m2d;

namespace {

constexpr int kT = 256;

struct MetricOut {
    double *xe, *ye, *xen, *yen, *xk, *yk, *xke, *yke, *Jp, *Je, *Jn, *alphae, *gamman, *betae,
        *betan;
};

struct RadiusOut {
    double *radius, *re, *rn, *rp;
};

__device__ __forceinline__ long at(int i, int j, int nx) { return (long)nx * (j - 1) + i - 1; }

// get_real_centroids_xy: opt 1 mean of the four corners, otherwise the
// area-weighted centroids of the triangles (SW, SE, NE) and (SW, NE, NW).
__global__ __launch_bounds__(kT) void k_centroids(int opt, int nx, int ny,
                                                 const double *__restrict__ x,
                                                 const double *__restrict__ y,
                                                 double *__restrict__ xp,
                                                 double *__restrict__ yp)
{
    const int i = blockIdx.x * kT + threadIdx.x + 1, j = blockIdx.y + 1;
    if (i > nx) return;
    const long np = at(i, j, nx);
    double cx = 0.0, cy = 0.0;
    if (i >= 2 && i <= nx - 1 && j >= 2 && j <= ny - 1) {
        const long nps = np - nx, npw = np - 1, npsw = nps - 1;
        if (opt == 1) {
            cx = (x[np] + x[npw] + x[npsw] + x[nps]) / 4.0;
            cy = (y[np] + y[npw] + y[npsw] + y[nps]) / 4.0;
        } else {
            const double A1 = 0.5 * (x[npsw] * (y[nps] - y[np]) + x[nps] * (y[np] - y[npsw]) +
                                     x[np] * (y[npsw] - y[nps]));
            const double A2 = 0.5 * (x[npsw] * (y[np] - y[npw]) + x[np] * (y[npw] - y[npsw]) +
                                     x[npw] * (y[npsw] - y[np]));
            cx = ((x[npsw] + x[nps] + x[np]) * A1 + (x[npsw] + x[np] + x[npw]) * A2) /
                 (3.0 * (A1 + A2));
            cy = ((y[npsw] + y[nps] + y[np]) * A1 + (y[npsw] + y[np] + y[npw]) * A2) /
                 (3.0 * (A1 + A2));
        }
    }
    xp[np] = cx;
    yp[np] = cy;
}

// csi-derivative at the east face of (i, j), 2 <= j <= ny-1: centroid
// difference inside, second-order one-sided differences from the face
// midpoints at i = 1 (with the volume (2, j)) and i = nx-1.
__device__ __forceinline__ double d_csi_east(const double *__restrict__ c,
                                             const double *__restrict__ cp, int i, int j, int nx)
{
    if (i == 1) {
        const long np = at(2, j, nx), nps = np - nx, npw = np - 1, npsw = nps - 1;
        const double fw = (c[npw] + c[npsw]) / 2.0, fe = (c[np] + c[nps]) / 2.0;
        return -3.0 * fw + 4.0 * cp[np] - fe;
    }
    const long np = at(i, j, nx);
    if (i == nx - 1) {
        const long nps = np - nx, npw = np - 1, npsw = nps - 1;
        const double fw = (c[npw] + c[npsw]) / 2.0, fe = (c[np] + c[nps]) / 2.0;
        return fw - 4.0 * cp[np] + 3.0 * fe;
    }
    return cp[np + 1] - cp[np];
}

// eta-derivative at the north face of (i, j), 2 <= i <= nx-1 (j = 1 with
// the volume (i, 2); j = ny-1 one-sided; centroid difference inside)
__device__ __forceinline__ double d_eta_north(const double *__restrict__ c,
                                              const double *__restrict__ cp, int i, int j, int nx,
                                              int ny)
{
    if (j == 1) {
        const long np = at(i, 2, nx), nps = np - nx, npw = np - 1, npsw = nps - 1;
        const double fs = (c[nps] + c[npsw]) / 2.0, fn = (c[np] + c[npw]) / 2.0;
        return -3.0 * fs + 4.0 * cp[np] - fn;
    }
    const long np = at(i, j, nx);
    if (j == ny - 1) {
        const long nps = np - nx, npw = np - 1, npsw = nps - 1;
        const double fs = (c[nps] + c[npsw]) / 2.0, fn = (c[np] + c[npw]) / 2.0;
        return fs - 4.0 * cp[np] + 3.0 * fn;
    }
    return cp[np + nx] - cp[np];
}

__global__ __launch_bounds__(kT) void k_metrics(int nx, int ny, const double *__restrict__ x,
                                               const double *__restrict__ y,
                                               const double *__restrict__ xp,
                                               const double *__restrict__ yp, MetricOut M)
{
    const int i = blockIdx.x * kT + threadIdx.x + 1, j = blockIdx.y + 1;
    if (i > nx) return;
    const long np = at(i, j, nx);
    double xe = 0.0, ye = 0.0, alphae = 0.0, xke = 0.0, yke = 0.0, betae = 0.0, Je = 0.0;
    double xk = 0.0, yk = 0.0, gamman = 0.0, xen = 0.0, yen = 0.0, betan = 0.0, Jn = 0.0;
    double Jp = 0.0;
    if (j >= 2 && j <= ny - 1 && i <= nx - 1) {  // east face
        xe = x[np] - x[np - nx];
        ye = y[np] - y[np - nx];
        alphae = xe * xe + ye * ye;
        xke = d_csi_east(x, xp, i, j, nx);
        yke = d_csi_east(y, yp, i, j, nx);
        betae = xke * xe + yke * ye;
        Je = fabs(1.0 / (xke * ye - xe * yke));
    }
    if (j <= ny - 1 && i >= 2 && i <= nx - 1) {  // north face
        xk = x[np] - x[np - 1];
        yk = y[np] - y[np - 1];
        gamman = xk * xk + yk * yk;
        xen = d_eta_north(x, xp, i, j, nx, ny);
        yen = d_eta_north(y, yp, i, j, nx, ny);
        betan = xk * xen + yk * yen;
        Jn = fabs(1.0 / (xk * yen - xen * yk));
    }
    if (i >= 2 && i <= nx - 1 && j >= 2 && j <= ny - 1) {  // volume centre
        const long nps = np - nx, npw = np - 1, npsw = nps - 1;
        double fw = (x[npw] + x[npsw]) / 2.0, fe = (x[np] + x[nps]) / 2.0;
        double fs = (x[nps] + x[npsw]) / 2.0, fn = (x[np] + x[npw]) / 2.0;
        const double xkp = fe - fw, xep = fn - fs;
        fw = (y[npw] + y[npsw]) / 2.0;
        fe = (y[np] + y[nps]) / 2.0;
        fs = (y[nps] + y[npsw]) / 2.0;
        fn = (y[np] + y[npw]) / 2.0;
        const double ykp = fe - fw, yep = fn - fs;
        Jp = fabs(1.0 / (xkp * yep - xep * ykp));
    }
    M.xe[np] = xe;
    M.ye[np] = ye;
    M.xen[np] = xen;
    M.yen[np] = yen;
    M.xk[np] = xk;
    M.yk[np] = yk;
    M.xke[np] = xke;
    M.yke[np] = yke;
    M.Jp[np] = Jp;
    M.Je[np] = Je;
    M.Jn[np] = Jn;
    M.alphae[np] = alphae;
    M.gamman[np] = gamman;
    M.betae[np] = betae;
    M.betan[np] = betan;
}

// radius at the centre of the east face (i = 1..nx-1, j = 2..ny-1) and of
// the north face (i = 2..nx-1, j = 1..ny-1); 1 elsewhere
__device__ __forceinline__ double r_east(const double *__restrict__ y, int i, int j, int nx,
                                         int ny)
{
    if (i >= 1 && i <= nx - 1 && j >= 2 && j <= ny - 1) {
        const long np = at(i, j, nx);
        return (y[np] + y[np - nx]) / 2.0;
    }
    return 1.0;
}

__device__ __forceinline__ double r_north(const double *__restrict__ y, int i, int j, int nx,
                                          int ny)
{
    if (i >= 2 && i <= nx - 1 && j >= 1 && j <= ny - 1) {
        const long np = at(i, j, nx);
        return (y[np] + y[np - 1]) / 2.0;
    }
    return 1.0;
}

// rp of a point that is not a corner: the centroid's y inside, the
// reflection of the neighbouring centroid about the face centre on the
// boundary
__device__ __forceinline__ double r_side(const double *__restrict__ y,
                                         const double *__restrict__ yp, int i, int j, int nx,
                                         int ny)
{
    if (i >= 2 && i <= nx - 1) {
        if (j == 1) {
            const double r = r_north(y, i, 1, nx, ny);
            return r - yp[at(i, 2, nx)] + r;
        }
        if (j == ny) {
            const double r = r_north(y, i, ny - 1, nx, ny);
            return r - yp[at(i, ny - 1, nx)] + r;
        }
        return yp[at(i, j, nx)];
    }
    if (i == 1) {
        const double r = r_east(y, 1, j, nx, ny);
        return r - yp[at(2, j, nx)] + r;
    }
    const double r = r_east(y, nx - 1, j, nx, ny);
    return r - yp[at(nx - 1, j, nx)] + r;
}

__global__ __launch_bounds__(kT) void k_radius(int coord, int nx, int ny,
                                              const double *__restrict__ y,
                                              const double *__restrict__ yp, RadiusOut R)
{
    const int i = blockIdx.x * kT + threadIdx.x + 1, j = blockIdx.y + 1;
    if (i > nx) return;
    const long np = at(i, j, nx);
    if (coord != 1) {
        R.radius[np] = 1.0;
        R.re[np] = 1.0;
        R.rn[np] = 1.0;
        R.rp[np] = 1.0;
        return;
    }
    R.radius[np] = y[np];
    R.re[np] = r_east(y, i, j, nx, ny);
    R.rn[np] = r_north(y, i, j, nx, ny);
    const bool ci = i == 1 || i == nx, cj = j == 1 || j == ny;
    double rp;
    if (!(ci && cj)) {
        rp = r_side(y, yp, i, j, nx, ny);
    } else {
        // corner: 4 y of the real corner node minus the three adjacent rp
        const int ii = i == 1 ? 2 : nx - 1, jj = j == 1 ? 2 : ny - 1;
        const long node = at(i == 1 ? 1 : nx - 1, j == 1 ? 1 : ny - 1, nx);
        const double a = r_side(y, yp, ii, j, nx, ny), b = r_side(y, yp, i, jj, nx, ny),
                     c = r_side(y, yp, ii, jj, nx, ny);
        rp = 4.0 * y[node] - a - b - c;  // row neighbour, column neighbour, diagonal
    }
    R.rp[np] = rp;
}

// device staging of host arrays: n_in inputs, n_out outputs of N values
struct Stage {
    double *buf = nullptr;
    ~Stage()
    {
        if (buf) (void)hipFree(buf);
    }
};

// checks the arguments; *dev: device arrays (else host arrays to stage)
int check_arrays(int nx, int ny, int n_in, const double *const *in, int n_out,
                 double *const *out, bool *dev_out)
{
    if (nx < 3 || ny < 3) return M2D_ERR_ARG;
    for (int k = 0; k < n_in; ++k)
        if (!in[k]) return M2D_ERR_ARG;
    for (int k = 0; k < n_out; ++k)
        if (!out[k]) return M2D_ERR_ARG;
    const bool dev = is_device_array(in[0]);
    for (int k = 0; k < n_in; ++k)
        if (is_device_array(in[k]) != dev) return M2D_ERR_ARG;
    for (int k = 0; k < n_out; ++k)
        if (is_device_array(out[k]) != dev) return M2D_ERR_ARG;
    *dev_out = dev;
    return M2D_OK;
}

template <typename Launch>
int with_arrays(int nx, int ny, int n_in, const double *const *in, int n_out,
                double *const *out, Launch launch)
{
    const double *din[4];
    double *dout[15];
    bool dev = false;
    const int rc = check_arrays(nx, ny, n_in, in, n_out, out, &dev);
    if (rc) return rc;
    const hipStream_t s = assembly_stream_current();
    const long N = (long)nx * ny;
    const dim3 grid((unsigned)((nx + kT - 1) / kT), (unsigned)ny);
    for (int k = 0; k < n_in; ++k) din[k] = in[k];
    for (int k = 0; k < n_out; ++k) dout[k] = out[k];
    if (dev) {  // enqueue only
        launch(grid, s, din, dout);
        return hipGetLastError() == hipSuccess ? M2D_OK : M2D_ERR_HIP;
    }
    Stage S;
    if (hipMalloc((void **)&S.buf, sizeof(double) * N * (n_in + n_out)) != hipSuccess)
        return M2D_ERR_HIP;
    for (int k = 0; k < n_in; ++k) {
        if (hipMemcpyAsync(S.buf + k * N, in[k], sizeof(double) * N, hipMemcpyHostToDevice, s) !=
            hipSuccess)
            return M2D_ERR_HIP;
        din[k] = S.buf + k * N;
    }
    for (int k = 0; k < n_out; ++k) dout[k] = S.buf + (n_in + k) * N;
    launch(grid, s, din, dout);
    if (hipGetLastError() != hipSuccess) return M2D_ERR_HIP;
    for (int k = 0; k < n_out; ++k)
        if (hipMemcpyAsync(out[k], dout[k], sizeof(double) * N, hipMemcpyDeviceToHost, s) !=
            hipSuccess)
            return M2D_ERR_HIP;
    return hipStreamSynchronize(s) == hipSuccess ? M2D_OK : M2D_ERR_HIP;
}

}  // namespace

void m2d_set_last_status_code(int status);

// get_real_centroids_xy, mod_grid.f90:237-300
extern "C" void m2d_get_real_centroids_xy(const int *opt, const int *nx, const int *ny,
                                          const double *x, const double *y, double *xp,
                                          double *yp)
{
    if (!opt || !nx || !ny) {
        m2d_set_last_status_code(M2D_ERR_ARG);
        return;
    }
    const int o = *opt, NX = *nx, NY = *ny;
    const double *in[] = {x, y};
    double *out[] = {xp, yp};
    const int rc = with_arrays(NX, NY, 2, in, 2, out,
                               [&](dim3 g, hipStream_t s, const double **i, double **w) {
                                   hipLaunchKernelGGL(k_centroids, g, dim3(kT), 0, s, o, NX, NY,
                                                      i[0], i[1], w[0], w[1]);
                               });
    m2d_set_last_status_code(rc);
}

// get_metrics, mod_grid.f90:304-558
extern "C" void m2d_get_metrics(const int *nx, const int *ny, const double *x, const double *y,
                                const double *xp, const double *yp, double *xe, double *ye,
                                double *xen, double *yen, double *xk, double *yk, double *xke,
                                double *yke, double *Jp, double *Je, double *Jn, double *alphae,
                                double *gamman, double *betae, double *betan)
{
    if (!nx || !ny) {
        m2d_set_last_status_code(M2D_ERR_ARG);
        return;
    }
    const int NX = *nx, NY = *ny;
    const double *in[] = {x, y, xp, yp};
    double *out[] = {xe, ye, xen, yen, xk, yk, xke, yke, Jp, Je, Jn, alphae, gamman, betae, betan};
    const int rc = with_arrays(NX, NY, 4, in, 15, out,
                               [&](dim3 g, hipStream_t s, const double **i, double **w) {
                                   MetricOut M = {w[0], w[1], w[2],  w[3],  w[4],
                                                  w[5], w[6], w[7],  w[8],  w[9],
                                                  w[10], w[11], w[12], w[13], w[14]};
                                   hipLaunchKernelGGL(k_metrics, g, dim3(kT), 0, s, NX, NY, i[0],
                                                      i[1], i[2], i[3], M);
                               });
    m2d_set_last_status_code(rc);
}

// get_radius, mod_grid.f90:562-723 (coord 1: axisymmetric, else all radii 1)
extern "C" void m2d_get_radius(const int *coord, const int *nx, const int *ny, const double *y,
                               const double *yp, double *radius, double *re, double *rn,
                               double *rp)
{
    if (!coord || !nx || !ny) {
        m2d_set_last_status_code(M2D_ERR_ARG);
        return;
    }
    const int c = *coord, NX = *nx, NY = *ny;
    const double *in[] = {y, yp};
    double *out[] = {radius, re, rn, rp};
    const int rc = with_arrays(NX, NY, 2, in, 4, out,
                               [&](dim3 g, hipStream_t s, const double **i, double **w) {
                                   RadiusOut R = {w[0], w[1], w[2], w[3]};
                                   hipLaunchKernelGGL(k_radius, g, dim3(kT), 0, s, c, NX, NY,
                                                      i[0], i[1], R);
                               });
    m2d_set_last_status_code(rc);
}
