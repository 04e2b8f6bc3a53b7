// tstep.hip -- the Scarborough time-step statistics of
// src/tstepper/mod_tstepper_scarborough.f90 on the GPU: the mean convergence
// coefficient of a 9- or 5-diagonal system (get_mean_cc_9d / _5d,
// :260-320), the first time step (tstepper_dt0_scarborough, :40-76) and the
// next one (tstepper_dt_scarborough, :80-123).
//
// The per-point coefficients (sum of |a_nb| over |a_P|, the sums left to
// right as the reference's SUM intrinsic adds them) are computed in parallel
// into a scratch array in the reference's loop order (i outer, j inner); one
// lane then adds them in that order, so the mean is the reference's bit for
// bit.  That chain of dependent additions bounds the call (about 10 cycles
// per point); tstepper_dt runs its four means concurrently.  The minimum of dt0 does not depend on the order and is a parallel
// reduction.  Entries take host or device arrays (all of one kind); they
// return the value, so they synchronise.
#include <math.h>

#include <vector>

#include "m2d_internal.h"
#include "seqsum.h"

using namespace m2d;

namespace {

constexpr int kT = 256;

// term of interior point (i, j) (1-based) -> t[(i-2)*(ny-2) + (j-2)]
template <int ND>
__global__ __launch_bounds__(kT) void k_cc_terms(const double *__restrict__ a, int nx, int ny,
                                                double *__restrict__ t)
{
    const int i = blockIdx.x * kT + threadIdx.x + 2, j = blockIdx.y + 2;
    if (i > nx - 1) return;
    const long N = (long)nx * ny, np = (long)nx * (j - 1) + i - 1;
    double s1, s2, ap;
    if constexpr (ND == 9) {
        s1 = fabs(a[np]) + fabs(a[N + np]) + fabs(a[2 * N + np]) + fabs(a[3 * N + np]);
        s2 = fabs(a[5 * N + np]) + fabs(a[6 * N + np]) + fabs(a[7 * N + np]) +
             fabs(a[8 * N + np]);
        ap = fabs(a[4 * N + np]);
    } else {
        s1 = fabs(a[np]) + fabs(a[N + np]);
        s2 = fabs(a[3 * N + np]) + fabs(a[4 * N + np]);
        ap = fabs(a[2 * N + np]);
    }
    t[(long)(i - 2) * (ny - 2) + (j - 2)] = (s1 + s2) / ap;
}

// m = t(q) + m over the terms in order, then m / n (mean) or m; one
// workgroup per array of n terms (t + blockIdx.x * n)
__global__ __launch_bounds__(kSeqThreads) void k_seq_mean(const double *__restrict__ t, long n,
                                                         double *__restrict__ out, int mean)
{
    __shared__ double lds[kSeqSlots * kSeqChunk];
    const double *__restrict__ tt = t + (long)blockIdx.x * n;
    const double m = seq_sum(tt, n, lds);
    if (threadIdx.x == 0) out[blockIdx.x] = mean ? m / (double)n : m;
}

// per-block minimum of 1/sqrt(Jp (u^2 + v^2)) over interior points
__global__ __launch_bounds__(kT) void k_dt0_partial(const double *__restrict__ Jp,
                                                   const double *__restrict__ u,
                                                   const double *__restrict__ v, int nx, int ny,
                                                   double *__restrict__ part)
{
    const int i = blockIdx.x * kT + threadIdx.x + 2, j = blockIdx.y + 2;
    double m = 1.0e10;
    if (i <= nx - 1) {
        const long np = (long)nx * (j - 1) + i - 1;
        const double aux = 1.0 / sqrt(Jp[np] * (u[np] * u[np] + v[np] * v[np]));
        if (m > aux) m = aux;
    }
    __shared__ double red[kT];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = kT / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o && red[threadIdx.x] > red[threadIdx.x + o])
            red[threadIdx.x] = red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kT) void k_min(const double *__restrict__ part, long n,
                                           double *__restrict__ out)
{
    __shared__ double red[kT];
    double m = 1.0e10;
    for (long q = threadIdx.x; q < n; q += kT)
        if (m > part[q]) m = part[q];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = kT / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o && red[threadIdx.x] > red[threadIdx.x + o])
            red[threadIdx.x] = red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] * 0.25;
}

// device copies of host inputs (or the inputs themselves), plus scratch
struct Staged {
    std::vector<const double *> in;
    double *buf = nullptr;
    ~Staged()
    {
        if (buf) (void)hipFree(buf);
    }
};

int stage(Staged &S, const std::vector<const double *> &ptrs, const std::vector<long> &lens,
          long scratch, double **scratch_out, hipStream_t s)
{
    bool dev = is_device_array(ptrs[0]);
    for (const double *p : ptrs)
        if (!p || is_device_array(p) != dev) return M2D_ERR_ARG;
    long tot = scratch;
    if (!dev)
        for (long n : lens) tot += n;
    if (hipMalloc((void **)&S.buf, sizeof(double) * (size_t)(tot > 0 ? tot : 1)) != hipSuccess)
        return M2D_ERR_HIP;
    long off = 0;
    for (size_t k = 0; k < ptrs.size(); ++k) {
        if (dev) {
            S.in.push_back(ptrs[k]);
            continue;
        }
        if (hipMemcpyAsync(S.buf + off, ptrs[k], sizeof(double) * lens[k], hipMemcpyHostToDevice,
                           s) != hipSuccess)
            return M2D_ERR_HIP;
        S.in.push_back(S.buf + off);
        off += lens[k];
    }
    *scratch_out = S.buf + off;
    return M2D_OK;
}

// means of K coefficient arrays (nd[k] diagonals each): the terms of all
// K into one scratch, then K concurrent in-order sums
int mean_cc(int nx, int ny, int K, const double *const *a, const int *nd, double *out)
{
    if (nx < 3 || ny < 3 || !out) return M2D_ERR_ARG;
    const hipStream_t s = assembly_stream_current();
    const long N = (long)nx * ny, n = (long)(nx - 2) * (ny - 2);
    std::vector<const double *> ptrs(a, a + K);
    std::vector<long> lens;
    for (int k = 0; k < K; ++k) lens.push_back(nd[k] * N);
    Staged S;
    double *t = nullptr;
    int rc = stage(S, ptrs, lens, K * n + K, &t, s);
    if (rc) return rc;
    mean_cc_enqueue(nx, ny, K, S.in.data(), nd, t, s);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(out, t + K * n, sizeof(double) * K, hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return M2D_ERR_HIP;
    return M2D_OK;
}

}  // namespace

void m2d::cc_terms_enqueue(int nx, int ny, int K, const double *const *a, const int *nd,
                           double *t, hipStream_t s)
{
    const long n = (long)(nx - 2) * (ny - 2);
    const dim3 grid((unsigned)((nx - 2 + kT - 1) / kT), (unsigned)(ny - 2));
    for (int k = 0; k < K; ++k) {
        if (nd[k] == 9)
            hipLaunchKernelGGL(k_cc_terms<9>, grid, dim3(kT), 0, s, a[k], nx, ny, t + k * n);
        else
            hipLaunchKernelGGL(k_cc_terms<5>, grid, dim3(kT), 0, s, a[k], nx, ny, t + k * n);
    }
}

void m2d::seq_sums_enqueue(const double *t, long n, int K, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_seq_mean, dim3(K), dim3(kSeqThreads), 0, s, t, n, out, 0);
}

void m2d::mean_cc_enqueue(int nx, int ny, int K, const double *const *a, const int *nd,
                          double *t, hipStream_t s)
{
    const long n = (long)(nx - 2) * (ny - 2);
    cc_terms_enqueue(nx, ny, K, a, nd, t, s);
    hipLaunchKernelGGL(k_seq_mean, dim3(K), dim3(kSeqThreads), 0, s, t, n, t + K * n, 1);
}

// tstepper_dt_scarborough's rule (mod_tstepper_scarborough.f90:80-123) on
// the four means: the largest against mincc and maxcc
double m2d::scarborough_dt(double dt, const double *cc, double h0, double mincc, double maxcc)
{
    double c = cc[0];
    for (int k = 1; k < 4; ++k)
        if (cc[k] > c) c = cc[k];
    double h;
    if (c < mincc)
        h = h0 * sqrt(1.0 - c / mincc);
    else if (maxcc < c)
        h = -h0 * sqrt(c / maxcc - 1.0);
    else
        h = 0.0;
    return dt * (1.0 + h);
}

void m2d_set_last_status_code(int status);

// get_mean_cc_9d (mod_tstepper_scarborough.f90:260-288)
extern "C" double m2d_get_mean_cc_9d(const int *nx, const int *ny, const double *a)
{
    double v = 0.0;
    const int nd = 9;
    const int rc = (!nx || !ny) ? M2D_ERR_ARG : mean_cc(*nx, *ny, 1, &a, &nd, &v);
    m2d_set_last_status_code(rc);
    return v;
}

// get_mean_cc_5d (:292-320)
extern "C" double m2d_get_mean_cc_5d(const int *nx, const int *ny, const double *a)
{
    double v = 0.0;
    const int nd = 5;
    const int rc = (!nx || !ny) ? M2D_ERR_ARG : mean_cc(*nx, *ny, 1, &a, &nd, &v);
    m2d_set_last_status_code(rc);
    return v;
}

// tstepper_dt0_scarborough (:40-76)
extern "C" double m2d_tstepper_dt0_scarborough(const int *nx_, const int *ny_, const double *Jp,
                                               const double *u, const double *v)
{
    double dt = 0.0;
    int rc = M2D_OK;
    if (!nx_ || !ny_ || *nx_ < 3 || *ny_ < 3 || !Jp || !u || !v) rc = M2D_ERR_ARG;
    if (!rc) {
        const int nx = *nx_, ny = *ny_;
        const long N = (long)nx * ny;
        const hipStream_t s = assembly_stream_current();
        const dim3 grid((unsigned)((nx - 2 + kT - 1) / kT), (unsigned)(ny - 2));
        const long np = (long)grid.x * grid.y;
        Staged S;
        double *part = nullptr;
        rc = stage(S, {Jp, u, v}, {N, N, N}, np + 1, &part, s);
        if (!rc) {
            hipLaunchKernelGGL(k_dt0_partial, grid, dim3(kT), 0, s, S.in[0], S.in[1], S.in[2],
                               nx, ny, part);
            hipLaunchKernelGGL(k_min, dim3(1), dim3(kT), 0, s, part, np, part + np);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(&dt, part + np, sizeof(double), hipMemcpyDeviceToHost, s) !=
                    hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                rc = M2D_ERR_HIP;
        }
    }
    m2d_set_last_status_code(rc);
    return dt;
}

// tstepper_dt_scarborough (:80-123); h0, mincc and maxcc are the module
// variables tstepper_scarborough_init reads from the parameter file.
extern "C" double m2d_tstepper_dt_scarborough(const int *nx, const int *ny, const double *dt,
                                              const double *au, const double *av,
                                              const double *at, const double *ap,
                                              const double *h0, const double *mincc,
                                              const double *maxcc)
{
    if (!nx || !ny || !dt || !h0 || !mincc || !maxcc) {
        m2d_set_last_status_code(M2D_ERR_ARG);
        return 0.0;
    }
    const double *a[4] = {au, av, at, ap};
    const int nd[4] = {9, 9, 9, 5};
    double cc[4] = {0.0, 0.0, 0.0, 0.0};
    const int rc = mean_cc(*nx, *ny, 4, a, nd, cc);
    m2d_set_last_status_code(rc);
    if (rc) return 0.0;
    return scarborough_dt(*dt, cc, *h0, *mincc, *maxcc);
}
