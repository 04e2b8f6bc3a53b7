// msi9.hip -- gfx950 kernels of the 9-diagonal MSI path.
//
// Reference routines: lu2d9 and fb2d9, src/solvers/msi2d9-v03.f90:42-432.
//
// Wavefront schedule.  The factorization and the forward sweep of point
// (i,j) read the W, SW, S and SE neighbours (msi2d9-v03.f90:357-366 and
// :155); the backward sweep reads E, NW, N and NE (:175).  Points with the
// same k = i + 2j are independent, so row j can process column i at step
// i + 2j.  One wave of 64 lanes owns a band of 64 consecutive rows, lane l
// holding row band*64+l: at local step t the lane works on column t - 2l.
// The S/SE/SW (or N/NW/NE) values come from lane l-1 (l+1) through a
// one-lane wave shuffle of its latest result plus two registers of history;
// the band's first (last) lane reads the neighbouring band's boundary row
// from global memory after waiting on that band's progress flag
// (agent-scope release/acquire, published every `publish_every` columns).
// Bands only ever wait on bands with a lower launch index, so the waits
// cannot deadlock under in-order workgroup dispatch, and every wait is
// bounded (kMaxSpin polls) so a fault in the protocol ends the kernel with
// an error flag instead of a hang.
//
// The residual (msi2d9-v03.f90:59-131) is a separate fully parallel kernel
// with a fixed-order block reduction of r.r; a one-block kernel finishes the
// 2-norm, records the history and applies the stopping rule (:258-259) on
// the device, so an iteration needs no host round trip.
#include <math.h>

#include "m2d_internal.h"
#include "msi_stencil.h"

namespace m2d {

__device__ __forceinline__ unsigned long long flag_load(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void flag_store(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long flag_word(unsigned gen, int progress)
{
    return ((unsigned long long)gen << 32) | (unsigned)progress;
}

// Wave-uniform, bounded wait until band `src` has published at least `need`
// columns in generation sy.gen.  Returns the progress seen, or -1 when the
// wait gave up (sets st->abort so every other band gives up at once).
__device__ int wait_progress(const SweepSync &sy, int src, int need)
{
    for (int spin = 0; spin < kMaxSpin; ++spin) {
        const unsigned long long f = flag_load(&sy.flags[src]);
        const unsigned g = (unsigned)__builtin_amdgcn_readfirstlane((int)(f >> 32));
        const int have = __builtin_amdgcn_readfirstlane((int)(f & 0xffffffffu));
        if (g == sy.gen && have >= need) return have;
        if (__builtin_amdgcn_readfirstlane(
                __hip_atomic_load(&sy.st->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
            return -1;
        __builtin_amdgcn_s_sleep(1);
    }
    __hip_atomic_store(&sy.st->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return -1;
}

template <typename T>
__device__ __forceinline__ T lane_up(T v)
{
    return __shfl_up(v, 1);
}

template <typename T>
__device__ __forceinline__ T lane_down(T v)
{
    return __shfl_down(v, 1);
}

template <typename T>
__device__ __forceinline__ bool bad_pivot(T d)
{
    return !(d == d) || d == T(0) || isinf(d);
}

// --------------------------------------------------------------------------
// lu2d9: incomplete factorization, wavefront over bands of 64 rows.
// --------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kWave) void k_factor(const T *__restrict__ a, T *__restrict__ dl,
                                                  T *__restrict__ du, int nx, int ny, T alpha0,
                                                  SweepSync sy, int xcd_stride)
{
    const int blk = blockIdx.x;
    if (blk % xcd_stride) return;
    const int band = blk / xcd_stride;
    const int nb = (ny + kWave - 1) / kWave;
    const int lane = threadIdx.x;
    const int j = band * kWave + lane;
    const int rows = min(kWave, ny - band * kWave);
    const long N = (long)nx * ny;
    const bool row_ok = lane < rows;
    const int tend = nx - 1 + 2 * (rows - 1);
    const bool pub = (lane == kWave - 1) && (band < nb - 1);
    const int K = sy.publish_every;
    const long rowbase = (long)j * nx;

    T cur[4] = {0, 0, 0, 0};  // this row's du at the previous column (W)
    T se[4] = {0, 0, 0, 0}, s[4] = {0, 0, 0, 0}, sw[4] = {0, 0, 0, 0};
    int have = (band == 0) ? nx : 0;
    int bad = 0;
    if (band > 0) {  // prime lane 0's history with column 0 of the row above
        have = wait_progress(sy, band - 1, 1);
        if (have < 0) return;
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 4; ++k) se[k] = du[k * N + rowbase - nx];
        }
    }

    for (int t = 0; t <= tend; ++t) {
        const int i = t - 2 * lane;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const T up = lane_up(cur[k]);
            sw[k] = s[k];
            s[k] = se[k];
            se[k] = up;
        }
        if (band > 0) {
            const int need = min(t + 2, nx);
            if (have < need) {
                have = wait_progress(sy, band - 1, need);
                if (have < 0) return;
            }
        }
        if (lane == 0) {
            const bool ok = band > 0 && t + 1 < nx;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                se[k] = ok ? du[k * N + rowbase - nx + t + 1] : T(0);
        }
        if (row_ok && i >= 0 && i < nx) {
            const long p = rowbase + i;
            T b[9], dw[4], ds[4], dse[4], dsw[4], L[5], U[4];
#pragma unroll
            for (int k = 0; k < 9; ++k) b[k] = a[k * N + p];
            const bool hw = i > 0, hs = j > 0, he = i + 1 < nx;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                dw[k] = hw ? cur[k] : T(0);
                ds[k] = hs ? s[k] : T(0);
                dse[k] = (hs && he) ? se[k] : T(0);
                dsw[k] = (hs && hw) ? sw[k] : T(0);
            }
            factor_point<T>(b, alpha_at<T>(i, j, nx, ny, alpha0), dw, ds, dse, dsw, L, U);
#pragma unroll
            for (int k = 0; k < 5; ++k) dl[k * N + p] = L[k];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                du[k * N + p] = U[k];
                cur[k] = U[k];
            }
            bad |= bad_pivot(L[0]);
            if (pub && ((i + 1) % K == 0 || i == nx - 1))
                flag_store(&sy.flags[band], flag_word(sy.gen, i + 1));
        }
    }
    if (bad) __hip_atomic_store(&sy.st->err_pivot, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// --------------------------------------------------------------------------
// fb2d9 forward sweep  L w = -r  (msi2d9-v03.f90:139-156).
// --------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kWave) void k_forward(const T *__restrict__ r,
                                                   const T *__restrict__ dl, T *__restrict__ w,
                                                   int nx, int ny, SweepSync sy, int xcd_stride)
{
    if (sy.st->done) return;
    const int blk = blockIdx.x;
    if (blk % xcd_stride) return;
    const int band = blk / xcd_stride;
    const int nb = (ny + kWave - 1) / kWave;
    const int lane = threadIdx.x;
    const int j = band * kWave + lane;
    const int rows = min(kWave, ny - band * kWave);
    const long N = (long)nx * ny;
    const bool row_ok = lane < rows;
    const int tend = nx - 1 + 2 * (rows - 1);
    const bool pub = (lane == kWave - 1) && (band < nb - 1);
    const int K = sy.publish_every;
    const long rowbase = (long)j * nx;

    T cur = 0, se = 0, s = 0, sw = 0;
    int have = (band == 0) ? nx : 0;
    if (band > 0) {  // prime lane 0's history with column 0 of the row above
        have = wait_progress(sy, band - 1, 1);
        if (have < 0) return;
        if (lane == 0) se = w[rowbase - nx];
    }

    for (int t = 0; t <= tend; ++t) {
        const int i = t - 2 * lane;
        const T up = lane_up(cur);
        sw = s;
        s = se;
        se = up;
        if (band > 0) {
            const int need = min(t + 2, nx);
            if (have < need) {
                have = wait_progress(sy, band - 1, need);
                if (have < 0) return;
            }
        }
        if (lane == 0) se = (band > 0 && t + 1 < nx) ? w[rowbase - nx + t + 1] : T(0);
        if (row_ok && i >= 0 && i < nx) {
            const long p = rowbase + i;
            const bool hw = i > 0, hs = j > 0, he = i + 1 < nx;
            const T v = forward_point<T>(r[p], dl[p], dl[N + p], dl[2 * N + p], dl[3 * N + p],
                                         dl[4 * N + p], hw ? cur : T(0),
                                         (hs && he) ? se : T(0), hs ? s : T(0),
                                         (hs && hw) ? sw : T(0));
            w[p] = v;
            cur = v;
            if (pub && ((i + 1) % K == 0 || i == nx - 1))
                flag_store(&sy.flags[band], flag_word(sy.gen, i + 1));
        }
    }
}

// --------------------------------------------------------------------------
// fb2d9 backward sweep  U z = w  and  var = z + var  (:158-181).  z
// overwrites w in place; bands run top to bottom (band = nb-1-launch index).
// --------------------------------------------------------------------------
template <typename T, typename TX>
__global__ __launch_bounds__(kWave) void k_backward(T *__restrict__ wz, const T *__restrict__ du,
                                                    TX *__restrict__ x, int nx, int ny,
                                                    SweepSync sy, int xcd_stride)
{
    if (sy.st->done) return;
    const int blk = blockIdx.x;
    if (blk % xcd_stride) return;
    const int nb = (ny + kWave - 1) / kWave;
    const int band = nb - 1 - blk / xcd_stride;
    const int lane = threadIdx.x;
    const int j = band * kWave + lane;
    const int rows = min(kWave, ny - band * kWave);
    const int last = rows - 1;
    const long N = (long)nx * ny;
    const bool row_ok = lane < rows;
    const int send = nx - 1 + 2 * last;
    const bool pub = (lane == 0) && (band > 0);
    const int K = sy.publish_every;
    const long rowbase = (long)j * nx;

    T cur = 0, nw = 0, n = 0, ne = 0;
    int have = (band == nb - 1) ? nx : 0;
    if (band < nb - 1) {  // prime the last lane's history with column nx-1 above
        have = wait_progress(sy, band + 1, 1);
        if (have < 0) return;
        if (lane == last) nw = wz[rowbase + nx + (nx - 1)];
    }

    for (int s = 0; s <= send; ++s) {
        const int i = nx - 1 - s + 2 * (last - lane);
        const T dn = lane_down(cur);
        ne = n;
        n = nw;
        nw = dn;
        if (band < nb - 1) {
            const int need = min(s + 2, nx);
            if (have < need) {
                have = wait_progress(sy, band + 1, need);
                if (have < 0) return;
            }
        }
        if (lane == last)
            nw = (band < nb - 1 && nx - 2 - s >= 0) ? wz[rowbase + nx + (nx - 2 - s)] : T(0);
        if (row_ok && i >= 0 && i < nx) {
            const long p = rowbase + i;
            const bool hn = j + 1 < ny, hw = i > 0, he = i + 1 < nx;
            const T v = backward_point<T>(wz[p], du[p], du[N + p], du[2 * N + p], du[3 * N + p],
                                          he ? cur : T(0), (hn && hw) ? nw : T(0),
                                          hn ? n : T(0), (hn && he) ? ne : T(0));
            wz[p] = v;
            x[p] = (TX)v + x[p];
            cur = v;
            if (pub && ((nx - i) % K == 0 || i == 0))
                flag_store(&sy.flags[band], flag_word(sy.gen, nx - i));
        }
    }
}

// --------------------------------------------------------------------------
// Residual r = A.x - rhs (2-D aware, reference term order) and per-block
// partial sums of r.r in a fixed reduction order.
// --------------------------------------------------------------------------
constexpr int kResThreads = 256;

template <typename TA, typename TX, typename TR>
__global__ __launch_bounds__(kResThreads) void k_residual(const TA *__restrict__ a,
                                                          const TX *__restrict__ x,
                                                          const TA *__restrict__ rhs,
                                                          TR *__restrict__ r, int nx, int ny,
                                                          double *__restrict__ partial,
                                                          const SolveState *st)
{
    if (st && st->done) return;
    const int i = blockIdx.x * kResThreads + threadIdx.x;
    const int j = blockIdx.y;
    const long N = (long)nx * ny;
    double acc = 0.0;
    if (i < nx) {
        const long p = (long)j * nx + i;
        const TR v = residual_point<TA, TX, TR>(a, N, x, (TR)rhs[p], i, j, nx, ny);
        r[p] = v;
        acc = (double)v * (double)v;
    }
    __shared__ double red[kResThreads];
    red[threadIdx.x] = acc;
    __syncthreads();
#pragma unroll
    for (int o = kResThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(long)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

int residual_blocks(int nx, int ny)
{
    return ((nx + kResThreads - 1) / kResThreads) * ny;
}

// One block: ||r||_2 from the partial sums, history, stopping rule of
// msi2d9-v03.f90:133,258-259 (k = 0 is the initial residual).
__global__ __launch_bounds__(256) void k_norm_finalize(const double *__restrict__ partial, int np,
                                                       SolveState *st, double *hist,
                                                       int hist_cap, int k, double gamma,
                                                       int nitm)
{
    if (st->done) return;
    __shared__ double red[256];
    double acc = 0.0;
    for (int q = threadIdx.x; q < np; q += 256) acc += partial[q];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double nrm = sqrt(red[0]);
    st->rn = nrm;
    if (k == 0) {
        st->r0 = nrm;
        st->gammarp = gamma * nrm;
        if (!isfinite(nrm)) {
            st->status = M2D_ERR_NONFINITE;
            st->done = 1;
            st->nit = 0;
        }
        return;
    }
    if (hist && k <= hist_cap) hist[k - 1] = nrm;
    if (st->abort) {
        st->status = M2D_ERR_SYNC_TIMEOUT;
        st->done = 1;
        st->nit = k;
    } else if (!isfinite(nrm)) {
        st->status = M2D_ERR_NONFINITE;
        st->done = 1;
        st->nit = k;
    } else if (nrm < st->gammarp) {
        st->done = 1;
        st->nit = k;
    } else if (k == nitm) {
        st->done = 1;
        st->nit = k;
        st->nitm_hit = 1;
    } else if (nitm <= 0 && k >= kIterCap) {
        st->status = M2D_ERR_ITER_CAP;
        st->done = 1;
        st->nit = k;
    }
}

// Count couplings towards neighbours outside the grid (must all be zero).
__global__ __launch_bounds__(256) void k_check_outside(const double *__restrict__ a, int nx,
                                                       int ny, SolveState *st)
{
    const long N = (long)nx * ny;
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    int bad = 0;
    if (p < N) {
        const int j = (int)(p / nx), i = (int)(p - (long)j * nx);
        const bool w = i == 0, e = i == nx - 1, s = j == 0, n = j == ny - 1;
        const bool out[9] = {s || w, s, s || e, w, false, e, n || w, n, n || e};
#pragma unroll
        for (int k = 0; k < 9; ++k) bad += (out[k] && a[k * N + p] != 0.0) ? 1 : 0;
    }
    if (__any(bad)) {
        if (bad) atomicAdd(&st->outside, bad);
    }
}

__global__ void k_zero_flags(unsigned long long *flags, int n, SolveState *st, int reset_state)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) flags[q] = 0ull;
    if (reset_state && q == 0) {
        st->done = 0;
        st->nit = 0;
        st->status = 0;
        st->nitm_hit = 0;
        st->abort = 0;
        st->err_pivot = 0;
        st->outside = 0;
        st->outer = 0;
        st->r0 = 0.0;
        st->gammarp = 0.0;
        st->rn = 0.0;
    }
}

template <typename TI, typename TO>
__global__ void k_convert(const TI *__restrict__ in, TO *__restrict__ out, long n)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = (TO)in[q];
}

// ---- launchers -------------------------------------------------------------
static inline int nbands(int ny) { return (ny + kWave - 1) / kWave; }

template <typename T>
void launch_factor(const T *a, T *dl, T *du, int nx, int ny, T alpha0, SweepSync sync,
                   int xcd_stride, hipStream_t s)
{
    hipLaunchKernelGGL(k_factor<T>, dim3(nbands(ny) * xcd_stride), dim3(kWave), 0, s, a, dl, du,
                       nx, ny, alpha0, sync, xcd_stride);
}

template <typename T>
void launch_forward(const T *r, const T *dl, T *w, int nx, int ny, SweepSync sync,
                    int xcd_stride, hipStream_t s)
{
    hipLaunchKernelGGL(k_forward<T>, dim3(nbands(ny) * xcd_stride), dim3(kWave), 0, s, r, dl, w,
                       nx, ny, sync, xcd_stride);
}

template <typename T, typename TX>
void launch_backward(T *wz, const T *du, TX *x, int nx, int ny, SweepSync sync, int xcd_stride,
                     hipStream_t s)
{
    hipLaunchKernelGGL((k_backward<T, TX>), dim3(nbands(ny) * xcd_stride), dim3(kWave), 0, s, wz,
                       du, x, nx, ny, sync, xcd_stride);
}

template <typename TA, typename TX, typename TR>
void launch_residual(const TA *a, const TX *x, const TA *rhs, TR *r, int nx, int ny,
                     double *partial, const SolveState *st, hipStream_t s)
{
    dim3 grid((nx + kResThreads - 1) / kResThreads, ny);
    hipLaunchKernelGGL((k_residual<TA, TX, TR>), grid, dim3(kResThreads), 0, s, a, x, rhs, r,
                       nx, ny, partial, st);
}

void launch_norm_finalize(const double *partial, int np, SolveState *st, double *hist,
                          int hist_cap, int k, double gamma, int nitm, hipStream_t s)
{
    hipLaunchKernelGGL(k_norm_finalize, dim3(1), dim3(256), 0, s, partial, np, st, hist,
                       hist_cap, k, gamma, nitm);
}

void launch_check_outside(const double *a, int nx, int ny, SolveState *st, hipStream_t s)
{
    const long N = (long)nx * ny;
    hipLaunchKernelGGL(k_check_outside, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, a,
                       nx, ny, st);
}

void launch_zero_flags(unsigned long long *flags, int n, SolveState *st, int reset_state,
                       hipStream_t s)
{
    const int m = n > 1 ? n : 1;
    hipLaunchKernelGGL(k_zero_flags, dim3((m + 255) / 256), dim3(256), 0, s, flags, n, st,
                       reset_state);
}

template <typename TI, typename TO>
void launch_convert(const TI *in, TO *out, long n, hipStream_t s)
{
    hipLaunchKernelGGL((k_convert<TI, TO>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       in, out, n);
}

template void launch_factor<double>(const double *, double *, double *, int, int, double,
                                    SweepSync, int, hipStream_t);
template void launch_factor<float>(const float *, float *, float *, int, int, float, SweepSync,
                                   int, hipStream_t);
template void launch_forward<double>(const double *, const double *, double *, int, int,
                                     SweepSync, int, hipStream_t);
template void launch_forward<float>(const float *, const float *, float *, int, int, SweepSync,
                                    int, hipStream_t);
template void launch_backward<double, double>(double *, const double *, double *, int, int,
                                              SweepSync, int, hipStream_t);
template void launch_backward<float, float>(float *, const float *, float *, int, int,
                                            SweepSync, int, hipStream_t);
template void launch_residual<double, double, double>(const double *, const double *,
                                                      const double *, double *, int, int,
                                                      double *, const SolveState *,
                                                      hipStream_t);
template void launch_residual<float, float, float>(const float *, const float *, const float *,
                                                   float *, int, int, double *,
                                                   const SolveState *, hipStream_t);
template void launch_convert<double, float>(const double *, float *, long, hipStream_t);
template void launch_convert<float, double>(const float *, double *, long, hipStream_t);

}  // namespace m2d
