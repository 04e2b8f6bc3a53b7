// msi9.hip -- fully parallel kernels of the 9-diagonal MSI path: residual
// r = A.var - rhs with the per-block partial sums of r.r (msi2d9-v03.f90:
// 59-131 and :183-255), whose last workgroup forms the norm and applies the
// stopping rule (:133, :258-259), the check of couplings towards neighbours
// outside the grid, and small state/conversion kernels.  The wavefront
// kernels are in sweep.hip.
//
// The stopping rule runs on the device, so an iteration needs no host round
// trip: once it fires, a flag turns the later kernels of the enqueued
// iterations into no-ops.
#include <assert.h>
#include <math.h>

#include "m2d_internal.h"
#include "msi_stencil.h"
#include "skew.h"
#include "wavefront.h"

namespace m2d {

// --------------------------------------------------------------------------
// Residual r = A.x - rhs (2-D aware, reference term order), per-block
// partial sums of r.r in a fixed reduction order, and the norm with the
// stopping rule, formed by the workgroup that finishes last.
// --------------------------------------------------------------------------
constexpr int kResThreads = 256;

// The partial sums travel as tagged words (wavefront.h: two 64-bit words
// per value, a 32-bit generation over 32 bits of payload, relaxed agent-scope
// stores and loads), so no workgroup waits on a fence or counts itself with
// an atomic (one counter taken by every workgroup serializes them).  Word
// 2 np of the buffer holds the generation of the last launch.
long residual_partial_words(int nx, int ny)
{
    return 2L * residual_blocks(nx, ny) + 1;
}

// One workgroup of 256 threads: ||r||_2 from the np partial sums (each
// thread adds every 256th partial, waiting for its tag, then a tree),
// history, stopping rule of msi2d9-v03.f90:133,258-259 (k = 0 is the
// initial residual).  A partial sum that never arrives ends the solve with
// M2D_ERR_SYNC_TIMEOUT.
__device__ void norm_rule_block(const u64 *__restrict__ pw, int np, unsigned tag, SolveState *st,
                                const NormRule &rule)
{
    const int k = rule.initial ? 0 : st->iter + 1;
    __shared__ double red[256];
    __shared__ int lost;
    if (threadIdx.x == 0) lost = 0;
    __syncthreads();
    // partials t, t + 256, ... in that order, loaded G at a time
    constexpr int G = 8;
    double acc = 0.0;
    for (int q0 = threadIdx.x; q0 < np; q0 += 256 * G) {
        Raw<double> v[G];
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (q0 + 256 * g < np) raw_load<double>(v[g], pw + 2L * (q0 + 256 * g));
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int q = q0 + 256 * g;
            if (q >= np) break;
            for (int spin = 0; !raw_ok<double>(v[g], tag); ++spin) {
                if (spin == kMaxSpin) {
                    lost = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                raw_load<double>(v[g], pw + 2L * q);
            }
            acc += raw_value(v[g]);
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double nrm = sqrt(red[0]);
    st->rn = nrm;
    st->iter = k;
    if (lost) {
        st->abort = 1;
        st->status = M2D_ERR_SYNC_TIMEOUT;
        st->done = 1;
        st->nit = k;
        return;
    }
    if (k == 0) {
        st->r0 = nrm;
        st->gammarp = rule.gamma * nrm;
        if (!isfinite(nrm)) {
            st->status = M2D_ERR_NONFINITE;
            st->done = 1;
            st->nit = 0;
        }
        return;
    }
    if (rule.hist && k <= rule.hist_cap) rule.hist[k - 1] = nrm;
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
    } else if (k == rule.nitm) {
        st->done = 1;
        st->nit = k;
        st->nitm_hit = 1;
    } else if (rule.nitm <= 0 && k >= kIterCap) {
        st->status = M2D_ERR_ITER_CAP;
        st->done = 1;
        st->nit = k;
    }
}

// Blocks are row segments of kResThreads columns, G = nbx * ny of them in
// the grid order (segment fastest).  Workgroup g runs on XCD g mod 8; it
// takes segment (g mod 8) * per + g / 8 (per = ceil(G / 8)), so each XCD
// works through a contiguous run of rows and the neighbour rows a stencil
// reads meet in its L2.  Partial sums keep the grid order.
//
// Every workgroup stores its partial sum under the launch's generation.  The
// grid's last workgroup (dispatched last; it holds one slot while it waits,
// so the others run whatever the dispatch order) then waits for every sum,
// adds them, applies the stopping rule and records the generation.
//
// Upd: the update var = z + var (msi2d9-v03.f90:181) comes first, in the
// same launch.  The iterate alternates between x (an even number of
// iterations done) and x2 (odd): the kernel reads one and writes the other,
// forming each stencil neighbour's updated value z + var where it reads it
// (z skewed by osk), the sum the update stores, so r is bit for bit the
// residual of the updated iterate.  settle returns an iterate left in x2.
template <typename TA, typename TX, typename TR, int ND, bool Upd>
__global__ __launch_bounds__(kResThreads) void k_residual(const TA *__restrict__ a,
                                                          TX *__restrict__ x, TX *__restrict__ x2,
                                                          const TR *__restrict__ z,
                                                          const TA *__restrict__ rhs,
                                                          TR *__restrict__ r, int nx, int ny,
                                                          u64 *__restrict__ pw, SolveState *st,
                                                          Skew osk, int nbx, int per,
                                                          NormRule rule)
{
    if (st->done) return;
    const int lg = (int)(blockIdx.x % 8u) * per + (int)(blockIdx.x / 8u);
    const int np = nbx * ny;
    // the previous launch's last workgroup wrote it; this launch's rewrites
    // it only after every workgroup's sum has arrived
    const unsigned tag = (unsigned)tag_load(pw + 2L * np) + 1u;
    if (lg < np) {  // uniform over the workgroup
        const int j = lg / nbx;
        const int i = (lg - j * nbx) * kResThreads + threadIdx.x;
        const long N = (long)nx * ny;
        double acc = 0.0;
        if (i < nx) {
            const long p = (long)j * nx + i;
            TR v;
            if constexpr (Upd) {
                const bool odd = st->iter & 1;
                const TX *__restrict__ xin = odd ? x2 : x;
                TX *__restrict__ xout = odd ? x : x2;
                auto xv = [&](int ii, int jj, long f) -> TR {
                    return z[osk.template of_point_c<64 / (int)sizeof(TR)>(ii, jj)] + (TR)xin[f];
                };
                v = ND == 5 ? residual_point5_f<TA, TR>(a, N, xv, (TR)rhs[p], i, j, nx, ny)
                            : residual_point_f<TA, TR>(a, N, xv, (TR)rhs[p], i, j, nx, ny);
                xout[p] = (TX)xv(i, j, p);
            } else {
                v = ND == 5 ? residual_point5<TA, TX, TR>(a, N, x, (TR)rhs[p], i, j, nx, ny)
                            : residual_point<TA, TX, TR>(a, N, x, (TR)rhs[p], i, j, nx, ny);
            }
            // skewed output when osk.S > 0 (chunk length 64 / sizeof(TR))
            r[osk.S ? osk.template of_point_c<64 / (int)sizeof(TR)>(i, j) : p] = v;
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
        if (threadIdx.x == 0) put_tagged(pw + 2L * lg, tag, red[0]);
    }
    if (blockIdx.x != gridDim.x - 1) return;
    norm_rule_block(pw, np, tag, st, rule);
    if (threadIdx.x == 0)
        __hip_atomic_store(pw + 2L * np, (u64)tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// After an MSI run: an iterate left in x2 (an odd number of iterations)
// goes back to x.
template <typename T>
__global__ __launch_bounds__(256) void k_settle(T *__restrict__ x, const T *__restrict__ x2,
                                                long n, const SolveState *st)
{
    if (!(st->iter & 1)) return;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) x[q] = x2[q];
}

int residual_blocks(int nx, int ny)
{
    return ((nx + kResThreads - 1) / kResThreads) * ny;
}

// Count couplings towards neighbours outside the grid (must all be zero);
// nd = 9 (msi2d9 slots) or 5 (S, W, P, E, N).
__global__ __launch_bounds__(256) void k_check_outside(const double *__restrict__ a, int nx,
                                                       int ny, SolveState *st, int nd)
{
    if (st->skip) return;
    const long N = (long)nx * ny;
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    int bad = 0;
    if (p < N) {
        const int j = (int)(p / nx), i = (int)(p - (long)j * nx);
        const bool w = i == 0, e = i == nx - 1, s = j == 0, n = j == ny - 1;
        if (nd == 5) {
            const bool out[5] = {s, w, false, e, n};
#pragma unroll
            for (int k = 0; k < 5; ++k) bad += (out[k] && a[k * N + p] != 0.0) ? 1 : 0;
        } else {
            const bool out[9] = {s || w, s, s || e, w, false, e, n || w, n, n || e};
#pragma unroll
            for (int k = 0; k < 9; ++k) bad += (out[k] && a[k * N + p] != 0.0) ? 1 : 0;
        }
    }
    if (__any(bad)) {
        if (bad) atomicAdd(&st->outside, bad);
    }
}

__global__ void k_reset_state(SolveState *st, int gated)
{
    if (threadIdx.x == 0) {
        if (gated && st->skip) return;
        st->skip = 0;
        st->done = 0;
        st->nit = 0;
        st->status = 0;
        st->nitm_hit = 0;
        st->abort = 0;
        st->err_pivot = 0;
        st->outside = 0;
        st->outer = 0;
        st->iter = 0;
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
template <typename TA, typename TX, typename TR>
void launch_residual_norm(const TA *a, const TX *x, const TA *rhs, TR *r, int nx, int ny,
                          u64 *partial, SolveState *st, const Skew *out_skew,
                          const NormRule &rule, hipStream_t s, int nd)
{
    const int nbx = (nx + kResThreads - 1) / kResThreads;
    const int per = (nbx * ny + 7) / 8;
    Skew osk{};
    if (out_skew) osk = *out_skew;
    assert(!out_skew || osk.S == 64 / (int)sizeof(TR));
    TX *xm = const_cast<TX *>(x);  // read only without the update
    if (nd == 5)
        hipLaunchKernelGGL((k_residual<TA, TX, TR, 5, false>), dim3(8 * per), dim3(kResThreads), 0,
                           s, a, xm, xm, (const TR *)nullptr, rhs, r, nx, ny, partial, st, osk,
                           nbx, per, rule);
    else
        hipLaunchKernelGGL((k_residual<TA, TX, TR, 9, false>), dim3(8 * per), dim3(kResThreads), 0,
                           s, a, xm, xm, (const TR *)nullptr, rhs, r, nx, ny, partial, st, osk,
                           nbx, per, rule);
}

template <typename T>
void launch_update_residual_norm(const T *a, T *x, T *x2, const T *z, const T *rhs, T *r,
                                 const Skew &sk, u64 *partial, SolveState *st,
                                 const NormRule &rule, hipStream_t s, int nd)
{
    const int nx = sk.nx, ny = sk.ny;
    const int nbx = (nx + kResThreads - 1) / kResThreads;
    const int per = (nbx * ny + 7) / 8;
    assert(sk.S == 64 / (int)sizeof(T));
    if (nd == 5)
        hipLaunchKernelGGL((k_residual<T, T, T, 5, true>), dim3(8 * per), dim3(kResThreads), 0, s,
                           a, x, x2, z, rhs, r, nx, ny, partial, st, sk, nbx, per, rule);
    else
        hipLaunchKernelGGL((k_residual<T, T, T, 9, true>), dim3(8 * per), dim3(kResThreads), 0, s,
                           a, x, x2, z, rhs, r, nx, ny, partial, st, sk, nbx, per, rule);
}

template <typename T>
void launch_settle(T *x, const T *x2, long n, const SolveState *st, hipStream_t s)
{
    hipLaunchKernelGGL(k_settle<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, x2, n,
                       st);
}

void launch_check_outside(const double *a, int nx, int ny, SolveState *st, hipStream_t s, int nd)
{
    const long N = (long)nx * ny;
    hipLaunchKernelGGL(k_check_outside, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, a,
                       nx, ny, st, nd);
}

void launch_reset_state(SolveState *st, hipStream_t s, bool gated)
{
    hipLaunchKernelGGL(k_reset_state, dim3(1), dim3(64), 0, s, st, gated ? 1 : 0);
}

template <typename TI, typename TO>
void launch_convert(const TI *in, TO *out, long n, hipStream_t s)
{
    hipLaunchKernelGGL((k_convert<TI, TO>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       in, out, n);
}

template void launch_residual_norm<double, double, double>(const double *, const double *,
                                                           const double *, double *, int, int,
                                                           u64 *, SolveState *, const Skew *,
                                                           const NormRule &, hipStream_t, int);
template void launch_update_residual_norm<double>(const double *, double *, double *,
                                                  const double *, const double *, double *,
                                                  const Skew &, u64 *, SolveState *,
                                                  const NormRule &, hipStream_t, int);
template void launch_update_residual_norm<float>(const float *, float *, float *, const float *,
                                                 const float *, float *, const Skew &, u64 *,
                                                 SolveState *, const NormRule &, hipStream_t,
                                                 int);
template void launch_settle<double>(double *, const double *, long, const SolveState *,
                                    hipStream_t);
template void launch_settle<float>(float *, const float *, long, const SolveState *, hipStream_t);
template void launch_residual_norm<float, float, float>(const float *, const float *,
                                                        const float *, float *, int, int,
                                                        u64 *, SolveState *, const Skew *,
                                                        const NormRule &, hipStream_t, int);
template void launch_convert<double, float>(const double *, float *, long, hipStream_t);
template void launch_convert<float, double>(const float *, double *, long, hipStream_t);

}  // namespace m2d
