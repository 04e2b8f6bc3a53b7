// batch.hip -- batched MSI: many independent 9-diagonal systems of uneven
// grid size solved together, one launch per phase for the whole batch.
//
// Every system keeps its own arrays (pooled), its own skewed layout,
// boundary words and solve state; a kernel finds its work item through a map
// of (system, band) or (system, block) pairs.  An MSI iteration of the batch
// is [forward sweep, backward sweep, update fused with the residual, norm +
// stopping rule] over every system that has not stopped yet; systems that stopped turn
// their part of each launch into a no-op.  Chunks of iterations are captured
// once into a hipGraph and replayed, with one read of the states per replay.
//
// Placement.  Workgroups are dealt to the eight XCDs round robin, so
// workgroup g runs on XCD g mod 8.  The "grouped" layout gives every system
// one XCD (longest-processing-time assignment by wavefront length times
// bands) and interleaves the XCDs' lists, so a system's bands exchange their
// boundary rows through one L2; the "spread" layouts list systems one after
// the other (natural order, or largest first).  An autotuner times the
// layouts on the caller's batch and keeps the fastest.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "batch.h"
#include "m2d_internal.h"
#include "msi_stencil.h"
#include "skew.h"

using namespace m2d;

namespace m2d {

constexpr int kPts = 256;  // points per workgroup of the point-wise kernels

// Point (i, j) of lane threadIdx.x in point block `blk` (kPts consecutive
// points in the reference order) of a system with nx columns.  One 32-bit
// division per workgroup (the block index is uniform), then at most
// kPts / nx + 1 row steps per lane; systems hold fewer than 2^27 points.
struct BlockPoint {
    int p, i, j;
};
__device__ __forceinline__ BlockPoint block_point(int blk, int nx)
{
    const int p0 = blk * kPts;
    int j = p0 / nx;
    int i = p0 - j * nx + (int)threadIdx.x;
    while (i >= nx) {
        i -= nx;
        ++j;
    }
    return {p0 + (int)threadIdx.x, i, j};
}

// r = A.x - rhs for the 256 points of block map[g].y of system map[g].x, and
// the block's partial sum of r.r (fixed reduction order).  skewed: write r
// in the system's skewed layout, else in the reference layout.  Entries with
// map[g].x < 0 fill the map up to a multiple of the XCD count.
template <typename T>
__global__ __launch_bounds__(kPts) void k_residual_batch(const BatchSys<T> *__restrict__ sys,
                                                         const int2 *__restrict__ map, int skewed)
{
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys<T> &d = sys[m.x];
    if (d.st->done) return;
    const int nx = d.sk.nx, ny = d.sk.ny;
    const int N = nx * ny;
    const BlockPoint q = block_point(m.y, nx);
    double acc = 0.0;
    if (q.p < N) {
        const T v = residual_point<T, T, T>(d.ra, N, d.rx, d.rrhs[q.p], q.i, q.j, nx, ny);
        d.r[skewed ? d.sk.template of_point_c<64 / (int)sizeof(T)>(q.i, q.j) : (long)q.p] = v;
        acc = (double)v * (double)v;
    }
    __shared__ double red[kPts];
    red[threadIdx.x] = acc;
    __syncthreads();
#pragma unroll
    for (int o = kPts / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) d.partial[m.y] = red[0];
}

// var = z + var (msi2d9-v03.f90:181) fused with the residual after it
// (:183-255) and its block partial sum.  During an MSI run the iterate
// alternates between rx (an even number of iterations done) and rx2 (odd):
// the kernel reads one and writes the other, and forms the updated value of
// each stencil neighbour where it reads it (z + var, the sum the update
// stores), so r is bit for bit the residual of the updated iterate.
// k_settle_batch returns a final iterate left in rx2 to rx.
template <typename T>
__global__ __launch_bounds__(kPts) void k_update_residual_batch(
    const BatchSys<T> *__restrict__ sys, const int2 *__restrict__ map)
{
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys<T> &d = sys[m.x];
    if (d.st->done) return;
    const bool odd = d.st->iter & 1;
    const T *__restrict__ xin = odd ? d.rx2 : d.rx;
    T *__restrict__ xout = odd ? d.rx : d.rx2;
    const T *__restrict__ z = d.w;
    const Skew sk = d.sk;
    const int nx = sk.nx, ny = sk.ny;
    const int N = nx * ny;
    constexpr int S = 64 / (int)sizeof(T);
    const BlockPoint q = block_point(m.y, nx);
    double acc = 0.0;
    if (q.p < N) {
        auto xv = [&](int ii, int jj, long f) -> T {
            return z[sk.template of_point_c<S>(ii, jj)] + xin[f];
        };
        const T v = residual_point_f<T, T>(d.ra, N, xv, d.rrhs[q.p], q.i, q.j, nx, ny);
        xout[q.p] = xv(q.i, q.j, q.p);
        d.r[sk.template of_point_c<S>(q.i, q.j)] = v;
        acc = (double)v * (double)v;
    }
    __shared__ double red[kPts];
    red[threadIdx.x] = acc;
    __syncthreads();
#pragma unroll
    for (int o = kPts / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) d.partial[m.y] = red[0];
}

// After an MSI run: a system that stopped after an odd number of iterations
// holds its iterate in rx2; copy it to rx.
template <typename T>
__global__ __launch_bounds__(kPts) void k_settle_batch(const BatchSys<T> *__restrict__ sys,
                                                       const int2 *__restrict__ map)
{
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys<T> &d = sys[m.x];
    if (!(d.st->iter & 1)) return;
    const int p = m.y * kPts + (int)threadIdx.x;
    if (p < d.sk.nx * d.sk.ny) d.rx[p] = d.rx2[p];
}

// One workgroup per system: ||r|| from the partial sums and the stopping
// rule of msi2d9-v03.f90:133,258-259 (same rule as the single-system residual's norm_rule_block).
template <typename T>
__global__ __launch_bounds__(256) void k_finalize_batch(const BatchSys<T> *__restrict__ sys,
                                                        int initial)
{
    const BatchSys<T> &d = sys[blockIdx.x];
    SolveState *st = d.st;
    if (st->done) return;
    const int k = initial ? 0 : st->iter + 1;
    __shared__ double red[256];
    double acc = 0.0;
    for (int q = threadIdx.x; q < d.np; q += 256) acc += d.partial[q];
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
    if (k == 0) {
        st->r0 = nrm;
        st->gammarp = d.gamma * nrm;
        if (!isfinite(nrm)) {
            st->status = M2D_ERR_NONFINITE;
            st->done = 1;
        }
        return;
    }
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
    } else if (k == d.nitm) {
        st->done = 1;
        st->nit = k;
        st->nitm_hit = 1;
    } else if (d.nitm <= 0 && k >= kIterCap) {
        st->status = M2D_ERR_ITER_CAP;
        st->done = 1;
        st->nit = k;
    }
}

// Fresh states; with `outer`, a state whose outer solve stopped starts done.
__global__ void k_reset_batch(SolveState *st, int n, const SolveState *outer)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    SolveState z;
    memset(&z, 0, sizeof(z));
    if (outer) z.done = outer[q].done;
    st[q] = z;
}

__global__ void k_clear_u64(u64 *p, long n)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = 0ull;
}

// Couplings towards neighbours outside the grid (must all be zero), counted
// into the system's state.
__global__ __launch_bounds__(kPts) void k_check_outside_batch(
    const BatchSys<double> *__restrict__ sys, const int2 *__restrict__ map)
{
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys<double> &d = sys[m.x];
    const int nx = d.sk.nx, ny = d.sk.ny;
    const int N = nx * ny;
    const BlockPoint q = block_point(m.y, nx);
    const int p = q.p, i = q.i, j = q.j;
    int bad = 0;
    if (p < N) {
        const bool w = i == 0, e = i == nx - 1, s = j == 0, n = j == ny - 1;
        const bool out[9] = {s || w, s, s || e, w, false, e, n || w, n, n || e};
#pragma unroll
        for (int k = 0; k < 9; ++k) bad += (out[k] && d.ra[k * N + p] != 0.0) ? 1 : 0;
    }
    if (__any(bad))
        if (bad) atomicAdd(&d.st->outside, bad);
}

// Mixed precision: FP32 copy of A (9 diagonals).
__global__ __launch_bounds__(kPts) void k_a32_batch(const BatchSys<double> *__restrict__ s64,
                                                    const BatchSys<float> *__restrict__ s32,
                                                    const int2 *__restrict__ map)
{
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const long N = (long)s64[m.x].sk.nx * s64[m.x].sk.ny;
    const long p = (long)m.y * kPts + threadIdx.x;
    if (p >= N) return;
    const double *a = s64[m.x].ra;
    float *a32 = const_cast<float *>(s32[m.x].ra);
#pragma unroll
    for (int k = 0; k < 9; ++k) a32[k * N + p] = (float)a[k * N + p];
}

// Refinement step: right-hand side -r (FP32) and c = 0 for the correction.
__global__ __launch_bounds__(kPts) void k_ir_prepare_batch(const BatchSys<double> *__restrict__ s64,
                                                           const BatchSys<float> *__restrict__ s32,
                                                           const int2 *__restrict__ map)
{
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys<double> &o = s64[m.x];
    if (o.st->done) return;
    const long N = (long)o.sk.nx * o.sk.ny;
    const long p = (long)m.y * kPts + threadIdx.x;
    if (p >= N) return;
    const_cast<float *>(s32[m.x].rrhs)[p] = (float)(-o.r[p]);
    s32[m.x].rx[p] = 0.f;
}

// x += c in FP64.
__global__ __launch_bounds__(kPts) void k_ir_update_batch(const BatchSys<double> *__restrict__ s64,
                                                          const BatchSys<float> *__restrict__ s32,
                                                          const int2 *__restrict__ map)
{
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys<double> &o = s64[m.x];
    if (o.st->done) return;
    const long N = (long)o.sk.nx * o.sk.ny;
    const long p = (long)m.y * kPts + threadIdx.x;
    if (p >= N) return;
    o.rx[p] = o.rx[p] + (double)s32[m.x].rx[p];
}

}  // namespace m2d

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
#define HIP_OK(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            fprintf(stderr, "mach2d_hip: %s failed: %s (%s:%d)\n", #call,     \
                    hipGetErrorString(e_), __FILE__, __LINE__);               \
            return M2D_ERR_HIP;                                               \
        }                                                                     \
    } while (0)

namespace {

struct Maps {
    int2 *d = nullptr;  // one allocation: sweep (forward/factor), backward, points, skew
    int nf = 0, nb = 0, np = 0, ns = 0;
    int2 *fwd = nullptr, *bwd = nullptr, *pts = nullptr, *skw = nullptr;
};

// One chunk of MSI iterations captured in a graph.
struct Chunk {
    hipGraphExec_t exec = nullptr;
    int iters = 0;
};

}  // namespace

struct m2d_batch {
    int n = 0;
    int precision = M2D_FP64;
    hipStream_t stream = nullptr;  // work stream (never the null stream: graphs are captured)
    hipStream_t own = nullptr;     // the batch's own stream, used unless one is set
    hipEvent_t ev_in = nullptr;    // orders the work after the legacy default stream
    std::vector<int> nx, ny;
    std::vector<long> N;
    std::vector<Skew> sk;  // layout of the sweep precision
    std::vector<void *> allocs;
    std::vector<BatchSys<double>> h64;
    std::vector<BatchSys<float>> h32;
    BatchSys<double> *d64 = nullptr;
    BatchSys<float> *d32 = nullptr;
    SolveState *st = nullptr, *st_in = nullptr;
    SolveState *h_st = nullptr, *h_in = nullptr;  // pinned mirrors
    u64 *words = nullptr;
    long nwords = 0;
    Maps maps;
    int layout = 0;  // placement (layout % 3), LDS-staged sweeps (layout >= 3)
    int chunk_iters = 8;
    Chunk graph;
    // inner FP32 MSI: refreshed from the FP64 residual every 50 inner
    // iterations (an outer step of a batch costs more than a single system's)
    double inner_gamma = 1e-4;
    int inner_nitm = 50;
    bool factored = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

template <typename T>
hipError_t balloc(m2d_batch *b, T **p, size_t count)
{
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, sizeof(T) * std::max<size_t>(count, 1));
    if (e == hipSuccess) {
        b->allocs.push_back(q);
        *p = (T *)q;
    }
    return e;
}

long blocks_of(long n, long per) { return (n + per - 1) / per; }

// Wavefront length of a system (dependent steps of one sweep) times bands:
// the work estimate used for placement.
long sweep_cost(const Skew &k) { return (long)(k.nst + 128L * (k.nb - 1)) * k.nb; }

// Sweep maps of a layout.  fwd: bands bottom to top per system; bwd: top to
// bottom.
void sweep_maps(const m2d_batch *b, int layout, std::vector<int2> &fwd, std::vector<int2> &bwd)
{
    fwd.clear();
    bwd.clear();
    std::vector<int> order(b->n);
    std::iota(order.begin(), order.end(), 0);
    if (layout >= 1)  // largest first
        std::stable_sort(order.begin(), order.end(), [&](int p, int q) {
            return sweep_cost(b->sk[p]) > sweep_cost(b->sk[q]);
        });
    if (layout != 1) {
        for (int q : order)
            for (int band = 0; band < b->sk[q].nb; ++band) {
                fwd.push_back(make_int2(q, band));
                bwd.push_back(make_int2(q, b->sk[q].nb - 1 - band));
            }
        return;
    }
    // grouped: each system on one XCD, LPT over the eight XCDs
    constexpr int kXcd = 8;
    std::vector<long> load(kXcd, 0);
    std::vector<std::vector<int>> lane(kXcd);
    for (int q : order) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[x] += sweep_cost(b->sk[q]);
        lane[x].push_back(q);
    }
    std::vector<std::vector<int2>> lf(kXcd), lb(kXcd);
    size_t depth = 0;
    for (int x = 0; x < kXcd; ++x) {
        for (int q : lane[x])
            for (int band = 0; band < b->sk[q].nb; ++band) {
                lf[x].push_back(make_int2(q, band));
                lb[x].push_back(make_int2(q, b->sk[q].nb - 1 - band));
            }
        depth = std::max(depth, lf[x].size());
    }
    for (size_t k = 0; k < depth; ++k)
        for (int x = 0; x < kXcd; ++x) {
            fwd.push_back(k < lf[x].size() ? lf[x][k] : make_int2(-1, 0));
            bwd.push_back(k < lb[x].size() ? lb[x][k] : make_int2(-1, 0));
        }
}

void drop_graph(m2d_batch *b)
{
    if (b->graph.exec) (void)hipGraphExecDestroy(b->graph.exec);
    b->graph.exec = nullptr;
    b->graph.iters = 0;
}

int build_maps(m2d_batch *b, int layout)
{
    std::vector<int2> fwd, bwd, pts, skw;
    sweep_maps(b, layout % 3, fwd, bwd);
    std::vector<int2> nat;
    for (int q = 0; q < b->n; ++q) {
        for (long k = 0; k < blocks_of(b->N[q], kPts); ++k) nat.push_back(make_int2(q, (int)k));
        for (long k = 0; k < blocks_of(b->sk[q].size(), 256); ++k)
            skw.push_back(make_int2(q, (int)k));
    }
    // Point blocks: XCD x (workgroups g with g mod 8 = x) takes the x-th
    // contiguous eighth of the list, so the rows a block's stencil reaches
    // (the blocks about nx / 256 before and after it) are mostly read on
    // the same XCD and meet in its L2.  Holes of the ragged end are (-1, 0).
    constexpr int kXcd = 8;
    const size_t per = (nat.size() + kXcd - 1) / kXcd;
    pts.assign(per * kXcd, make_int2(-1, 0));
    for (size_t k = 0; k < nat.size(); ++k) pts[(k % per) * kXcd + k / per] = nat[k];
    Maps &m = b->maps;
    if (m.d) (void)hipFree(m.d);
    m.d = nullptr;
    const size_t tot = fwd.size() + bwd.size() + pts.size() + skw.size();
    HIP_OK(hipMalloc((void **)&m.d, sizeof(int2) * tot));
    std::vector<int2> all;
    all.reserve(tot);
    all.insert(all.end(), fwd.begin(), fwd.end());
    all.insert(all.end(), bwd.begin(), bwd.end());
    all.insert(all.end(), pts.begin(), pts.end());
    all.insert(all.end(), skw.begin(), skw.end());
    HIP_OK(hipMemcpy(m.d, all.data(), sizeof(int2) * tot, hipMemcpyHostToDevice));
    m.nf = (int)fwd.size();
    m.nb = (int)bwd.size();
    m.np = (int)pts.size();
    m.ns = (int)skw.size();
    m.fwd = m.d;
    m.bwd = m.fwd + m.nf;
    m.pts = m.bwd + m.nb;
    m.skw = m.pts + m.np;
    b->layout = layout;
    drop_graph(b);
    return M2D_OK;
}

int upload_desc(m2d_batch *b)
{
    if (!b->h64.empty())
        HIP_OK(hipMemcpyAsync(b->d64, b->h64.data(), sizeof(BatchSys<double>) * b->n,
                              hipMemcpyHostToDevice, b->stream));
    if (!b->h32.empty())
        HIP_OK(hipMemcpyAsync(b->d32, b->h32.data(), sizeof(BatchSys<float>) * b->n,
                              hipMemcpyHostToDevice, b->stream));
    return M2D_OK;
}

template <typename T>
const BatchSys<T> *dsys(const m2d_batch *b);
template <>
const BatchSys<double> *dsys<double>(const m2d_batch *b)
{
    return b->d64;
}
template <>
const BatchSys<float> *dsys<float>(const m2d_batch *b)
{
    return b->d32;
}

// Reset the MSI states of precision T (outer: start stopped systems done) and
// clear the boundary words.
template <typename T>
void begin_msi(m2d_batch *b, SolveState *st, const SolveState *outer)
{
    hipLaunchKernelGGL(k_reset_batch, dim3((unsigned)blocks_of(b->n, 64)), dim3(64), 0, b->stream,
                       st, b->n, outer);
    hipLaunchKernelGGL(k_clear_u64, dim3((unsigned)blocks_of(b->nwords, 256)), dim3(256), 0,
                       b->stream, b->words, b->nwords);
}

template <typename T>
void enqueue_iteration(m2d_batch *b)
{
    const BatchSys<T> *d = dsys<T>(b);
    const Maps &m = b->maps;
    launch_forward_batch<T>(d, m.fwd, m.nf, b->stream, b->layout >= 3);
    launch_backward_batch<T>(d, m.bwd, m.nb, b->stream, b->layout >= 3);
    hipLaunchKernelGGL(k_update_residual_batch<T>, dim3(m.np), dim3(kPts), 0, b->stream, d,
                       m.pts);
    hipLaunchKernelGGL(k_finalize_batch<T>, dim3(b->n), dim3(256), 0, b->stream, d, 0);
}

// Capture chunk_iters iterations of precision T once; replays stay valid
// while the maps and the descriptor arrays stay in place (kernels read the
// descriptors and the iteration counters from device memory).
template <typename T>
int chunk_graph(m2d_batch *b)
{
    if (b->graph.exec && b->graph.iters == b->chunk_iters) return M2D_OK;
    drop_graph(b);
    hipGraph_t g = nullptr;
    HIP_OK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < b->chunk_iters; ++k) enqueue_iteration<T>(b);
    HIP_OK(hipStreamEndCapture(b->stream, &g));
    hipError_t e = hipGraphInstantiate(&b->graph.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_OK(e);
    b->graph.iters = b->chunk_iters;
    return M2D_OK;
}

int read_states(m2d_batch *b, const SolveState *dev, SolveState *host)
{
    HIP_OK(hipMemcpyAsync(host, dev, sizeof(SolveState) * b->n, hipMemcpyDeviceToHost, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    return M2D_OK;
}

bool all_done(const m2d_batch *b, const SolveState *h)
{
    for (int q = 0; q < b->n; ++q)
        if (!h[q].done) return false;
    return true;
}

// MSI in precision T on every system from its current iterate, until every
// state stopped; states end up in `host`.
template <typename T>
int run_msi(m2d_batch *b, SolveState *st, SolveState *host, const SolveState *outer)
{
    begin_msi<T>(b, st, outer);
    const BatchSys<T> *d = dsys<T>(b);
    hipLaunchKernelGGL(k_residual_batch<T>, dim3(b->maps.np), dim3(kPts), 0, b->stream, d,
                       b->maps.pts, 1);
    hipLaunchKernelGGL(k_finalize_batch<T>, dim3(b->n), dim3(256), 0, b->stream, d, 1);
    HIP_OK(hipGetLastError());
    int rc = chunk_graph<T>(b);
    if (rc) return rc;
    for (long rep = 0;; ++rep) {
        HIP_OK(hipGraphLaunch(b->graph.exec, b->stream));
        rc = read_states(b, st, host);
        if (rc) return rc;
        if (all_done(b, host)) break;
        if (rep * b->chunk_iters > 2L * kIterCap) return M2D_ERR_ITER_CAP;  // not reached
    }
    hipLaunchKernelGGL(k_settle_batch<T>, dim3(b->maps.np), dim3(kPts), 0, b->stream, d,
                       b->maps.pts);
    HIP_OK(hipGetLastError());
    return M2D_OK;
}

template <typename T>
int factor_msi(m2d_batch *b)
{
    const BatchSys<T> *d = dsys<T>(b);
    const Maps &m = b->maps;
    launch_to_skew_batch<T>(d, m.skw, m.ns, b->stream);
    launch_factor_batch<T>(d, m.fwd, m.nf, b->stream);
    launch_prepare_batch<T>(d, m.skw, m.ns, b->stream);
    HIP_OK(hipGetLastError());
    return M2D_OK;
}

void fill(const m2d_batch *b, int q, const SolveState &s, int nit, int outer, float ms,
          m2d_status *out)
{
    if (!out) return;
    out[q].status = s.status;
    out[q].nit = nit;
    out[q].nitm_reached = s.nitm_hit;
    out[q].outer = outer;
    out[q].r0 = s.r0;
    out[q].rn = s.rn;
    out[q].gammarp = s.gammarp;
    out[q].device_ms = ms;
}

int set_io(m2d_batch *b, const double *const *d_a, const double *const *d_rhs,
           double *const *d_var, double gamma, int nitm, const double *gammas = nullptr)
{
    if (b->stream == b->own) {  // inputs may come from the legacy default stream
        HIP_OK(hipEventRecord(b->ev_in, nullptr));
        HIP_OK(hipStreamWaitEvent(b->stream, b->ev_in, 0));
    }
    for (int q = 0; q < b->n; ++q) {
        if (!d_a[q] || (d_rhs && !d_rhs[q]) || (d_var && !d_var[q])) return M2D_ERR_ARG;
        BatchSys<double> &o = b->h64[q];
        o.ra = d_a[q];
        if (d_rhs) o.rrhs = d_rhs[q];
        if (d_var) o.rx = d_var[q];
        o.gamma = gammas ? gammas[q] : gamma;
        o.nitm = nitm;
    }
    return upload_desc(b);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int m2d_batch_destroy(m2d_batch *b)
{
    if (!b) return M2D_OK;
    drop_graph(b);
    for (void *p : b->allocs) (void)hipFree(p);
    if (b->maps.d) (void)hipFree(b->maps.d);
    if (b->h_st) (void)hipHostFree(b->h_st);
    if (b->h_in) (void)hipHostFree(b->h_in);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    if (b->own) (void)hipStreamDestroy(b->own);
    delete b;
    return M2D_OK;
}

extern "C" int m2d_batch_create(int count, const int *nx, const int *ny, int precision,
                                m2d_batch **out)
{
    if (!out || count < 1 || !nx || !ny || (precision != M2D_FP64 && precision != M2D_FP32_IR))
        return M2D_ERR_ARG;
    for (int q = 0; q < count; ++q)
        if (nx[q] < 3 || ny[q] < 3 || (long)nx[q] * ny[q] > (1L << 31) / 16) return M2D_ERR_ARG;
    m2d_batch *b = new m2d_batch();
    b->n = count;
    b->precision = precision;
    const bool fp64 = precision == M2D_FP64;
    const int vb = fp64 ? 8 : 4;
    long L = 0, NN = 0, NP = 0;
    for (int q = 0; q < count; ++q) {
        b->nx.push_back(nx[q]);
        b->ny.push_back(ny[q]);
        b->N.push_back((long)nx[q] * ny[q]);
        b->sk.push_back(skew_for(nx[q], ny[q], vb));
        L += b->sk[q].size();
        NN += b->N[q];
        NP += blocks_of(b->N[q], kPts);
        b->nwords += boundary_words(nx[q], ny[q]) + scratch_words(ny[q]);
    }
    bool ok = true;
    double *part = nullptr;
    ok = ok && balloc(b, &b->st, count) == hipSuccess && balloc(b, &b->st_in, count) == hipSuccess;
    ok = ok && balloc(b, &b->words, b->nwords) == hipSuccess;
    ok = ok && balloc(b, &part, NP) == hipSuccess;
    ok = ok && balloc(b, &b->d64, count) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&b->h_st, sizeof(SolveState) * count) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&b->h_in, sizeof(SolveState) * count) == hipSuccess;
    ok = ok && hipEventCreate(&b->ev0) == hipSuccess && hipEventCreate(&b->ev1) == hipSuccess &&
         hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) == hipSuccess &&
         hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking) == hipSuccess;
    b->stream = b->own;
    if (!ok) {
        m2d_batch_destroy(b);
        return M2D_ERR_HIP;
    }
    // skewed pool of the sweep precision: a 9, dl 6, du 4, r, w
    double *p64 = nullptr, *r64 = nullptr, *alt64 = nullptr;
    float *p32 = nullptr, *ref32 = nullptr;
    if (fp64) {
        ok = balloc(b, &p64, 21 * L) == hipSuccess && balloc(b, &alt64, NN) == hipSuccess;
    } else {
        ok = balloc(b, &p32, 21 * L) == hipSuccess && balloc(b, &ref32, 12 * NN) == hipSuccess &&
             balloc(b, &r64, NN) == hipSuccess && balloc(b, &b->d32, count) == hipSuccess;
    }
    if (!ok) {
        m2d_batch_destroy(b);
        return M2D_ERR_HIP;
    }
    if (fp64)
        ok = hipMemset(p64, 0, sizeof(double) * 21 * L) == hipSuccess;
    else
        ok = hipMemset(p32, 0, sizeof(float) * 21 * L) == hipSuccess;
    b->h64.resize(count);
    if (!fp64) b->h32.resize(count);
    long ls = 0, ln = 0, lp = 0, lw = 0;
    for (int q = 0; q < count && ok; ++q) {
        const Skew &k = b->sk[q];
        const long Lq = k.size(), Nq = b->N[q];
        const int np = (int)blocks_of(Nq, kPts);
        BatchSys<double> &o = b->h64[q];
        memset(&o, 0, sizeof(o));
        o.sk = fp64 ? k : skew_for(nx[q], ny[q], 8);
        o.st = b->st + q;
        o.partial = part + lp;
        o.np = np;
        if (fp64) {
            double *base = p64 + 21 * ls;
            o.a = base;
            o.dl = base + 9 * Lq;
            o.du = base + 15 * Lq;
            o.r = base + 19 * Lq;
            o.w = base + 20 * Lq;
            o.rx2 = alt64 + ln;
            o.bnd = b->words + lw;
            o.scratch = o.bnd + boundary_words(nx[q], ny[q]);
        } else {
            o.r = r64 + ln;  // outer residual, reference layout
            BatchSys<float> &i = b->h32[q];
            memset(&i, 0, sizeof(i));
            float *base = p32 + 21 * ls;
            i.sk = k;
            i.a = base;
            i.dl = base + 9 * Lq;
            i.du = base + 15 * Lq;
            i.r = base + 19 * Lq;
            i.w = base + 20 * Lq;
            i.bnd = b->words + lw;
            i.scratch = i.bnd + boundary_words(nx[q], ny[q]);
            i.st = b->st_in + q;
            float *rf = ref32 + 12 * ln;
            i.ra = rf;             // FP32 A, 9 diagonals
            i.rrhs = rf + 9 * Nq;  // -r in FP32
            i.rx = rf + 10 * Nq;   // correction c
            i.rx2 = rf + 11 * Nq;  // its alternate
            i.partial = o.partial;
            i.np = np;
            i.gamma = b->inner_gamma;
            i.nitm = b->inner_nitm;
        }
        ls += Lq;
        ln += Nq;
        lp += np;
        lw += boundary_words(nx[q], ny[q]) + scratch_words(ny[q]);
    }
    if (!ok || build_maps(b, 0) != M2D_OK) {
        m2d_batch_destroy(b);
        return M2D_ERR_HIP;
    }
    *out = b;
    return M2D_OK;
}

// stream: a non-null stream for the batch's work; nullptr restores the
// batch's own stream (which first waits for the legacy default stream).
extern "C" int m2d_batch_set_stream(m2d_batch *b, void *stream)
{
    if (!b) return M2D_ERR_ARG;
    hipStream_t s = stream ? (hipStream_t)stream : b->own;
    if (s != b->stream) drop_graph(b);
    b->stream = s;
    return M2D_OK;
}

extern "C" int m2d_batch_set_inner(m2d_batch *b, double inner_gamma, int inner_nitm)
{
    if (!b || !(inner_gamma > 0.0) || inner_nitm < 1) return M2D_ERR_ARG;
    b->inner_gamma = inner_gamma;
    b->inner_nitm = inner_nitm;
    for (auto &i : b->h32) {
        i.gamma = inner_gamma;
        i.nitm = inner_nitm;
    }
    return M2D_OK;
}

extern "C" int m2d_batch_set_tuning(m2d_batch *b, int layout, int chunk_iters)
{
    if (!b || layout < 0 || layout > 5 || chunk_iters < 1 || chunk_iters > 64) return M2D_ERR_ARG;
    b->chunk_iters = chunk_iters;
    if (layout != b->layout) return build_maps(b, layout);
    return M2D_OK;
}

extern "C" int m2d_batch_get_tuning(m2d_batch *b, int *layout, int *chunk_iters)
{
    if (!b || !layout || !chunk_iters) return M2D_ERR_ARG;
    *layout = b->layout;
    *chunk_iters = b->chunk_iters;
    return M2D_OK;
}

extern "C" int m2d_batch_factor(m2d_batch *b, const double *const *d_a)
{
    if (!b || !d_a) return M2D_ERR_ARG;
    int rc = set_io(b, d_a, nullptr, nullptr, 0.0, 0);
    if (rc) return rc;
    const Maps &m = b->maps;
    // every state fresh (outer and inner), boundary words cleared
    begin_msi<double>(b, b->st, nullptr);
    hipLaunchKernelGGL(k_reset_batch, dim3((unsigned)blocks_of(b->n, 64)), dim3(64), 0, b->stream,
                       b->st_in, b->n, (const SolveState *)nullptr);
    hipLaunchKernelGGL(k_check_outside_batch, dim3(m.np), dim3(kPts), 0, b->stream, b->d64,
                       m.pts);
    if (b->precision == M2D_FP64) {
        rc = factor_msi<double>(b);
    } else {
        rc = upload_desc(b);
        if (!rc) {
            hipLaunchKernelGGL(k_a32_batch, dim3(m.np), dim3(kPts), 0, b->stream, b->d64, b->d32,
                               m.pts);
            rc = factor_msi<float>(b);
        }
    }
    if (rc) return rc;
    rc = read_states(b, b->st, b->h_st);
    if (!rc) rc = read_states(b, b->st_in, b->h_in);
    if (rc) return rc;
    int worst = M2D_OK;
    for (int q = 0; q < b->n && !worst; ++q) {
        const SolveState &s = b->precision == M2D_FP64 ? b->h_st[q] : b->h_in[q];
        if (b->h_st[q].outside) worst = M2D_ERR_OUTSIDE_COUPLING;
        else if (s.abort) worst = M2D_ERR_SYNC_TIMEOUT;
        else if (s.err_pivot) worst = M2D_ERR_NONFINITE;
    }
    b->factored = worst == M2D_OK || worst == M2D_ERR_NONFINITE;
    return worst;
}

static int batch_solve(m2d_batch *b, const double *const *d_a, const double *const *d_rhs,
                       double *const *d_var, double gamma, const double *gammas, int nitm,
                       m2d_status *st);

extern "C" int m2d_batch_solve(m2d_batch *b, const double *const *d_a,
                               const double *const *d_rhs, double *const *d_var, double gamma,
                               int nitm, m2d_status *st)
{
    return batch_solve(b, d_a, d_rhs, d_var, gamma, nullptr, nitm, st);
}

extern "C" int m2d_batch_solve_each(m2d_batch *b, const double *const *d_a,
                                    const double *const *d_rhs, double *const *d_var,
                                    const double *gammas, int nitm, m2d_status *st)
{
    if (!gammas) return M2D_ERR_ARG;
    return batch_solve(b, d_a, d_rhs, d_var, 0.0, gammas, nitm, st);
}

static int batch_solve(m2d_batch *b, const double *const *d_a, const double *const *d_rhs,
                       double *const *d_var, double gamma, const double *gammas, int nitm,
                       m2d_status *st)
{
    if (!b || !d_a || !d_rhs || !d_var) return M2D_ERR_ARG;
    if (!b->factored) return M2D_ERR_ARG;
    int rc = set_io(b, d_a, d_rhs, d_var, gamma, nitm, gammas);
    if (rc) return rc;
    HIP_OK(hipEventRecord(b->ev0, b->stream));
    std::vector<int> inner_total(b->n, 0);
    if (b->precision == M2D_FP64) {
        rc = run_msi<double>(b, b->st, b->h_st, nullptr);
        if (rc) return rc;
    } else {
        // outer FP64 residual and stopping rule; inner FP32 MSI per step
        const Maps &m = b->maps;
        hipLaunchKernelGGL(k_reset_batch, dim3((unsigned)blocks_of(b->n, 64)), dim3(64), 0,
                           b->stream, b->st, b->n, (const SolveState *)nullptr);
        hipLaunchKernelGGL(k_residual_batch<double>, dim3(m.np), dim3(kPts), 0, b->stream, b->d64,
                           m.pts, 0);
        hipLaunchKernelGGL(k_finalize_batch<double>, dim3(b->n), dim3(256), 0, b->stream, b->d64,
                           1);
        HIP_OK(hipGetLastError());
        rc = read_states(b, b->st, b->h_st);
        if (rc) return rc;
        const int kmax = nitm > 0 ? nitm : kIterCap;
        for (int k = 1; k <= kmax && !all_done(b, b->h_st); ++k) {
            hipLaunchKernelGGL(k_ir_prepare_batch, dim3(m.np), dim3(kPts), 0, b->stream, b->d64,
                               b->d32, m.pts);
            rc = run_msi<float>(b, b->st_in, b->h_in, b->st);
            if (rc) return rc;
            for (int q = 0; q < b->n; ++q)
                if (!b->h_st[q].done) inner_total[q] += b->h_in[q].nit;
            hipLaunchKernelGGL(k_ir_update_batch, dim3(m.np), dim3(kPts), 0, b->stream, b->d64,
                               b->d32, m.pts);
            hipLaunchKernelGGL(k_residual_batch<double>, dim3(m.np), dim3(kPts), 0, b->stream,
                               b->d64, m.pts, 0);
            hipLaunchKernelGGL(k_finalize_batch<double>, dim3(b->n), dim3(256), 0, b->stream,
                               b->d64, 0);
            HIP_OK(hipGetLastError());
            rc = read_states(b, b->st, b->h_st);
            if (rc) return rc;
            for (int q = 0; q < b->n; ++q) {
                const int s = b->h_in[q].status;
                if (!b->h_st[q].done && (s == M2D_ERR_NONFINITE || s == M2D_ERR_SYNC_TIMEOUT))
                    return s;
            }
        }
    }
    HIP_OK(hipEventRecord(b->ev1, b->stream));
    HIP_OK(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    int worst = M2D_OK;
    for (int q = 0; q < b->n; ++q) {
        const SolveState &s = b->h_st[q];
        if (b->precision == M2D_FP64)
            fill(b, q, s, s.nit, 0, ms, st);
        else
            fill(b, q, s, inner_total[q], s.nit, ms, st);
        if (s.status && !worst) worst = s.status;
    }
    return worst;
}

// Autotuner: time `iters` MSI iterations (in the batch's sweep precision,
// after a warm-up one) of every layout on the caller's systems, from a zero
// iterate held in scratch memory (in refinement mode: the first correction
// solve), with the stopping rule off; keep the fastest layout.  ms_out (3
// entries, optional): time per iteration of each layout.
extern "C" int m2d_batch_tune(m2d_batch *b, const double *const *d_a, const double *const *d_rhs,
                              int iters, int *layout_out, double *ms_out)
{
    if (!b || !d_a || !d_rhs || iters < 1) return M2D_ERR_ARG;
    if (!b->factored) return M2D_ERR_ARG;
    const bool fp64 = b->precision == M2D_FP64;
    std::vector<double *> x(b->n, nullptr);
    int rc = M2D_OK;
    for (int q = 0; q < b->n && !rc; ++q)
        if (hipMalloc((void **)&x[q], sizeof(double) * b->N[q]) != hipSuccess) rc = M2D_ERR_HIP;
    const int keep_chunk = b->chunk_iters;
    double best = 1e300;
    int best_layout = b->layout;
    for (int layout = 0; layout < 6 && !rc; ++layout) {
        rc = build_maps(b, layout);
        double per = 0.0;
        for (int pass = 0; pass < 2 && !rc; ++pass) {  // pass 0 warms up
            const int n_it = pass == 0 ? 1 : iters;
            b->chunk_iters = n_it;
            rc = fp64 ? chunk_graph<double>(b) : chunk_graph<float>(b);  // not timed
            for (int q = 0; q < b->n && !rc; ++q)
                if (hipMemsetAsync(x[q], 0, sizeof(double) * b->N[q], b->stream) != hipSuccess)
                    rc = M2D_ERR_HIP;
            if (!rc) rc = set_io(b, d_a, d_rhs, x.data(), 0.0, n_it);
            if (rc) break;
            if (!fp64) {  // right-hand side of the first correction, inner rule off
                const Maps &m = b->maps;
                hipLaunchKernelGGL(k_reset_batch, dim3((unsigned)blocks_of(b->n, 64)), dim3(64),
                                   0, b->stream, b->st, b->n, (const SolveState *)nullptr);
                hipLaunchKernelGGL(k_residual_batch<double>, dim3(m.np), dim3(kPts), 0,
                                   b->stream, b->d64, m.pts, 0);
                hipLaunchKernelGGL(k_ir_prepare_batch, dim3(m.np), dim3(kPts), 0, b->stream,
                                   b->d64, b->d32, m.pts);
                for (auto &i : b->h32) {
                    i.gamma = 0.0;
                    i.nitm = n_it;
                }
                rc = upload_desc(b);
                if (rc) break;
            }
            if (hipEventRecord(b->ev0, b->stream) != hipSuccess) rc = M2D_ERR_HIP;
            if (!rc)
                rc = fp64 ? run_msi<double>(b, b->st, b->h_st, nullptr)
                          : run_msi<float>(b, b->st_in, b->h_in, nullptr);
            float ms = 0.f;
            if (!rc && (hipEventRecord(b->ev1, b->stream) != hipSuccess ||
                        hipEventSynchronize(b->ev1) != hipSuccess ||
                        hipEventElapsedTime(&ms, b->ev0, b->ev1) != hipSuccess))
                rc = M2D_ERR_HIP;
            per = ms / n_it;
        }
        if (ms_out) ms_out[layout] = per;
        if (!rc && per < best) {
            best = per;
            best_layout = layout;
        }
    }
    for (auto &i : b->h32) {  // restore the inner rule
        i.gamma = b->inner_gamma;
        i.nitm = b->inner_nitm;
    }
    for (double *p : x)
        if (p) (void)hipFree(p);
    b->chunk_iters = keep_chunk;
    if (!rc) rc = build_maps(b, best_layout);
    if (!rc) rc = upload_desc(b);
    if (layout_out) *layout_out = best_layout;
    return rc;
}
