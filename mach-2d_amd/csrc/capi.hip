// capi.hip -- the C ABI (include/mach2d_hip.h): solver handles, the device
// side MSI iteration loop, and the reference-signature entry points.
//
// A handle owns the factor and sweep vectors in the wavefront layout of
// skew.h (a skewed copy of A for the factorization, dl, du, r, w/z).  One MSI
// iteration is [forward sweep, backward sweep, var = z + var, residual with
// partial sums of r.r, norm and stopping rule]; the stopping rule runs on the
// device and, once it fires, turns the remaining enqueued kernels into
// no-ops, so the host enqueues iterations in chunks and reads the state once
// per chunk instead of once per iteration.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "m2d_internal.h"
#include "msi_stencil.h"
#include "skew.h"
#include "wavefront.h"

using namespace m2d;

#define HIP_OK(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            fprintf(stderr, "mach2d_hip: %s failed: %s (%s:%d)\n", #call,     \
                    hipGetErrorString(e_), __FILE__, __LINE__);               \
            return M2D_ERR_HIP;                                               \
        }                                                                     \
    } while (0)

// Kernel ids of the timing facility: factor, forward, backward, residual;
// +4 for the FP32 kernels of mixed precision.
enum { KT_FACTOR = 0, KT_FORWARD = 1, KT_BACKWARD = 2, KT_RESIDUAL = 3, KT_FP32 = 4 };

// Skewed buffers of one precision.
template <typename T>
struct Work {
    Skew sk{};
    T *a = nullptr;   // 9 arrays: skewed copy of A (factor input)
    T *dl = nullptr;  // 6 arrays: dl1..dl5, 1/dl1
    T *du = nullptr;  // 4 arrays
    T *r = nullptr;   // residual
    T *w = nullptr;   // forward result, then z
    T *x2 = nullptr;  // the iterate after an odd number of iterations (reference layout)
    void release()
    {
        T *b[] = {a, dl, du, r, w, x2};
        for (T *p : b)
            if (p) (void)hipFree(p);
        a = dl = du = r = w = x2 = nullptr;
    }
    bool alloc(int nx, int ny)
    {
        sk = skew_for(nx, ny, (int)sizeof(T));
        const long L = sk.size();
        const size_t z = sizeof(T) * (size_t)L;
        bool ok = hipMalloc((void **)&a, 9 * z) == hipSuccess &&
                  hipMalloc((void **)&dl, 6 * z) == hipSuccess &&
                  hipMalloc((void **)&du, 4 * z) == hipSuccess &&
                  hipMalloc((void **)&r, z) == hipSuccess && hipMalloc((void **)&w, z) == hipSuccess &&
                  hipMalloc((void **)&x2, sizeof(T) * (size_t)nx * ny) == hipSuccess;
        // padding entries are never consumed; start them finite.  dl5 and
        // du4 stay zero for a 5-diagonal system (its factorization writes
        // neither)
        if (ok)
            ok = hipMemset(r, 0, z) == hipSuccess && hipMemset(w, 0, z) == hipSuccess &&
                 hipMemset(dl, 0, 6 * z) == hipSuccess && hipMemset(du, 0, 4 * z) == hipSuccess;
        return ok;
    }
};

// Device control block of the pressure-correction loop
// (m2d_intflow_pl_loop): the per-call scalars the captured iterations read
// and the exit test's results.
struct PlCtl {
    double dt, beta;     // read by the assembly kernels (sc[0], sc[1])
    double p_avg, tolm;  // exit test: max |pl| / p_avg < tolm
    double mx;           // max |pl| of the latest iteration
    double normpl;       // max |pl| / p_avg of the latest iteration that ended normally
    int itp;             // iterations run
    int overrun;         // a p' solve was still iterating after its enqueued iterations
    int maxit;           // most MSI iterations of one p' solve in this call
    int pad;
};

// One captured batch of pressure-correction iterations and what it was
// captured for (pointers, loop parameters and the solver's launch settings).
struct PlGraph {
    m2d_pl_loop_fields f;
    int modvis, nitm, iters, xcd_stride, staged, solve_cap;
    double tol;
    hipGraphExec_t exec;
};

struct m2d_system {
    int nx = 0, ny = 0;
    long N = 0;
    int precision = M2D_FP64;
    // diagonals of A: 9 (msi2d9 slots) or 5 (msi2d5 slots S, W, P, E, N; the
    // factor is that of the zero-corner embedding, the iterations use the
    // 5-diagonal residual and sweeps)
    int nd = 9;
    hipStream_t stream = nullptr;
    // 8: the bands of the system on one XCD (workgroups are dealt to the
    // XCDs round robin, every eighth runs a band); 1: on consecutive XCDs.
    // The default (m2d_create_stencil) follows measurement: 8 from five
    // bands up (1026x514 solves 5-6 % faster), 1 below (258x130 time step
    // 1-2 % faster)
    int xcd_stride = 1;
    // sweep operands through the LDS ring (m2d_set_staging); off by default:
    // a single system's working set sits in the last-level cache, where the
    // register ring's prefetch suffices and the ring's extra LDS pass costs
    bool staged = false;
    int chunk = 8;
    // inner FP32 MSI: short runs between FP64 residuals; measured on the
    // 1026x514 system, a refresh every 25 inner iterations needs 650 inner
    // iterations in all (plain FP64: 646), every 200 needs 734
    double inner_gamma = 1e-4;
    int inner_nitm = 25;

    Work<double> w64;
    Work<float> w32;
    // mixed precision: FP32 A (reference layout), outer FP64 residual, FP32
    // correction and its right-hand side
    float *a32 = nullptr, *c32 = nullptr, *rr32 = nullptr;
    double *r64 = nullptr;

    u64 *partial = nullptr;  // residual partial sums (tagged words)
    int np = 0;
    u64 *bnd = nullptr;  // boundary words, then per-lane scratch words
    long nbnd = 0;
    SolveState *st = nullptr;     // device
    SolveState *st_in = nullptr;  // device, inner solve (mixed precision)
    SolveState *h_st = nullptr;   // pinned host mirror
    // run_msi's pipelined state reads: two pinned states after consecutive
    // chunks of iterations and their events (null: one chunk at a time)
    SolveState *h_chunk = nullptr;
    hipEvent_t ev_chunk[2] = {nullptr, nullptr};
    bool factored = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    // optional per-kernel event timing (m2d_enable_kernel_timing)
    struct Rec {
        int kernel, k;
        hipEvent_t a, b;
    };
    bool timing = false;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    std::vector<Rec> recs;
    double kt_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    long kt_n[8] = {0, 0, 0, 0, 0, 0, 0, 0};

    // pressure-correction loop: control block (device) and its pinned
    // mirror, the stream batches are captured on, the captured batches,
    // the previous call's iteration count (the next call's first batch)
    // and the last call's host reads and batches; the p' solves' enqueued
    // iterations (setting: 0 from the previous call's longest solve, > 0
    // fixed, < 0 all nitm_p), the previous call's longest solve, the last
    // call's enqueued iterations and reruns, and the saved fields a rerun
    // starts from
    PlCtl *pl = nullptr, *h_pl = nullptr;
    hipStream_t cap = nullptr;
    std::vector<PlGraph> plg;
    bool pl_graph = true;
    int pl_last = 0, pl_reads = 0, pl_batches = 0;
    int pl_cap_set = 0, pl_maxit = 0, pl_cap_used = 0, pl_reruns = 0;
    double *pl_save = nullptr;

    // time step (m2d_intflow_time_step): a second stream for the in-order
    // statistics that run beside the step, its events, the device scratch
    // of those sums and the pinned copy of their results (p_avg, 4 means)
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_dep = nullptr;
    double *ts_scr = nullptr, *h_ts = nullptr;
};

static hipEvent_t pool_event(m2d_system *s)
{
    if (s->pool_used == s->pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        s->pool.push_back(e);
    }
    return s->pool[s->pool_used++];
}

// Bracket the launches made by fn() with events when timing is on.
template <typename F>
static void timed(m2d_system *s, int kernel, int k, F fn)
{
    if (!s->timing) {
        fn();
        return;
    }
    hipEvent_t a = pool_event(s), b = pool_event(s);
    if (!a || !b) {
        fn();
        return;
    }
    (void)hipEventRecord(a, s->stream);
    fn();
    (void)hipEventRecord(b, s->stream);
    s->recs.push_back({kernel, k, a, b});
}

// Accumulate the recorded launches of iterations k <= nit (launches after
// the stopping rule fired are no-ops and are not counted).
static void harvest(m2d_system *s, int nit)
{
    if (!s->timing) return;
    for (const auto &r : s->recs) {
        if (r.k > nit) continue;
        float ms = 0.f;
        if (hipEventSynchronize(r.b) == hipSuccess &&
            hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            s->kt_ms[r.kernel] += ms;
            s->kt_n[r.kernel] += 1;
        }
    }
    s->recs.clear();
    s->pool_used = 0;
}

static thread_local m2d_status g_last = {0, 0, 0, 0, 0.0, 0.0, 0.0, 0.0};

static void set_last(int status)
{
    memset(&g_last, 0, sizeof(g_last));
    g_last.status = status;
}

void m2d_set_last_status_code(int status) { set_last(status); }

extern "C" const char *m2d_version(void) { return "mach2d_hip 0.2 (gfx950)"; }

extern "C" int m2d_last_status(m2d_status *out)
{
    if (!out) return M2D_ERR_ARG;
    *out = g_last;
    return M2D_OK;
}

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
template <typename T>
static hipError_t dalloc(T **p, long n)
{
    return hipMalloc((void **)p, sizeof(T) * (size_t)std::max<long>(n, 1));
}

extern "C" int m2d_destroy(m2d_system *s)
{
    if (!s) return M2D_OK;
    s->w64.release();
    s->w32.release();
    void *bufs[] = {s->a32, s->c32, s->rr32, s->r64, s->partial, s->bnd, s->st, s->st_in};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (hipEvent_t e : s->pool) (void)hipEventDestroy(e);
    for (PlGraph &g : s->plg) (void)hipGraphExecDestroy(g.exec);
    if (s->cap) (void)hipStreamDestroy(s->cap);
    if (s->aux) (void)hipStreamDestroy(s->aux);
    if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
    if (s->ev_join) (void)hipEventDestroy(s->ev_join);
    if (s->ev_dep) (void)hipEventDestroy(s->ev_dep);
    if (s->ts_scr) (void)hipFree(s->ts_scr);
    if (s->h_ts) (void)hipHostFree(s->h_ts);
    if (s->pl) (void)hipFree(s->pl);
    if (s->h_pl) (void)hipHostFree(s->h_pl);
    if (s->pl_save) (void)hipFree(s->pl_save);
    if (s->h_st) (void)hipHostFree(s->h_st);
    if (s->h_chunk) (void)hipHostFree(s->h_chunk);
    for (hipEvent_t e : s->ev_chunk)
        if (e) (void)hipEventDestroy(e);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    delete s;
    return M2D_OK;
}

extern "C" int m2d_create_stencil(int nx, int ny, int precision, int ndiag, m2d_system **out)
{
    if (!out || nx < 3 || ny < 3 || (precision != M2D_FP64 && precision != M2D_FP32_IR) ||
        (ndiag != 5 && ndiag != 9))
        return M2D_ERR_ARG;
    if ((long)nx * ny > (1L << 31) / 16) return M2D_ERR_ARG;
    m2d_system *s = new m2d_system();
    s->nd = ndiag;
    s->nx = nx;
    s->ny = ny;
    s->N = (long)nx * ny;
    s->xcd_stride = (ny + 63) / 64 >= 5 ? 8 : 1;
    s->precision = precision;
    s->np = residual_blocks(nx, ny);
    s->nbnd = boundary_words(nx, ny) + scratch_words(ny);
    const long N = s->N;
    const long npw = residual_partial_words(nx, ny);
    bool ok = dalloc(&s->partial, npw) == hipSuccess &&
              hipMemset(s->partial, 0, sizeof(u64) * (size_t)npw) == hipSuccess &&
              dalloc(&s->bnd, s->nbnd) == hipSuccess &&
              dalloc(&s->st, 1) == hipSuccess && dalloc(&s->st_in, 1) == hipSuccess &&
              dalloc(&s->r64, N) == hipSuccess;
    if (ok && precision == M2D_FP64) ok = s->w64.alloc(nx, ny);
    if (ok && precision == M2D_FP32_IR) {
        ok = s->w32.alloc(nx, ny) && dalloc(&s->a32, (long)ndiag * N) == hipSuccess &&
             dalloc(&s->c32, N) == hipSuccess && dalloc(&s->rr32, N) == hipSuccess;
    }
    ok = ok && hipHostMalloc((void **)&s->h_st, sizeof(SolveState), hipHostMallocDefault) ==
                   hipSuccess;
    ok = ok && hipHostMalloc((void **)&s->h_chunk, 2 * sizeof(SolveState),
                             hipHostMallocDefault) == hipSuccess;
    for (hipEvent_t &e : s->ev_chunk)
        ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreate(&s->ev0) == hipSuccess && hipEventCreate(&s->ev1) == hipSuccess;
    if (!ok) {
        m2d_destroy(s);
        return M2D_ERR_HIP;
    }
    *out = s;
    return M2D_OK;
}

extern "C" int m2d_create(int nx, int ny, int precision, m2d_system **out)
{
    return m2d_create_stencil(nx, ny, precision, 9, out);
}

extern "C" int m2d_set_stream(m2d_system *s, void *stream)
{
    if (!s) return M2D_ERR_ARG;
    s->stream = (hipStream_t)stream;
    return M2D_OK;
}

extern "C" int m2d_set_tuning(m2d_system *s, int publish_every, int xcd_stride)
{
    // publish_every is accepted for interface stability: the wavefront
    // kernels publish once per chunk of steps.
    if (!s || publish_every < 1 || (xcd_stride != 1 && xcd_stride != 8)) return M2D_ERR_ARG;
    s->xcd_stride = xcd_stride;
    return M2D_OK;
}

extern "C" int m2d_set_staging(m2d_system *s, int staged)
{
    if (!s || (staged != 0 && staged != 1)) return M2D_ERR_ARG;
    s->staged = staged != 0;
    return M2D_OK;
}

extern "C" int m2d_set_inner(m2d_system *s, double inner_gamma, int inner_nitm)
{
    if (!s || !(inner_gamma > 0.0) || inner_nitm < 1) return M2D_ERR_ARG;
    s->inner_gamma = inner_gamma;
    s->inner_nitm = inner_nitm;
    return M2D_OK;
}

static Link make_link(m2d_system *s, SolveState *st, int phase)
{
    Link y;
    y.bnd = s->bnd;
    y.scratch = s->bnd + boundary_words(s->nx, s->ny);
    y.st = st;
    y.phase = phase;
    y.gen = 0;
    return y;
}

static int read_state(m2d_system *s, SolveState *dev)
{
    HIP_OK(hipMemcpyAsync(s->h_st, dev, sizeof(SolveState), hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    return M2D_OK;
}

// Reset the solve state and clear the boundary words: generations restart
// with every solve (they follow the device iteration counter).
static void begin(m2d_system *s, SolveState *st, bool gated = false)
{
    launch_reset_state(st, s->stream, gated);
    launch_clear_boundary(s->bnd, s->nx, s->ny, s->stream);
}

extern "C" int m2d_factor(m2d_system *s, const double *d_a)
{
    if (!s || !d_a) return M2D_ERR_ARG;
    begin(s, s->st);
    launch_check_outside(d_a, s->nx, s->ny, s->st, s->stream, s->nd);
    if (s->precision == M2D_FP64) {
        Work<double> &W = s->w64;
        if (s->nd == 5)
            launch_to_skew5<double, double>(d_a, W.a, W.sk, s->stream);
        else
            launch_to_skew<double, double>(d_a, W.a, 9, W.sk, s->stream, kP);
        timed(s, KT_FACTOR, 0, [&] {
            launch_factor<double>(W.a, W.dl, W.du, 0.5, W.sk, make_link(s, s->st, 0), s->xcd_stride,
                                  s->stream, s->nd == 5);
        });
        launch_prepare_factor<double>(W.dl, W.du, W.sk, s->stream, s->st);
    } else {
        Work<float> &W = s->w32;
        launch_convert<double, float>(d_a, s->a32, (long)s->nd * s->N, s->stream);
        if (s->nd == 5)
            launch_to_skew5<double, float>(d_a, W.a, W.sk, s->stream);
        else
            launch_to_skew<double, float>(d_a, W.a, 9, W.sk, s->stream, kP);
        timed(s, KT_FACTOR + KT_FP32, 0, [&] {
            launch_factor<float>(W.a, W.dl, W.du, 0.5f, W.sk, make_link(s, s->st, 0),
                                 s->xcd_stride, s->stream, s->nd == 5);
        });
        launch_prepare_factor<float>(W.dl, W.du, W.sk, s->stream, s->st);
    }
    HIP_OK(hipGetLastError());
    int rc = read_state(s, s->st);
    if (rc) return rc;
    harvest(s, 0);
    if (s->h_st->outside) return M2D_ERR_OUTSIDE_COUPLING;
    if (s->h_st->abort) return M2D_ERR_SYNC_TIMEOUT;
    s->factored = true;
    if (s->h_st->err_pivot) return M2D_ERR_NONFINITE;
    return M2D_OK;
}

extern "C" int m2d_get_factor(m2d_system *s, double *d_dl, double *d_du)
{
    if (!s || s->precision != M2D_FP64 || !s->factored || !d_dl || !d_du) return M2D_ERR_ARG;
    // 5 diagonals: dl1..dl4 and du1..du3 (msi2d5's dl(nij,4), du(nij,3))
    launch_from_skew<double>(s->w64.dl, d_dl, s->nd == 5 ? 4 : 5, s->w64.sk, s->stream);
    launch_from_skew<double>(s->w64.du, d_du, s->nd == 5 ? 3 : 4, s->w64.sk, s->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(s->stream));
    return M2D_OK;
}

// Install an externally computed FP64 factor (fb2d9 takes dl, du as input).
static int put_factor(m2d_system *s, const double *d_dl, const double *d_du)
{
    const long L = s->w64.sk.size();
    if (s->nd == 5) {  // dl5 and du4 of the embedding are zero (and never read)
        HIP_OK(hipMemsetAsync(s->w64.dl + 4 * L, 0, sizeof(double) * L, s->stream));
        HIP_OK(hipMemsetAsync(s->w64.du + 3 * L, 0, sizeof(double) * L, s->stream));
    }
    launch_to_skew<double, double>(d_dl, s->w64.dl, s->nd == 5 ? 4 : 5, s->w64.sk, s->stream);
    launch_to_skew<double, double>(d_du, s->w64.du, s->nd == 5 ? 3 : 4, s->w64.sk, s->stream);
    launch_prepare_factor<double>(s->w64.dl, s->w64.du, s->w64.sk, s->stream);
    HIP_OK(hipGetLastError());
    s->factored = true;
    return M2D_OK;
}

// One MSI iteration (k >= 1) in precision T.
template <typename T>
static void enqueue_iteration(m2d_system *s, Work<T> &W, SolveState *st, const T *a, const T *rhs,
                              T *x, double *hist, int hist_cap, int k, double gamma, int nitm)
{
    const int off = sizeof(T) == 4 ? KT_FP32 : 0;
    if (!s->staged) {  // both sweeps in one launch, timed as the forward sweep
        timed(s, KT_FORWARD + off, k, [&] {
            launch_sweeps<T>(W.r, W.dl, W.w, W.du, W.sk, make_link(s, st, 1), s->xcd_stride,
                             s->stream, s->nd == 5);
        });
    } else {  // the staged rings of both sweeps do not fit one workgroup's LDS
        timed(s, KT_FORWARD + off, k, [&] {
            launch_forward_staged<T>(W.r, W.dl, W.w, W.sk, make_link(s, st, 1), s->xcd_stride,
                                     s->stream, s->nd == 5);
        });
        timed(s, KT_BACKWARD + off, k, [&] {
            launch_backward_staged<T>(W.w, W.du, W.sk, make_link(s, st, 2), s->xcd_stride,
                                      s->stream, s->nd == 5);
        });
    }
    timed(s, KT_RESIDUAL + off, k, [&] {
        launch_update_residual_norm<T>(a, x, W.x2, W.w, rhs, W.r, W.sk, s->partial, st,
                                       NormRule{hist, hist_cap, 0, gamma, nitm}, s->stream,
                                       s->nd);
    });
}

// fb2d9 on the device, in precision T, from the current x (reference
// layout); returns after the stopping rule fired (state in s->h_st).
// Iterations go out in chunks (2, 4, ... up to s->chunk), each followed by
// a copy of the state; the next chunk is enqueued before the host waits for
// the previous one's state, so the device never waits for the host, and a
// chunk enqueued after the stopping rule fired turns into no-ops.
template <typename T>
static int run_msi(m2d_system *s, Work<T> &W, SolveState *st, const T *a, const T *rhs, T *x,
                   double gamma, int nitm, double *hist, int hist_cap)
{
    begin(s, st);
    launch_residual_norm<T, T, T>(a, x, rhs, W.r, s->nx, s->ny, s->partial, st, &W.sk,
                                  NormRule{nullptr, 0, 1, gamma, nitm}, s->stream, s->nd);
    const int kmax = nitm > 0 ? nitm : kIterCap;
    const bool piped = s->h_chunk && s->ev_chunk[0] && s->ev_chunk[1];
    int k = 1;
    int chunk = 2;  // grows geometrically up to s->chunk
    auto enqueue_chunk = [&](int slot) -> int {
        const int kend = std::min(kmax, k + chunk - 1);
        for (; k <= kend; ++k)
            enqueue_iteration<T>(s, W, st, a, rhs, x, hist, hist_cap, k, gamma, nitm);
        chunk = std::min(chunk * 2, s->chunk);
        HIP_OK(hipMemcpyAsync(s->h_chunk + slot, st, sizeof(SolveState), hipMemcpyDeviceToHost,
                              s->stream));
        HIP_OK(hipEventRecord(s->ev_chunk[slot], s->stream));
        return M2D_OK;
    };
    if (!piped) {
        while (k <= kmax) {
            const int kend = std::min(kmax, k + chunk - 1);
            for (; k <= kend; ++k)
                enqueue_iteration<T>(s, W, st, a, rhs, x, hist, hist_cap, k, gamma, nitm);
            HIP_OK(hipGetLastError());
            int rc = read_state(s, st);
            if (rc) return rc;
            if (s->h_st->done) break;
            chunk = std::min(chunk * 2, s->chunk);
        }
    } else {
        // chunks in flight, oldest first in slot head, the next in head ^ 1
        int head = 0, inflight = 0;
        for (;;) {
            while (k <= kmax && inflight < 2) {  // one chunk queued behind the oldest
                const int rc = enqueue_chunk(head ^ inflight);
                if (rc) return rc;
                ++inflight;
            }
            if (inflight == 0) break;
            HIP_OK(hipGetLastError());
            HIP_OK(hipEventSynchronize(s->ev_chunk[head]));
            *s->h_st = s->h_chunk[head];
            head ^= 1;
            --inflight;
            if (s->h_st->done) break;
        }
    }
    launch_settle<T>(x, W.x2, s->N, st, s->stream);
    HIP_OK(hipGetLastError());
    harvest(s, s->h_st->nit);
    return M2D_OK;
}

static void fill_status(m2d_system *s, m2d_status *st, float ms)
{
    if (!st) return;
    st->status = s->h_st->status;
    st->nit = s->h_st->nit;
    st->nitm_reached = s->h_st->nitm_hit;
    st->outer = s->h_st->outer;
    st->r0 = s->h_st->r0;
    st->rn = s->h_st->rn;
    st->gammarp = s->h_st->gammarp;
    st->device_ms = ms;
}

static int solve_fp64(m2d_system *s, const double *d_a, const double *d_rhs, double *d_var,
                      double gamma, int nitm, double *d_hist, int hist_cap, m2d_status *out)
{
    HIP_OK(hipEventRecord(s->ev0, s->stream));
    int rc = run_msi<double>(s, s->w64, s->st, d_a, d_rhs, d_var, gamma, nitm, d_hist, hist_cap);
    if (rc) return rc;
    HIP_OK(hipEventRecord(s->ev1, s->stream));
    HIP_OK(hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    fill_status(s, out, ms);
    return s->h_st->status;
}

// ---------------------------------------------------------------------------
// Mixed precision: FP32 factor and FP32 MSI sweeps inside FP64 iterative
// refinement.  Outer step: r = A x - rhs in FP64 (the stopping rule of fb2d9
// on this FP64 residual), then the correction c solves A c = -r
// approximately by FP32 MSI iterations from c = 0 (A, its factor, residuals
// and sweeps in FP32) until its own residual drops by inner_gamma, then
// x += c in FP64.
// ---------------------------------------------------------------------------
__global__ void k_ir_prepare(const double *__restrict__ r64, float *__restrict__ rr32,
                             float *__restrict__ c32, long n, const SolveState *st)
{
    if (st->done) return;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) {
        rr32[q] = (float)(-r64[q]);
        c32[q] = 0.f;
    }
}

__global__ void k_ir_update(double *__restrict__ x, const float *__restrict__ c32, long n,
                            const SolveState *st)
{
    if (st->done) return;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) x[q] = x[q] + (double)c32[q];
}

static int solve_fp32_ir(m2d_system *s, const double *d_a, const double *d_rhs, double *d_var,
                         double gamma, int nitm, double *d_hist, int hist_cap, m2d_status *out)
{
    HIP_OK(hipEventRecord(s->ev0, s->stream));
    const long N = s->N;
    const unsigned nblk = (unsigned)((N + 255) / 256);
    launch_reset_state(s->st, s->stream);
    launch_residual_norm<double, double, double>(d_a, d_var, d_rhs, s->r64, s->nx, s->ny,
                                                 s->partial, s->st, nullptr,
                                                 NormRule{nullptr, 0, 1, gamma, nitm}, s->stream,
                                                 s->nd);
    HIP_OK(hipGetLastError());
    int rc = read_state(s, s->st);
    if (rc) return rc;
    SolveState outer = *s->h_st;
    int total_inner = 0;
    const int kmax = nitm > 0 ? nitm : kIterCap;
    for (int k = 1; k <= kmax && !outer.done; ++k) {
        hipLaunchKernelGGL(k_ir_prepare, dim3(nblk), dim3(256), 0, s->stream, s->r64, s->rr32,
                           s->c32, N, s->st);
        rc = run_msi<float>(s, s->w32, s->st_in, s->a32, s->rr32, s->c32, s->inner_gamma,
                            s->inner_nitm, nullptr, 0);
        if (rc) return rc;
        total_inner += s->h_st->nit;
        const int inner_status = s->h_st->status;
        hipLaunchKernelGGL(k_ir_update, dim3(nblk), dim3(256), 0, s->stream, d_var, s->c32, N,
                           s->st);
        launch_residual_norm<double, double, double>(d_a, d_var, d_rhs, s->r64, s->nx, s->ny,
                                                     s->partial, s->st, nullptr,
                                                     NormRule{d_hist, hist_cap, 0, gamma, nitm},
                                                     s->stream, s->nd);
        HIP_OK(hipGetLastError());
        rc = read_state(s, s->st);
        if (rc) return rc;
        outer = *s->h_st;
        if (inner_status == M2D_ERR_NONFINITE || inner_status == M2D_ERR_SYNC_TIMEOUT) {
            outer.status = inner_status;
            outer.done = 1;
        }
    }
    HIP_OK(hipEventRecord(s->ev1, s->stream));
    HIP_OK(hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    *s->h_st = outer;
    s->h_st->outer = outer.nit;
    s->h_st->nit = total_inner;
    fill_status(s, out, ms);
    return outer.status;
}

extern "C" int m2d_residual_norm(m2d_system *s, const double *d_a, const double *d_rhs,
                                 const double *d_var, double *out)
{
    if (!s || !d_a || !d_rhs || !d_var || !out) return M2D_ERR_ARG;
    launch_reset_state(s->st, s->stream);
    launch_residual_norm<double, double, double>(d_a, d_var, d_rhs, s->r64, s->nx, s->ny,
                                                 s->partial, s->st, nullptr,
                                                 NormRule{nullptr, 0, 1, 1.0, 1}, s->stream,
                                                 s->nd);
    HIP_OK(hipGetLastError());
    const int rc = read_state(s, s->st);
    if (rc) return rc;
    *out = s->h_st->r0;
    return M2D_OK;
}

// Per-block partial sums of (A g).b, (A g).(A g), g.b and g.(A g), blocks of
// 256 points in index order (the host adds them in block order).
__global__ __launch_bounds__(256) void k_guess_dots(const double *__restrict__ a,
                                                    const double *__restrict__ g,
                                                    const double *__restrict__ b, int nx, int ny,
                                                    double *__restrict__ part, int nd)
{
    const long N = (long)nx * ny;
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    double ab = 0.0, aa = 0.0, gb = 0.0, ga = 0.0;
    if (p < N) {
        const int j = (int)(p / nx), i = (int)(p - (long)j * nx);
        const double v = nd == 5
                             ? residual_point5<double, double, double>(a, N, g, 0.0, i, j, nx, ny)
                             : residual_point<double, double, double>(a, N, g, 0.0, i, j, nx, ny);
        ab = v * b[p];
        aa = v * v;
        gb = g[p] * b[p];
        ga = g[p] * v;
    }
    __shared__ double red[4][256];
    red[0][threadIdx.x] = ab;
    red[1][threadIdx.x] = aa;
    red[2][threadIdx.x] = gb;
    red[3][threadIdx.x] = ga;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) part[4 * blockIdx.x + k] = red[k][0];
}

__global__ void k_scale(double *__restrict__ g, long n, double alpha)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) g[q] = alpha * g[q];
}

// Scale g in place: energy = 0 by the residual-minimising alpha
// (A g . rhs) / (A g . A g), energy = 1 by the alpha that minimises the
// energy functional (alpha g).A(alpha g)/2 - (alpha g).rhs, i.e.
// (g . rhs) / (g . A g), which weighs the error's smooth components (those
// MSI removes slowest) where the residual norm weighs the oscillatory ones.
static int guess_scale(m2d_system *s, const double *d_a, const double *d_rhs, double *d_g,
                       double *alpha_out, int energy)
{
    if (!s || !d_a || !d_rhs || !d_g) return M2D_ERR_ARG;
    const long N = s->N;
    const unsigned nblk = (unsigned)((N + 255) / 256);
    hipLaunchKernelGGL(k_guess_dots, dim3(nblk), dim3(256), 0, s->stream, d_a, d_g, d_rhs, s->nx,
                       s->ny, s->r64, s->nd);
    std::vector<double> part(4 * (size_t)nblk);
    HIP_OK(hipMemcpyAsync(part.data(), s->r64, sizeof(double) * part.size(),
                          hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    double d[4] = {0.0, 0.0, 0.0, 0.0};
    for (unsigned k = 0; k < nblk; ++k)
        for (int q = 0; q < 4; ++q) d[q] += part[4 * k + q];
    const double num = energy ? d[2] : d[0], den = energy ? d[3] : d[1];
    const double alpha = den > 0.0 && isfinite(num) && isfinite(den) ? num / den : 0.0;
    hipLaunchKernelGGL(k_scale, dim3(nblk), dim3(256), 0, s->stream, d_g, N, alpha);
    HIP_OK(hipGetLastError());
    if (alpha_out) *alpha_out = alpha;
    return M2D_OK;
}

// per-block partial sums of |b - A g|^2 (blocks of 256 points in index order)
__global__ __launch_bounds__(256) void k_resid_sq(const double *__restrict__ a,
                                                  const double *__restrict__ g,
                                                  const double *__restrict__ b, int nx, int ny,
                                                  double *__restrict__ part, int nd)
{
    const long N = (long)nx * ny;
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    double rr = 0.0;
    if (p < N) {
        const int j = (int)(p / nx), i = (int)(p - (long)j * nx);
        const double v = nd == 5
                             ? residual_point5<double, double, double>(a, N, g, b[p], i, j, nx, ny)
                             : residual_point<double, double, double>(a, N, g, b[p], i, j, nx, ny);
        rr = v * v;
    }
    __shared__ double red[256];
    red[threadIdx.x] = rr;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// guess_scale for `count` systems of the handle's grid with two host reads
// in all (the dots of every system, then the residual norms of the scaled
// guesses): alpha_out[q], and rnorm_out[q] = ||rhs_q - A_q g_q|| after the
// scaling (block partials added in block order).
extern "C" int m2d_guess_prepare_many(m2d_system *s, int count, const double *const *d_a,
                                      const double *const *d_rhs, double *const *d_g,
                                      int energy, double *alpha_out, double *rnorm_out)
{
    if (!s || count < 1 || !d_a || !d_rhs || !d_g || !alpha_out || !rnorm_out)
        return M2D_ERR_ARG;
    for (int q = 0; q < count; ++q)
        if (!d_a[q] || !d_rhs[q] || !d_g[q]) return M2D_ERR_ARG;
    const long N = s->N;
    const unsigned nblk = (unsigned)((N + 255) / 256);
    double *part = nullptr;
    HIP_OK(hipMalloc((void **)&part, sizeof(double) * 4 * (size_t)nblk * count));
    std::vector<double> h(4 * (size_t)nblk * count);
    int rc = M2D_OK;
    for (int q = 0; q < count; ++q)
        hipLaunchKernelGGL(k_guess_dots, dim3(nblk), dim3(256), 0, s->stream, d_a[q], d_g[q],
                           d_rhs[q], s->nx, s->ny, part + 4 * (size_t)nblk * q, s->nd);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                       s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        rc = M2D_ERR_HIP;
    for (int q = 0; q < count && !rc; ++q) {
        double d[4] = {0.0, 0.0, 0.0, 0.0};
        const double *pq = h.data() + 4 * (size_t)nblk * q;
        for (unsigned k = 0; k < nblk; ++k)
            for (int c = 0; c < 4; ++c) d[c] += pq[4 * k + c];
        const double num = energy ? d[2] : d[0], den = energy ? d[3] : d[1];
        const double alpha = den > 0.0 && isfinite(num) && isfinite(den) ? num / den : 0.0;
        alpha_out[q] = alpha;
        hipLaunchKernelGGL(k_scale, dim3(nblk), dim3(256), 0, s->stream, d_g[q], N, alpha);
        hipLaunchKernelGGL(k_resid_sq, dim3(nblk), dim3(256), 0, s->stream, d_a[q], d_g[q],
                           d_rhs[q], s->nx, s->ny, part + (size_t)nblk * q, s->nd);
    }
    if (!rc && (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(h.data(), part, sizeof(double) * (size_t)nblk * count,
                               hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
                hipStreamSynchronize(s->stream) != hipSuccess))
        rc = M2D_ERR_HIP;
    for (int q = 0; q < count && !rc; ++q) {
        double r2 = 0.0;
        for (unsigned k = 0; k < nblk; ++k) r2 += h[(size_t)nblk * q + k];
        rnorm_out[q] = sqrt(r2);
    }
    (void)hipFree(part);
    return rc;
}

extern "C" int m2d_guess_scale(m2d_system *s, const double *d_a, const double *d_rhs,
                               double *d_g, double *alpha_out)
{
    return guess_scale(s, d_a, d_rhs, d_g, alpha_out, 0);
}

extern "C" int m2d_guess_scale_energy(m2d_system *s, const double *d_a, const double *d_rhs,
                                      double *d_g, double *alpha_out)
{
    return guess_scale(s, d_a, d_rhs, d_g, alpha_out, 1);
}

extern "C" int m2d_solve(m2d_system *s, const double *d_a, const double *d_rhs, double *d_var,
                         double gamma, int nitm, double *d_hist, int hist_cap, m2d_status *st)
{
    if (!s || !d_a || !d_rhs || !d_var || (d_hist && hist_cap < 1)) return M2D_ERR_ARG;
    if (!s->factored) return M2D_ERR_ARG;
    if (s->precision == M2D_FP64)
        return solve_fp64(s, d_a, d_rhs, d_var, gamma, nitm, d_hist, hist_cap, st);
    return solve_fp32_ir(s, d_a, d_rhs, d_var, gamma, nitm, d_hist, hist_cap, st);
}

// ---------------------------------------------------------------------------
// Device-resident pressure-correction loop of the internal-flow case
// ---------------------------------------------------------------------------
// max |x| over n values into out[0] (one workgroup, fixed reduction order);
// no-op while *gate is set.
// max |x| over n values into *out, which holds 0 on entry (k_pl_start, then
// k_pl_check after each use): per-block maxima combined by an unsigned
// 64-bit atomic maximum of their bit patterns, which order as the values do
// for non-negative doubles (fmax drops NaN and the maxima start at 0, so
// none is NaN); the maximum does not depend on the order.
constexpr int kAbsmaxThreads = 256;
__global__ __launch_bounds__(kAbsmaxThreads) void k_absmax(const double *__restrict__ x, long n,
                                                           double *__restrict__ out,
                                                           const int *gate)
{
    if (*gate) return;
    __shared__ double red[kAbsmaxThreads];
    double m = 0.0;
    for (long q = (long)blockIdx.x * kAbsmaxThreads + threadIdx.x; q < n;
         q += (long)gridDim.x * kAbsmaxThreads)
        m = fmax(m, fabs(x[q]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = kAbsmaxThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicMax(reinterpret_cast<unsigned long long *>(out),
                  (unsigned long long)__double_as_longlong(red[0]));
}

// Start of a pressure-correction loop: iterations armed, none run.
// pl = 0 (mach2d_main.f90:257) by every thread; the per-call scalars into
// the control block and the loop armed by the first.
__global__ __launch_bounds__(256) void k_pl_start(SolveState *st, PlCtl *c, double *pl, long n,
                                                  double dt, double beta, double p_avg,
                                                  double tolm)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e < n) pl[e] = 0.0;
    if (e != 0) return;
    st->skip = 0;
    c->dt = dt;
    c->beta = beta;
    c->p_avg = p_avg;
    c->tolm = tolm;
    c->itp = 0;
    c->mx = 0.0;
    c->normpl = 0.0;
    c->overrun = 0;
    c->maxit = 0;
}

// After a p' solve enqueued with fewer iterations than nitm_p: a solve
// whose stopping rule has not fired ran out of enqueued iterations.  The
// loop ends there (every later kernel no-ops) and the host runs the call
// again from the saved fields with all nitm_p iterations enqueued.
__global__ void k_pl_solve_check(SolveState *st, PlCtl *c)
{
    if (threadIdx.x != 0 || st->skip) return;
    if (!st->done) {
        c->overrun = 1;
        st->skip = 1;
        st->done = 1;
    }
}

// The fields the loop updates in place (ue, ve, un, vn, Uce, Vcn, roe, ron,
// ro, p) to (save) or from the handle's copy, one launch.
struct PlFields10 {
    double *f[10];
};
constexpr int kPlSaved = 10;
__global__ __launch_bounds__(256) void k_pl_fields_copy(PlFields10 p, double *save, long n,
                                                        int to_save)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    double *slot = save + (long)blockIdx.y * n;
    double *fld = p.f[blockIdx.y];
    if (to_save)
        slot[e] = fld[e];
    else
        fld[e] = slot[e];
}

static void pl_fields_copy(const m2d_pl_loop_fields *f, double *save, long n, bool to_save,
                           hipStream_t s)
{
    const PlFields10 p = {{f->ue, f->ve, f->un, f->vn, f->Uce, f->Vcn, f->roe, f->ron, f->ro,
                           f->p}};
    hipLaunchKernelGGL(k_pl_fields_copy, dim3((unsigned)((n + 255) / 256), kPlSaved), dim3(256),
                       0, s, p, save, n, to_save ? 1 : 0);
}

// Exit test at the end of one iteration (mach2d_main.f90:306-308): count
// it; a solver error, or normpl = max|pl| / p_avg below tolm, ends the loop
// (st->skip: the iterations enqueued after this one are no-ops and the
// solver state stays that of this iteration's solve).
__global__ void k_pl_check(SolveState *st, PlCtl *c)
{
    if (threadIdx.x != 0 || st->skip) return;
    c->itp += 1;
    if (st->nit > c->maxit) c->maxit = st->nit;
    const bool err = st->outside || st->err_pivot || st->abort ||
                     st->status == M2D_ERR_NONFINITE || st->status == M2D_ERR_SYNC_TIMEOUT;
    if (!err) c->normpl = c->mx / c->p_avg;
    c->mx = 0.0;  // k_absmax's next maximum starts from 0
    if (err || c->normpl < c->tolm) {
        st->skip = 1;
        st->done = 1;
    }
}

// lu2d5 / lu2d9 enqueued without reading the state (errors land in it);
// gated: a no-op once a loop around it has ended (st->skip).
static void enqueue_factor64(m2d_system *s, const double *d_a, bool gated = false)
{
    Work<double> &W = s->w64;
    begin(s, s->st, gated);
    launch_check_outside(d_a, s->nx, s->ny, s->st, s->stream, s->nd);
    if (s->nd == 5)
        launch_to_skew5<double, double>(d_a, W.a, W.sk, s->stream);
    else
        launch_to_skew<double, double>(d_a, W.a, 9, W.sk, s->stream, kP);
    launch_factor<double>(W.a, W.dl, W.du, 0.5, W.sk, make_link(s, s->st, 0), s->xcd_stride,
                          s->stream, s->nd == 5);
    launch_prepare_factor<double>(W.dl, W.du, W.sk, s->stream, s->st);
}

// fb2d5 / fb2d9 with at most nitm iterations, enqueued without host reads:
// the device stopping rule turns the iterations after it into no-ops.
// enq < nitm enqueues only the first enq of them (the stopping rule keeps
// nitm).
static void enqueue_solve64(m2d_system *s, const double *a, const double *rhs, double *x,
                            double gamma, int nitm, int enq = -1)
{
    Work<double> &W = s->w64;
    launch_residual_norm<double, double, double>(a, x, rhs, W.r, s->nx, s->ny, s->partial,
                                                 s->st, &W.sk, NormRule{nullptr, 0, 1, gamma, nitm},
                                                 s->stream, s->nd);
    const int n = enq < 0 ? nitm : std::min(enq, nitm);
    for (int k = 1; k <= n; ++k)
        enqueue_iteration<double>(s, W, s->st, a, rhs, x, nullptr, 0, k, gamma, nitm);
    launch_settle<double>(x, W.x2, s->N, s->st, s->stream);
}

// `count` iterations of the pressure-correction loop (mach2d_main.f90:261-309)
// enqueued on s->stream with no host read: every kernel is gated by
// s->st->skip, which the exit test of the iteration that ends the loop sets.
// The assembly calls read dt and beta from the control block.  Each p'
// solve gets solve_cap of its nitm_p iterations (checked by
// k_pl_solve_check when fewer).
static int enqueue_pl_iterations(m2d_system *s, const m2d_pl_loop_fields *f, int modvis,
                                 int nitm_p, double tol_p, int count, int solve_cap)
{
    const int nx = s->nx, ny = s->ny;
    const double dt = 0.0, beta = 0.0;  // placeholders: the kernels read s->pl
    auto ok = [&]() {
        m2d_status q;
        m2d_last_status(&q);
        return q.status == M2D_OK;
    };
    const hipStream_t keep = assembly_stream(s->stream);
    assembly_gate(&s->st->skip, &s->pl->dt);
    int rc = M2D_OK;
    for (int k = 0; k < count && !rc; ++k) {
        m2d_get_p_coefficients(&nx, &ny, &dt, f->rp, f->re, f->rn, f->Jp, f->Uce, f->Vcn,
                               f->roe, f->ron, f->g, f->de, f->dn, f->ap);
        if (ok()) m2d_intflow_set_bcpl(&nx, &ny, f->x, f->xp, f->pl, f->plin, f->ap, f->bp);
        if (ok())
            m2d_get_p_source(&nx, &ny, &dt, f->rp, f->re, f->rn, f->Jp, f->roe, f->ron, f->ro,
                             f->roa, f->Uce, f->Vcn, f->de, f->dn, f->g, f->pl, f->bp);
        if (!ok()) {
            rc = M2D_ERR_ARG;
            break;
        }
        enqueue_factor64(s, f->ap, true);
        enqueue_solve64(s, f->ap, f->bp, f->pl, tol_p, nitm_p, solve_cap);
        if (solve_cap < nitm_p)
            hipLaunchKernelGGL(k_pl_solve_check, dim3(1), dim3(64), 0, s->stream, s->st, s->pl);
        m2d_get_velocity_correction_at_faces_with_pl(&nx, &ny, f->xe, f->ye, f->xk, f->yk,
                                                     f->due, f->dve, f->dun, f->dvn, f->de,
                                                     f->dn, f->pl, f->ue, f->ve, f->un, f->vn,
                                                     f->Uce, f->Vcn);
        if (ok())
            m2d_intflow_velocities_at_boundary_faces(&nx, &ny, &modvis, f->x, f->xp, f->xe,
                                                     f->ye, f->xk, f->yk, f->u, f->v, f->ue,
                                                     f->ve, f->un, f->vn, f->Uce, f->Vcn);
        if (ok()) m2d_get_pressure_correction_with_pl(&nx, &ny, f->pl, f->p);
        if (ok()) m2d_get_density_correction_with_pl(&nx, &ny, f->pl, f->g, f->ro);
        if (ok()) m2d_get_density_at_faces(&nx, &ny, &beta, f->ro, f->Uce, f->Vcn, f->roe, f->ron);
        if (!ok()) {
            rc = M2D_ERR_ARG;
            break;
        }
        hipLaunchKernelGGL(k_absmax,
                           dim3((unsigned)std::min<long>(1024, (s->N + kAbsmaxThreads - 1) /
                                                                   kAbsmaxThreads)),
                           dim3(kAbsmaxThreads), 0, s->stream, f->pl, s->N, &s->pl->mx,
                           &s->st->skip);
        hipLaunchKernelGGL(k_pl_check, dim3(1), dim3(64), 0, s->stream, s->st, s->pl);
        if (hipGetLastError() != hipSuccess) rc = M2D_ERR_HIP;
    }
    assembly_gate(nullptr, nullptr);
    assembly_stream(keep);
    return rc;
}

// The captured batch of `count` iterations for these arguments: captured on
// s->cap and instantiated the first time, replayed after that (the per-call
// scalars live in the control block, so a batch serves every call with the
// same arrays and loop parameters).
static int pl_graph_for(m2d_system *s, const m2d_pl_loop_fields *f, int modvis, int nitm_p,
                        double tol_p, int count, int solve_cap, hipGraphExec_t *out)
{
    for (const PlGraph &g : s->plg)
        if (!memcmp(&g.f, f, sizeof(*f)) && g.modvis == modvis && g.nitm == nitm_p &&
            g.tol == tol_p && g.iters == count && g.xcd_stride == s->xcd_stride &&
            g.staged == (int)s->staged && g.solve_cap == solve_cap) {
            *out = g.exec;
            return M2D_OK;
        }
    if (!s->cap) HIP_OK(hipStreamCreateWithFlags(&s->cap, hipStreamNonBlocking));
    const hipStream_t user = s->stream;
    const bool timing = s->timing;
    s->stream = s->cap;
    s->timing = false;
    HIP_OK(hipStreamBeginCapture(s->cap, hipStreamCaptureModeRelaxed));
    int rc = enqueue_pl_iterations(s, f, modvis, nitm_p, tol_p, count, solve_cap);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s->cap, &g);
    s->stream = user;
    s->timing = timing;
    if (!rc && e != hipSuccess) rc = M2D_ERR_HIP;
    PlGraph pg;
    pg.f = *f;
    pg.modvis = modvis;
    pg.nitm = nitm_p;
    pg.tol = tol_p;
    pg.iters = count;
    pg.xcd_stride = s->xcd_stride;
    pg.staged = (int)s->staged;
    pg.solve_cap = solve_cap;
    pg.exec = nullptr;
    if (!rc && hipGraphInstantiate(&pg.exec, g, nullptr, nullptr, 0) != hipSuccess)
        rc = M2D_ERR_HIP;
    if (g) (void)hipGraphDestroy(g);
    if (rc) return rc;
    if (s->plg.size() >= 8) {  // a few shapes of batch at most; drop the oldest
        (void)hipGraphExecDestroy(s->plg.front().exec);
        s->plg.erase(s->plg.begin());
    }
    s->plg.push_back(pg);
    *out = pg.exec;
    return M2D_OK;
}

extern "C" int m2d_set_pl_graph(m2d_system *s, int use_graph)
{
    if (!s || (use_graph != 0 && use_graph != 1)) return M2D_ERR_ARG;
    s->pl_graph = use_graph != 0;
    return M2D_OK;
}

extern "C" int m2d_pl_loop_info(const m2d_system *s, int *host_reads, int *batches)
{
    if (!s || !host_reads || !batches) return M2D_ERR_ARG;
    *host_reads = s->pl_reads;
    *batches = s->pl_batches;
    return M2D_OK;
}

extern "C" int m2d_intflow_pl_loop(m2d_system *s, const m2d_pl_loop_fields *f, int modvis,
                                   double dt, double beta, double p_avg, double tolm, int itpmax,
                                   int nitm_p, double tol_p, int *itp_out, double *normpl_out)
{
    if (!s || !f || s->nd != 5 || s->precision != M2D_FP64 || itpmax < 1 || nitm_p < 1 ||
        !(p_avg != 0.0))
        return M2D_ERR_ARG;
    const void *need[] = {f->x,   f->xp,  f->xe,  f->ye,  f->xk,  f->yk,  f->rp,  f->re,
                          f->rn,  f->Jp,  f->due, f->dve, f->dun, f->dvn, f->de,  f->dn,
                          f->g,   f->roa, f->u,   f->v,   f->plin, f->ue, f->ve,  f->un,
                          f->vn,  f->Uce, f->Vcn, f->roe, f->ron, f->ro,  f->p,   f->pl,
                          f->ap,  f->bp};
    for (const void *p : need)
        if (!p) return M2D_ERR_ARG;
    if (!s->pl) {
        HIP_OK(hipMalloc((void **)&s->pl, sizeof(PlCtl)));
        HIP_OK(hipHostMalloc((void **)&s->h_pl, sizeof(PlCtl), hipHostMallocDefault));
    }
    // MSI iterations enqueued per p' solve: by default the previous call's
    // longest solve plus kPlCapMargin, rounded up to a multiple of 4 (few
    // distinct captured batches); the first call on a handle enqueues all
    // nitm_p.  A solve that needs more ends the loop early, and the call
    // runs again from the fields saved here with every iteration enqueued,
    // so the results are those of nitm_p enqueued iterations either way.
    constexpr int kPlCapMargin = 2;
    int cap = nitm_p;
    if (s->pl_cap_set > 0)
        cap = std::min(nitm_p, s->pl_cap_set);
    else if (s->pl_cap_set == 0 && s->pl_maxit > 0)
        cap = std::min(nitm_p, (s->pl_maxit + kPlCapMargin + 3) / 4 * 4);
    s->pl_cap_used = cap;
    s->pl_reruns = 0;
    if (cap < nitm_p) {
        if (!s->pl_save)
            HIP_OK(hipMalloc((void **)&s->pl_save, sizeof(double) * kPlSaved * s->N));
        pl_fields_copy(f, s->pl_save, s->N, true, s->stream);
        HIP_OK(hipGetLastError());
    }
    PlCtl *h = s->h_pl;
    s->pl_reads = s->pl_batches = 0;
    int rc = M2D_OK;
    // one run of the loop with `solve_cap` iterations enqueued per solve
    auto run_loop = [&](int solve_cap) -> int {
        // pl = 0 (:257) and the per-call scalars, ahead of the iterations
        memset(h, 0, sizeof(*h));
        hipLaunchKernelGGL(k_pl_start, dim3((unsigned)((s->N + 255) / 256)), dim3(256), 0,
                           s->stream, s->st, s->pl, f->pl, s->N, dt, beta, p_avg, tolm);
        if (hipGetLastError() != hipSuccess) return M2D_ERR_HIP;
        // batches of iterations, each followed by one read of the control
        // block and the solver state.  The first batch is as long as the
        // previous call's loop, so a steady run reads the host once per
        // call; the ones after it hold one iteration each (an iteration past
        // the exit costs its launches, about 18 + 2 solve_cap of them, as
        // no-ops).
        int r = M2D_OK, run = 0;
        const int first = std::min(itpmax, std::max(1, s->pl_last));
        bool ended = false;
        while (!r && !ended && run < itpmax) {
            const int count = run == 0 ? first : 1;
            if (s->pl_graph) {
                hipGraphExec_t exec = nullptr;
                r = pl_graph_for(s, f, modvis, nitm_p, tol_p, count, solve_cap, &exec);
                if (!r && hipGraphLaunch(exec, s->stream) != hipSuccess) r = M2D_ERR_HIP;
            } else {
                r = enqueue_pl_iterations(s, f, modvis, nitm_p, tol_p, count, solve_cap);
            }
            if (r) break;
            run += count;
            ++s->pl_batches;
            if (hipMemcpyAsync(h, s->pl, sizeof(PlCtl), hipMemcpyDeviceToHost, s->stream) !=
                hipSuccess) {
                r = M2D_ERR_HIP;
                break;
            }
            r = read_state(s, s->st);  // also completes the copy of the control block
            ++s->pl_reads;
            if (r) break;
            const SolveState &st = *s->h_st;
            if (st.outside) r = M2D_ERR_OUTSIDE_COUPLING;
            else if (st.err_pivot || st.status == M2D_ERR_NONFINITE) r = M2D_ERR_NONFINITE;
            else if (st.abort || st.status == M2D_ERR_SYNC_TIMEOUT) r = M2D_ERR_SYNC_TIMEOUT;
            ended = st.skip != 0;
        }
        return r;
    };
    rc = run_loop(cap);
    if (!rc && h->overrun) {
        pl_fields_copy(f, s->pl_save, s->N, false, s->stream);
        rc = hipGetLastError() == hipSuccess ? run_loop(nitm_p) : M2D_ERR_HIP;
        s->pl_reruns = 1;
    }
    const int itp = std::min(std::max(h->itp, 1), itpmax);
    if (!rc) {
        s->pl_last = h->itp;
        s->pl_maxit = h->maxit;
    }
    if (itp_out) *itp_out = itp;
    if (normpl_out) *normpl_out = h->normpl;
    set_last(rc);
    return rc;
}

extern "C" int m2d_set_pl_solve_cap(m2d_system *s, int cap)
{
    if (!s) return M2D_ERR_ARG;
    s->pl_cap_set = cap;
    return M2D_OK;
}

extern "C" int m2d_pl_loop_cap_info(const m2d_system *s, int *cap, int *reruns)
{
    if (!s || !cap || !reruns) return M2D_ERR_ARG;
    *cap = s->pl_cap_used;
    *reruns = s->pl_reruns;
    return M2D_OK;
}

// ---------------------------------------------------------------------------
// Reference-signature entry points (host arrays).  A per-thread cache keeps
// the device buffers and an FP64 handle for the last grid size.
// ---------------------------------------------------------------------------
struct HostCache {
    int nx = 0, ny = 0, nd = 0;
    m2d_system *sys = nullptr;
    double *a = nullptr, *dl = nullptr, *du = nullptr, *rhs = nullptr, *var = nullptr;
    ~HostCache() { release(); }
    void release()
    {
        double *bufs[] = {a, dl, du, rhs, var};
        for (double *b : bufs)
            if (b) (void)hipFree(b);
        a = dl = du = rhs = var = nullptr;
        if (sys) m2d_destroy(sys);
        sys = nullptr;
        nx = ny = nd = 0;
    }
    int ensure(int nx_, int ny_, int nd_)
    {
        if (sys && nx == nx_ && ny == ny_ && nd == nd_) return M2D_OK;
        release();
        int rc = m2d_create_stencil(nx_, ny_, M2D_FP64, nd_, &sys);
        if (rc) return rc;
        const long N = (long)nx_ * ny_;
        if (dalloc(&a, 9 * N) || dalloc(&dl, 5 * N) || dalloc(&du, 4 * N) || dalloc(&rhs, N) ||
            dalloc(&var, N)) {
            release();
            return M2D_ERR_HIP;
        }
        nx = nx_;
        ny = ny_;
        nd = nd_;
        return M2D_OK;
    }
};
static thread_local HostCache g_cache;

static bool sizes_ok(int nij, int ni, int nj)
{
    return ni >= 3 && nj >= 3 && (long)ni * nj == (long)nij;
}

// lu2d9 / lu2d5 on host arrays: b has nd diagonals, dl nd == 9 ? 5 : 4 and
// du nd == 9 ? 4 : 3.
static void host_lu(const double *b, double *dl, double *du, const int *nij, const int *ni,
                    const int *nj, int nd)
{
    if (!b || !dl || !du || !nij || !ni || !nj || !sizes_ok(*nij, *ni, *nj)) {
        set_last(M2D_ERR_ARG);
        return;
    }
    const long N = *nij;
    const int nl = nd == 9 ? 5 : 4, nu = nd == 9 ? 4 : 3;
    int rc = g_cache.ensure(*ni, *nj, nd);
    if (rc) {
        set_last(rc);
        return;
    }
    m2d_system *s = g_cache.sys;
    if (hipMemcpy(g_cache.a, b, sizeof(double) * nd * N, hipMemcpyHostToDevice) != hipSuccess) {
        set_last(M2D_ERR_HIP);
        return;
    }
    rc = m2d_factor(s, g_cache.a);
    set_last(rc);
    if (rc == M2D_OK || rc == M2D_ERR_NONFINITE) {
        int rc2 = m2d_get_factor(s, g_cache.dl, g_cache.du);
        if (rc2 ||
            hipMemcpy(dl, g_cache.dl, sizeof(double) * nl * N, hipMemcpyDeviceToHost) !=
                hipSuccess ||
            hipMemcpy(du, g_cache.du, sizeof(double) * nu * N, hipMemcpyDeviceToHost) != hipSuccess)
            set_last(M2D_ERR_HIP);
    }
}

extern "C" void m2d_lu2d9(const double *b, double *dl, double *du, const int *nij,
                          const int *ni, const int *nj)
{
    host_lu(b, dl, du, nij, ni, nj, 9);
}

extern "C" void m2d_lu2d5(const double *b, double *dl, double *du, const int *nij,
                          const int *ni, const int *nj)
{
    host_lu(b, dl, du, nij, ni, nj, 5);
}

static void host_fb(const double *b, const double *dl, const double *du, int nj, int ni, int nij,
                    double *var, double gamma, int nitm, const double *rhs, bool factor_here,
                    int nd)
{
    const long N = nij;
    const int nl = nd == 9 ? 5 : 4, nu = nd == 9 ? 4 : 3;
    int rc = g_cache.ensure(ni, nj, nd);
    if (rc) {
        set_last(rc);
        return;
    }
    m2d_system *s = g_cache.sys;
    bool ok = hipMemcpy(g_cache.a, b, sizeof(double) * nd * N, hipMemcpyHostToDevice) ==
                  hipSuccess &&
              hipMemcpy(g_cache.rhs, rhs, sizeof(double) * N, hipMemcpyHostToDevice) ==
                  hipSuccess &&
              hipMemcpy(g_cache.var, var, sizeof(double) * N, hipMemcpyHostToDevice) ==
                  hipSuccess;
    if (!ok) {
        set_last(M2D_ERR_HIP);
        return;
    }
    if (factor_here) {
        rc = m2d_factor(s, g_cache.a);
        if (rc && rc != M2D_ERR_NONFINITE) {
            set_last(rc);
            return;
        }
    } else {
        ok = hipMemcpy(g_cache.dl, dl, sizeof(double) * nl * N, hipMemcpyHostToDevice) ==
                 hipSuccess &&
             hipMemcpy(g_cache.du, du, sizeof(double) * nu * N, hipMemcpyHostToDevice) ==
                 hipSuccess;
        if (!ok) {
            set_last(M2D_ERR_HIP);
            return;
        }
        // fb2d9 / fb2d5 do not factor; still reject couplings across the grid edge
        begin(s, s->st);
        launch_check_outside(g_cache.a, ni, nj, s->st, s->stream, nd);
        rc = read_state(s, s->st);
        if (!rc && s->h_st->outside) rc = M2D_ERR_OUTSIDE_COUPLING;
        if (!rc) rc = put_factor(s, g_cache.dl, g_cache.du);
        if (rc) {
            set_last(rc);
            return;
        }
    }
    m2d_status st;
    memset(&st, 0, sizeof(st));
    rc = m2d_solve(s, g_cache.a, g_cache.rhs, g_cache.var, gamma, nitm, nullptr, 0, &st);
    g_last = st;
    g_last.status = rc;
    if (rc == M2D_ERR_HIP) return;
    if (hipMemcpy(var, g_cache.var, sizeof(double) * N, hipMemcpyDeviceToHost) != hipSuccess)
        g_last.status = M2D_ERR_HIP;
}

extern "C" void m2d_fb2d9(const double *b, const double *dl, const double *du, const int *nj,
                          const int *ni, const int *nij, double *var, const double *gamma,
                          const int *nitm, const double *rhs)
{
    if (!b || !dl || !du || !nj || !ni || !nij || !var || !gamma || !nitm || !rhs ||
        !sizes_ok(*nij, *ni, *nj)) {
        set_last(M2D_ERR_ARG);
        return;
    }
    host_fb(b, dl, du, *nj, *ni, *nij, var, *gamma, *nitm, rhs, false, 9);
}

extern "C" void m2d_fb2d5(const double *b, const double *dl, const double *du, const int *nj,
                          const int *ni, const int *nij, double *var, const double *gamma,
                          const int *nitm, const double *rhs)
{
    if (!b || !dl || !du || !nj || !ni || !nij || !var || !gamma || !nitm || !rhs ||
        !sizes_ok(*nij, *ni, *nj)) {
        set_last(M2D_ERR_ARG);
        return;
    }
    host_fb(b, dl, du, *nj, *ni, *nij, var, *gamma, *nitm, rhs, false, 5);
}

extern "C" void m2d_solve_msi9d(const int *nx, const int *ny, const int *nitm, const double *tol,
                                const double *a, const double *b, double *sol)
{
    if (!nx || !ny || !nitm || !tol || !a || !b || !sol || *nx < 3 || *ny < 3) {
        set_last(M2D_ERR_ARG);
        return;
    }
    host_fb(a, nullptr, nullptr, *ny, *nx, *nx * *ny, sol, *tol, *nitm, b, true, 9);
}

extern "C" void m2d_solve_msi5d(const int *nx, const int *ny, const int *nitm, const double *tol,
                                const double *a, const double *b, double *sol)
{
    if (!nx || !ny || !nitm || !tol || !a || !b || !sol || *nx < 3 || *ny < 3) {
        set_last(M2D_ERR_ARG);
        return;
    }
    host_fb(a, nullptr, nullptr, *ny, *nx, *nx * *ny, sol, *tol, *nitm, b, true, 5);
}

// ---------------------------------------------------------------------------
// class_solver_msi9d%solve on device arrays without allocation.  Every
// device buffer the solve needs beyond the caller's a, b and sol (the skewed
// copy of A, the factor, r, w/z, the second iterate, the residual partial sums, the boundary
// words and the solve state) is carved from one file-scope device array.
// Each GPU holds its own copy of the array from the library's load until
// the process ends, and every call on that GPU reuses it: calls that
// overlap on one GPU (two host threads, or two streams) race on it.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocBytes = size_t(48) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocDevices = 64;
static unsigned char *g_scratch_dev[kNoAllocDevices];  // device address of `scratch`, per device
static std::mutex g_scratch_mu;

// Device address of this device's `scratch` (looked up once per device, with
// that device current).  In host code the array's name is not a device
// address; only this pointer is.
static int scratch_address(unsigned char **out)
{
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocDevices) return M2D_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = (unsigned char *)p;
    }
    *out = g_scratch_dev[dev];
    return M2D_OK;
}

// Byte offsets of the solve's buffers inside `scratch` (each 256-byte
// aligned); returns the bytes needed.
struct NoAllocLayout {
    size_t a, dl, du, r, w, x2, partial, bnd, st, total;
};
static NoAllocLayout noalloc_layout(int nx, int ny)
{
    auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
    const size_t L = (size_t)skew_for(nx, ny, (int)sizeof(double)).size() * sizeof(double);
    NoAllocLayout o;
    o.a = 0;
    o.dl = o.a + up(9 * L);
    o.du = o.dl + up(6 * L);
    o.r = o.du + up(4 * L);
    o.w = o.r + up(L);
    o.x2 = o.w + up(L);
    o.partial = o.x2 + up(sizeof(double) * (size_t)nx * ny);
    o.bnd = o.partial + up(sizeof(u64) * (size_t)residual_partial_words(nx, ny));
    o.st = o.bnd + up(sizeof(u64) * (size_t)(boundary_words(nx, ny) + scratch_words(ny)));
    o.total = o.st + up(sizeof(SolveState));
    return o;
}

extern "C" int m2d_solve_msi9d_noalloc(const int *nx, const int *ny, const int *nitm,
                                       const double *tol, const double *d_a, const double *d_b,
                                       double *d_sol, void *stream)
{
    if (!nx || !ny || !nitm || !tol || !d_a || !d_b || !d_sol || *nx < 3 || *ny < 3 ||
        (long)*nx * *ny > (1L << 31) / 16) {
        set_last(M2D_ERR_ARG);
        return M2D_ERR_ARG;
    }
    const NoAllocLayout o = noalloc_layout(*nx, *ny);
    if (o.total > kNoAllocBytes) {
        fprintf(stderr,
                "mach2d_hip: m2d_solve_msi9d_noalloc: a %dx%d grid needs %zu bytes of solver "
                "memory, more than the %zu bytes the library holds for it\n",
                *nx, *ny, o.total, kNoAllocBytes);
        set_last(M2D_ERR_ARG);
        return M2D_ERR_ARG;
    }
    unsigned char *base = nullptr;
    int rc = scratch_address(&base);
    if (rc) {
        set_last(rc);
        return rc;
    }
    // a host-side view of the carved buffers; it owns nothing and is never
    // passed to m2d_destroy
    m2d_system s;
    SolveState h_st;
    s.nx = *nx;
    s.ny = *ny;
    s.xcd_stride = (*ny + 63) / 64 >= 5 ? 8 : 1;  // as m2d_create_stencil
    s.N = (long)*nx * *ny;
    s.nd = 9;
    s.precision = M2D_FP64;
    s.stream = (hipStream_t)stream;
    s.np = residual_blocks(*nx, *ny);
    s.nbnd = boundary_words(*nx, *ny) + scratch_words(*ny);
    s.w64.sk = skew_for(*nx, *ny, (int)sizeof(double));
    s.w64.a = (double *)(base + o.a);
    s.w64.dl = (double *)(base + o.dl);
    s.w64.du = (double *)(base + o.du);
    s.w64.r = (double *)(base + o.r);
    s.w64.w = (double *)(base + o.w);
    s.w64.x2 = (double *)(base + o.x2);
    s.partial = (u64 *)(base + o.partial);
    s.bnd = (u64 *)(base + o.bnd);
    s.st = (SolveState *)(base + o.st);
    s.h_st = &h_st;
    // the array keeps the last call's contents: the padding entries of r and
    // w (read by the sweeps, never written by the residual) start at zero
    const size_t L = (size_t)s.w64.sk.size() * sizeof(double);
    HIP_OK(hipMemsetAsync(s.w64.r, 0, L, s.stream));
    HIP_OK(hipMemsetAsync(s.w64.w, 0, L, s.stream));
    // the residual's tagged partial sums start from generation zero
    HIP_OK(hipMemsetAsync(s.partial, 0, sizeof(u64) * (size_t)residual_partial_words(*nx, *ny),
                          s.stream));
    enqueue_factor64(&s, d_a);
    HIP_OK(hipGetLastError());
    rc = read_state(&s, s.st);
    if (!rc && h_st.outside) rc = M2D_ERR_OUTSIDE_COUPLING;
    if (!rc && h_st.abort) rc = M2D_ERR_SYNC_TIMEOUT;
    if (rc) {
        set_last(rc);
        return rc;
    }
    rc = run_msi<double>(&s, s.w64, s.st, d_a, d_b, d_sol, *tol, *nitm, nullptr, 0);
    if (rc) {
        set_last(rc);
        return rc;
    }
    m2d_status st;
    memset(&st, 0, sizeof(st));
    fill_status(&s, &st, 0.f);
    g_last = st;
    return st.status;
}

// Batched entry: every system keeps its own handle, factor and arrays.
extern "C" int m2d_enable_kernel_timing(m2d_system *s, int on)
{
    if (!s) return M2D_ERR_ARG;
    s->timing = on != 0;
    for (int k = 0; k < 8; ++k) {
        s->kt_ms[k] = 0.0;
        s->kt_n[k] = 0;
    }
    s->recs.clear();
    s->pool_used = 0;
    return M2D_OK;
}

extern "C" int m2d_kernel_timing(m2d_system *s, int kernel, double *total_ms, long *launches)
{
    if (!s || kernel < 0 || kernel > 7 || !total_ms || !launches) return M2D_ERR_ARG;
    *total_ms = s->kt_ms[kernel];
    *launches = s->kt_n[kernel];
    return M2D_OK;
}

// ---------------------------------------------------------------------------
// Device-resident time step of the internal-flow case
// ---------------------------------------------------------------------------
namespace m2d {
int norm_l1_device(const double *a, const double *b, const double *var, int nx, int ny,
                   double *d_work, double *h_out, hipStream_t s);
int norm_l1_work_size(int nx, int ny);
int norm_l1_enqueue(const double *a, const double *b, const double *var, int nx, int ny,
                    double *d_work, double *d_res, hipStream_t s);
void norm5_enqueue(const double *var, const double *b, const double *a, int nx, int ny,
                   double *d_scr, double *d_res, hipStream_t s);
int time_cycle_enqueue(int NX, int NY, double g, double rg, double po, double t0,
                       const double *x, const double *xp, const double *u, double *uin,
                       double *vin, double *pin, double *Tin, double *Mw, double *plin,
                       double *pina, double *p, double *d_scr, double *h_buf, hipStream_t s);
int tdma_device_many(int count, int nx, int ny, int nitm, const double *const *a,
                     const double *const *b, double *const *x, hipStream_t s);
int tdma_enqueue_many(int count, int nx, int ny, int nitm, const double *const *a,
                      const double *const *b, double *const *x, hipStream_t s);
}  // namespace m2d

// dst[k] = src[k] for the K arrays of n values of a list, one launch
// (grid row k copies array k)
struct CopyList {
    const double *src[9];
    double *dst[9];
};
__global__ __launch_bounds__(256) void k_copy_list(CopyList c, long n)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e < n) c.dst[blockIdx.y][e] = c.src[blockIdx.y][e];
}

// g = 1 / (Rg T) (mach2d_main.f90:153)
__global__ __launch_bounds__(256) void k_g_nodes(const double *__restrict__ T, double Rg, long n,
                                                 double *__restrict__ g)
{
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q < n) g[q] = 1.0 / (Rg * T[q]);
}

// T from the conservation of total enthalpy (mach2d_main.f90:352)
__global__ __launch_bounds__(256) void k_T_nodes(const double *__restrict__ u,
                                                 const double *__restrict__ v, double Href,
                                                 double Cpref, long n, double *__restrict__ T)
{
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q < n) T[q] = (Href - (u[q] * u[q] + v[q] * v[q]) / 2.0) / Cpref;
}

// solver9d%solve (mod_solver_msi.f90:96-133 or mod_solver_tdma.f90) on the
// 9-diagonal system a x = b from the current x
static int solve9(m2d_system *s9, const m2d_time_step_params *P, const double *a,
                  const double *b, double *x, int nx, int ny, hipStream_t stream, hipEvent_t dep)
{
    if (P->solver9d == 1) return m2d_tdma2d9_device(nx, ny, P->nitm_u, a, b, x, stream);
    // the 9-diagonal handle has its own stream: it waits for the system
    // assembled on `stream` through `dep` (no host wait); read_state below
    // completes the solve before `stream` uses x
    if (hipEventRecord(dep, stream) != hipSuccess ||
        hipStreamWaitEvent(s9->stream, dep, 0) != hipSuccess)
        return M2D_ERR_HIP;
    enqueue_factor64(s9, a);
    enqueue_solve64(s9, a, b, x, P->tol_u, P->nitm_u);
    int rc = read_state(s9, s9->st);
    if (rc) return rc;
    const SolveState &st = *s9->h_st;
    if (st.outside) return M2D_ERR_OUTSIDE_COUPLING;
    if (st.err_pivot || st.status == M2D_ERR_NONFINITE) return M2D_ERR_NONFINITE;
    if (st.abort || st.status == M2D_ERR_SYNC_TIMEOUT) return M2D_ERR_SYNC_TIMEOUT;
    return M2D_OK;
}

// The time step's side stream and its scratch: the residual norms' work
// (norm_l1_work_size), then the in-order sums': the average of p (n + 2
// doubles), the terms of the four mean coefficients and of the mass check
// (5 n) and their five sums, n = (nx-2)(ny-2), then kTsNormSlots sums of the
// step's residual norms, then the inflow update's 3 (ny-2) values; h_ts
// receives p's average, the five sums and the norms' sums, and holds the
// inflow update's host values.  Created with the first step on the handle.
constexpr int kTsNormSlots = 8;
static int ts_side_ready(m2d_system *s)
{
    if (s->aux) return M2D_OK;
    const long n = (long)(s->nx - 2) * (s->ny - 2), nr = 3L * (s->ny - 2);
    HIP_OK(hipMalloc((void **)&s->ts_scr,
                     sizeof(double) * (size_t)(norm_l1_work_size(s->nx, s->ny) +
                                               average_scratch(s->nx, s->ny) + 5 * n + 5 +
                                               kTsNormSlots + nr)));
    HIP_OK(hipHostMalloc((void **)&s->h_ts, sizeof(double) * (size_t)(6 + kTsNormSlots + nr),
                         hipHostMallocDefault));
    HIP_OK(hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&s->ev_dep, hipEventDisableTiming));
    HIP_OK(hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking));
    return M2D_OK;
}

extern "C" int m2d_intflow_time_step(m2d_system *s9, m2d_system *s5,
                                     const m2d_time_step_fields *f,
                                     const m2d_time_step_params *P, double *dt_io,
                                     m2d_time_step_result *res)
{
    // every parameter is checked here, before the step changes any field
    if (!s5 || !f || !P || !dt_io || s5->nd != 5 || s5->precision != M2D_FP64 ||
        P->itmmax < 1 || P->nitm_u < 1 || P->itpmax < 1 || P->nitm_p < 1 ||
        (P->modvis != 0 && P->modvis != 1) || (P->solver9d != 0 && P->solver9d != 1))
        return M2D_ERR_ARG;
    if (P->solver9d == 0 && (!s9 || s9->nd != 9 || s9->precision != M2D_FP64 ||
                             s9->nx != s5->nx || s9->ny != s5->ny))
        return M2D_ERR_ARG;
    const bool euler = P->modvis == 0;
    const void *need[] = {f->x,    f->xp,   f->xe,   f->ye,   f->xk,   f->yk,  f->xke, f->yke,
                          f->xen,  f->yen,  f->rp,   f->re,   f->rn,   f->Jp,  f->Je,  f->Jn,
                          f->alphae, f->betae, f->betan, f->gamman, f->vle, f->vln, f->vlp,
                          f->u,    f->v,    f->ue,   f->ve,   f->un,   f->vn,  f->Uce, f->Vcn,
                          f->p,    f->T,    f->ro,   f->roe,  f->ron,  f->pl,  f->ua,  f->va,
                          f->uea,  f->vea,  f->una,  f->vna,  f->pa,   f->Ta,  f->roa, f->au,
                          f->av,   f->at,   f->ap,   f->bu,   f->bv,   f->bt,  f->bp,  f->cup,
                          f->sup,  f->cvp,  f->svp,  f->due,  f->dve,  f->dun, f->dvn, f->de,
                          f->dn,   f->g,    f->uin,  f->vin,  f->pin,  f->Tin, f->Mw,  f->plin,
                          f->pina, f->Tbe,  f->Tbw,  f->Tbn,  f->Tbs};
    for (const void *q : need)
        if (!q) return M2D_ERR_ARG;
    if (!euler && (!f->cp || !f->ke || !f->kn || !f->fbe || !f->fbn)) return M2D_ERR_ARG;
    set_last(M2D_OK);  // the stage checks below read the thread's last status
    const int nx = s5->nx, ny = s5->ny;
    const long N = s5->N;
    const hipStream_t st = s5->stream;
    const hipStream_t keep = assembly_stream(st);
    const int modvis = P->modvis, coord = P->coord;
    double dt = *dt_io, beta = P->beta;
    int rc = M2D_OK;
    auto ok = [&]() {
        m2d_status q;
        m2d_last_status(&q);
        if (q.status && !rc) rc = q.status;
        return rc == M2D_OK;
    };
    const unsigned nb = (unsigned)((N + 255) / 256);
    if (ts_side_ready(s5) != M2D_OK) {
        assembly_stream(keep);
        return M2D_ERR_HIP;
    }
    double *const work = s5->ts_scr;
    double *const side = work + norm_l1_work_size(nx, ny);
    const long nin = (long)(nx - 2) * (ny - 2);
    double norm = 0.0, rmass = 0.0, normpl = 0.0;
    int itp = 0, stage = 1;
    // The residual norms the step adds up (mach2d_main.f90:229-236, 311,
    // 380): their sums are enqueued into device slots as the step goes and
    // read once with the step's last wait, then added in the same order.
    // kind 9: norm_l1_9d relative (two sums, slots k and k+1); kind 5:
    // norm_l1_5d (one sum).  A new mass cycle starts the list again.
    double *const nres = side + average_scratch(nx, ny) + 5 * nin + 5;
    int nkind[kTsNormSlots], nslot[kTsNormSlots], nops = 0, nused = 0;
    auto norm9_relative = [&](const double *a, const double *b, const double *x) {
        if (rc) return;
        rc = norm_l1_enqueue(a, b, x, nx, ny, work, nres + nused, st);
        nkind[nops] = 9;
        nslot[nops++] = nused;
        nused += 2;
    };
    auto norm5 = [&](const double *a, const double *b, const double *x) {
        if (rc) return;
        norm5_enqueue(x, b, a, nx, ny, work, nres + nused, st);
        if (hipGetLastError() != hipSuccess) rc = M2D_ERR_HIP;
        nkind[nops] = 5;
        nslot[nops++] = nused;
        nused += 1;
    };
    // time levels (mach2d_main.f90:96-104), one launch
    {
        const CopyList cl = {{f->ro, f->T, f->p, f->u, f->v, f->ue, f->ve, f->un, f->vn},
                             {f->roa, f->Ta, f->pa, f->ua, f->va, f->uea, f->vea, f->una, f->vna}};
        hipLaunchKernelGGL(k_copy_list, dim3(nb, 9), dim3(256), 0, st, cl, N);
        if (hipGetLastError() != hipSuccess) rc = M2D_ERR_HIP;
    }
    // inflow update and fictitious pressures (:107)
    stage = 2;
    if (!rc)
        rc = time_cycle_enqueue(nx, ny, P->gamma, P->Rg, P->P0, P->T0, f->x, f->xp, f->u, f->uin,
                                f->vin, f->pin, f->Tin, f->Mw, f->plin, f->pina, f->p,
                                nres + kTsNormSlots, s5->h_ts + 6 + kTsNormSlots, st);
    // get_average(p) (:151) on the side stream, beside the u and v stages:
    // only the p' loop changes p, and it waits for the value
    // The fork is recorded here, at p after the inflow update; the side
    // stream's calls are issued once the u and v stages are queued, so the
    // device is not left without work while the host makes them.
    double p_avg = 0.0;
    bool avg_forked = false, avg_pending = false;
    if (ok()) stage = 3;
    if (!rc) {
        if (hipEventRecord(s5->ev_fork, st) != hipSuccess) rc = M2D_ERR_HIP;
        avg_forked = !rc;
    }
    auto avg_enqueue = [&]() {
        if (!avg_forked || rc) return;
        avg_forked = false;
        double *scr = side;
        if (hipStreamWaitEvent(s5->aux, s5->ev_fork, 0) != hipSuccess) {
            rc = M2D_ERR_HIP;
            return;
        }
        average_enqueue(nx, ny, f->p, scr, s5->aux);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(s5->h_ts, scr, sizeof(double), hipMemcpyDeviceToHost, s5->aux) !=
                hipSuccess ||
            hipEventRecord(s5->ev_join, s5->aux) != hipSuccess) {
            rc = M2D_ERR_HIP;
            return;
        }
        avg_pending = true;
    };
    if (ok()) {
        hipLaunchKernelGGL(k_g_nodes, dim3(nb), dim3(256), 0, st, f->T, P->Rg, N, f->g);
        if (hipGetLastError() != hipSuccess) rc = M2D_ERR_HIP;
    }
    // The in-order sums that end the step run on the side stream: the mass
    // check's (:346-348), whose terms are taken before the densities change,
    // and, for Euler steps, where au, av, at, ap are final once the last p'
    // loop has run, the four mean coefficients of the next time step (:394).
    // Their terms are taken on st right after the last p' loop (neither
    // reads u or v, which the rest of that cycle changes), so the sums, the
    // step's longest tail, start as early as they can.  Scratch: terms of
    // the four means at t[0, 4n), of the mass check at t[4n, 5n), the sums
    // at t[5n, 5n + 5).
    double *tsum = side + average_scratch(nx, ny);
    bool side_forked = false, side_pending = false;
    auto sums_enqueue = [&]() {
        if (side_forked || rc) return;
        side_forked = true;
        if (euler) {
            const double *a4[4] = {f->au, f->av, f->at, f->ap};
            const int nd4[4] = {9, 9, 9, 5};
            cc_terms_enqueue(nx, ny, 4, a4, nd4, tsum, st);
        }
        const double *mi[10] = {f->rp, f->re, f->rn,  f->Jp,  f->ro,
                                f->roa, f->ron, f->roe, f->Uce, f->Vcn};
        mass_terms_enqueue(nx, ny, dt, mi, tsum + 4 * nin, st);
        const int first = euler ? 0 : 4, K = 5 - first;
        if (hipGetLastError() != hipSuccess || hipEventRecord(s5->ev_fork, st) != hipSuccess ||
            hipStreamWaitEvent(s5->aux, s5->ev_fork, 0) != hipSuccess) {
            rc = M2D_ERR_HIP;
            return;
        }
        seq_sums_enqueue(tsum + first * nin, nin, K, tsum + 5 * nin + first, s5->aux);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(s5->h_ts + 1 + first, tsum + 5 * nin + first, sizeof(double) * K,
                           hipMemcpyDeviceToHost, s5->aux) != hipSuccess ||
            hipEventRecord(s5->ev_join, s5->aux) != hipSuccess) {
            rc = M2D_ERR_HIP;
            return;
        }
        side_pending = true;
    };
    for (int itm = 1; itm <= P->itmmax && !rc; ++itm) {  // mass cycle (:155-344)
        nops = nused = 0;  // norm = 0
        stage = 4;
        m2d_get_u_coefficients(&nx, &ny, &modvis, &dt, f->rp, f->re, f->rn, f->Jp, f->Je, f->Jn,
                               f->ye, f->yk, f->alphae, f->betae, f->betan, f->gamman, f->vle,
                               f->vln, f->roe, f->ron, f->roa, f->Uce, f->Vcn, f->au);
        if (ok())
            m2d_get_u_source(&nx, &ny, &modvis, &beta, &dt, f->rp, f->re, f->rn, f->xe, f->ye,
                             f->xk, f->yk, f->xke, f->yke, f->xen, f->yen, f->Jp, f->Je, f->Jn,
                             f->roe, f->ron, f->roa, f->p, f->vle, f->vln, f->Uce, f->Vcn,
                             f->ua, f->u, f->v, f->cup, f->sup, f->bu);
        if (ok())
            m2d_intflow_set_bcu(&nx, &ny, &modvis, f->x, f->xp, f->xk, f->yk, f->u, f->v, f->au,
                                f->bu);
        if (ok())
            m2d_get_v_coefficients(&nx, &ny, &coord, &modvis, &dt, f->rp, f->re, f->rn, f->Jp,
                                   f->Je, f->Jn, f->xe, f->xk, f->alphae, f->betae, f->betan,
                                   f->gamman, f->vle, f->vln, f->vlp, f->roe, f->ron, f->roa,
                                   f->Uce, f->Vcn, f->av);
        if (ok())
            m2d_get_v_source(&nx, &ny, &coord, &modvis, &beta, &dt, f->rp, f->re, f->rn, f->xe,
                             f->ye, f->xk, f->yk, f->xke, f->yke, f->xen, f->yen, f->Jp, f->Je,
                             f->Jn, f->roe, f->ron, f->roa, f->p, f->vle, f->vln, f->Uce,
                             f->Vcn, f->va, f->u, f->v, f->cvp, f->svp, f->bv);
        if (ok())
            m2d_intflow_set_bcv(&nx, &ny, &modvis, f->x, f->xp, f->xk, f->yk, f->u, f->v, f->av,
                                f->bv);
        if (ok())
            m2d_get_internal_simplec_coefficients(&nx, &ny, f->re, f->rn, f->xe, f->ye, f->xk,
                                                  f->yk, f->au, f->av, f->due, f->dve, f->dun,
                                                  f->dvn, f->de, f->dn);
        if (ok())
            m2d_intflow_boundary_simplec(&nx, &ny, &modvis, f->due, f->dve, f->dun, f->dvn,
                                         f->de, f->dn);
        if (!ok()) break;
        avg_enqueue();
        if (rc) break;
        stage = 5;
        if (P->solver9d == 1) {
            // TDMA: u and v are independent systems (both assembled above),
            // solved in the same launches; each result is its solo one
            const double *as[2] = {f->au, f->av}, *bs[2] = {f->bu, f->bv};
            double *xs[2] = {f->u, f->v};
            rc = tdma_enqueue_many(2, nx, ny, P->nitm_u, as, bs, xs, st);
            if (rc) break;
            norm9_relative(f->au, f->bu, f->u);
            norm9_relative(f->av, f->bv, f->v);
            if (rc) break;
        } else {
            rc = solve9(s9, P, f->au, f->bu, f->u, nx, ny, st, s5->ev_dep);
            norm9_relative(f->au, f->bu, f->u);
            if (rc) break;
            stage = 6;
            rc = solve9(s9, P, f->av, f->bv, f->v, nx, ny, st, s5->ev_dep);
            norm9_relative(f->av, f->bv, f->v);
            if (rc) break;
        }
        stage = 7;
        m2d_get_velocities_at_internal_faces(&nx, &ny, &dt, f->rp, f->re, f->rn, f->xe, f->ye,
                                             f->xk, f->yk, f->xen, f->yen, f->xke, f->yke,
                                             f->Jp, f->cup, f->cvp, f->sup, f->svp, f->au, f->av,
                                             f->roa, f->p, f->u, f->v, f->uea, f->vea, f->una,
                                             f->vna, f->ue, f->ve, f->un, f->vn, f->Uce, f->Vcn);
        if (ok())
            m2d_intflow_velocities_at_boundary_faces(&nx, &ny, &modvis, f->x, f->xp, f->xe,
                                                     f->ye, f->xk, f->yk, f->u, f->v, f->ue,
                                                     f->ve, f->un, f->vn, f->Uce, f->Vcn);
        if (!ok()) break;
        stage = 8;
        m2d_pl_loop_fields L = {f->x,   f->xp,  f->xe,  f->ye,  f->xk,  f->yk,  f->rp,
                                f->re,  f->rn,  f->Jp,  f->due, f->dve, f->dun, f->dvn,
                                f->de,  f->dn,  f->g,   f->roa, f->u,   f->v,   f->plin,
                                f->ue,  f->ve,  f->un,  f->vn,  f->Uce, f->Vcn, f->roe,
                                f->ron, f->ro,  f->p,   f->pl,  f->ap,  f->bp};
        avg_enqueue();  // (already issued after the u and v stages)
        if (rc) break;
        if (avg_pending) {
            if (hipEventSynchronize(s5->ev_join) != hipSuccess) {
                rc = M2D_ERR_HIP;
                break;
            }
            p_avg = s5->h_ts[0];
            avg_pending = false;
        }
        rc = m2d_intflow_pl_loop(s5, &L, modvis, dt, beta, p_avg, P->tolm, P->itpmax,
                                 P->nitm_p, P->tol_p, &itp, &normpl);
        if (rc) break;
        if (itm == P->itmmax) sums_enqueue();
        if (rc) break;
        stage = 9;
        norm5(f->ap, f->bp, f->pl);
        if (ok())
            m2d_get_u_v_at_real_nodes_with_pl(&nx, &ny, f->xe, f->ye, f->xk, f->yk, f->rp, f->pl,
                                              f->au, f->av, f->u, f->v);
        if (ok())
            m2d_intflow_extrapolate_u_v_to_fictitious(&nx, &ny, &modvis, f->x, f->xp, f->xk,
                                                      f->yk, f->u, f->v);
        ok();
    }
    if (!rc) stage = 10;
    sums_enqueue();  // (issued after the last p' loop)
    if (!rc) stage = 11;
    if (!rc && euler) {  // Euler, constant properties: T from H (:349-362)
        hipLaunchKernelGGL(k_T_nodes, dim3(nb), dim3(256), 0, st, f->u, f->v, P->Href, P->Cpref,
                           N, f->T);
        if (hipGetLastError() != hipSuccess) rc = M2D_ERR_HIP;
        if (!rc)
            m2d_intflow_boundary_T_from_H_conservation(&nx, &ny, &P->Cpref, &P->Href, f->Tin,
                                                       f->u, f->ue, f->un, f->v, f->ve, f->vn,
                                                       f->T, f->Tbe, f->Tbw, f->Tbn, f->Tbs);
        ok();
    } else if (!rc) {  // T equation (:363-380)
        m2d_get_T_coefficients_and_source(
            &nx, &ny, &modvis, &beta, &dt, f->rp, f->re, f->rn, f->xe, f->ye, f->xk, f->yk,
            f->xke, f->yke, f->xen, f->yen, f->alphae, f->betae, f->betan, f->gamman, f->Jp,
            f->Je, f->Jn, f->roe, f->ron, f->ro, f->roa, f->p, f->pa, f->cp, f->vle, f->vln,
            f->ke, f->kn, f->Uce, f->Vcn, f->u, f->v, f->ue, f->ve, f->un, f->vn, f->ua, f->va,
            f->T, f->Ta, f->fbe, f->fbn, f->at, f->bt);
        if (ok()) m2d_intflow_set_bcT(&nx, &ny, &P->Twall, f->x, f->xp, f->T, f->Tin, f->at, f->bt);
        if (ok()) rc = solve9(s9, P, f->at, f->bt, f->T, nx, ny, st, s5->ev_dep);
        norm9_relative(f->at, f->bt, f->T);
    }
    if (!rc) stage = 12;
    if (!rc) {  // state equation and face densities (:386-391)
        m2d_get_density_at_nodes(&nx, &ny, &P->Rg, f->p, f->T, f->ro);
        if (ok()) m2d_get_density_at_faces(&nx, &ny, &beta, f->ro, f->Uce, f->Vcn, f->roe, f->ron);
        ok();
    }
    if (!rc && side_pending) {
        if (hipEventSynchronize(s5->ev_join) != hipSuccess)
            rc = M2D_ERR_HIP;
        else
            rmass = s5->h_ts[5];
    }
    if (!rc) stage = 13;
    if (!rc && euler) {  // next time step (:394), means from the side stream
        // mean = sum / n, the division get_mean_cc makes, here on the host
        double cc[4];
        for (int k = 0; k < 4; ++k) cc[k] = s5->h_ts[1 + k] / (double)nin;
        *dt_io = scarborough_dt(dt, cc, P->h0, P->mincc, P->maxcc);
    } else if (!rc) {  // next time step (:394)
        const double nd = m2d_tstepper_dt_scarborough(&nx, &ny, &dt, f->au, f->av, f->at, f->ap,
                                                      &P->h0, &P->mincc, &P->maxcc);
        if (ok()) *dt_io = nd;
    }
    if (!rc && nused &&
        hipMemcpyAsync(s5->h_ts + 6, nres, sizeof(double) * nused, hipMemcpyDeviceToHost, st) !=
            hipSuccess)
        rc = M2D_ERR_HIP;
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = M2D_ERR_HIP;
    if (!rc) {  // the norms, added in the step's order
        const double eps = 10.0 * 2.220446049250313080847e-16;  // 10*epsilon(1d0)
        const double *o = s5->h_ts + 6;
        for (int k = 0; k < nops; ++k) {
            const double *q = o + nslot[k];
            if (nkind[k] == 9)
                norm = (q[1] > eps) ? norm + q[0] / q[1] : norm + q[0];
            else
                norm = norm + q[0];
        }
    }
    // nothing of the step stays in flight on the side stream, on any path
    if (s5->aux && hipStreamSynchronize(s5->aux) != hipSuccess && !rc) rc = M2D_ERR_HIP;
    assembly_stream(keep);
    if (res) {
        res->norm = norm;
        res->rmass = rmass;
        res->normpl = normpl;
        res->itp = itp;
        res->stage = rc ? stage : 0;
    }
    set_last(rc);
    return rc;
}
