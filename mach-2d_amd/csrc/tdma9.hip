// tdma9.hip -- gfx950 kernels of the alternating-direction line TDMA for
// 9-diagonal systems (reference: src/solvers/tdma2d9.f90:38-322) and of the
// L1 residual norms (src/solvers/mod_solvers.f90:128-276).
//
// One TDMA iteration is two line-Jacobi passes.  x-pass: every grid row is a
// tridiagonal system in its W/E couplings with the other six neighbours
// taken from x as it was before the pass; y-pass: every column, S/N
// implicit, explicit neighbours from the x-pass result.  The Thomas
// recurrences of a line,
//
//     den(t) = a_P + a_lo P(t-1),  P(t) = -a_up / den(t)      (t >= 1)
//     Q(t)   = (D(t) - a_lo Q(t-1)) / den(t),  x(t) = P(t) x(t+1) + Q(t)
//
// (den(0) = a_P, Q(0) = D(0) / a_P), keep the reference's operation order
// (bit-identical results), but the work is split by what it depends on:
//
//   * P and den depend on A only: one chain per line, once per call
//     (k_tdma_pre), for all nitm iterations, with 1/den, so that the
//     quotient of Q is taken exactly through the reciprocal (two Markstein
//     corrections, div_by_rcp) instead of a full division on the chain;
//   * D, the right-hand side with the explicit neighbours, depends on x but
//     not on the recurrence: one thread per point (k_tdma_d);
//   * what is left per iteration is the chain of Q and the back
//     substitution (k_tdma_line): one lane per line, a subtraction, a
//     multiplication and a division per step forward, a multiplication and
//     an addition back.
//
// A chain runs on lane l < R of the first wave of its workgroup, one line
// per lane.  Its operands stream through an LDS ring: the wave itself
// issues direct-to-LDS loads (no registers) three chunks ahead and waits on
// its vector-memory counter for the chunk it needs.  Its results (Q, then x
// over it; den and P in k_tdma_pre) stay in an LDS tile that every thread
// of the workgroup stores once the chain has ended, so the chain issues no
// global store (a store ahead of a load in the in-order vector-memory
// counter would hold up the wait for that load).  The operands are read in
// a "pass layout" where the lanes' lines are adjacent in memory: element
// (t, l) of line l at step t sits at t*LP + l (LP = the pass's lines padded
// to a multiple of 64), written once per call by k_tdma_setup (the
// coefficients) and per iteration by k_tdma_d (D).  Measured on one
// MI355X at 258x130 (u and v together): an x-pass chain step (forward and
// back) costs ~110 ns and a y-pass step ~150 ns, against ~25 ns of
// dependent FP64 arithmetic (6 cycles per dependent operation): the LDS
// operand traffic and the ring's load issue take the rest.  Several
// systems of one grid size (u and v of a SIMPLEC iteration) run in the
// same launches.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "m2d_internal.h"
#include "msi_stencil.h"

namespace m2d {

// LDS of a chain workgroup, in doubles (144 KB of the 160 KB of a gfx950 CU)
constexpr int kTdmaLds = 18432;
// systems per launch (kernel-argument struct)
constexpr int kTdmaMaxSys = 8;
// Operand ring of a chain: kTdmaSlots chunks (kTdmaSlots - 1 in flight), a
// chunk being kTdmaChunk doubles of each streamed array, i.e. 512/R steps
// of the workgroup's R lines, loaded by kTdmaNI 16-byte direct-to-LDS loads
// per array (64 lanes x 2 doubles each).
constexpr int kTdmaSlots = 4;
constexpr int kTdmaNI = 4;
constexpr int kTdmaChunk = 128 * kTdmaNI;
// longest line (grid dimension) the chain kernels take
constexpr int kTdmaMaxLen = 2048;

// Lines of a pass padded to a multiple of 64 (the pass layout's row pitch).
__host__ __device__ inline long tdma_lp(int lines) { return ((long)lines + 63) / 64 * 64; }

// Lanes (lines) of a chain workgroup: the largest R in 64, 32, ..., 2 whose
// `tiles` LDS tiles fit beside a ring of `na` arrays; 0 if none does.  A
// tile holds R+1 columns of tdma_rows(len, R) rows.  With R = 32 it has
// 64-R extra rows, and idle lane q >= R writes the padding column of row
// t + q - R, its own address (the idle lanes' writes do not collide, and
// need no branch); with R < 32 the idle lanes' writes are masked off.
__host__ __device__ constexpr int tdma_rows(int len, int R) { return R >= 32 ? len + 64 - R : len; }
// the tile writes of lanes beyond R are masked off (see tdma_rows)
template <int R>
__device__ __forceinline__ bool tdma_writes(int lane)
{
    return R >= 32 || lane < R;
}
__host__ __device__ inline int tdma_lanes(int len, int tiles, int na)
{
    const int budget = kTdmaLds - kTdmaSlots * na * kTdmaChunk;
    int R = 64;
    while (R > 2 && (R + 1) * tiles * tdma_rows(len, R) > budget) R >>= 1;
    return (R + 1) * tiles * tdma_rows(len, R) <= budget ? R : 0;
}
constexpr int kTdmaPreArrays = 3, kTdmaPreTiles = 2;    // k_tdma_pre
constexpr int kTdmaLineArrays = 4, kTdmaLineTiles = 1;  // k_tdma_line

struct TdmaSys {
    const double *a, *b;
    double *x;
    double *w;  // tdma_scratch(nx, ny) doubles
};
struct TdmaBatch {
    TdmaSys s[kTdmaMaxSys];
    int count, nx, ny;
};

// Per-system scratch: for each pass, seven arrays in the pass layout
// (element (t, l) of line l at step t at t*LP + l; x-pass lines are rows,
// LP = tdma_lp(ny), steps the nx columns; y-pass lines are columns, LP =
// tdma_lp(nx), steps the ny rows): the coefficients a_lo, a_P, a_up (x:
// a_W, a_P, a_E; y: a_S, a_P, a_N), den, 1/den, P and D; padding lines hold
// a_P = 1, a_lo = a_up = D = 0.  Then the x-pass result (reference layout).
enum TdmaArr { kTLo = 0, kTMp, kTUp, kTDen, kTRd, kTP, kTD, kTArrays };

__host__ __device__ inline long tdma_pass_size(int nx, int ny, bool x)
{
    return x ? tdma_lp(ny) * nx : tdma_lp(nx) * ny;
}
__host__ __device__ inline long tdma_scratch(int nx, int ny)
{
    const long N = (long)nx * ny;
    return kTArrays * (tdma_pass_size(nx, ny, true) + tdma_pass_size(nx, ny, false)) +
           (N + 1) / 2 * 2;
}
__device__ __forceinline__ double *tdma_arr(const TdmaBatch &B, const TdmaSys &S, bool x, int k)
{
    const long Sx = tdma_pass_size(B.nx, B.ny, true), Sy = tdma_pass_size(B.nx, B.ny, false);
    return S.w + (x ? (long)k * Sx : kTArrays * Sx + (long)k * Sy);
}
__device__ __forceinline__ double *tdma_tmp(const TdmaBatch &B, const TdmaSys &S)
{
    return S.w + kTArrays * (tdma_pass_size(B.nx, B.ny, true) + tdma_pass_size(B.nx, B.ny, false));
}

// ---------------------------------------------------------------------------
// Parallel cyclic reduction of the line systems (m2d_set_tdma_mode(1)): the
// same line-Jacobi passes, each line's tridiagonal system
//
//     lo(t) x(t-1) + d(t) x(t) + up(t) x(t+1) = D(t)   (lo(0), up(n-1) unused,
//                                                       as in the recurrence)
//
// solved by reduction instead of the Thomas recurrence: step s = 1, 2, 4, ...
// combines row t with rows t-s and t+s (k1 = lo(t)/d(t-s), k2 = up(t)/d(t+s))
// until every row stands alone, x(t) = D'(t)/d'(t).  What depends on A only
// (k1, k2 of every step, 1/d') is computed once per call (k_tdma_pcr_setup);
// an iteration forms each line's D and applies the steps to it in one
// workgroup per line (k_tdma_pcr_line): ceil(log2 n) rounds of two FMAs per
// point with the line in LDS, instead of a recurrence of n dependent steps.  The result is the same solution up to rounding (not the
// reference's rounding).  Lines of up to kPcrMaxLen points.
// ---------------------------------------------------------------------------
constexpr int kPcrMaxLen = 400;  // the line's D and factors fit 64 KB of LDS
// a line's workgroup: one lane per point, ceil(n / 64) waves
constexpr int kPcrMaxThreads = (kPcrMaxLen + 63) / 64 * 64;
inline unsigned pcr_threads(int n) { return (unsigned)((n + 63) / 64 * 64); }
__host__ __device__ inline int pcr_steps(int n)
{
    int k = 0;
    while ((1 << k) < n) ++k;
    return k;
}
// per system, after the recurrence's scratch: for the x-pass (lines = ny,
// n = nx) then the y-pass (lines = nx, n = ny), line-major:
// K [line][step][2][n] (k1, k2), R [line][n] (1/d')
__host__ __device__ inline long pcr_pass_size(int nx, int ny, bool x)
{
    const int n = x ? nx : ny;
    return (long)nx * ny * (2 * pcr_steps(n) + 1);
}
__host__ __device__ inline long pcr_scratch(int nx, int ny)
{
    return pcr_pass_size(nx, ny, true) + pcr_pass_size(nx, ny, false);
}
__device__ __forceinline__ double *pcr_base(const TdmaBatch &B, const TdmaSys &S, bool x)
{
    return S.w + tdma_scratch(B.nx, B.ny) + (x ? 0 : pcr_pass_size(B.nx, B.ny, true));
}
__device__ __forceinline__ double *pcr_k(const TdmaBatch &B, const TdmaSys &S, bool x, int line)
{
    const int n = x ? B.nx : B.ny;
    return pcr_base(B, S, x) + (long)line * 2 * pcr_steps(n) * n;
}
__device__ __forceinline__ double *pcr_r(const TdmaBatch &B, const TdmaSys &S, bool x)
{
    const int n = x ? B.nx : B.ny;
    return pcr_base(B, S, x) + (long)B.nx * B.ny * 2 * pcr_steps(n);
}

// Coefficients of both passes in their pass layouts (padding lines
// included, with their D), once per call.
__global__ __launch_bounds__(256) void k_tdma_setup(TdmaBatch B)
{
    const int nx = B.nx, ny = B.ny;
    const long N = (long)nx * ny;
    const long Sx = tdma_pass_size(nx, ny, true), Sy = tdma_pass_size(nx, ny, false);
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const int q = (int)(e / (Sx + Sy));
    if (q >= B.count) return;
    const TdmaSys S = B.s[q];
    long r = e - (long)q * (Sx + Sy);
    const bool x = r < Sx;
    if (!x) r -= Sx;
    const long LP = x ? tdma_lp(ny) : tdma_lp(nx);
    const int t = (int)(r / LP), l = (int)(r - (long)t * LP);
    double lo = 0.0, mp = 1.0, up = 0.0;
    if (l < (x ? ny : nx)) {
        const long p = x ? (long)l * nx + t : (long)t * nx + l;
        lo = S.a[(x ? kW : kS) * N + p];
        mp = S.a[kP * N + p];
        up = S.a[(x ? kE : kN) * N + p];
    } else {
        tdma_arr(B, S, x, kTD)[r] = 0.0;
    }
    tdma_arr(B, S, x, kTLo)[r] = lo;
    tdma_arr(B, S, x, kTMp)[r] = mp;
    tdma_arr(B, S, x, kTUp)[r] = up;
}

// D of the x-pass (tdma2d9.f90:79-173) at point (i, j), from the pass input x
__device__ __forceinline__ double d_xpass(const double *__restrict__ a,
                                          const double *__restrict__ b,
                                          const double *__restrict__ x, long N, int nx, int ny,
                                          int i, int j)
{
    const long p = (long)j * nx + i;
    const double bp = b[p];
    if (i == 0) {
        if (j == 0) return bp - a[kN * N + p] * x[p + nx] - a[kNE * N + p] * x[p + nx + 1];
        if (j == ny - 1) return bp - a[kS * N + p] * x[p - nx] - a[kSE * N + p] * x[p - nx + 1];
        return bp - a[kS * N + p] * x[p - nx] - a[kSE * N + p] * x[p - nx + 1] -
               a[kN * N + p] * x[p + nx] - a[kNE * N + p] * x[p + nx + 1];
    }
    if (i < nx - 1) {
        if (j == 0)
            return bp - a[kNW * N + p] * x[p + nx - 1] - a[kN * N + p] * x[p + nx] -
                   a[kNE * N + p] * x[p + nx + 1];
        if (j == ny - 1)
            return bp - a[kSW * N + p] * x[p - nx - 1] - a[kS * N + p] * x[p - nx] -
                   a[kSE * N + p] * x[p - nx + 1];
        return bp - a[kSW * N + p] * x[p - nx - 1] - a[kS * N + p] * x[p - nx] -
               a[kSE * N + p] * x[p - nx + 1] - a[kNW * N + p] * x[p + nx - 1] -
               a[kN * N + p] * x[p + nx] - a[kNE * N + p] * x[p + nx + 1];
    }
    if (j == 0) return bp - a[kNW * N + p] * x[p + nx - 1] - a[kN * N + p] * x[p + nx];
    if (j == ny - 1) return bp - a[kSW * N + p] * x[p - nx - 1] - a[kS * N + p] * x[p - nx];
    return bp - a[kSW * N + p] * x[p - nx - 1] - a[kS * N + p] * x[p - nx] -
           a[kNW * N + p] * x[p + nx - 1] - a[kN * N + p] * x[p + nx];
}

// D of the y-pass (tdma2d9.f90:200-296) at point (i, j), from the x-pass result x
__device__ __forceinline__ double d_ypass(const double *__restrict__ a,
                                          const double *__restrict__ b,
                                          const double *__restrict__ x, long N, int nx, int ny,
                                          int i, int j)
{
    const long p = (long)j * nx + i;
    const double bp = b[p];
    if (j == 0) {
        if (i == 0) return bp - a[kE * N + p] * x[p + 1] - a[kNE * N + p] * x[p + nx + 1];
        if (i == nx - 1) return bp - a[kW * N + p] * x[p - 1] - a[kNW * N + p] * x[p + nx - 1];
        return bp - a[kW * N + p] * x[p - 1] - a[kNW * N + p] * x[p + nx - 1] -
               a[kE * N + p] * x[p + 1] - a[kNE * N + p] * x[p + nx + 1];
    }
    if (j < ny - 1) {
        if (i == 0)
            return bp - a[kSE * N + p] * x[p - nx + 1] - a[kE * N + p] * x[p + 1] -
                   a[kNE * N + p] * x[p + nx + 1];
        if (i == nx - 1)
            return bp - a[kSW * N + p] * x[p - nx - 1] - a[kW * N + p] * x[p - 1] -
                   a[kNW * N + p] * x[p + nx - 1];
        return bp - a[kSW * N + p] * x[p - nx - 1] - a[kW * N + p] * x[p - 1] -
               a[kNW * N + p] * x[p + nx - 1] - a[kSE * N + p] * x[p - nx + 1] -
               a[kE * N + p] * x[p + 1] - a[kNE * N + p] * x[p + nx + 1];
    }
    if (i == 0) return bp - a[kSE * N + p] * x[p - nx + 1] - a[kE * N + p] * x[p + 1];
    if (i == nx - 1) return bp - a[kSW * N + p] * x[p - nx - 1] - a[kW * N + p] * x[p - 1];
    return bp - a[kSW * N + p] * x[p - nx - 1] - a[kW * N + p] * x[p - 1] -
           a[kSE * N + p] * x[p - nx + 1] - a[kE * N + p] * x[p + 1];
}

// D of one pass for every real point of every system, in the pass layout
// (padding lines keep the zero of k_tdma_setup).  x-pass: input x (the
// iteration's); y-pass: input the x-pass result.
template <bool XPASS>
__global__ __launch_bounds__(256) void k_tdma_d(TdmaBatch B)
{
    const int nx = B.nx, ny = B.ny;
    const long N = (long)nx * ny;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const int q = (int)(e / N);
    if (q >= B.count) return;
    const long p = e - (long)q * N;
    const int j = (int)(p / nx), i = (int)(p - (long)j * nx);
    const TdmaSys S = B.s[q];
    double *D = tdma_arr(B, S, XPASS, kTD);
    if (XPASS)
        D[(long)i * tdma_lp(ny) + j] = d_xpass(S.a, S.b, S.x, N, nx, ny, i, j);
    else
        D[(long)j * tdma_lp(nx) + i] = d_ypass(S.a, S.b, tdma_tmp(B, S), N, nx, ny, i, j);
}

typedef __attribute__((address_space(1))) const void tdma_gvoid;
typedef __attribute__((address_space(3))) void tdma_lvoid;
// ring slots as LDS pointers: a generic pointer cast to LDS per load costs a
// null test and a select each
typedef __attribute__((address_space(3))) double tdma_ldouble;

// s_waitcnt vmcnt(N) alone (expcnt and lgkmcnt left alone)
template <int N>
__device__ __forceinline__ void tdma_wait_vmem()
{
    static_assert(N >= 0 && N <= 63, "vmcnt is six bits");
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// One chunk of NA arrays into its ring slot, issued by the 64 lanes of the
// chain wave: load ii of array a brings 128 doubles, lane q the two at flat
// position f = 128 ii + 2q of the chunk, i.e. step s = f / R, lines
// l0 + f % R and the next one (adjacent in the pass layout: 16 bytes,
// aligned since LP, l0 and f % R are even).  step(s) maps the chunk's step
// index to the line step (clamped into the line by the caller's mapping).
template <int NA, int R, typename STEP>
__device__ __forceinline__ void tdma_issue(tdma_ldouble *slot, const double *const (&arr)[NA], long LP,
                                           int l0, int lane, STEP step)
{
#pragma unroll
    for (int ii = 0; ii < kTdmaNI; ++ii) {
        const int f = 128 * ii + 2 * lane;
        const long g = (long)step(f / R) * LP + l0 + (f % R);
#pragma unroll
        for (int a = 0; a < NA; ++a)
            __builtin_amdgcn_global_load_lds((tdma_gvoid *)(arr[a] + g),
                                             (tdma_lvoid *)(slot + a * kTdmaChunk + 128 * ii), 16,
                                             0, 0);
    }
}

// tdma_issue for a chunk that lies wholly inside the line, where step(s) is
// t0 + s (forward) or t0 - s (backward): the address of load ii of every
// array is the lane's fixed offset goff[ii] = (+-(f / R)) * LP + l0 + f % R
// plus the wave-uniform t0 * LP, so no lane recomputes a 64-bit product.
template <int NA>
__device__ __forceinline__ void tdma_issue_at(tdma_ldouble *slot, const double *const (&arr)[NA],
                                              long base, const long (&goff)[kTdmaNI])
{
#pragma unroll
    for (int ii = 0; ii < kTdmaNI; ++ii) {
        const long g = base + goff[ii];
#pragma unroll
        for (int a = 0; a < NA; ++a)
            __builtin_amdgcn_global_load_lds((tdma_gvoid *)(arr[a] + g),
                                             (tdma_lvoid *)(slot + a * kTdmaChunk + 128 * ii), 16,
                                             0, 0);
    }
}
template <int R>
__device__ __forceinline__ void tdma_lane_offsets(long (&goff)[kTdmaNI], long LP, int l0,
                                                  int lane, int dir)
{
#pragma unroll
    for (int ii = 0; ii < kTdmaNI; ++ii) {
        const int f = 128 * ii + 2 * lane;
        goff[ii] = (long)(dir * (f / R)) * LP + l0 + (f % R);
    }
}

// Block b of a chain launch over one pass -> (system, first line)
struct TdmaBlock {
    int q, l0, len;
    long LP;
};
template <int R>
__device__ __forceinline__ TdmaBlock tdma_block(const TdmaBatch &B, bool x, int b)
{
    TdmaBlock k;
    k.len = x ? B.nx : B.ny;
    k.LP = tdma_lp(x ? B.ny : B.nx);
    const int per = (int)(k.LP / R);
    k.q = b / per;
    k.l0 = (b - k.q * per) * R;
    return k;
}

// P, den and 1/den of every line of one pass (once per call).  The chain
// den(t) = a_P + a_lo P(t-1) (den(0) = a_P), P(t) = -a_up / den(t) runs on
// lane l < R of wave 0 with its operands streamed through the LDS ring; den
// and P go to LDS tiles, and after the chain every thread of the
// workgroup stores den, 1/den (computed there, off the chain) and P.
// 1/den (a correctly rounded quotient) lets the iterations take Q's
// quotient exactly through the reciprocal (div_by_rcp).
template <bool XPASS, int R>
__device__ __forceinline__ void tdma_pre_body(const TdmaBatch &B, int blk, double *lds)
{
    constexpr int NA = kTdmaPreArrays, RS = R + 1, K = kTdmaChunk / R;
    const TdmaBlock k = tdma_block<R>(B, XPASS, blk);
    if (k.q >= B.count) return;
    const TdmaSys S = B.s[k.q];
    const int len = k.len;
    const long T = (long)RS * tdma_rows(len, R);
    double *ring = lds, *tden = lds + kTdmaSlots * NA * kTdmaChunk, *tP = tden + T;
    tdma_ldouble *lring = (tdma_ldouble *)lds;
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        const int wl = lane < R ? lane : R + (lane - R) * RS;  // idle lanes (R >= 32)
        const int rl = lane < R ? lane : 0;
        const double *const arr[NA] = {tdma_arr(B, S, XPASS, kTLo), tdma_arr(B, S, XPASS, kTMp),
                                       tdma_arr(B, S, XPASS, kTUp)};
        const int nch = (len + K - 1) / K;
        auto issue = [&](int c) {
            tdma_issue<NA, R>(lring + (c % kTdmaSlots) * NA * kTdmaChunk, arr, k.LP, k.l0, lane,
                              [&](int s) { return min(c * K + s, len - 1); });
        };
        for (int c = 0; c < kTdmaSlots - 1; ++c) issue(c);
        double Pm = 0.0;
        for (int c = 0; c < nch; ++c) {
            issue(c + kTdmaSlots - 1);
            tdma_wait_vmem<(kTdmaSlots - 1) * NA * kTdmaNI>();
            const double *sl = ring + (c % kTdmaSlots) * NA * kTdmaChunk;
            const int t0 = c * K, ns = min(K, len - t0);
            for (int s = 0; s < ns; ++s) {
                const double lo = sl[s * R + rl], mp = sl[kTdmaChunk + s * R + rl],
                             up = sl[2 * kTdmaChunk + s * R + rl];
                const int t = t0 + s;
                const double den = t == 0 ? mp : mp + lo * Pm;
                Pm = -up / den;
                if (tdma_writes<R>(lane)) {
                    tden[t * RS + wl] = den;
                    tP[t * RS + wl] = Pm;
                }
            }
        }
        tdma_wait_vmem<0>();
    }
    __syncthreads();
    double *gden = tdma_arr(B, S, XPASS, kTDen), *grd = tdma_arr(B, S, XPASS, kTRd),
           *gP = tdma_arr(B, S, XPASS, kTP);
    for (int e = threadIdx.x; e < R * len; e += 256) {
        const int t = e / R, lane = e % R;
        const long g = (long)t * k.LP + k.l0 + lane;
        const double den = tden[t * RS + lane];
        gden[g] = den;
        grd[g] = 1.0 / den;
        gP[g] = tP[t * RS + lane];
    }
}

// Both passes' P chains in one launch: blocks [0, nbx) the x-pass lines (RX
// per block), the rest the y-pass lines (RY per block).
template <int RX, int RY>
__global__ __launch_bounds__(256) void k_tdma_pre(TdmaBatch B, int nbx)
{
    __shared__ double lds[kTdmaLds];
    if ((int)blockIdx.x < nbx)
        tdma_pre_body<true, RX>(B, blockIdx.x, lds);
    else
        tdma_pre_body<false, RY>(B, blockIdx.x - nbx, lds);
}

// Q chain and back substitution of every line of one pass (each
// iteration); the result goes to the reference layout: x-pass into the
// scratch's x-pass result, y-pass into the system's x.  The operands a_lo,
// den, 1/den, D (forward) and P (backward) stream through the LDS ring; Q
// and then x over it stay in an LDS tile, stored by every thread once the
// chain has ended.
template <bool XPASS, int R>
__global__ __launch_bounds__(256) void k_tdma_line(TdmaBatch B)
{
    __shared__ double lds[kTdmaLds];
    constexpr int NA = kTdmaLineArrays, RS = R + 1, K = kTdmaChunk / R;
    const TdmaBlock k = tdma_block<R>(B, XPASS, blockIdx.x);
    if (k.q >= B.count) return;
    const TdmaSys S = B.s[k.q];
    const int len = k.len;
    double *ring = lds, *tile = lds + kTdmaSlots * NA * kTdmaChunk;
    tdma_ldouble *lring = (tdma_ldouble *)lds;
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        const int wl = lane < R ? lane : R + (lane - R) * RS;  // idle lanes (R >= 32)
        const int rl = lane < R ? lane : 0;
        // forward: Q(0) = D(0) / den(0) (den(0) = a_P), then
        // Q(t) = (D(t) - a_lo Q(t-1)) / den(t), each quotient exact through 1/den
        {
            const double *const arr[NA] = {tdma_arr(B, S, XPASS, kTLo),
                                           tdma_arr(B, S, XPASS, kTDen),
                                           tdma_arr(B, S, XPASS, kTRd), tdma_arr(B, S, XPASS, kTD)};
            const int nch = (len + K - 1) / K;
            long goff[kTdmaNI];
            tdma_lane_offsets<R>(goff, k.LP, k.l0, lane, 1);
            auto issue = [&](int c) {
                tdma_ldouble *slot = lring + (c % kTdmaSlots) * NA * kTdmaChunk;
                if (c * K + K <= len)
                    tdma_issue_at<NA>(slot, arr, (long)(c * K) * k.LP, goff);
                else
                    tdma_issue<NA, R>(slot, arr, k.LP, k.l0, lane,
                                      [&](int s) { return min(c * K + s, len - 1); });
            };
            for (int c = 0; c < kTdmaSlots - 1; ++c) issue(c);
            double Qm = 0.0;
            for (int c = 0; c < nch; ++c) {
                issue(c + kTdmaSlots - 1);
                tdma_wait_vmem<(kTdmaSlots - 1) * NA * kTdmaNI>();
                const double *sl = ring + (c % kTdmaSlots) * NA * kTdmaChunk;
                const double *slo = sl, *sden = sl + kTdmaChunk, *srd = sl + 2 * kTdmaChunk,
                             *sD = sl + 3 * kTdmaChunk;
                const int t0 = c * K;
                if (c > 0 && t0 + K <= len) {
                    // full chunk past step 0, G steps at a time: the operands
                    // of group g+1 are read from LDS while group g computes
                    constexpr int G = 4, NG = K / G;
                    double lo[2][G], dn[2][G], rd[2][G], dv[2][G];
                    auto fetch = [&](int b, int s0) {
#pragma unroll
                        for (int u = 0; u < G; ++u) {
                            const int o = (s0 + u) * R + rl;
                            lo[b][u] = slo[o];
                            dn[b][u] = sden[o];
                            rd[b][u] = srd[o];
                            dv[b][u] = sD[o];
                        }
                    };
                    fetch(0, 0);
#pragma unroll
                    for (int g = 0; g < NG; ++g) {
                        if (g + 1 < NG) fetch((g + 1) & 1, (g + 1) * G);
                        const int b = g & 1;
#pragma unroll
                        for (int u = 0; u < G; ++u) {
                            Qm = div_by_rcp<double>(dv[b][u] - lo[b][u] * Qm, dn[b][u], rd[b][u]);
                            if (tdma_writes<R>(lane)) tile[(t0 + g * G + u) * RS + wl] = Qm;
                        }
                    }
                } else {
                    const int ns = min(K, len - t0);
                    for (int s = 0; s < ns; ++s) {
                        const int o = s * R + rl, t = t0 + s;
                        Qm = t == 0 ? div_by_rcp<double>(sD[o], sden[o], srd[o])
                                    : div_by_rcp<double>(sD[o] - slo[o] * Qm, sden[o], srd[o]);
                        if (tdma_writes<R>(lane)) tile[t * RS + wl] = Qm;
                    }
                }
            }
            tdma_wait_vmem<0>();  // the ring is reused below
        }
        // back substitution: x(len-1) = Q(len-1) (already in the tile), then
        // x(t) = P(t) x(t+1) + Q(t) (x-pass, tdma2d9.f90:175-193) or
        // x(t) = x(t+1) P(t) + Q(t) (y-pass, :298-316), t = len-2 .. 0,
        // P streamed in descending steps
        if (len > 1) {
            const double *const arr[1] = {tdma_arr(B, S, XPASS, kTP)};
            const int nst = len - 1, nch = (nst + K - 1) / K;
            long goff[kTdmaNI];
            tdma_lane_offsets<R>(goff, k.LP, k.l0, lane, -1);
            auto issue = [&](int c) {
                tdma_ldouble *slot = lring + (c % kTdmaSlots) * kTdmaChunk;
                if (len - 2 - c * K - (K - 1) >= 0)
                    tdma_issue_at<1>(slot, arr, (long)(len - 2 - c * K) * k.LP, goff);
                else
                    tdma_issue<1, R>(slot, arr, k.LP, k.l0, lane,
                                     [&](int s) { return max(len - 2 - c * K - s, 0); });
            };
            for (int c = 0; c < kTdmaSlots - 1; ++c) issue(c);
            double xv = tile[(len - 1) * RS + rl];
            for (int c = 0; c < nch; ++c) {
                issue(c + kTdmaSlots - 1);
                tdma_wait_vmem<(kTdmaSlots - 1) * kTdmaNI>();
                const double *sP = ring + (c % kTdmaSlots) * kTdmaChunk;
                const int ns = min(K, nst - c * K);
                int t = len - 2 - c * K;
                if (ns == K) {  // G steps at a time, group g+1's reads beside group g
                    constexpr int G = 4, NG = K / G;
                    double pv[2][G], qv[2][G];
                    auto fetch = [&](int b, int s0) {
#pragma unroll
                        for (int u = 0; u < G; ++u) {
                            pv[b][u] = sP[(s0 + u) * R + rl];
                            qv[b][u] = tile[(t - s0 - u) * RS + rl];
                        }
                    };
                    fetch(0, 0);
#pragma unroll
                    for (int g = 0; g < NG; ++g) {
                        if (g + 1 < NG) fetch((g + 1) & 1, (g + 1) * G);
                        const int b = g & 1;
#pragma unroll
                        for (int u = 0; u < G; ++u) {
                            xv = XPASS ? pv[b][u] * xv + qv[b][u] : xv * pv[b][u] + qv[b][u];
                            if (tdma_writes<R>(lane)) tile[(t - g * G - u) * RS + wl] = xv;
                        }
                    }
                } else {
                    for (int s = 0; s < ns; ++s) {
                        const double pv = sP[s * R + rl], qv = tile[(t - s) * RS + rl];
                        xv = XPASS ? pv * xv + qv : xv * pv + qv;
                        if (tdma_writes<R>(lane)) tile[(t - s) * RS + wl] = xv;
                    }
                }
            }
            tdma_wait_vmem<0>();
        }
    }
    __syncthreads();
    const int LS = XPASS ? B.ny : B.nx;
    const int lanes = min(R, LS - k.l0);  // lines of the grid (the rest is padding)
    if (lanes <= 0) return;
    if (XPASS) {  // rows l0.. into the x-pass result, row by row
        double *out = tdma_tmp(B, S);
        for (int e = threadIdx.x; e < lanes * len; e += 256) {
            const int lane = e / len, t = e - lane * len;
            out[(long)(k.l0 + lane) * B.nx + t] = tile[t * RS + lane];
        }
    } else {  // columns l0.. into x
        for (int e = threadIdx.x; e < lanes * len; e += 256) {
            const int t = e / lanes, lane = e - t * lanes;
            S.x[(long)t * B.nx + k.l0 + lane] = tile[t * RS + lane];
        }
    }
}

// One workgroup per (system, line) of a pass: the reduction's factors from
// A, in LDS, double-buffered per step.
template <bool XPASS>
__global__ __launch_bounds__(kPcrMaxThreads) void k_tdma_pcr_setup(TdmaBatch B)
{
    __shared__ double buf[2][3][kPcrMaxLen];
    const int nx = B.nx, ny = B.ny, n = XPASS ? nx : ny, lines = XPASS ? ny : nx;
    const long N = (long)nx * ny;
    const int q = (int)blockIdx.x / lines, line = (int)blockIdx.x - q * lines;
    if (q >= B.count) return;
    const TdmaSys S = B.s[q];
    for (int t = threadIdx.x; t < n; t += (int)blockDim.x) {
        const long p = XPASS ? (long)line * nx + t : (long)t * nx + line;
        buf[0][0][t] = t > 0 ? S.a[(XPASS ? kW : kS) * N + p] : 0.0;
        buf[0][1][t] = S.a[kP * N + p];
        buf[0][2][t] = t < n - 1 ? S.a[(XPASS ? kE : kN) * N + p] : 0.0;
    }
    __syncthreads();
    double *K = pcr_k(B, S, XPASS, line);
    const int ns = pcr_steps(n);
    int cur = 0;
    for (int k = 0, s = 1; k < ns; ++k, s <<= 1, cur ^= 1) {
        const double *lo = buf[cur][0], *d = buf[cur][1], *up = buf[cur][2];
        for (int t = threadIdx.x; t < n; t += (int)blockDim.x) {
            const bool hl = t - s >= 0, hu = t + s < n;
            const double k1 = hl ? lo[t] / d[t - s] : 0.0;
            const double k2 = hu ? up[t] / d[t + s] : 0.0;
            double dn = d[t];
            if (hl) dn = fma(-k1, up[t - s], dn);
            if (hu) dn = fma(-k2, lo[t + s], dn);
            buf[cur ^ 1][0][t] = hl ? -k1 * lo[t - s] : 0.0;
            buf[cur ^ 1][1][t] = dn;
            buf[cur ^ 1][2][t] = hu ? -k2 * up[t + s] : 0.0;
            K[(2L * k) * n + t] = k1;
            K[(2L * k + 1) * n + t] = k2;
        }
        __syncthreads();
    }
    double *R = pcr_r(B, S, XPASS) + (long)line * n;
    for (int t = threadIdx.x; t < n; t += (int)blockDim.x) R[t] = 1.0 / buf[cur][1][t];
}

// One workgroup per (system, line) of a pass, each iteration: every step's
// factors into LDS while the line's D is formed (the x-pass from the
// iteration's x, along a row; the y-pass from the x-pass result, down a
// column), the steps applied, x = D' / d'.  x-pass into the scratch's x-pass
// result, y-pass into x.
template <bool XPASS>
__global__ __launch_bounds__(kPcrMaxThreads) void k_tdma_pcr_line(TdmaBatch B)
{
    extern __shared__ double lds[];
    const int nx = B.nx, ny = B.ny, n = XPASS ? nx : ny, lines = XPASS ? ny : nx;
    const int q = (int)blockIdx.x / lines, line = (int)blockIdx.x - q * lines;
    if (q >= B.count) return;
    const TdmaSys S = B.s[q];
    const int ns = pcr_steps(n);
    double *rhs[2] = {lds, lds + n}, *kl = lds + 2 * n;
    const double *K = pcr_k(B, S, XPASS, line);
    // the factors (ns * n 16-byte vectors, 16-byte aligned: 2 ns n doubles a
    // line, even) straight into LDS, every load in flight at once: each wave
    // brings 64 vectors per round (lane q the wave's q-th), a wave wholly
    // past the end none; lanes past the end re-read the last vector into the
    // slack of 64 vectors after it
    const int tid = threadIdx.x, nv = ns * n;
    const int w0 = tid & ~63;  // the wave's first thread
    for (int v0 = w0; v0 < nv; v0 += (int)blockDim.x)
        __builtin_amdgcn_global_load_lds((tdma_gvoid *)(K + 2L * min(v0 + (tid & 63), nv - 1)),
                                         (tdma_lvoid *)(kl + 2L * v0), 16, 0, 0);
    const long N = (long)nx * ny;
    for (int t = tid; t < n; t += (int)blockDim.x)
        rhs[0][t] = XPASS ? d_xpass(S.a, S.b, S.x, N, nx, ny, t, line)
                          : d_ypass(S.a, S.b, tdma_tmp(B, S), N, nx, ny, line, t);
    tdma_wait_vmem<0>();
    __syncthreads();
    int cur = 0;
    for (int k = 0, s = 1; k < ns; ++k, s <<= 1, cur ^= 1) {
        const double *r = rhs[cur], *k1 = kl + 2L * k * n, *k2 = k1 + n;
        for (int t = threadIdx.x; t < n; t += (int)blockDim.x) {
            double v = r[t];
            if (t - s >= 0) v = fma(-k1[t], r[t - s], v);
            if (t + s < n) v = fma(-k2[t], r[t + s], v);
            rhs[cur ^ 1][t] = v;
        }
        __syncthreads();
    }
    const double *R = pcr_r(B, S, XPASS) + (long)line * n;
    for (int t = threadIdx.x; t < n; t += (int)blockDim.x) {
        const double x = rhs[cur][t] * R[t];
        if (XPASS)
            tdma_tmp(B, S)[(long)line * nx + t] = x;
        else
            S.x[(long)t * nx + line] = x;
    }
}

static unsigned tdma_setup_blocks(const TdmaBatch &B)
{
    const long per = tdma_pass_size(B.nx, B.ny, true) + tdma_pass_size(B.nx, B.ny, false);
    return (unsigned)(((long)B.count * per + 255) / 256);
}
static unsigned tdma_point_blocks(const TdmaBatch &B)
{
    return (unsigned)(((long)B.count * B.nx * B.ny + 255) / 256);
}
static unsigned tdma_chain_blocks(const TdmaBatch &B, bool x, int R)
{
    return (unsigned)(B.count * (tdma_lp(x ? B.ny : B.nx) / R));
}

// Launch KERNEL<R...> for the runtime lane count r (a power of two 2..64).
#define TDMA_R_SWITCH(r, CALL) \
    switch (r) {               \
    case 64: CALL(64); break;  \
    case 32: CALL(32); break;  \
    case 16: CALL(16); break;  \
    case 8: CALL(8); break;    \
    case 4: CALL(4); break;    \
    default: CALL(2); break;   \
    }

template <int RX>
static void tdma_launch_pre(const TdmaBatch &B, int ry, hipStream_t s)
{
    const unsigned nbx = tdma_chain_blocks(B, true, RX);
#define TDMA_PRE(RY)                                                                            \
    hipLaunchKernelGGL((k_tdma_pre<RX, RY>), dim3(nbx + tdma_chain_blocks(B, false, RY)),       \
                       dim3(256), 0, s, B, (int)nbx)
    TDMA_R_SWITCH(ry, TDMA_PRE)
#undef TDMA_PRE
}

// tdma2d9 on up to kTdmaMaxSys systems of one grid size by reduction: the
// factors of both passes once, then nitm iterations of [x lines, y lines].
static void tdma_enqueue_pcr(const TdmaBatch &B, int nitm, hipStream_t s)
{
    const dim3 Wx(pcr_threads(B.nx)), Wy(pcr_threads(B.ny));
    const unsigned bx = (unsigned)(B.count * B.ny), by = (unsigned)(B.count * B.nx);
    hipLaunchKernelGGL(k_tdma_pcr_setup<true>, dim3(bx), Wx, 0, s, B);
    hipLaunchKernelGGL(k_tdma_pcr_setup<false>, dim3(by), Wy, 0, s, B);
    // D double-buffered, then the factors with a wave's slack of 128 doubles
    const size_t lx = sizeof(double) * ((size_t)(2 + 2 * pcr_steps(B.nx)) * B.nx + 128),
                 ly = sizeof(double) * ((size_t)(2 + 2 * pcr_steps(B.ny)) * B.ny + 128);
    for (int it = 0; it < nitm; ++it) {
        hipLaunchKernelGGL(k_tdma_pcr_line<true>, dim3(bx), Wx, lx, s, B);
        hipLaunchKernelGGL(k_tdma_pcr_line<false>, dim3(by), Wy, ly, s, B);
    }
}

static void tdma_enqueue(const TdmaBatch &B, int nitm, hipStream_t s)
{
    if (nitm < 1) return;
    const dim3 T(256);
    hipLaunchKernelGGL(k_tdma_setup, dim3(tdma_setup_blocks(B)), T, 0, s, B);
    const int prx = tdma_lanes(B.nx, kTdmaPreTiles, kTdmaPreArrays),
              pry = tdma_lanes(B.ny, kTdmaPreTiles, kTdmaPreArrays);
#define TDMA_PRE_X(RX) tdma_launch_pre<RX>(B, pry, s)
    TDMA_R_SWITCH(prx, TDMA_PRE_X)
#undef TDMA_PRE_X
    const int rx = tdma_lanes(B.nx, kTdmaLineTiles, kTdmaLineArrays),
              ry = tdma_lanes(B.ny, kTdmaLineTiles, kTdmaLineArrays);
    for (int it = 0; it < nitm; ++it) {
        hipLaunchKernelGGL(k_tdma_d<true>, dim3(tdma_point_blocks(B)), T, 0, s, B);
#define TDMA_LX(R) \
    hipLaunchKernelGGL((k_tdma_line<true, R>), dim3(tdma_chain_blocks(B, true, R)), T, 0, s, B)
        TDMA_R_SWITCH(rx, TDMA_LX)
#undef TDMA_LX
        hipLaunchKernelGGL(k_tdma_d<false>, dim3(tdma_point_blocks(B)), T, 0, s, B);
#define TDMA_LY(R) \
    hipLaunchKernelGGL((k_tdma_line<false, R>), dim3(tdma_chain_blocks(B, false, R)), T, 0, s, B)
        TDMA_R_SWITCH(ry, TDMA_LY)
#undef TDMA_LY
    }
}

// Scratch of the device entry points, kept per host thread and grown on
// demand (one allocation serves every later call of that size or smaller).
struct TdmaScratchCache {
    double *p = nullptr;
    size_t n = 0;
    ~TdmaScratchCache()
    {
        if (p) (void)hipFree(p);
    }
    double *get(size_t need)
    {
        if (need > n) {
            if (p) (void)hipFree(p);
            p = nullptr;
            n = 0;
            if (hipMalloc((void **)&p, sizeof(double) * need) != hipSuccess) return nullptr;
            n = need;
        }
        return p;
    }
};
static thread_local TdmaScratchCache g_tdma_scratch;
// 0: the reference's recurrence (bit for bit); 1: parallel cyclic reduction
static thread_local int g_tdma_mode = 0;

// Grid dimensions the chain kernels take: 3 .. kTdmaMaxLen (each pass's
// line fits the LDS tiles beside the operand ring).
bool tdma_size_ok(int nx, int ny)
{
    return nx >= 3 && ny >= 3 && nx <= kTdmaMaxLen && ny <= kTdmaMaxLen &&
           tdma_lanes(nx, kTdmaPreTiles, kTdmaPreArrays) >= 2 &&
           tdma_lanes(ny, kTdmaPreTiles, kTdmaPreArrays) >= 2 &&
           tdma_lanes(nx, kTdmaLineTiles, kTdmaLineArrays) >= 2 &&
           tdma_lanes(ny, kTdmaLineTiles, kTdmaLineArrays) >= 2;
}

// tdma_device_many without the final wait: the solutions are complete
// when `s` reaches them (the caller's later work on `s` sees them).  The
// scratch is this host thread's, reused by its next call: a next call on
// another stream must wait for `s` first (the time step, its one caller,
// ends with a wait on `s`).
int tdma_enqueue_many(int count, int nx, int ny, int nitm, const double *const *a,
                      const double *const *b, double *const *x, hipStream_t s)
{
    if (count < 1 || !tdma_size_ok(nx, ny) || !a || !b || !x) return M2D_ERR_ARG;
    for (int q = 0; q < count; ++q)
        if (!a[q] || !b[q] || !x[q]) return M2D_ERR_ARG;
    if (nitm < 1) return M2D_OK;
    const bool pcr = g_tdma_mode == 1 && nx <= kPcrMaxLen && ny <= kPcrMaxLen;
    const long W = tdma_scratch(nx, ny) + (pcr ? pcr_scratch(nx, ny) : 0);
    const int per = std::min(count, kTdmaMaxSys);
    double *w = g_tdma_scratch.get((size_t)W * per);
    if (!w) return M2D_ERR_HIP;
    for (int q0 = 0; q0 < count; q0 += kTdmaMaxSys) {
        TdmaBatch B;
        memset(&B, 0, sizeof(B));
        B.count = std::min(kTdmaMaxSys, count - q0);
        B.nx = nx;
        B.ny = ny;
        for (int q = 0; q < B.count; ++q)
            B.s[q] = {a[q0 + q], b[q0 + q], x[q0 + q], w + W * q};
        if (pcr)
            tdma_enqueue_pcr(B, nitm, s);
        else
            tdma_enqueue(B, nitm, s);
        // the next group reuses the scratch: wait for this one
        if (q0 + kTdmaMaxSys < count && hipStreamSynchronize(s) != hipSuccess) return M2D_ERR_HIP;
    }
    return hipGetLastError() == hipSuccess ? M2D_OK : M2D_ERR_HIP;
}

int tdma_device_many(int count, int nx, int ny, int nitm, const double *const *a,
                     const double *const *b, double *const *x, hipStream_t s)
{
    const int rc = tdma_enqueue_many(count, nx, ny, nitm, a, b, x, s);
    if (rc) return rc;
    return hipStreamSynchronize(s) == hipSuccess ? M2D_OK : M2D_ERR_HIP;
}

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
// The two sums of the L1 norm enqueued on `s` into d_res[0..1] (partials in
// d_work, norm_l1_work_size doubles).
int norm_l1_enqueue(const double *a, const double *b, const double *var, int nx, int ny,
                    double *d_work, double *d_res, hipStream_t s)
{
    const int gx = (nx - 2 + 255) / 256, gy = ny - 2;
    const int np = gx * gy;
    double *pa = d_work, *pb = d_work + np;
    hipLaunchKernelGGL(k_norm_l1_partial, dim3(gx, gy), dim3(256), 0, s, a, b, var, nx, ny, pa,
                       pb);
    hipLaunchKernelGGL(k_sum2, dim3(1), dim3(256), 0, s, pa, pb, np, d_res);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        fprintf(stderr, "mach2d_hip: norm kernels failed to launch: %s\n", hipGetErrorString(e));
        return M2D_ERR_HIP;
    }
    return M2D_OK;
}

int norm_l1_device(const double *a, const double *b, const double *var, int nx, int ny,
                   double *d_work, double *h_out, hipStream_t s)
{
    const int np = ((nx - 2 + 255) / 256) * (ny - 2);
    double *res = d_work + 2 * np;
    const int rc = norm_l1_enqueue(a, b, var, nx, ny, d_work, res, s);
    if (rc) return rc;
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

extern "C" int m2d_set_tdma_mode(int mode)
{
    if (mode != 0 && mode != 1) return M2D_ERR_ARG;
    g_tdma_mode = mode;
    return M2D_OK;
}

extern "C" int m2d_tdma2d9_device(int nx, int ny, int nitm, const double *d_a, const double *d_b,
                                  double *d_x, void *stream)
{
    return tdma_device_many(1, nx, ny, nitm, &d_a, &d_b, &d_x, (hipStream_t)stream);
}

extern "C" int m2d_tdma2d9_device_many(int count, int nx, int ny, int nitm,
                                       const double *const *d_a, const double *const *d_b,
                                       double *const *d_x, void *stream)
{
    return tdma_device_many(count, nx, ny, nitm, d_a, d_b, d_x, (hipStream_t)stream);
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
