// m2d_internal.h -- device-side state, synchronisation helpers and kernel
// launch declarations shared by the gfx950 translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mach2d_hip.h"

namespace m2d {

constexpr int kWave = 64;            // gfx950 wavefront width
constexpr int kMaxSpin = 1 << 22;    // bound of every inter-workgroup wait
constexpr int kIterCap = 1000000;    // nitm <= 0: safety cap on iterations

// Per-solve device state (one per system; lives in device memory).
struct SolveState {
    int done;        // 1 once the stopping rule fired (kernels then no-op)
    int nit;         // iterations run
    int status;      // M2D_* code
    int nitm_hit;    // loop ended on nit == nitm
    int abort;       // an inter-workgroup wait timed out
    int err_pivot;   // factorization met a zero / non-finite pivot
    int outside;     // number of non-zero couplings to outside neighbours
    int outer;       // refinement steps (mixed precision)
    int iter;        // iterations completed (drives the launch generations)
    int skip;        // set when a loop enqueued ahead of its exit test has
                     // ended: the gated reset keeps the state, the
                     // factorization and every kernel of the loop no-op
    double r0;       // ||r0||
    double gammarp;  // gamma * ||r0||
    double rn;       // latest ||r||
    double pad;
};

struct Link;
struct Skew;

// ---- wavefront kernels and layout (sweep.hip) --------------------------------
long boundary_words(int nx, int ny);
long scratch_words(int ny);
Skew skew_for(int nx, int ny, int value_bytes);
void launch_clear_boundary(unsigned long long *bnd, int nx, int ny, hipStream_t s);
template <typename T>
// five: lu2d5 on the embedding (A's S, W, P, E, N arrays; dl1..dl4, du1..du3)
void launch_factor(const T *a_sk, T *dl, T *du, T alpha0, const Skew &sk, Link lk,
                   int xcd_stride, hipStream_t s, bool five = false);
// the LDS-staged sweeps, one launch each; five: the 5-diagonal sweeps
// (msi2d5) on the same arrays (dl5 and du4 unused)
template <typename T>
void launch_forward_staged(const T *r, const T *dl, T *w, const Skew &sk, Link lk,
                           int xcd_stride, hipStream_t s, bool five = false);
template <typename T>
void launch_backward_staged(T *wz, const T *du, const Skew &sk, Link lk, int xcd_stride,
                            hipStream_t s, bool five = false);
// both sweeps of one iteration in one launch (w from r, then z over w)
template <typename T>
void launch_sweeps(const T *r, const T *dl, T *w, const T *du, const Skew &sk, Link lk,
                   int xcd_stride, hipStream_t s, bool five = false);
template <typename TI, typename TO>
// unit: array whose padding entries get one (the diagonal of A), or -1
void launch_to_skew(const TI *src, TO *dst, int nk, const Skew &sk, hipStream_t s,
                    int unit = -1);
// 5 diagonals (S, W, P, E, N), reference layout -> the 9 skewed arrays of
// their zero-corner embedding (SW, SE, NW, NE zero)
template <typename TI, typename TO>
void launch_to_skew5(const TI *src, TO *dst, const Skew &sk, hipStream_t s);
// dl: 6 arrays (dl1..dl5, then 1/dl1, written here); padding made finite.
template <typename T>
// st: a zero or non-finite pivot sets st->err_pivot; no-op while st->skip
// is set (nullptr: neither)
void launch_prepare_factor(T *dl, T *du, const Skew &sk, hipStream_t s,
                           SolveState *st = nullptr);
template <typename T>
void launch_from_skew(const T *src, T *dst, int nk, const Skew &sk, hipStream_t s);

// ---- batched wavefront and layout kernels (sweep.hip) -------------------------
template <typename T>
struct BatchSys;
template <typename T>
void launch_factor_batch(const BatchSys<T> *sys, const int2 *map, int nblk, hipStream_t s);
template <typename T>
void launch_forward_batch(const BatchSys<T> *sys, const int2 *map, int nblk, hipStream_t s,
                          bool staged, bool five = false);
template <typename T>
void launch_backward_batch(const BatchSys<T> *sys, const int2 *map, int nblk, hipStream_t s,
                           bool staged, bool five = false);
template <typename T>
void launch_to_skew_batch(const BatchSys<T> *sys, const int2 *map, int nblk, hipStream_t s);
template <typename T>
void launch_prepare_batch(const BatchSys<T> *sys, const int2 *map, int nblk, hipStream_t s);

// ---- parallel kernels (msi9.hip) ---------------------------------------------
// nd: 9 diagonals (msi2d9 slots) or 5 (msi2d5 slots S, W, P, E, N)
// The norm and stopping rule that follow a residual: initial = 1 for the
// residual of the incoming var (k = 0), otherwise the residual after
// iteration k = st->iter + 1, which becomes st->iter; ||r|| of iteration k
// goes to hist[k-1] for k <= hist_cap.
struct NormRule {
    double *hist;
    int hist_cap;
    int initial;
    double gamma;
    int nitm;
};
// r = A.x - rhs, its norm and the stopping rule, in one launch; partial
// holds residual_partial_words(nx, ny) words, zero when first used
template <typename TA, typename TX, typename TR>
void launch_residual_norm(const TA *a, const TX *x, const TA *rhs, TR *r, int nx, int ny,
                          unsigned long long *partial, SolveState *st, const Skew *out_skew,
                          const NormRule &rule, hipStream_t s, int nd = 9);
// the update var = z + var (z skewed by sk) and then the residual, its norm
// and the stopping rule, in one launch; the iterate alternates between x
// (even number of iterations done) and x2 (odd), launch_settle returns it
template <typename T>
void launch_update_residual_norm(const T *a, T *x, T *x2, const T *z, const T *rhs, T *r,
                                 const Skew &sk, unsigned long long *partial, SolveState *st,
                                 const NormRule &rule, hipStream_t s, int nd = 9);
template <typename T>
void launch_settle(T *x, const T *x2, long n, const SolveState *st, hipStream_t s);
int residual_blocks(int nx, int ny);
long residual_partial_words(int nx, int ny);
void launch_check_outside(const double *a, int nx, int ny, SolveState *st, hipStream_t s,
                          int nd = 9);
// gated: leave the state as it is when st->skip is set (a plain reset
// clears skip too)
void launch_reset_state(SolveState *st, hipStream_t s, bool gated = false);

// ---- coefficient assembly (assembly.hip) -------------------------------------
// Stream of this thread's device-array assembly calls; returns the previous one.
hipStream_t assembly_stream(hipStream_t s);
hipStream_t assembly_stream_current();
// This thread's device-array assembly calls pass gate and sc to their
// kernels (nullptr, nullptr: off): a kernel returns at once while *gate is
// non-zero, and reads dt = sc[0], beta = sc[1] from device memory.
void assembly_gate(const int *gate, const double *sc);

// ---- in-order statistics of the time step (tstep.hip, intflow.hip) -----------
// The K mean convergence coefficients of device arrays a[k] (nd[k] = 9 or 5
// diagonals) enqueued on s: scratch t of K*(nx-2)*(ny-2) + K doubles, the
// means at its end.
void mean_cc_enqueue(int nx, int ny, int K, const double *const *a, const int *nd, double *t,
                     hipStream_t s);
// the terms of those means only (t + k*n for array k), and K in-order sums of
// consecutive arrays of n terms (t + k*n) into out[k], each on its own
// workgroup of one launch
void cc_terms_enqueue(int nx, int ny, int K, const double *const *a, const int *nd, double *t,
                      hipStream_t s);
void seq_sums_enqueue(const double *t, long n, int K, double *out, hipStream_t s);
// check_mass_conservation's terms (in = rp, re, rn, Jp, ro, roa, ron, roe,
// Uce, Vcn; device arrays) into t[0 .. (nx-2)(ny-2)), in the reference's
// order, on s; their in-order sum is rmass
void mass_terms_enqueue(int nx, int ny, double dt, const double *const *in, double *t,
                        hipStream_t s);
// tstepper_dt_scarborough's step rule on the four means cc (u, v, T, p')
double scarborough_dt(double dt, const double *cc, double h0, double mincc, double maxcc);
// get_average of the device field F enqueued on s: the mean lands in scr[0]
// (scratch of average_scratch(nx, ny) doubles)
long average_scratch(int nx, int ny);
void average_enqueue(int nx, int ny, const double *F, double *scr, hipStream_t s);
// true for device memory (false for host memory and unknown pointers)
bool is_device_array(const void *p);
template <typename TI, typename TO>
void launch_convert(const TI *in, TO *out, long n, hipStream_t s);

}  // namespace m2d
