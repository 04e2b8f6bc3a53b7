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
    double r0;       // ||r0||
    double gammarp;  // gamma * ||r0||
    double rn;       // latest ||r||
    double pad;
};

// Progress flags of a wavefront launch: one 64-bit word per band,
// (generation << 32) | columns published.  Zeroed at the start of a solve.
struct SweepSync {
    unsigned long long *flags;
    SolveState *st;
    unsigned gen;
    int publish_every;
};

// ---- launchers (msi9.hip) --------------------------------------------------
template <typename T>
void launch_factor(const T *a, T *dl, T *du, int nx, int ny, T alpha0, SweepSync sync,
                   int xcd_stride, hipStream_t s);
template <typename T>
void launch_forward(const T *r, const T *dl, T *w, int nx, int ny, SweepSync sync,
                    int xcd_stride, hipStream_t s);
template <typename T, typename TX>
void launch_backward(T *wz, const T *du, TX *x, int nx, int ny, SweepSync sync,
                     int xcd_stride, hipStream_t s);
template <typename TA, typename TX, typename TR>
void launch_residual(const TA *a, const TX *x, const TA *rhs, TR *r, int nx, int ny,
                     double *partial, const SolveState *st, hipStream_t s);
int residual_blocks(int nx, int ny);
void launch_norm_finalize(const double *partial, int np, SolveState *st, double *hist,
                          int hist_cap, int k, double gamma, int nitm, hipStream_t s);
void launch_check_outside(const double *a, int nx, int ny, SolveState *st, hipStream_t s);
void launch_zero_flags(unsigned long long *flags, int n, SolveState *st, int reset_state,
                       hipStream_t s);
template <typename TI, typename TO>
void launch_convert(const TI *in, TO *out, long n, hipStream_t s);

}  // namespace m2d
