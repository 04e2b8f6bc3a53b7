// seqsum.h -- in-order sum of n terms, for results that must equal a
// sequential loop bit for bit (the reference's left-to-right sums).
//
// The chain of dependent FP64 additions is the whole cost (about 5 cycles
// each on gfx950), so one wave does it and keeps it fed.  The terms reach
// LDS in chunks of kSeqChunk by direct-to-LDS loads (16 bytes per lane, no
// registers) issued kSeqSlots - 1 chunks ahead, which covers the latency of
// memory another XCD wrote; every lane then adds the terms in order from
// LDS, with lane-uniform 16-byte reads (a broadcast) issued a few groups of
// terms ahead of the additions that use them.  Call from a workgroup of
// kSeqThreads (one wave); every lane returns the sum.
#pragma once

#include <hip/hip_runtime.h>

namespace m2d {

constexpr int kSeqThreads = 64;
constexpr int kSeqChunk = 512, kSeqSlots = 4, kSeqLoads = kSeqChunk / 128;

typedef __attribute__((address_space(1))) const void seq_gvoid;
typedef __attribute__((address_space(3))) void seq_lvoid;

template <int N>
__device__ __forceinline__ void seq_wait_vmem()
{
    static_assert(N >= 0 && N <= 63, "vmcnt is six bits");
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// m + s[a] + s[a+1] + ... + s[b-1], left to right (s in LDS, a, b
// wave-uniform): groups of 8 read as 16-byte pairs, three groups ahead
__device__ __forceinline__ double seq_add_lds(const double *s, int a, int b, double m)
{
    constexpr int G = 8;
    for (; a < b && (a & 1); ++a) m = s[a] + m;  // to an even start (16-byte reads)
    if (b - a >= 4 * G) {
        double A[G], B[G], C[G];
        auto rd = [&](double(&X)[G], int q) {
#pragma unroll
            for (int i = 0; i < G; i += 2) {
                const double2 v = *(const double2 *)(s + q + i);
                X[i] = v.x;
                X[i + 1] = v.y;
            }
        };
        auto ad = [&](const double(&X)[G]) {
#pragma unroll
            for (int i = 0; i < G; ++i) m = X[i] + m;
        };
        rd(A, a);
        rd(B, a + G);
        rd(C, a + 2 * G);
        for (; a + 6 * G <= b; a += 3 * G) {
            ad(A);
            __builtin_amdgcn_sched_barrier(0);
            rd(A, a + 3 * G);
            __builtin_amdgcn_sched_barrier(0);
            ad(B);
            __builtin_amdgcn_sched_barrier(0);
            rd(B, a + 4 * G);
            __builtin_amdgcn_sched_barrier(0);
            ad(C);
            __builtin_amdgcn_sched_barrier(0);
            rd(C, a + 5 * G);
            __builtin_amdgcn_sched_barrier(0);
        }
        ad(A);
        ad(B);
        ad(C);
        a += 3 * G;
    }
    for (; a < b; ++a) m = s[a] + m;
    return m;
}

// p[0] + p[1] + ... + p[n-1], added left to right; lds: kSeqSlots *
// kSeqChunk doubles of LDS.  The loads start at p rounded down to 16
// bytes (the memory before a misaligned p belongs to the same array: every
// caller's arrays are 256-byte aligned) and never read past the last
// 16-byte pair that holds p[n-1].
__device__ __forceinline__ double seq_sum(const double *__restrict__ p, long n, double *lds)
{
    const int lane = threadIdx.x & 63;
    const int sh = (int)(((unsigned long)p >> 3) & 1);  // p[0] at position sh of the stream
    const double *q = p - sh;
    const long nt = n + sh;                            // stream length
    const long last = (nt - 1) & ~1L;                  // first term of the last pair
    const long nch = (nt + kSeqChunk - 1) / kSeqChunk;
    auto issue = [&](long c) {
        double *slot = lds + (c % kSeqSlots) * kSeqChunk;
#pragma unroll
        for (int ii = 0; ii < kSeqLoads; ++ii) {
            long e = c * kSeqChunk + 128 * ii + 2 * lane;
            e = e < last ? e : last;
            __builtin_amdgcn_global_load_lds((seq_gvoid *)(q + e), (seq_lvoid *)(slot + 128 * ii),
                                             16, 0, 0);
        }
    };
    for (long c = 0; c < kSeqSlots - 1; ++c) issue(c);
    double m = 0.0;
    for (long c = 0; c < nch; ++c) {
        issue(c + kSeqSlots - 1);
        seq_wait_vmem<(kSeqSlots - 1) * kSeqLoads>();
        const double *slot = lds + (c % kSeqSlots) * kSeqChunk;
        const long c0 = c * kSeqChunk;
        const int a = c == 0 ? sh : 0;
        const int b = (int)(nt - c0 < kSeqChunk ? nt - c0 : kSeqChunk);
        m = seq_add_lds(slot, a, b, m);
    }
    seq_wait_vmem<0>();
    return m;
}

}  // namespace m2d
