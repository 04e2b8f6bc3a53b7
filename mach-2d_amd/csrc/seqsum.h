// seqsum.h -- in-order sum of n terms by one lane, for results that must
// equal a sequential loop bit for bit (the reference's left-to-right sums).
//
// The chain of dependent additions is the whole cost, so lane 0 of wave 0
// only adds, reading the terms from LDS, while
// the other waves of the workgroup stage the next chunk from HBM.  Call from
// every thread of a 256-thread workgroup; returns the sum on thread 0.
#pragma once

#include <hip/hip_runtime.h>

namespace m2d {

constexpr int kSeqThreads = 256, kSeqChunk = 2048;

// term(q), q in [0, n), gives the q-th term of the sum
template <typename Term>
__device__ double seq_sum(Term term, long n)
{
    __shared__ double buf[2][kSeqChunk];
    const int tid = threadIdx.x;
    const long nch = (n + kSeqChunk - 1) / kSeqChunk;
    for (long q = tid; q < kSeqChunk && q < n; q += kSeqThreads) buf[0][q] = term(q);
    __syncthreads();
    double m = 0.0;
    for (long c = 0; c < nch; ++c) {
        const int cur = (int)(c & 1);
        const long base = c * kSeqChunk;
        const int len = (int)min((long)kSeqChunk, n - base);
        if (tid >= 64) {
            const long nb = base + kSeqChunk;
            if (nb < n) {
                const int nlen = (int)min((long)kSeqChunk, n - nb);
                for (int q = tid - 64; q < nlen; q += kSeqThreads - 64)
                    buf[cur ^ 1][q] = term(nb + q);
            }
        } else if (tid == 0) {
            // groups of G terms: the reads of the next group are in flight
            // while the G additions of this one run (G additions outlast
            // the LDS latency)
            constexpr int G = 16;
            const double *b = buf[cur];
            int q = 0;
            if (len >= G) {
                double v[G], w[G];
#pragma unroll
                for (int k = 0; k < G; ++k) v[k] = b[k];
                for (; q + 2 * G <= len; q += G) {
#pragma unroll
                    for (int k = 0; k < G; ++k) w[k] = b[q + G + k];
#pragma unroll
                    for (int k = 0; k < G; ++k) m = v[k] + m;
#pragma unroll
                    for (int k = 0; k < G; ++k) v[k] = w[k];
                }
#pragma unroll
                for (int k = 0; k < G; ++k) m = v[k] + m;
                q += G;
            }
            for (; q < len; ++q) m = b[q] + m;
        }
        __syncthreads();
    }
    return m;
}

}  // namespace m2d
