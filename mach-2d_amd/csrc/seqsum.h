// seqsum.h -- in-order sum of n terms, for results that must equal a
// sequential loop bit for bit (the reference's left-to-right sums).
//
// The chain of dependent FP64 additions is the whole cost (about 10 cycles
// per term on gfx950, reads included), so one wave does it and keeps it fed.  The terms reach
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

// m + s[0] + ... + s[kSeqChunk-1] in order, for a whole chunk: three groups
// of 8 terms in flight in fixed registers (v160..v207), each group read as
// four lane-uniform 16-byte LDS reads and waited for by count, so no wait
// ever drains the reads of the groups behind it (the compiler's own waits
// drain the LDS counter at every loop back edge).  The reads of group g+3
// follow the additions of group g as a block (spread between the dependent
// additions they measured 25 % slower).  Every read is waited for inside the
// block.  Addition is commutative, so adding t to m is m + t.
__device__ __forceinline__ double seq_add_chunk(const double *slot, double m)
{
    static_assert(kSeqChunk == 512, "the block below adds 64 groups of 8 terms");
    unsigned a = (unsigned)(unsigned long)(const seq_lvoid *)slot;  // LDS byte address
    unsigned k;
    asm volatile(
        "s_waitcnt lgkmcnt(0)\n"
        "ds_read_b128 v[160:163], %[a] offset:0\n"
        "ds_read_b128 v[164:167], %[a] offset:16\n"
        "ds_read_b128 v[168:171], %[a] offset:32\n"
        "ds_read_b128 v[172:175], %[a] offset:48\n"
        "ds_read_b128 v[176:179], %[a] offset:64\n"
        "ds_read_b128 v[180:183], %[a] offset:80\n"
        "ds_read_b128 v[184:187], %[a] offset:96\n"
        "ds_read_b128 v[188:191], %[a] offset:112\n"
        "ds_read_b128 v[192:195], %[a] offset:128\n"
        "ds_read_b128 v[196:199], %[a] offset:144\n"
        "ds_read_b128 v[200:203], %[a] offset:160\n"
        "ds_read_b128 v[204:207], %[a] offset:176\n"
        "s_mov_b32 %[k], 20\n"
        "1:\n"
        "s_waitcnt lgkmcnt(8)\n"
        "v_add_f64 %[m], v[160:161], %[m]\n"
        "v_add_f64 %[m], v[162:163], %[m]\n"
        "v_add_f64 %[m], v[164:165], %[m]\n"
        "v_add_f64 %[m], v[166:167], %[m]\n"
        "v_add_f64 %[m], v[168:169], %[m]\n"
        "v_add_f64 %[m], v[170:171], %[m]\n"
        "v_add_f64 %[m], v[172:173], %[m]\n"
        "v_add_f64 %[m], v[174:175], %[m]\n"
        "ds_read_b128 v[160:163], %[a] offset:192\n"
        "ds_read_b128 v[164:167], %[a] offset:208\n"
        "ds_read_b128 v[168:171], %[a] offset:224\n"
        "ds_read_b128 v[172:175], %[a] offset:240\n"
        "s_waitcnt lgkmcnt(8)\n"
        "v_add_f64 %[m], v[176:177], %[m]\n"
        "v_add_f64 %[m], v[178:179], %[m]\n"
        "v_add_f64 %[m], v[180:181], %[m]\n"
        "v_add_f64 %[m], v[182:183], %[m]\n"
        "v_add_f64 %[m], v[184:185], %[m]\n"
        "v_add_f64 %[m], v[186:187], %[m]\n"
        "v_add_f64 %[m], v[188:189], %[m]\n"
        "v_add_f64 %[m], v[190:191], %[m]\n"
        "ds_read_b128 v[176:179], %[a] offset:256\n"
        "ds_read_b128 v[180:183], %[a] offset:272\n"
        "ds_read_b128 v[184:187], %[a] offset:288\n"
        "ds_read_b128 v[188:191], %[a] offset:304\n"
        "s_waitcnt lgkmcnt(8)\n"
        "v_add_f64 %[m], v[192:193], %[m]\n"
        "v_add_f64 %[m], v[194:195], %[m]\n"
        "v_add_f64 %[m], v[196:197], %[m]\n"
        "v_add_f64 %[m], v[198:199], %[m]\n"
        "v_add_f64 %[m], v[200:201], %[m]\n"
        "v_add_f64 %[m], v[202:203], %[m]\n"
        "v_add_f64 %[m], v[204:205], %[m]\n"
        "v_add_f64 %[m], v[206:207], %[m]\n"
        "ds_read_b128 v[192:195], %[a] offset:320\n"
        "ds_read_b128 v[196:199], %[a] offset:336\n"
        "ds_read_b128 v[200:203], %[a] offset:352\n"
        "ds_read_b128 v[204:207], %[a] offset:368\n"
        "v_add_u32 %[a], 192, %[a]\n"
        "s_sub_u32 %[k], %[k], 1\n"
        "s_cmp_lg_u32 %[k], 0\n"
        "s_cbranch_scc1 1b\n"
        "s_waitcnt lgkmcnt(8)\n"
        "v_add_f64 %[m], v[160:161], %[m]\n"
        "v_add_f64 %[m], v[162:163], %[m]\n"
        "v_add_f64 %[m], v[164:165], %[m]\n"
        "v_add_f64 %[m], v[166:167], %[m]\n"
        "v_add_f64 %[m], v[168:169], %[m]\n"
        "v_add_f64 %[m], v[170:171], %[m]\n"
        "v_add_f64 %[m], v[172:173], %[m]\n"
        "v_add_f64 %[m], v[174:175], %[m]\n"
        "ds_read_b128 v[160:163], %[a] offset:192\n"
        "ds_read_b128 v[164:167], %[a] offset:208\n"
        "ds_read_b128 v[168:171], %[a] offset:224\n"
        "ds_read_b128 v[172:175], %[a] offset:240\n"
        "s_waitcnt lgkmcnt(8)\n"
        "v_add_f64 %[m], v[176:177], %[m]\n"
        "v_add_f64 %[m], v[178:179], %[m]\n"
        "v_add_f64 %[m], v[180:181], %[m]\n"
        "v_add_f64 %[m], v[182:183], %[m]\n"
        "v_add_f64 %[m], v[184:185], %[m]\n"
        "v_add_f64 %[m], v[186:187], %[m]\n"
        "v_add_f64 %[m], v[188:189], %[m]\n"
        "v_add_f64 %[m], v[190:191], %[m]\n"
        "s_waitcnt lgkmcnt(4)\n"
        "v_add_f64 %[m], v[192:193], %[m]\n"
        "v_add_f64 %[m], v[194:195], %[m]\n"
        "v_add_f64 %[m], v[196:197], %[m]\n"
        "v_add_f64 %[m], v[198:199], %[m]\n"
        "v_add_f64 %[m], v[200:201], %[m]\n"
        "v_add_f64 %[m], v[202:203], %[m]\n"
        "v_add_f64 %[m], v[204:205], %[m]\n"
        "v_add_f64 %[m], v[206:207], %[m]\n"
        "s_waitcnt lgkmcnt(0)\n"
        "v_add_f64 %[m], v[160:161], %[m]\n"
        "v_add_f64 %[m], v[162:163], %[m]\n"
        "v_add_f64 %[m], v[164:165], %[m]\n"
        "v_add_f64 %[m], v[166:167], %[m]\n"
        "v_add_f64 %[m], v[168:169], %[m]\n"
        "v_add_f64 %[m], v[170:171], %[m]\n"
        "v_add_f64 %[m], v[172:173], %[m]\n"
        "v_add_f64 %[m], v[174:175], %[m]\n"
        : [m] "+v"(m), [a] "+v"(a), [k] "=&s"(k)
        :
        : "v160", "v161", "v162", "v163", "v164", "v165", "v166", "v167", "v168", "v169", "v170", "v171", "v172", "v173", "v174", "v175", "v176", "v177", "v178", "v179", "v180", "v181", "v182", "v183", "v184", "v185", "v186", "v187", "v188", "v189", "v190", "v191", "v192", "v193", "v194", "v195", "v196", "v197", "v198", "v199", "v200", "v201", "v202", "v203", "v204", "v205", "v206", "v207", "scc", "memory");
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
        if (a == 0 && b == kSeqChunk)
            m = seq_add_chunk(slot, m);
        else
            m = seq_add_lds(slot, a, b, m);
    }
    seq_wait_vmem<0>();
    return m;
}

}  // namespace m2d
