// wavefront.h -- device helpers of the pipelined wavefront kernels.
//
// Inter-band exchange without fences.  A band's boundary row travels to the
// neighbouring band through a small buffer of 64-bit words, each carrying a
// 32-bit launch generation in its high half and 32 bits of payload in its
// low half (an FP64 value takes two words).  Aligned 64-bit accesses are
// single-copy atomic, so a consumer that sees the current generation in
// every word of a value has the value; stale words carry an older
// generation.  No release/acquire pair (and so no wait for all outstanding
// memory operations of the wave) is needed: the producer issues relaxed
// agent-scope stores, the consumer polls with relaxed agent-scope loads.
// Generations restart at 1 with every solve, after the buffer is cleared.
#pragma once

#include <hip/hip_runtime.h>

#include "m2d_internal.h"

namespace m2d {

// wave_shr:1 -- lane l receives lane l-1's value (lane 0 receives 0).
__device__ __forceinline__ int dpp_shr1(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false);
}
// wave_shl:1 -- lane l receives lane l+1's value (lane 63 receives 0).
__device__ __forceinline__ int dpp_shl1(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false);
}

__device__ __forceinline__ double lane_from_below(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = dpp_shr1((int)(b & 0xffffffff)), hi = dpp_shr1((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float lane_from_below(float v)
{
    return __int_as_float(dpp_shr1(__float_as_int(v)));
}
__device__ __forceinline__ double lane_from_above(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = dpp_shl1((int)(b & 0xffffffff)), hi = dpp_shl1((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float lane_from_above(float v)
{
    return __int_as_float(dpp_shl1(__float_as_int(v)));
}

// Same shifts, with the lane that has no source (lane 0 for the shift from
// below, lane 63 from above) receiving `edge` instead of 0: the DPP move
// keeps the destination's old value there (bound_ctrl off).
__device__ __forceinline__ int dpp_shr1_or(int v, int edge)
{
    return __builtin_amdgcn_update_dpp(edge, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ int dpp_shl1_or(int v, int edge)
{
    return __builtin_amdgcn_update_dpp(edge, v, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ double lane_from_below_or(double v, double edge)
{
    const long long b = __double_as_longlong(v), e = __double_as_longlong(edge);
    const int lo = dpp_shr1_or((int)(b & 0xffffffff), (int)(e & 0xffffffff));
    const int hi = dpp_shr1_or((int)(b >> 32), (int)(e >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float lane_from_below_or(float v, float edge)
{
    return __int_as_float(dpp_shr1_or(__float_as_int(v), __float_as_int(edge)));
}
__device__ __forceinline__ double lane_from_above_or(double v, double edge)
{
    const long long b = __double_as_longlong(v), e = __double_as_longlong(edge);
    const int lo = dpp_shl1_or((int)(b & 0xffffffff), (int)(e & 0xffffffff));
    const int hi = dpp_shl1_or((int)(b >> 32), (int)(e >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float lane_from_above_or(float v, float edge)
{
    return __int_as_float(dpp_shl1_or(__float_as_int(v), __float_as_int(edge)));
}

// row_shl:N -- lane l receives lane l+N's value within its row of 16 lanes
// (lanes whose source lies past the row end receive 0); for lane 0 this is
// lane N's value, N <= 15.
template <int N>
__device__ __forceinline__ double lane_plus(double v)
{
    static_assert(N >= 1 && N <= 15, "row_shl takes 1..15");
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), 0x100 + N, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x100 + N, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int N>
__device__ __forceinline__ float lane_plus(float v)
{
    static_assert(N >= 1 && N <= 15, "row_shl takes 1..15");
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x100 + N, 0xf, 0xf, true));
}

// Value of lane `src` (wave-uniform constant) broadcast to every lane.
__device__ __forceinline__ double read_lane(double v, int src)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float read_lane(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

typedef unsigned long long u64;

__device__ __forceinline__ void tag_store(u64 *p, unsigned gen, unsigned payload)
{
    __hip_atomic_store(p, ((u64)gen << 32) | payload, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 tag_load(const u64 *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Words per value in the boundary buffer.
template <typename T>
struct TagWords;
template <>
struct TagWords<double> {
    static constexpr int n = 2;
};
template <>
struct TagWords<float> {
    static constexpr int n = 1;
};

__device__ __forceinline__ void put_tagged(u64 *p, unsigned gen, double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    tag_store(p, gen, (unsigned)(b & 0xffffffffu));
    tag_store(p + 1, gen, (unsigned)(b >> 32));
}
__device__ __forceinline__ void put_tagged(u64 *p, unsigned gen, float v)
{
    tag_store(p, gen, (unsigned)__float_as_int(v));
}

// Raw words of one tagged value (loaded ahead of use).
template <typename T>
struct Raw {
    u64 w[TagWords<T>::n];
};

template <typename T>
__device__ __forceinline__ void raw_load(Raw<T> &r, const u64 *p)
{
#pragma unroll
    for (int k = 0; k < TagWords<T>::n; ++k) r.w[k] = tag_load(p + k);
}
template <typename T>
__device__ __forceinline__ bool raw_ok(const Raw<T> &r, unsigned gen)
{
    bool ok = true;
#pragma unroll
    for (int k = 0; k < TagWords<T>::n; ++k) ok = ok && (unsigned)(r.w[k] >> 32) == gen;
    return ok;
}
__device__ __forceinline__ double raw_value(const Raw<double> &r)
{
    return __longlong_as_double((long long)((r.w[1] << 32) | (r.w[0] & 0xffffffffull)));
}
__device__ __forceinline__ float raw_value(const Raw<float> &r)
{
    return __int_as_float((int)(r.w[0] & 0xffffffffull));
}

// Exchange descriptor of one wavefront launch.
struct Link {
    u64 *bnd;          // boundary words: [slot][column][words per column]
    u64 *scratch;      // 8 private words per lane (stores that publish nothing)
    SolveState *st;
    int phase;         // 0 factorization, 1 forward sweep, 2 backward sweep
    unsigned gen;      // this launch's generation, set by the kernel itself
};

// Generation of a launch, from the device iteration counter, so that a
// captured launch sequence can be replayed: factorization 1, forward sweep
// of iteration iter+1: 2*iter+3, backward sweep: 2*iter+4.
__device__ __forceinline__ unsigned launch_gen(const Link &lk)
{
    if (lk.phase == 0) return 1u;
    const unsigned it = (unsigned)__builtin_amdgcn_readfirstlane(lk.st->iter);
    return 2u * it + 2u + (unsigned)lk.phase;
}

}  // namespace m2d
