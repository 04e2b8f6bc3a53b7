// sweep.hip -- pipelined wavefront kernels of the 9-diagonal MSI:
// factorization lu2d9 (src/solvers/msi2d9-v03.f90:277-432) and the forward /
// backward sweeps of fb2d9 (:139-181), plus the layout conversions and the
// update var = z + var (:181).
//
// Schedule.  The factorization and the forward sweep of point (i,j) read the
// W, SW, S and SE neighbours (:357-366, :155), the backward sweep E, NW, N
// and NE (:175), so points with equal k = i + 2j are independent.  A wave of
// 64 lanes owns a band of 64 rows; lane l (row 64*band + l) works on column
// i at stored step u = i + 2l + pad (layout in skew.h).  S/SE/SW (N/NW/NE)
// values come from lane l-1 (l+1) through a one-lane DPP shift of its latest
// result and two registers of history.
//
// Chunks.  A kernel advances F steps per chunk.  Every lane loads its F
// consecutive values of each input array with 16-byte loads one chunk ahead,
// keeps its F results in registers and stores them with 16-byte stores after
// the chunk, so the loop body has a fixed, small number of vector memory
// operations and the only wait per chunk is for loads issued a chunk
// earlier.  The sweeps reload each 16-byte vector of the chunk buffer as
// soon as its steps are done, so one buffer serves both the current and the
// next chunk.
//
// Padding.  Steps whose column (or row) is outside the grid compute on
// padding.  In the sweeps the padding holds r = 0, dl1 = 1, 1/dl1 = 1 (in
// the scaled FP32 rows ny = 0) and zero couplings (k_prepare_factor), so
// padding results are finite and no
// neighbour value needs a mask: every factor entry that couples a point to
// a neighbour outside the grid is an exact zero (the couplings of A to
// outside neighbours are zero, and the factorization keeps them zero), and
// zero times a finite value adds nothing.  The factorization still masks.
//
// Bands.  A band's edge row travels to the neighbouring band through
// generation-tagged words (wavefront.h), F columns per chunk: lanes 0..F-1
// of the producer store the values of the edge lane (every lane stores; the
// others write a private scratch word, which keeps the store count fixed);
// lanes 0..F of the consumer load the F+1 columns its next chunk needs, one
// chunk ahead, and the edge lane takes them with lane reads.  A band only
// waits on a band with a lower launch index, so the waits cannot deadlock
// under in-order workgroup dispatch, and every wait is bounded.
#include <math.h>

#include <type_traits>

#include "batch.h"
#include "m2d_internal.h"
#include "msi_stencil.h"
#include "skew.h"
#include "wavefront.h"

namespace m2d {

template <typename T, int F>
__device__ __forceinline__ void vload(T (&dst)[F], const T *__restrict__ src)
{
    static_assert((F * sizeof(T)) % 16 == 0, "a chunk is a whole number of 16-byte words");
    const uint4 *p = reinterpret_cast<const uint4 *>(src);
#pragma unroll
    for (int q = 0; q < (int)(F * sizeof(T) / 16); ++q) {
        const uint4 v = p[q];
        __builtin_memcpy(&dst[q * (int)(16 / sizeof(T))], &v, 16);
    }
}

template <typename T, int F>
__device__ __forceinline__ void vstore(T *__restrict__ dst, const T (&src)[F])
{
    uint4 *p = reinterpret_cast<uint4 *>(dst);
#pragma unroll
    for (int q = 0; q < (int)(F * sizeof(T) / 16); ++q) {
        uint4 v;
        __builtin_memcpy(&v, &src[q * (int)(16 / sizeof(T))], 16);
        p[q] = v;
    }
}

// Same, at base + ofs with a wave-uniform base and a 32-bit lane offset.
template <typename T, int F>
__device__ __forceinline__ void vstore(T *__restrict__ base, unsigned ofs, const T (&src)[F])
{
#pragma unroll
    for (int q = 0; q < (int)(F * sizeof(T) / 16); ++q) {
        uint4 v;
        __builtin_memcpy(&v, &src[q * (int)(16 / sizeof(T))], 16);
        *reinterpret_cast<uint4 *>(base + (ofs + q * (unsigned)(16 / sizeof(T)))) = v;
    }
}

template <typename T>
__device__ __forceinline__ bool bad_pivot(T d)
{
    return !(d == d) || d == T(0) || isinf(d);
}

// Slow path of a boundary wait: sleep, reload, re-check, bounded.  Returns
// false when it gave up or another band did (abort flag).
template <typename FOK, typename FRE>
__device__ bool slow_wait(const Link &lk, FOK ok, FRE reload)
{
    for (int spin = 0; spin < kMaxSpin; ++spin) {
        if (__hip_atomic_load(&lk.st->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return false;
        __builtin_amdgcn_s_sleep(1);
        reload();
        if (__all(ok())) return true;
    }
    __hip_atomic_store(&lk.st->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return false;
}

// Value of lane `sel` of a per-lane array held by one lane: lane q < F gets
// element q of `v` as held by lane `src`.
template <typename T, int F>
__device__ __forceinline__ T spread_from_lane(const T (&v)[F], int src, int lane)
{
    T out = read_lane(v[0], src);
#pragma unroll
    for (int q = 1; q < F; ++q) {
        const T x = read_lane(v[q], src);
        out = lane == q ? x : out;
    }
    return out;
}

// Half-chunk buffers of the sweeps.  The F steps of a chunk are two halves
// of H = F/2 steps; three buffers hold halves h, h+1 and h+2, and the loads
// of half h+2 are issued at the start of half h into the buffer half h-1
// used, so every load has a chunk of compute to arrive and never overwrites
// a live value.  The loop body covers three chunks, after which the buffer
// roles repeat.  Chunk length equals the skew segment (F == sk.S), so a
// chunk starts at sk.chunk_base.
template <typename T, int NA, int H>
struct Halves {
    T v[NA][H];
};

// Row form of the sweeps.  FP64 sweeps reproduce the reference's rounding
// (its operation order, exact division); FP32 sweeps only serve as the inner
// solver of the FP64 refinement, so their rows are pre-scaled by -1/dl1 and
// their sums fused and reordered to shorten the dependent chain per step.
template <typename T>
struct ScaledRows {
    static constexpr bool value = sizeof(T) == 4;
};

// Operand arrays of the sweeps.  9 diagonals: forward r, dl1..dl5, 1/dl1
// (FP32 scaled rows: r, ny, ne2..ne5), backward w, du1..du4.  5 diagonals
// (msi2d5): the same arrays without dl5 (ne5) and du4, which the zero-corner
// embedding makes exactly zero.  Forward operand k >= 1 is dl array fslot(k).
template <typename T, bool Five>
struct SweepOps {
    static constexpr bool kScaled = ScaledRows<T>::value;
    static constexpr int NF = (kScaled ? 6 : 7) - (Five ? 1 : 0);
    static constexpr int NB = Five ? 4 : 5;
    static constexpr int fslot(int k) { return (!kScaled && Five && k == NF - 1) ? 5 : k - 1; }
};

// One forward step from the operands of element t of a half buffer.
template <bool Five, typename T, int NA, int H>
__device__ __forceinline__ T forward_step(const Halves<T, NA, H> &cb, int t, T cur, T se, T sv,
                                          T sw)
{
    if constexpr (ScaledRows<T>::value) {
        if constexpr (Five)
            return forward_point5_scaled<T>(cb.v[0][t], cb.v[1][t], cb.v[2][t], cb.v[3][t],
                                            cb.v[4][t], cur, se, sv);
        else
            return forward_point_scaled<T>(cb.v[0][t], cb.v[1][t], cb.v[2][t], cb.v[3][t],
                                           cb.v[4][t], cb.v[5][t], cur, se, sv, sw);
    } else {
        if constexpr (Five)
            return forward_point5_y<T>(cb.v[0][t], cb.v[1][t], cb.v[NA - 1][t], cb.v[2][t],
                                       cb.v[3][t], cb.v[4][t], cur, se, sv);
        else
            return forward_point_y<T>(cb.v[0][t], cb.v[1][t], cb.v[NA - 1][t], cb.v[2][t],
                                      cb.v[3][t], cb.v[4][t], cb.v[5][t], cur, se, sv, sw);
    }
}

// One backward step from the operands of element qh of a half buffer.
template <bool Five, typename T, int NA, int H>
__device__ __forceinline__ T backward_step(const Halves<T, NA, H> &cb, int qh, T cur, T nw, T nv,
                                           T ne)
{
    if constexpr (ScaledRows<T>::value) {
        if constexpr (Five)
            return backward_point5_fused<T>(cb.v[0][qh], cb.v[1][qh], cb.v[2][qh], cb.v[3][qh],
                                            cur, nw, nv);
        else
            return backward_point_fused<T>(cb.v[0][qh], cb.v[1][qh], cb.v[2][qh], cb.v[3][qh],
                                           cb.v[4][qh], cur, nw, nv, ne);
    } else {
        if constexpr (Five)
            return backward_point5<T>(cb.v[0][qh], cb.v[1][qh], cb.v[2][qh], cb.v[3][qh], cur, nw,
                                      nv);
        else
            return backward_point<T>(cb.v[0][qh], cb.v[1][qh], cb.v[2][qh], cb.v[3][qh],
                                     cb.v[4][qh], cur, nw, nv, ne);
    }
}

// Results of the sweeps leave through LDS.  A sweep workgroup has two
// waves: wave 0 computes, wave 1 stores.  After each chunk the computing
// wave writes its F results per lane into one of two ring slots and meets
// the storing wave at a barrier; the storing wave then writes the chunk to
// HBM and publishes the edge row to the neighbouring band while the
// computing wave goes on with the next chunk.  The computing wave thus
// issues no global stores: on this architecture a store counts in the same
// in-order counter as the loads and its registers stay busy until the store
// completes, so stores in the computing wave made it wait, chunk after
// chunk, for the prefetch loads behind them.  `stop` tells the storing wave
// that the computing wave has given up (a bounded boundary wait failed);
// both waves pass the same number of barriers on every path.
template <typename T, int F>
struct OutRing {
    T v[2][kWave * F];
    int stop;
};

// Barrier between the two waves of a sweep workgroup, after this wave's LDS
// accesses have completed.  Only LDS is ordered: the wait leaves the vector
// memory counter alone (vmcnt 63, expcnt 7, lgkmcnt 0), so loads in flight
// stay in flight.  The signal fences keep the compiler from moving memory
// accesses across the barrier.
__device__ __forceinline__ void handoff()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// A plain LDS read after the barrier (handoff's fences keep it there); an
// atomic read would make the compiler wait for every load in flight first.
template <typename Ring>
__device__ __forceinline__ bool ring_stopped(Ring &ring)
{
    return ring.stop != 0;
}

// The storing wave of a sweep: for chunk c = 0.. (after barrier c) writes
// ring slot c&1 to dst at chunk `stored(c)` and publishes the edge lane's F
// values to the neighbouring band: lane q < F takes element `elem(q)` of
// lane `edge` and publishes it as column col(c, q) when `ok(col)`.
template <typename T, int F, typename FSTORED, typename FCOL>
__device__ void store_chunks(T *__restrict__ dst, const Skew &sk, int band, int nf,
                             OutRing<T, F> &ring, u64 *bout, bool has_out, unsigned gen,
                             int edge, bool reversed, FSTORED stored, FCOL col_of)
{
    const int lane = threadIdx.x & (kWave - 1);
    const unsigned lofs = (unsigned)lane * F;
    constexpr int NW = TagWords<T>::n;
    for (int c = 0; c < nf; ++c) {
        handoff();
        if (ring_stopped(ring)) return;
        const T *slot = ring.v[c & 1];
        T vals[F];
        vload<T, F>(vals, slot + lofs);
        vstore<T, F>(dst + sk.chunk_base(band, stored(c), 0), lofs, vals);
        const int q = lane < F ? lane : 0;
        const T pv = slot[edge * F + (reversed ? F - 1 - q : q)];
        const int col = col_of(c, lane);
        if (has_out && lane < F && col >= 0 && col < sk.nx)
            put_tagged(bout + (long)col * NW, gen, pv);
    }
}

// Loads of a half chunk with padding masked off (the batch's sweeps, whose
// small systems are mostly padding): buffer loads, whose out-of-range
// offsets read zeros without a memory request.  A lane whose half lies
// wholly outside its row's columns (or below the last row) gets that offset
// and so zeros where the arrays hold the padding values; the zero padding
// values that follow meet only zero couplings, as the stored ones do.
constexpr int kNoOffset = int(0x80000000u);
__device__ __forceinline__ __amdgpu_buffer_rsrc_t array_rsrc(const void *base, long elems,
                                                             int size)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0,
                                             (int)(elems * size), 0x00020000);
}
template <typename T, int NA, int H>
__device__ __forceinline__ void masked_half(Halves<T, NA, H> &b,
                                            const __amdgpu_buffer_rsrc_t (&rs)[NA], int ofs)
{
    constexpr int V = 16 / (int)sizeof(T);
#pragma unroll
    for (int k = 0; k < NA; ++k)
#pragma unroll
        for (int q = 0; q < H / V; ++q) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs[k], ofs + q * 16, 0, 0);
            __builtin_memcpy(&b.v[k][q * V], &v, 16);
        }
}

// --------------------------------------------------------------------------
// forward sweep  L w = -r  (msi2d9-v03.f90:139-156); r, dl (dl1..dl5, then
// 1/dl1), w in the skewed layout.
// --------------------------------------------------------------------------
// Mask: padding loads masked off (masked_half).
template <typename T, int F, bool Five, bool Mask = false>
__device__ __forceinline__ void forward_band(const T *__restrict__ r, const T *__restrict__ dl,
                                             T *__restrict__ w, const Skew &sk, const Link &lk,
                                             int band)
{
    __shared__ OutRing<T, F> ring;
    const int lane = threadIdx.x & (kWave - 1);
    const int nx = sk.nx, pad = sk.pad(band);
    const int nf = sk.nst / F;  // chunks of F steps
    const long L = sk.size();
    const bool has_up = band > 0, has_down = band < sk.nb - 1;
    constexpr int NW = TagWords<T>::n;
    if (threadIdx.x >= kWave) {  // the storing wave; edge = lane 63, columns c*F + q - pad - 126
        store_chunks<T, F>(
            w, sk, band, nf, ring, lk.bnd + (long)band * nx * NW, has_down, lk.gen, kWave - 1,
            false, [](int c) { return c; },
            [=](int c, int q) { return c * F + q - pad - 2 * (kWave - 1); });
        return;
    }
    constexpr int H = F / 2;
    constexpr int V = 16 / (int)sizeof(T);  // values per 16-byte vector
    // a lane's offset in a chunk block; chunk bases are wave-uniform, so the
    // vector accesses take an SGPR base and this fixed VGPR offset
    const unsigned lofs = (unsigned)lane * F;
    typedef SweepOps<T, Five> Ops;
    constexpr int NA = Ops::NF;
    typedef Halves<T, NA, H> Buf;  // r, dl1..dl5, 1/dl1 (scaled: r, ny, ne2..ne5)
    const u64 *bin = lk.bnd + (long)(has_up ? band - 1 : 0) * nx * NW;

    Buf X, Y, Z;
    Raw<T> bn{};
    T bv[F + 1];
    T cur = 0, se = 0, sv = 0, sw = 0;

    __amdgpu_buffer_rsrc_t rs[NA];
    const bool row_real = band * kWave + lane < sk.ny;
    if constexpr (Mask) {
#pragma unroll
        for (int k = 0; k < NA; ++k)
            rs[k] = array_rsrc(k == 0 ? r : dl + (long)Ops::fslot(k) * L, L, (int)sizeof(T));
    }
    auto load_half = [&](Buf &b, int h) {  // half h = chunk h/2, elements (h%2)*H..
        const int c = min(h >> 1, nf - 1);
        const long e = sk.chunk_base(band, c, 0) + (h & 1) * H;
        if constexpr (Mask) {
            const int i0 = c * F + (h & 1) * H - 2 * lane - pad;  // column of the first step
            const bool on = row_real && i0 < nx && i0 + H > 0;
            masked_half<T, NA, H>(b, rs, on ? (int)((e + lofs) * sizeof(T)) : kNoOffset);
            return;
        }
#pragma unroll
        for (int k = 0; k < NA; ++k) {
            const T *p = (k == 0 ? r : dl + (long)Ops::fslot(k) * L) + e;
#pragma unroll
            for (int q = 0; q < H / V; ++q) {
                const uint4 v = *reinterpret_cast<const uint4 *>(p + (lofs + q * V));
                __builtin_memcpy(&b.v[k][q * V], &v, 16);
            }
        }
    };
    auto bcol = [&](int c) { return c * F - pad + lane; };  // column held by this lane
    auto bneed = [&](int c) {
        const int col = bcol(c);
        return has_up && lane <= F && col >= 0 && col < nx;
    };
    auto load_bnd = [&](Raw<T> &b, int c) {
        raw_load<T>(b, bin + (long)min(max(bcol(c), 0), nx - 1) * NW);
    };
    // Lane q of chunk c's boundary holds column bcol(c); columns outside the
    // grid (and every column of the bottom band) read as zero.
    auto take = [&](const Raw<T> &b, int c) {
        const T val = bneed(c) ? raw_value(b) : T(0);
#pragma unroll
        for (int q = 0; q <= F; ++q) bv[q] = read_lane(val, q);
    };
    // Wait for chunk c's boundary words and take them.  The slow path reloads
    // into a temporary and takes from it, so no register it loads is live
    // after the two paths join and the fast path's wait covers only loads
    // issued before the chunk's data loads.
    auto sync = [&](const Raw<T> &b, int c) -> bool {
        if (!has_up || __all(!bneed(c) || raw_ok<T>(b, lk.gen))) {
            take(b, c);
            return true;
        }
        Raw<T> t{};
        const bool ok = slow_wait(lk, [&] { return !bneed(c) || raw_ok<T>(t, lk.gen); },
                                  [&] { load_bnd(t, c); });
        take(t, c);
        return ok;
    };
    // Giving up: tell the storing wave, meet it at the barrier it waits at.
    auto bail = [&]() -> bool {
        if (lane == 0) ring.stop = 1;
        handoff();
        return false;
    };
    // Chunk c: halves 2c (in h0) and 2c+1 (in h1); loads halves 2c+2 into h2
    // and 2c+3 into h0.
    auto chunk = [&](Buf &h0, Buf &h1, Buf &h2, int c) -> bool {
        load_bnd(bn, c + 1);  // chunk c+1's boundary words, before any data load
        T outv[F];
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            Buf &cb = hh == 0 ? h0 : h1;
            // the loads of a half stay where they are: the scheduler would
            // otherwise sink them below independent arithmetic and cut the
            // distance the ring gives them
            __builtin_amdgcn_sched_barrier(0);
            load_half(hh == 0 ? h2 : h0, 2 * c + 2 + hh);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = 0; t < H; ++t) {
                const int s = hh * H + t;
                const T up = lane_from_below_or(cur, bv[s + 1]);  // lane 0: column i+1 above
                sw = sv;
                sv = se;
                se = up;
                const T v = forward_step<Five>(cb, t, cur, se, sv, sw);
                outv[s] = v;
                cur = v;
            }
        }
        vstore<T, F>(ring.v[c & 1], lofs, outv);
        handoff();
        // chunk nf has no columns in the grid, so its sync returns at once;
        // running it unconditionally keeps the boundary load of every chunk
        // complete before the loop's back edge
        return sync(bn, c + 1) || bail();
    };

    if (lane == 0) ring.stop = 0;
    load_half(X, 0);
    load_half(Y, 1);
    load_bnd(bn, 0);
    if (!sync(bn, 0)) {
        bail();
        return;
    }
    se = lane == 0 ? bv[0] : T(0);  // column -pad above (used only if pad = 0)
    for (int c = 0; c < nf; c += 3) {
        if (!chunk(X, Y, Z, c) || c + 1 >= nf) break;
        if (!chunk(Z, X, Y, c + 1) || c + 2 >= nf) break;
        if (!chunk(Y, Z, X, c + 2)) break;
    }
}

// --------------------------------------------------------------------------
// backward sweep  U z = w  (msi2d9-v03.f90:158-176): chunks in reverse
// stored order, z over w in place; bands top to bottom.
// --------------------------------------------------------------------------
// Mask: padding loads masked off (masked_half).
template <typename T, int F, bool Five, bool Mask = false>
__device__ __forceinline__ void backward_band(T *__restrict__ wz, const T *__restrict__ du,
                                              const Skew &sk, const Link &lk, int band)
{
    __shared__ OutRing<T, F> ring;
    const int lane = threadIdx.x & (kWave - 1);
    const int nx = sk.nx, pad = sk.pad(band);
    const int nf = sk.nst / F;
    const long L = sk.size();
    const bool has_above = band < sk.nb - 1, has_below = band > 0;
    constexpr int NW = TagWords<T>::n;
    if (threadIdx.x >= kWave) {  // the storing wave; edge = lane 0, step q = column
                                 // sc*F + F-1-q - pad of stored chunk sc = nf-1-c
        store_chunks<T, F>(
            wz, sk, band, nf, ring, lk.bnd + (long)(has_below ? band - 1 : 0) * nx * NW,
            has_below, lk.gen, 0, true, [=](int c) { return nf - 1 - c; },
            [=](int c, int q) { return (nf - 1 - c) * F + F - 1 - q - pad; });
        return;
    }
    constexpr int H = F / 2;
    constexpr int V = 16 / (int)sizeof(T);
    const unsigned lofs = (unsigned)lane * F;  // lane offset in a chunk block
    constexpr int NA = SweepOps<T, Five>::NB;
    typedef Halves<T, NA, H> Buf;  // w, du1..du4
    const u64 *bin = lk.bnd + (long)(has_above ? band : 0) * nx * NW;

    Buf X, Y, Z;
    Raw<T> bn{};
    T bv[F + 1];
    T cur = 0, nw = 0, nv = 0, ne = 0;

    __amdgpu_buffer_rsrc_t rs[NA];
    const bool row_real = band * kWave + lane < sk.ny;
    if constexpr (Mask) {
#pragma unroll
        for (int k = 0; k < NA; ++k)
            rs[k] = array_rsrc(k == 0 ? (const T *)wz : du + (long)(k - 1) * L, L, (int)sizeof(T));
    }
    // backward half h: stored chunk nf-1-h/2, its upper half first
    auto load_half = [&](Buf &b, int h) {
        const int sc = max(nf - 1 - (h >> 1), 0);
        const long e = sk.chunk_base(band, sc, 0) + (1 - (h & 1)) * H;
        if constexpr (Mask) {
            const int i0 = sc * F + (1 - (h & 1)) * H - 2 * lane - pad;
            const bool on = row_real && i0 < nx && i0 + H > 0;
            masked_half<T, NA, H>(b, rs, on ? (int)((e + lofs) * sizeof(T)) : kNoOffset);
            return;
        }
#pragma unroll
        for (int k = 0; k < NA; ++k) {
            const T *p = (k == 0 ? wz : du + (long)(k - 1) * L) + e;
#pragma unroll
            for (int q = 0; q < H / V; ++q) {
                const uint4 v = *reinterpret_cast<const uint4 *>(p + (lofs + q * V));
                __builtin_memcpy(&b.v[k][q * V], &v, 16);
            }
        }
    };
    auto bcol = [&](int c) { return nx - 1 - c * F - lane; };
    auto bneed = [&](int c) {
        const int col = bcol(c);
        return has_above && lane <= F && col >= 0 && col < nx;
    };
    auto load_bnd = [&](Raw<T> &b, int c) {
        raw_load<T>(b, bin + (long)min(max(bcol(c), 0), nx - 1) * NW);
    };
    auto take = [&](const Raw<T> &b, int c) {
        const T val = bneed(c) ? raw_value(b) : T(0);
#pragma unroll
        for (int q = 0; q <= F; ++q) bv[q] = read_lane(val, q);
    };
    auto sync = [&](const Raw<T> &b, int c) -> bool {
        if (!has_above || __all(!bneed(c) || raw_ok<T>(b, lk.gen))) {
            take(b, c);
            return true;
        }
        Raw<T> t{};
        const bool ok = slow_wait(lk, [&] { return !bneed(c) || raw_ok<T>(t, lk.gen); },
                                  [&] { load_bnd(t, c); });
        take(t, c);
        return ok;
    };
    auto bail = [&]() -> bool {
        if (lane == 0) ring.stop = 1;
        handoff();
        return false;
    };
    auto chunk = [&](Buf &h0, Buf &h1, Buf &h2, int c) -> bool {
        load_bnd(bn, c + 1);
        T outv[F];
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            Buf &cb = hh == 0 ? h0 : h1;
            __builtin_amdgcn_sched_barrier(0);  // loads stay in place, as in the forward sweep
            load_half(hh == 0 ? h2 : h0, 2 * c + 2 + hh);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = 0; t < H; ++t) {
                const int s = hh * H + t;
                const int q = F - 1 - s, qh = H - 1 - t;  // element in the chunk, in the half
                // lane 63: column i-1 of the band above; the top lane of a
                // partial band reads a padding row above it (zero couplings)
                const T dn = lane_from_above_or(cur, bv[s + 1]);
                ne = nv;
                nv = nw;
                nw = dn;
                const T v = backward_step<Five>(cb, qh, cur, nw, nv, ne);
                outv[q] = v;
                cur = v;
            }
        }
        vstore<T, F>(ring.v[c & 1], lofs, outv);
        handoff();
        return sync(bn, c + 1) || bail();
    };

    if (lane == 0) ring.stop = 0;
    load_half(X, 0);
    load_half(Y, 1);
    load_bnd(bn, 0);
    if (!sync(bn, 0)) {
        bail();
        return;
    }
    nw = lane == kWave - 1 ? bv[0] : T(0);  // column nx-1 of the row above
    for (int c = 0; c < nf; c += 3) {
        if (!chunk(X, Y, Z, c) || c + 1 >= nf) break;
        if (!chunk(Z, X, Y, c + 1) || c + 2 >= nf) break;
        if (!chunk(Y, Z, X, c + 2)) break;
    }
}

// --------------------------------------------------------------------------
// LDS-staged sweeps.  As above, wave 0 computes and wave 1 moves data, but
// wave 1 now also loads: it copies every chunk's inputs from HBM straight
// into an LDS ring of kStageSlots chunk slots (direct-to-LDS loads, no
// registers), kStageSlots - 2 chunks ahead of the computing wave, which reads
// its operands from LDS.  The computing wave then issues no global data
// access at all except the boundary words.  Loads of a chunk are complete
// before the barrier that hands the chunk over: the moving wave waits on its
// own vector-memory counter, whose count of later operations is fixed
// because every iteration issues the same number of stores and loads (loads
// past the last chunk re-read the last chunk into the free slot; stores of
// lanes that publish nothing go to a private scratch word).
//
// Barrier protocol (k = 0 .. nf): after barrier k the computing wave works
// on chunk k, and chunks k and k+1 are complete in LDS; before barrier k+1 it
// has written chunk k's results into out-slot k&1.  Between barriers k and
// k+1 the moving wave stores chunk k-1, loads chunk k+D-1 into the slot
// chunk k-1 used, and waits for chunk k+2.
// --------------------------------------------------------------------------
constexpr int kStageSlots = 4;

template <typename T, int F, int NA>
struct StagedRing {
    static constexpr int NQ = F * (int)sizeof(T) / 16;  // 16-byte vectors per lane and array
    uint4 in[kStageSlots][NA * NQ * kWave];            // [slot][(a * NQ + q) * 64 + lane]
    T out[2][kWave * F];
    int stop;
};

typedef __attribute__((address_space(1))) const void global_cvoid;
typedef __attribute__((address_space(3))) void local_void;

// s_waitcnt vmcnt(N) alone (expcnt and lgkmcnt left alone).
template <int N>
__device__ __forceinline__ void wait_vmem()
{
    static_assert(N >= 0 && N <= 63, "vmcnt is six bits");
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// The moving wave of a staged sweep.  base(a, c) is the wave-uniform start of
// array a's block of processing chunk c; stored(c) the stored chunk index
// of processing chunk c (for the output); the edge and publication
// arguments are those of store_chunks.
template <typename T, int F, int NA, typename FBASE, typename FSTORED, typename FCOL>
__device__ void move_chunks(StagedRing<T, F, NA> &R, T *__restrict__ dst, const Skew &sk, int band,
                            int nf, u64 *bout, bool has_out, u64 *scratch, unsigned gen, int edge,
                            bool reversed, FBASE base, FSTORED stored, FCOL col_of)
{
    constexpr int D = kStageSlots, NQ = StagedRing<T, F, NA>::NQ;
    constexpr int V = 16 / (int)sizeof(T);
    constexpr int NW = TagWords<T>::n;
    constexpr int LPC = NA * NQ;  // loads per chunk
    constexpr int SPC = NQ + NW;  // stores per chunk
    const int lane = threadIdx.x & (kWave - 1);
    const unsigned lofs = (unsigned)lane * F;
    auto stage = [&](int c) {  // all NA*NQ vectors of chunk min(c, nf-1) into slot c % D
        const int cc = min(c, nf - 1);
        uint4 *slot = R.in[c % D];
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const T *g = base(a, cc);
#pragma unroll
            for (int q = 0; q < NQ; ++q)
                __builtin_amdgcn_global_load_lds((global_cvoid *)(g + (lofs + q * V)),
                                                 (local_void *)(slot + (a * NQ + q) * kWave), 16,
                                                 0, 0);
        }
    };
    auto store = [&](int c) {  // chunk c's results, from out-slot c & 1
        const T *slot = R.out[c & 1];
        T vals[F];
        vload<T, F>(vals, slot + lofs);
        vstore<T, F>(dst + sk.chunk_base(band, stored(c), 0), lofs, vals);
        const int q = lane < F ? lane : 0;
        const T pv = slot[edge * F + (reversed ? F - 1 - q : q)];
        const int col = col_of(c, lane);
        const bool ok = has_out && lane < F && col >= 0 && col < sk.nx;
        put_tagged(ok ? bout + (long)col * NW : scratch, gen, pv);
    };
    for (int c = 0; c < D - 1; ++c) stage(c);
    wait_vmem<(D - 3) * LPC>();  // chunks 0 and 1 complete
    handoff();
    if (ring_stopped(R)) return;
    for (int k = 0; k < nf; ++k) {
        if (k >= 1) store(k - 1);
        stage(k + D - 1);
        if (k == 0)
            wait_vmem<(D - 3) * LPC>();  // chunk 2 (from the prologue) complete
        else
            wait_vmem<(D - 3) * (LPC + SPC)>();  // chunk k+2 complete
        handoff();
        if (ring_stopped(R)) return;
    }
    store(nf - 1);
}

template <typename T, int F, int NA, int H>
__device__ __forceinline__ void read_half(Halves<T, NA, H> &b, const StagedRing<T, F, NA> &R,
                                          int c, int qh0, int lane)
{
    constexpr int NQ = StagedRing<T, F, NA>::NQ, V = 16 / (int)sizeof(T);
    const uint4 *slot = R.in[c % kStageSlots];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int q = 0; q < H / V; ++q) {
            const uint4 v = slot[(a * NQ + qh0 + q) * kWave + lane];
            __builtin_memcpy(&b.v[a][q * V], &v, 16);
        }
}

// forward sweep, staged (same arithmetic and results as forward_band).
template <typename T, int F, bool Five>
__device__ __forceinline__ void forward_band_staged(const T *__restrict__ r,
                                                    const T *__restrict__ dl, T *__restrict__ w,
                                                    const Skew &sk, const Link &lk, int band)
{
    typedef SweepOps<T, Five> Ops;
    constexpr int NA = Ops::NF;
    constexpr int H = F / 2, NQ = StagedRing<T, F, NA>::NQ;
    __shared__ StagedRing<T, F, NA> R;
    const int lane = threadIdx.x & (kWave - 1);
    const int nx = sk.nx, pad = sk.pad(band);
    const int nf = sk.nst / F;
    const long L = sk.size();
    const bool has_up = band > 0, has_down = band < sk.nb - 1;
    constexpr int NW = TagWords<T>::n;
    if (threadIdx.x >= kWave) {
        move_chunks<T, F, NA>(
            R, w, sk, band, nf, lk.bnd + (long)band * nx * NW, has_down,
            lk.scratch + ((long)band * kWave + lane) * 8, lk.gen, kWave - 1, false,
            [=](int a, int c) {
                return (a == 0 ? r : dl + (long)Ops::fslot(a) * L) + sk.chunk_base(band, c, 0);
            },
            [](int c) { return c; },
            [=](int c, int q) { return c * F + q - pad - 2 * (kWave - 1); });
        return;
    }
    typedef Halves<T, NA, H> Buf;
    const unsigned lofs = (unsigned)lane * F;
    const u64 *bin = lk.bnd + (long)(has_up ? band - 1 : 0) * nx * NW;
    Buf A, B;
    Raw<T> bn{};
    T bv[F + 1];
    T cur = 0, se = 0, sv = 0, sw = 0;
    auto bcol = [&](int c) { return c * F - pad + lane; };
    auto bneed = [&](int c) {
        const int col = bcol(c);
        return has_up && lane <= F && col >= 0 && col < nx;
    };
    auto load_bnd = [&](Raw<T> &b, int c) {
        raw_load<T>(b, bin + (long)min(max(bcol(c), 0), nx - 1) * NW);
    };
    auto take = [&](const Raw<T> &b, int c) {
        const T val = bneed(c) ? raw_value(b) : T(0);
#pragma unroll
        for (int q = 0; q <= F; ++q) bv[q] = read_lane(val, q);
    };
    auto sync = [&](const Raw<T> &b, int c) -> bool {
        if (!has_up || __all(!bneed(c) || raw_ok<T>(b, lk.gen))) {
            take(b, c);
            return true;
        }
        Raw<T> t{};
        const bool ok = slow_wait(lk, [&] { return !bneed(c) || raw_ok<T>(t, lk.gen); },
                                  [&] { load_bnd(t, c); });
        take(t, c);
        return ok;
    };
    auto bail = [&]() {
        if (lane == 0) R.stop = 1;
        handoff();
    };
    auto steps = [&](const Buf &cb, T(&outv)[F], int hh) {
#pragma unroll
        for (int t = 0; t < H; ++t) {
            const int s = hh * H + t;
            const T up = lane_from_below_or(cur, bv[s + 1]);
            sw = sv;
            sv = se;
            se = up;
            const T v = forward_step<Five>(cb, t, cur, se, sv, sw);
            outv[s] = v;
            cur = v;
        }
    };

    if (lane == 0) R.stop = 0;
    load_bnd(bn, 0);
    if (!sync(bn, 0)) {
        bail();
        return;
    }
    se = lane == 0 ? bv[0] : T(0);
    handoff();  // barrier 0: chunks 0 and 1 in LDS
    read_half<T, F, NA, H>(A, R, 0, 0, lane);
    for (int c = 0; c < nf; ++c) {
        load_bnd(bn, c + 1);
        T outv[F];
        read_half<T, F, NA, H>(B, R, c, NQ / 2, lane);
        steps(A, outv, 0);
        read_half<T, F, NA, H>(A, R, c + 1, 0, lane);
        steps(B, outv, 1);
        vstore<T, F>(R.out[c & 1], lofs, outv);
        if (!sync(bn, c + 1)) {
            bail();
            return;
        }
        handoff();  // barrier c+1
    }
}

// backward sweep, staged (same arithmetic and results as backward_band).
// Processing chunk c is stored chunk nf-1-c, visited from its upper half down.
template <typename T, int F, bool Five>
__device__ __forceinline__ void backward_band_staged(T *__restrict__ wz, const T *__restrict__ du,
                                                     const Skew &sk, const Link &lk, int band)
{
    constexpr int NA = SweepOps<T, Five>::NB, H = F / 2, NQ = StagedRing<T, F, NA>::NQ;
    __shared__ StagedRing<T, F, NA> R;
    const int lane = threadIdx.x & (kWave - 1);
    const int nx = sk.nx, pad = sk.pad(band);
    const int nf = sk.nst / F;
    const long L = sk.size();
    const bool has_above = band < sk.nb - 1, has_below = band > 0;
    constexpr int NW = TagWords<T>::n;
    if (threadIdx.x >= kWave) {
        move_chunks<T, F, NA>(
            R, wz, sk, band, nf, lk.bnd + (long)(has_below ? band - 1 : 0) * nx * NW, has_below,
            lk.scratch + ((long)band * kWave + lane) * 8, lk.gen, 0, true,
            [=](int a, int c) {
                return (a == 0 ? (const T *)wz : du + (long)(a - 1) * L) +
                       sk.chunk_base(band, nf - 1 - c, 0);
            },
            [=](int c) { return nf - 1 - c; },
            [=](int c, int q) { return (nf - 1 - c) * F + F - 1 - q - pad; });
        return;
    }
    typedef Halves<T, NA, H> Buf;
    const unsigned lofs = (unsigned)lane * F;
    const u64 *bin = lk.bnd + (long)(has_above ? band : 0) * nx * NW;
    Buf A, B;
    Raw<T> bn{};
    T bv[F + 1];
    T cur = 0, nw = 0, nv = 0, ne = 0;
    auto bcol = [&](int c) { return nx - 1 - c * F - lane; };
    auto bneed = [&](int c) {
        const int col = bcol(c);
        return has_above && lane <= F && col >= 0 && col < nx;
    };
    auto load_bnd = [&](Raw<T> &b, int c) {
        raw_load<T>(b, bin + (long)min(max(bcol(c), 0), nx - 1) * NW);
    };
    auto take = [&](const Raw<T> &b, int c) {
        const T val = bneed(c) ? raw_value(b) : T(0);
#pragma unroll
        for (int q = 0; q <= F; ++q) bv[q] = read_lane(val, q);
    };
    auto sync = [&](const Raw<T> &b, int c) -> bool {
        if (!has_above || __all(!bneed(c) || raw_ok<T>(b, lk.gen))) {
            take(b, c);
            return true;
        }
        Raw<T> t{};
        const bool ok = slow_wait(lk, [&] { return !bneed(c) || raw_ok<T>(t, lk.gen); },
                                  [&] { load_bnd(t, c); });
        take(t, c);
        return ok;
    };
    auto bail = [&]() {
        if (lane == 0) R.stop = 1;
        handoff();
    };
    auto steps = [&](const Buf &cb, T(&outv)[F], int hh) {
#pragma unroll
        for (int t = 0; t < H; ++t) {
            const int s = hh * H + t;
            const int q = F - 1 - s, qh = H - 1 - t;
            const T dn = lane_from_above_or(cur, bv[s + 1]);
            ne = nv;
            nv = nw;
            nw = dn;
            const T v = backward_step<Five>(cb, qh, cur, nw, nv, ne);
            outv[q] = v;
            cur = v;
        }
    };

    if (lane == 0) R.stop = 0;
    load_bnd(bn, 0);
    if (!sync(bn, 0)) {
        bail();
        return;
    }
    nw = lane == kWave - 1 ? bv[0] : T(0);
    handoff();
    read_half<T, F, NA, H>(A, R, 0, NQ / 2, lane);  // upper half first
    for (int c = 0; c < nf; ++c) {
        load_bnd(bn, c + 1);
        T outv[F];
        read_half<T, F, NA, H>(B, R, c, 0, lane);
        steps(A, outv, 0);
        read_half<T, F, NA, H>(A, R, c + 1, NQ / 2, lane);
        steps(B, outv, 1);
        vstore<T, F>(R.out[c & 1], lofs, outv);
        if (!sync(bn, c + 1)) {
            bail();
            return;
        }
        handoff();
    }
}

// --------------------------------------------------------------------------
// lu2d9 factorization: A (skewed copy) -> dl, du (skewed).  The edge row
// carries the four du values of each column.
//
// The skewed copy of A holds a unit diagonal at padding entries (columns
// outside the grid, rows beyond the last one), where alpha is 0 as well, so
// a padding point's factor is dl1 = 1 and every other entry a zero: its du
// values are the zeros the reference's edge formulas assume for a missing
// neighbour, and the step needs no masks.  Pivots are checked afterwards,
// by prepare_entry.  Per step the wave issues the point's arithmetic, three
// divisions (dl4, dl2, 1/dl1; the four du quotients through 1/dl1) and two
// lane shifts per du value; the band below's edge values reach lane 0 by a
// row-local lane shift once per chunk.  Results leave through LDS as the
// sweeps' do: wave 0 computes and writes each half chunk of dl and du into
// one of two ring slots, wave 1 stores the chunk and publishes the band's
// last row (the du values of each column) to the band above, so the
// computing wave issues no global store (whose registers would stay busy
// until it completes, behind the operand loads in the in-order counter).
// --------------------------------------------------------------------------
template <typename T, int F, int NO>
struct FactorRing {
    T v[2][NO][kWave * F];
    int stop;
};

// fn(integral_constant<int, I>) for I = 0 .. N-1: the step index as a
// compile-time constant (DPP controls are immediates)
template <int I, int N, typename Fn>
__device__ __forceinline__ void static_for(Fn &&fn)
{
    if constexpr (I < N) {
        fn(std::integral_constant<int, I>{});
        static_for<I + 1, N>(fn);
    }
}

template <typename T, int F, bool Five>
__device__ __forceinline__ void factor_band(const T *__restrict__ a, T *__restrict__ dl,
                                            T *__restrict__ du, T alpha0, const Skew &sk,
                                            const Link &lk, int band)
{
    // Five (lu2d5): the S, W, P, E, N arrays of the skewed 9-array copy in,
    // dl1..dl4 and du1..du3 out (dl5 and du4 of the embedding are zero and
    // are neither read nor written)
    constexpr int NA = Five ? 5 : 9, NL = Five ? 4 : 5, NU = Five ? 3 : 4, NO = NL + NU;
    constexpr int S = 2 * F;  // skew chunk: 64 bytes per lane and array
    constexpr int H = F / 2;  // results go to the ring per 16-byte half chunk
    __shared__ FactorRing<T, F, NO> ring;
    const int lane = threadIdx.x & (kWave - 1);
    const int nx = sk.nx, ny = sk.ny, pad = sk.pad(band);
    const int nf = sk.nst / F;
    const long L = sk.size();
    const bool has_up = band > 0, has_down = band < sk.nb - 1;
    constexpr int NW = TagWords<T>::n;
    constexpr int CW = 4 * NW;
    const unsigned lofs = (unsigned)lane * F;
    if (threadIdx.x >= kWave) {  // the storing wave
        u64 *bout = lk.bnd + (long)band * nx * CW;
        for (int c = 0; c < nf; ++c) {
            handoff();
            if (ring_stopped(ring)) return;
            const long e = sk.template index_c<S>(band, lane, c * F);
#pragma unroll
            for (int k = 0; k < NO; ++k) {
                T vals[F];
                vload<T, F>(vals, ring.v[c & 1][k] + lofs);
                vstore<T, F>((k < NL ? dl + k * L : du + (k - NL) * L) + e, vals);
            }
            // lane q < F: lane 63's du values of step q, column c*F + q - pad - 126
            const int col = c * F + lane - pad - 2 * (kWave - 1);
            if (has_down && lane < F && col >= 0 && col < nx) {
#pragma unroll
                for (int k = 0; k < NU; ++k)
                    put_tagged(bout + (long)col * CW + k * NW, lk.gen,
                               ring.v[c & 1][NL + k][(kWave - 1) * F + lane]);
            }
        }
        return;
    }
    const int j = band * kWave + lane;
    const u64 *bin = lk.bnd + (long)(has_up ? band - 1 : 0) * nx * CW;
    // alpha of this row's real volumes (alpha_at: 1 <= i <= nx-2 as well)
    const T arow = (j >= 1 && j <= ny - 2) ? alpha0 : T(0);
    auto aslot = [](int k) { return Five ? (k == 0 ? kS : k == 1 ? kW : k == 2 ? kP : k == 3 ? kE : kN) : k; };

    T A[NA][F], B[NA][F];
    Raw<T> bn[4] = {};
    T val[NU];
    T cur[NU] = {}, se[NU] = {}, sv[NU] = {}, sw[NU] = {};

    auto load_data = [&](T(&buf)[NA][F], int c) {
        const long e = sk.template index_c<S>(band, lane, min(c, nf - 1) * F);
#pragma unroll
        for (int k = 0; k < NA; ++k) vload<T, F>(buf[k], a + aslot(k) * L + e);
    };
    auto bcol = [&](int c) { return c * F - pad + lane; };
    auto bneed = [&](int c) {
        const int col = bcol(c);
        return has_up && lane <= F && col >= 0 && col < nx;
    };
    auto load_bnd = [&](Raw<T>(&b)[4], int c) {
        const u64 *p = bin + (long)min(max(bcol(c), 0), nx - 1) * CW;
#pragma unroll
        for (int k = 0; k < NU; ++k) raw_load<T>(b[k], p + k * NW);
    };
    auto okall = [&](const Raw<T>(&b)[4]) {
        bool ok = true;
#pragma unroll
        for (int k = 0; k < NU; ++k) ok = ok && raw_ok<T>(b[k], lk.gen);
        return ok;
    };
    auto ready = [&](Raw<T>(&b)[4], int c) -> bool {
        if (!has_up || __all(!bneed(c) || okall(b))) return true;
        Raw<T> t[4] = {};
        const bool ok =
            slow_wait(lk, [&] { return !bneed(c) || okall(t); }, [&] { load_bnd(t, c); });
#pragma unroll
        for (int k = 0; k < NU; ++k) b[k] = t[k];
        return ok;
    };
    // lane q <= F holds column c*F - pad + q of the band below (zero where
    // there is none); at step s lane 0 takes lane s+1's value
    auto take = [&](const Raw<T>(&b)[4], int c) {
#pragma unroll
        for (int k = 0; k < NU; ++k) val[k] = bneed(c) ? raw_value(b[k]) : T(0);
    };
    // Giving up: tell the storing wave, meet it at the barrier it waits at.
    auto bail = [&]() -> bool {
        if (lane == 0) ring.stop = 1;
        handoff();
        return false;
    };
    // operands are loaded Ahead chunks before use: two for lu2d5's five
    // arrays (a ring of three register buffers), one for nine (whose
    // registers do not allow a third buffer)
    constexpr int Ahead = Five ? 2 : 1;
    auto chunk = [&](T(&cb)[NA][F], T(&nbuf)[NA][F], int c) -> bool {
        load_data(nbuf, c + Ahead);
        load_bnd(bn, c + 1);
        T oL[NL][F], oU[NU][F];
        const int i0 = c * F - pad - 2 * lane;
        auto &slot = ring.v[c & 1];
        static_for<0, F>([&](auto St) {
            constexpr int s = decltype(St)::value;
            static_assert(s + 1 <= 15, "lane 0's edge value comes by row_shl");
#pragma unroll
            for (int k = 0; k < NU; ++k) {
                const T up = lane_from_below_or(cur[k], lane_plus<s + 1>(val[k]));
                sw[k] = sv[k];
                sv[k] = se[k];
                se[k] = up;
            }
            const T al = (unsigned)(i0 + s - 1) < (unsigned)(nx - 2) ? arow : T(0);
            T bp[NA], Lv[NL], Uv[NU];
#pragma unroll
            for (int k = 0; k < NA; ++k) bp[k] = cb[k][s];
            if constexpr (Five)
                factor_point5_y<T>(bp, al, cur, sv, se, Lv, Uv);
            else
                factor_point_y<T>(bp, al, cur, sv, se, sw, Lv, Uv);
#pragma unroll
            for (int k = 0; k < NL; ++k) oL[k][s] = Lv[k];
#pragma unroll
            for (int k = 0; k < NU; ++k) {
                oU[k][s] = Uv[k];
                cur[k] = Uv[k];
            }
            if constexpr (s % H == H - 1) {
                constexpr int h = s / H;
#pragma unroll
                for (int k = 0; k < NL; ++k)
                    vstore<T, H>(slot[k] + lofs + h * H,
                                 reinterpret_cast<const T(&)[H]>(oL[k][h * H]));
#pragma unroll
                for (int k = 0; k < NU; ++k)
                    vstore<T, H>(slot[NL + k] + lofs + h * H,
                                 reinterpret_cast<const T(&)[H]>(oU[k][h * H]));
            }
        });
        handoff();
        if (c + 1 < nf) {
            if (!ready(bn, c + 1)) return bail();
            take(bn, c + 1);
        }
        return true;
    };

    if (lane == 0) ring.stop = 0;
    load_data(A, 0);
    if constexpr (Five) load_data(B, 1);
    load_bnd(bn, 0);
    if (!ready(bn, 0)) {
        bail();
        return;
    }
    take(bn, 0);
#pragma unroll
    for (int k = 0; k < NU; ++k) se[k] = lane == 0 ? val[k] : T(0);
    if constexpr (Five) {
        // chunk c computes from buffer c mod 3 and loads chunk c + 2 into
        // the buffer chunk c - 1 used
        T C[NA][F];
        for (int c = 0; c < nf; c += 3) {
            if (!chunk(A, C, c) || c + 1 >= nf) break;
            if (!chunk(B, A, c + 1) || c + 2 >= nf) break;
            if (!chunk(C, B, c + 2)) break;
        }
    } else {
        for (int c = 0; c < nf; c += 2) {
            if (!chunk(A, B, c) || c + 1 >= nf) break;
            if (!chunk(B, A, c + 1)) break;
        }
    }
}

// --------------------------------------------------------------------------
// Single-system kernels: workgroup g (g % xcd_stride == 0) runs band
// g / xcd_stride, so with xcd_stride 8 every band lands on the same XCD
// (workgroups are dealt to the XCDs round robin) and shares its L2.
// --------------------------------------------------------------------------
// Both sweeps of one iteration in one launch: each band's workgroup runs its
// forward sweep, then its backward sweep, which waits only for the band
// above, whose forward sweep ended after this band's.  Boundary slot s
// carries band s's forward edge to band s+1 and band s+1's backward edge to
// band s, each under its own sweep's generation; band s+1 writes its
// backward words only after its forward sweep has read every forward word of
// the slot, so no forward word lands on a backward one.
template <typename T, int F, bool Five>
__global__ __launch_bounds__(2 * kWave) void k_sweeps(const T *__restrict__ r,
                                                  const T *__restrict__ dl, T *__restrict__ w,
                                                  const T *__restrict__ du, Skew sk, Link lk0,
                                                  int xcd_stride)
{
    if (lk0.st->done) return;
    const int blk = blockIdx.x;
    if (blk % xcd_stride) return;
    const int band = blk / xcd_stride;
    Link lk = lk0;
    lk.phase = 1;
    lk.gen = launch_gen(lk);
    forward_band<T, F, Five>(r, dl, w, sk, lk, band);
    // the band's w is stored (the storing wave's last chunk included) before
    // the backward sweep loads it: the two waves share the CU, so waiting for
    // the stores to complete is enough (no cache write-back); a band that
    // gave up skips the backward sweep, on both waves alike
    __shared__ int stop;
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (threadIdx.x == 0)
        stop = __hip_atomic_load(&lk0.st->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (stop) return;
    lk.phase = 2;
    lk.gen = launch_gen(lk);
    backward_band<T, F, Five>(w, du, sk, lk, band);
}

template <typename T, int F, bool Five>
__global__ __launch_bounds__(2 * kWave) void k_forward_staged(const T *__restrict__ r,
                                                          const T *__restrict__ dl,
                                                          T *__restrict__ w, Skew sk, Link lk0,
                                                          int xcd_stride)
{
    if (lk0.st->done) return;
    const int blk = blockIdx.x;
    if (blk % xcd_stride) return;
    Link lk = lk0;
    lk.gen = launch_gen(lk0);
    forward_band_staged<T, F, Five>(r, dl, w, sk, lk, blk / xcd_stride);
}

template <typename T, int F, bool Five>
__global__ __launch_bounds__(2 * kWave) void k_backward_staged(T *__restrict__ wz,
                                                           const T *__restrict__ du, Skew sk,
                                                           Link lk0, int xcd_stride)
{
    if (lk0.st->done) return;
    const int blk = blockIdx.x;
    if (blk % xcd_stride) return;
    Link lk = lk0;
    lk.gen = launch_gen(lk0);
    backward_band_staged<T, F, Five>(wz, du, sk, lk, sk.nb - 1 - blk / xcd_stride);
}

template <typename T, int F, bool Five>
__global__ __launch_bounds__(2 * kWave) void k_factor(const T *__restrict__ a, T *__restrict__ dl,
                                                  T *__restrict__ du, T alpha0, Skew sk, Link lk0,
                                                  int xcd_stride)
{
    const int blk = blockIdx.x;
    if (blk % xcd_stride || lk0.st->skip) return;
    Link lk = lk0;
    lk.gen = launch_gen(lk0);
    factor_band<T, F, Five>(a, dl, du, alpha0, sk, lk, blk / xcd_stride);
}

// --------------------------------------------------------------------------
// Batched kernels: workgroup g runs band map[g].y of system map[g].x (x < 0:
// an idle slot).  The map lists the bands of every system in dependency
// order (bottom to top for the factorization and the forward sweep, top to
// bottom for the backward sweep), so a band again waits only on workgroups
// with lower indices.
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ Link batch_link(const BatchSys<T> &d, int phase)
{
    Link lk;
    lk.bnd = d.bnd;
    lk.scratch = d.scratch;
    lk.st = d.st;
    lk.phase = phase;
    lk.gen = 0;
    lk.gen = launch_gen(lk);
    return lk;
}

template <typename T, int F, bool Five>
__global__ __launch_bounds__(2 * kWave) void k_forward_batch(const BatchSys<T> *__restrict__ sys,
                                                         const int2 *__restrict__ map)
{
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys<T> &d = sys[m.x];
    if (d.st->done) return;
    forward_band<T, F, Five, true>(d.r, d.dl, d.w, d.sk, batch_link(d, 1), m.y);
}

template <typename T, int F, bool Five>
__global__ __launch_bounds__(2 * kWave) void k_backward_batch(const BatchSys<T> *__restrict__ sys,
                                                          const int2 *__restrict__ map)
{
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys<T> &d = sys[m.x];
    if (d.st->done) return;
    backward_band<T, F, Five, true>(d.w, d.du, d.sk, batch_link(d, 2), m.y);
}

template <typename T, int F, bool Five>
__global__ __launch_bounds__(2 * kWave) void k_forward_batch_staged(
    const BatchSys<T> *__restrict__ sys, const int2 *__restrict__ map)
{
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys<T> &d = sys[m.x];
    if (d.st->done) return;
    forward_band_staged<T, F, Five>(d.r, d.dl, d.w, d.sk, batch_link(d, 1), m.y);
}

template <typename T, int F, bool Five>
__global__ __launch_bounds__(2 * kWave) void k_backward_batch_staged(
    const BatchSys<T> *__restrict__ sys, const int2 *__restrict__ map)
{
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys<T> &d = sys[m.x];
    if (d.st->done) return;
    backward_band_staged<T, F, Five>(d.w, d.du, d.sk, batch_link(d, 2), m.y);
}

template <typename T, int F>
__global__ __launch_bounds__(2 * kWave) void k_factor_batch(const BatchSys<T> *__restrict__ sys,
                                                        const int2 *__restrict__ map)
{
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;
    const BatchSys<T> &d = sys[m.x];
    factor_band<T, F, false>(d.a, d.dl, d.du, T(0.5), d.sk, batch_link(d, 0), m.y);
}

// --------------------------------------------------------------------------
// layout conversions and the update
// --------------------------------------------------------------------------
// Grid point of skewed entry e: i, j (0-based); real = inside the grid.
struct SkewPoint {
    int i, j;
    bool real;
};
__device__ __forceinline__ SkewPoint skew_point(const Skew &sk, long e)
{
    const int s = (int)(e % sk.S);
    const long t1 = e / sk.S;
    const int l = (int)(t1 % 64);
    const long t2 = t1 / 64;
    const int c = (int)(t2 % sk.nch);
    const int b = (int)(t2 / sk.nch);
    SkewPoint q;
    q.j = b * 64 + l;
    q.i = c * sk.S + s - sk.pad(b) - 2 * l;
    q.real = q.j < sk.ny && q.i >= 0 && q.i < sk.nx;
    return q;
}

// nk arrays, reference layout (stride nx*ny) -> skewed entry e (stride
// sk.size()); padding entries get zero, and one in array `unit` (the
// diagonal of A: the factorization's unit padding; -1: none).
template <typename TI, typename TO>
__device__ __forceinline__ void to_skew_entry(const TI *__restrict__ src, TO *__restrict__ dst,
                                              int nk, const Skew &sk, long e, int unit)
{
    const long L = sk.size();
    const SkewPoint q = skew_point(sk, e);
    const long N = (long)sk.nx * sk.ny;
    const long p = q.real ? (long)q.j * sk.nx + q.i : 0;
    for (int k = 0; k < nk; ++k)
        dst[k * L + e] = q.real ? (TO)src[k * N + p] : (k == unit ? TO(1) : TO(0));
}

// Sweep-ready factor at skewed entry e: dl[5] = 1/dl1 at grid points; at
// padding entries (outside the grid or beyond the last row) dl1 = 1/dl1 = 1
// and every coupling zero, so the sweeps compute finite padding values.
// Returns true for a grid point whose pivot dl1 is zero or not finite.
template <typename T>
__device__ __forceinline__ bool prepare_entry(T *__restrict__ dl, T *__restrict__ du,
                                              const Skew &sk, long e)
{
    const long L = sk.size();
    if (skew_point(sk, e).real) {
        const bool bad = bad_pivot(dl[e]);
        if constexpr (ScaledRows<T>::value) {  // ny = -1/dl1, ne_k = ny * dl_k
            const T ny = -(T(1) / dl[e]);
            dl[e] = ny;
#pragma unroll
            for (int k = 1; k < 5; ++k) dl[k * L + e] = dl[k * L + e] * ny;
        } else {
            dl[5 * L + e] = T(1) / dl[e];
        }
        return bad;
    }
    dl[e] = ScaledRows<T>::value ? T(0) : T(1);
    dl[5 * L + e] = T(1);
#pragma unroll
    for (int k = 1; k < 5; ++k) dl[k * L + e] = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k) du[k * L + e] = T(0);
    return false;
}

// a zero or non-finite pivot of the factorization: st->err_pivot
__device__ __forceinline__ void flag_pivot(bool bad, SolveState *st)
{
    if (__any(bad) && bad && st)
        __hip_atomic_store(&st->err_pivot, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename TI, typename TO>
__global__ __launch_bounds__(256) void k_to_skew(const TI *__restrict__ src, TO *__restrict__ dst,
                                                 int nk, Skew sk, int unit)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e < sk.size()) to_skew_entry<TI, TO>(src, dst, nk, sk, e, unit);
}

// 5 diagonals (S, W, P, E, N), reference layout -> the 9 skewed arrays of
// the zero-corner embedding: slot k9 takes 5-diagonal slot from5[k9] or 0.
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void k_to_skew5(const TI *__restrict__ src, TO *__restrict__ dst,
                                                  Skew sk)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= sk.size()) return;
    constexpr int from5[9] = {-1, k5S, -1, k5W, k5P, k5E, -1, k5N, -1};
    const long L = sk.size();
    const SkewPoint q = skew_point(sk, e);
    const long N = (long)sk.nx * sk.ny;
    const long p = q.real ? (long)q.j * sk.nx + q.i : 0;
#pragma unroll
    for (int k = 0; k < 9; ++k)
        dst[k * L + e] = q.real ? (from5[k] >= 0 ? (TO)src[from5[k] * N + p] : TO(0))
                                : (k == kP ? TO(1) : TO(0));
}

template <typename T>
__global__ __launch_bounds__(256) void k_prepare_factor(T *__restrict__ dl, T *__restrict__ du,
                                                        Skew sk, SolveState *st)
{
    if (st && st->skip) return;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    flag_pivot(e < sk.size() && prepare_entry<T>(dl, du, sk, e), st);
}

// Batched: workgroup g covers skewed entries map[g].y*256 .. of system
// map[g].x.  A (reference layout, precision T) -> its skewed copy.
template <typename T>
__global__ __launch_bounds__(256) void k_to_skew_batch(const BatchSys<T> *__restrict__ sys,
                                                       const int2 *__restrict__ map)
{
    const int2 m = map[blockIdx.x];
    const BatchSys<T> &d = sys[m.x];
    const long e = (long)m.y * 256 + threadIdx.x;
    if (e < d.sk.size()) to_skew_entry<T, T>(d.ra, d.a, 9, d.sk, e, kP);
}

template <typename T>
__global__ __launch_bounds__(256) void k_prepare_batch(const BatchSys<T> *__restrict__ sys,
                                                       const int2 *__restrict__ map)
{
    const int2 m = map[blockIdx.x];
    const BatchSys<T> &d = sys[m.x];
    const long e = (long)m.y * 256 + threadIdx.x;
    flag_pivot(e < d.sk.size() && prepare_entry<T>(d.dl, d.du, d.sk, e), d.st);
}

// skewed -> reference layout, one thread per point.
template <typename T>
__global__ __launch_bounds__(256) void k_from_skew(const T *__restrict__ src, T *__restrict__ dst,
                                                   int nk, Skew sk)
{
    const long N = (long)sk.nx * sk.ny;
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p >= N) return;
    const int j = (int)(p / sk.nx), i = (int)(p - (long)j * sk.nx);
    const long e = sk.of_point(i, j);
    const long L = sk.size();
    for (int k = 0; k < nk; ++k) dst[k * N + p] = src[k * L + e];
}

// Clear the boundary words (generations restart with every solve).
__global__ void k_clear_words(u64 *p, long n)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = 0ull;
}

// ---- launchers -------------------------------------------------------------
static inline unsigned blocks_for(long n) { return (unsigned)((n + 255) / 256); }

long boundary_words(int nx, int ny) { return (long)((ny + 63) / 64) * nx * 8; }
long scratch_words(int ny) { return (long)((ny + 63) / 64) * 64 * 8; }

void launch_clear_boundary(u64 *bnd, int nx, int ny, hipStream_t s)
{
    const long n = boundary_words(nx, ny);
    hipLaunchKernelGGL(k_clear_words, dim3(blocks_for(n)), dim3(256), 0, s, bnd, n);
}

template <typename T>
constexpr int sweep_steps()
{
    return 64 / (int)sizeof(T);  // 64 bytes per lane and array per chunk
}
template <typename T>
constexpr int factor_steps()
{
    return 32 / (int)sizeof(T);
}

Skew skew_for(int nx, int ny, int value_bytes) { return Skew::make(nx, ny, 64 / value_bytes); }

template <typename T>
void launch_factor(const T *a_sk, T *dl, T *du, T alpha0, const Skew &sk, Link lk, int xcd_stride,
                   hipStream_t s, bool five)
{
    constexpr int F = factor_steps<T>();
    const dim3 g(sk.nb * xcd_stride), b(2 * kWave);  // computing and storing wave
    if (five)
        hipLaunchKernelGGL((k_factor<T, F, true>), g, b, 0, s, a_sk, dl, du, alpha0, sk, lk,
                           xcd_stride);
    else
        hipLaunchKernelGGL((k_factor<T, F, false>), g, b, 0, s, a_sk, dl, du, alpha0, sk, lk,
                           xcd_stride);
}

template <typename T>
void launch_forward_staged(const T *r, const T *dl, T *w, const Skew &sk, Link lk,
                           int xcd_stride, hipStream_t s, bool five)
{
    constexpr int F = sweep_steps<T>();
    const dim3 g(sk.nb * xcd_stride), b(2 * kWave);
    if (five)
        hipLaunchKernelGGL((k_forward_staged<T, F, true>), g, b, 0, s, r, dl, w, sk, lk, xcd_stride);
    else
        hipLaunchKernelGGL((k_forward_staged<T, F, false>), g, b, 0, s, r, dl, w, sk, lk,
                           xcd_stride);
}

template <typename T>
void launch_backward_staged(T *wz, const T *du, const Skew &sk, Link lk, int xcd_stride,
                            hipStream_t s, bool five)
{
    constexpr int F = sweep_steps<T>();
    const dim3 g(sk.nb * xcd_stride), b(2 * kWave);
    if (five)
        hipLaunchKernelGGL((k_backward_staged<T, F, true>), g, b, 0, s, wz, du, sk, lk, xcd_stride);
    else
        hipLaunchKernelGGL((k_backward_staged<T, F, false>), g, b, 0, s, wz, du, sk, lk,
                           xcd_stride);
}

template <typename T>
void launch_sweeps(const T *r, const T *dl, T *w, const T *du, const Skew &sk, Link lk,
                   int xcd_stride, hipStream_t s, bool five)
{
    constexpr int F = sweep_steps<T>();
    const dim3 g(sk.nb * xcd_stride), b(2 * kWave);
    if (five)
        hipLaunchKernelGGL((k_sweeps<T, F, true>), g, b, 0, s, r, dl, w, du, sk, lk, xcd_stride);
    else
        hipLaunchKernelGGL((k_sweeps<T, F, false>), g, b, 0, s, r, dl, w, du, sk, lk, xcd_stride);
}

template <typename T>
void launch_prepare_factor(T *dl, T *du, const Skew &sk, hipStream_t s, SolveState *st)
{
    hipLaunchKernelGGL(k_prepare_factor<T>, dim3(blocks_for(sk.size())), dim3(256), 0, s, dl, du,
                       sk, st);
}

template <typename T>
void launch_factor_batch(const BatchSys<T> *sys, const int2 *map, int nblk, hipStream_t s)
{
    hipLaunchKernelGGL((k_factor_batch<T, factor_steps<T>()>), dim3(nblk), dim3(2 * kWave), 0,
                       s, sys, map);
}
template <typename T>
void launch_forward_batch(const BatchSys<T> *sys, const int2 *map, int nblk, hipStream_t s,
                          bool staged, bool five)
{
    constexpr int F = sweep_steps<T>();
    const dim3 g(nblk), b(2 * kWave);
    if (staged && five)
        hipLaunchKernelGGL((k_forward_batch_staged<T, F, true>), g, b, 0, s, sys, map);
    else if (staged)
        hipLaunchKernelGGL((k_forward_batch_staged<T, F, false>), g, b, 0, s, sys, map);
    else if (five)
        hipLaunchKernelGGL((k_forward_batch<T, F, true>), g, b, 0, s, sys, map);
    else
        hipLaunchKernelGGL((k_forward_batch<T, F, false>), g, b, 0, s, sys, map);
}
template <typename T>
void launch_backward_batch(const BatchSys<T> *sys, const int2 *map, int nblk, hipStream_t s,
                           bool staged, bool five)
{
    constexpr int F = sweep_steps<T>();
    const dim3 g(nblk), b(2 * kWave);
    if (staged && five)
        hipLaunchKernelGGL((k_backward_batch_staged<T, F, true>), g, b, 0, s, sys, map);
    else if (staged)
        hipLaunchKernelGGL((k_backward_batch_staged<T, F, false>), g, b, 0, s, sys, map);
    else if (five)
        hipLaunchKernelGGL((k_backward_batch<T, F, true>), g, b, 0, s, sys, map);
    else
        hipLaunchKernelGGL((k_backward_batch<T, F, false>), g, b, 0, s, sys, map);
}
template <typename T>
void launch_to_skew_batch(const BatchSys<T> *sys, const int2 *map, int nblk, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_skew_batch<T>, dim3(nblk), dim3(256), 0, s, sys, map);
}
template <typename T>
void launch_prepare_batch(const BatchSys<T> *sys, const int2 *map, int nblk, hipStream_t s)
{
    hipLaunchKernelGGL(k_prepare_batch<T>, dim3(nblk), dim3(256), 0, s, sys, map);
}

template <typename TI, typename TO>
void launch_to_skew(const TI *src, TO *dst, int nk, const Skew &sk, hipStream_t s, int unit)
{
    hipLaunchKernelGGL((k_to_skew<TI, TO>), dim3(blocks_for(sk.size())), dim3(256), 0, s, src, dst,
                       nk, sk, unit);
}

template <typename TI, typename TO>
void launch_to_skew5(const TI *src, TO *dst, const Skew &sk, hipStream_t s)
{
    hipLaunchKernelGGL((k_to_skew5<TI, TO>), dim3(blocks_for(sk.size())), dim3(256), 0, s, src, dst,
                       sk);
}

template <typename T>
void launch_from_skew(const T *src, T *dst, int nk, const Skew &sk, hipStream_t s)
{
    hipLaunchKernelGGL(k_from_skew<T>, dim3(blocks_for((long)sk.nx * sk.ny)), dim3(256), 0, s, src,
                       dst, nk, sk);
}


#define M2D_INST(T)                                                                           \
    template void launch_factor<T>(const T *, T *, T *, T, const Skew &, Link, int, hipStream_t, \
                                   bool);                                                     \
    template void launch_forward_staged<T>(const T *, const T *, T *, const Skew &, Link, int,   \
                                           hipStream_t, bool);                                 \
    template void launch_backward_staged<T>(T *, const T *, const Skew &, Link, int,            \
                                            hipStream_t, bool);                                \
    template void launch_sweeps<T>(const T *, const T *, T *, const T *, const Skew &, Link,    \
                                   int, hipStream_t, bool);                                    \
    template void launch_from_skew<T>(const T *, T *, int, const Skew &, hipStream_t);          \
    template void launch_prepare_factor<T>(T *, T *, const Skew &, hipStream_t, SolveState *);  \
    template void launch_factor_batch<T>(const BatchSys<T> *, const int2 *, int, hipStream_t);  \
    template void launch_forward_batch<T>(const BatchSys<T> *, const int2 *, int, hipStream_t,  \
                                          bool, bool);                                        \
    template void launch_backward_batch<T>(const BatchSys<T> *, const int2 *, int, hipStream_t, \
                                           bool, bool);                                       \
    template void launch_to_skew_batch<T>(const BatchSys<T> *, const int2 *, int, hipStream_t); \
    template void launch_prepare_batch<T>(const BatchSys<T> *, const int2 *, int, hipStream_t); \

M2D_INST(double)
M2D_INST(float)
template void launch_to_skew<double, double>(const double *, double *, int, const Skew &,
                                             hipStream_t, int);
template void launch_to_skew<double, float>(const double *, float *, int, const Skew &,
                                            hipStream_t, int);
template void launch_to_skew5<double, double>(const double *, double *, const Skew &, hipStream_t);
template void launch_to_skew5<double, float>(const double *, float *, const Skew &, hipStream_t);

}  // namespace m2d
