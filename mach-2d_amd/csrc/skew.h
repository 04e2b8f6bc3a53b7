// skew.h -- the wavefront ("skewed") storage layout of the solver's own
// arrays (factor dl/du, residual r, sweep vectors w/z, a copy of A).
//
// A wave of 64 lanes owns a band of 64 grid rows; lane l (row j = 64b + l)
// works on column i at stored step u = i + 2l + pad(b).  Steps are grouped in
// chunks of S; within a chunk the S consecutive steps of one lane are
// contiguous, lanes follow each other:
//
//     index(b, l, u) = ((b * nch + u / S) * 64 + l) * S + u % S
//
// so a lane reads or writes the S values of a chunk as S*sizeof(T) contiguous
// bytes (16-byte vector accesses), and one chunk of one array is a
// contiguous 64*S*sizeof(T) block.  Every band has nst = roundup(nx+126, S)
// stored steps; the real steps of band b (nx + 2(rows-1) of them) are placed
// at the end (front padding pad(b)), so the backward sweep, which visits the
// band's steps in exactly reverse order, walks the same chunks backwards.
// Padding entries (steps whose column is outside 0..nx-1, lanes beyond the
// last row) are read and written like real ones; their values only ever
// meet zero couplings.
#pragma once

#if defined(__HIPCC__)
#define SKEW_HD __host__ __device__ __forceinline__
#else
#define SKEW_HD inline
#endif

namespace m2d {

struct Skew {
    int nx, ny, nb, S, nst, nch;

    static Skew make(int nx, int ny, int S)
    {
        Skew k;
        k.nx = nx;
        k.ny = ny;
        k.nb = (ny + 63) / 64;
        k.S = S;
        k.nst = ((nx + 126 + S - 1) / S) * S;
        k.nch = k.nst / S;
        return k;
    }
    SKEW_HD int rows(int b) const { return ny - 64 * b < 64 ? ny - 64 * b : 64; }
    SKEW_HD int pad(int b) const { return nst - (nx + 2 * (rows(b) - 1)); }
    SKEW_HD long index(int b, int l, int u) const
    {
        return (((long)b * nch + u / S) * 64 + l) * S + u % S;
    }
    // index for a chunk length known at compile time (S_ == S)
    template <int S_>
    SKEW_HD long index_c(int b, int l, int u) const
    {
        return (((long)b * nch + (long)((unsigned)u / S_)) * 64 + l) * S_ + (long)((unsigned)u % S_);
    }
    SKEW_HD long chunk_base(int b, int c, int l) const { return (((long)b * nch + c) * 64 + l) * S; }
    SKEW_HD long of_point(int i, int j) const
    {
        const int b = j >> 6, l = j & 63;
        return index(b, l, i + 2 * l + pad(b));
    }
    // of_point for a chunk length known at compile time (S_ == S): shifts and
    // masks instead of two integer divisions
    template <int S_>
    SKEW_HD long of_point_c(int i, int j) const
    {
        const int b = j >> 6, l = j & 63;
        const unsigned u = (unsigned)(i + 2 * l + pad(b));
        return (((long)b * nch + (long)(u / S_)) * 64 + l) * S_ + (long)(u % S_);
    }
    SKEW_HD long size() const { return (long)nb * nch * 64 * S; }
};

}  // namespace m2d
