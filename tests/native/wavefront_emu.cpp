// wavefront_emu.cpp -- CPU emulation of the gfx950 wavefront kernels
// (mach-2d_amd/csrc/msi9.hip: k_factor, k_forward, k_backward).
//
// Test infrastructure.  Each band (one wave of 64 lanes) is a state machine
// stepping through exactly the kernel's loop: the one-lane shuffle of the
// previous step's result, the history shift, the boundary-row read of the
// first/last lane, the per-point formula (shared header msi_stencil.h) and
// the progress publication.  A scheduler interleaves bands in a seeded
// random order and lets a band take a step only when the wait condition the
// kernel would spin on holds, so the emulation checks that
//   * every boundary-row value a band reads was covered by a publication of
//     its producer band (else: abort with a message),
//   * the protocol cannot deadlock (some band can always move),
//   * every array index stays inside its buffer,
// and produces the same numbers the kernels compute, for a bitwise
// comparison with the oracle.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <vector>

#include "msi_stencil.h"

using namespace m2d;

namespace {

constexpr int W64 = 64;

struct Fail {
    char msg[256];
};
Fail g_fail;
bool g_failed = false;

void fail(const char *m, long a = 0, long b = 0)
{
    if (!g_failed) snprintf(g_fail.msg, sizeof(g_fail.msg), "%s (%ld, %ld)", m, a, b);
    g_failed = true;
}

template <typename T>
struct Buf {
    std::vector<T> v;
    const char *name;
    Buf(size_t n, const char *nm) : v(n, T(0)), name(nm) {}
    T &at(long k)
    {
        if (k < 0 || k >= (long)v.size()) {
            fail(name, k, (long)v.size());
            static T dummy;
            return dummy;
        }
        return v[(size_t)k];
    }
};

// Progress publications: progress[band] = columns published.
struct Sync {
    std::vector<int> progress;
    int K;
};

// ---- factor ----------------------------------------------------------------
struct FactorBand {
    int band, nb, rows, tend, t = 0;
    double cur[W64][4], se[W64][4], s[W64][4], sw[W64][4];
    FactorBand(int b, int nb_, int nx, int ny) : band(b), nb(nb_)
    {
        rows = std::min(W64, ny - b * W64);
        tend = nx - 1 + 2 * (rows - 1);
        memset(cur, 0, sizeof(cur));
        memset(se, 0, sizeof(se));
        memset(s, 0, sizeof(s));
        memset(sw, 0, sizeof(sw));
    }
    bool done() const { return t > tend; }
    bool can_step(const Sync &sy, int nx) const
    {
        if (band == 0) return true;
        return sy.progress[band - 1] >= std::min(t + 2, nx);
    }
    void step(Buf<double> &a, Buf<double> &dl, Buf<double> &du, int nx, int ny, double alpha0,
              Sync &sy)
    {
        const long N = (long)nx * ny;
        if (t == 0 && band > 0) {  // the kernel primes before its loop
            if (sy.progress[band - 1] < 1) fail("factor: unpublished prime", band, 0);
            for (int k = 0; k < 4; ++k) se[0][k] = du.at(k * N + (long)(band * W64) * nx - nx);
        }
        double up[W64][4];
        for (int l = 0; l < W64; ++l)
            for (int k = 0; k < 4; ++k) up[l][k] = (l == 0) ? cur[0][k] : cur[l - 1][k];
        for (int l = 0; l < W64; ++l)
            for (int k = 0; k < 4; ++k) {
                sw[l][k] = s[l][k];
                s[l][k] = se[l][k];
                se[l][k] = up[l][k];
            }
        {
            const int j = band * W64;
            const bool ok = band > 0 && t + 1 < nx;
            for (int k = 0; k < 4; ++k) {
                if (ok) {
                    if (sy.progress[band - 1] < t + 2) fail("factor: unpublished read", band, t);
                    se[0][k] = du.at(k * N + (long)j * nx - nx + t + 1);
                } else {
                    se[0][k] = 0.0;
                }
            }
        }
        for (int l = 0; l < rows; ++l) {
            const int i = t - 2 * l, j = band * W64 + l;
            if (i < 0 || i >= nx) continue;
            const long p = (long)j * nx + i;
            double b[9], dw[4], ds[4], dse[4], dsw[4], L[5], U[4];
            for (int k = 0; k < 9; ++k) b[k] = a.at(k * N + p);
            const bool hw = i > 0, hs = j > 0, he = i + 1 < nx;
            for (int k = 0; k < 4; ++k) {
                dw[k] = hw ? cur[l][k] : 0.0;
                ds[k] = hs ? s[l][k] : 0.0;
                dse[k] = (hs && he) ? se[l][k] : 0.0;
                dsw[k] = (hs && hw) ? sw[l][k] : 0.0;
            }
            factor_point<double>(b, alpha_at<double>(i, j, nx, ny, alpha0), dw, ds, dse, dsw, L,
                                 U);
            for (int k = 0; k < 5; ++k) dl.at(k * N + p) = L[k];
            for (int k = 0; k < 4; ++k) {
                du.at(k * N + p) = U[k];
                cur[l][k] = U[k];
            }
            if (l == W64 - 1 && band < nb - 1 && ((i + 1) % sy.K == 0 || i == nx - 1))
                sy.progress[band] = i + 1;
        }
        ++t;
    }
};

// ---- forward ---------------------------------------------------------------
struct FwdBand {
    int band, nb, rows, tend, t = 0;
    double cur[W64], se[W64], s[W64], sw[W64];
    FwdBand(int b, int nb_, int nx, int ny) : band(b), nb(nb_)
    {
        rows = std::min(W64, ny - b * W64);
        tend = nx - 1 + 2 * (rows - 1);
        for (int l = 0; l < W64; ++l) cur[l] = se[l] = s[l] = sw[l] = 0.0;
    }
    bool done() const { return t > tend; }
    bool can_step(const Sync &sy, int nx) const
    {
        return band == 0 || sy.progress[band - 1] >= std::min(t + 2, nx);
    }
    void step(Buf<double> &r, Buf<double> &dl, Buf<double> &w, int nx, int ny, Sync &sy)
    {
        const long N = (long)nx * ny;
        if (t == 0 && band > 0) {
            if (sy.progress[band - 1] < 1) fail("forward: unpublished prime", band, 0);
            se[0] = w.at((long)(band * W64) * nx - nx);
        }
        double up[W64];
        for (int l = 0; l < W64; ++l) up[l] = (l == 0) ? cur[0] : cur[l - 1];
        for (int l = 0; l < W64; ++l) {
            sw[l] = s[l];
            s[l] = se[l];
            se[l] = up[l];
        }
        if (band > 0 && t + 1 < nx) {
            if (sy.progress[band - 1] < t + 2) fail("forward: unpublished read", band, t);
            se[0] = w.at((long)(band * W64) * nx - nx + t + 1);
        } else {
            se[0] = 0.0;
        }
        for (int l = 0; l < rows; ++l) {
            const int i = t - 2 * l, j = band * W64 + l;
            if (i < 0 || i >= nx) continue;
            const long p = (long)j * nx + i;
            const bool hw = i > 0, hs = j > 0, he = i + 1 < nx;
            const double v = forward_point<double>(
                r.at(p), dl.at(p), dl.at(N + p), dl.at(2 * N + p), dl.at(3 * N + p),
                dl.at(4 * N + p), hw ? cur[l] : 0.0, (hs && he) ? se[l] : 0.0, hs ? s[l] : 0.0,
                (hs && hw) ? sw[l] : 0.0);
            w.at(p) = v;
            cur[l] = v;
            if (l == W64 - 1 && band < nb - 1 && ((i + 1) % sy.K == 0 || i == nx - 1))
                sy.progress[band] = i + 1;
        }
        ++t;
    }
};

// ---- backward --------------------------------------------------------------
struct BwdBand {
    int band, nb, rows, last, send, s_ = 0;
    double cur[W64], nw[W64], n[W64], ne[W64];
    BwdBand(int b, int nb_, int nx, int ny) : band(b), nb(nb_)
    {
        rows = std::min(W64, ny - b * W64);
        last = rows - 1;
        send = nx - 1 + 2 * last;
        for (int l = 0; l < W64; ++l) cur[l] = nw[l] = n[l] = ne[l] = 0.0;
    }
    bool done() const { return s_ > send; }
    bool can_step(const Sync &sy, int nx) const
    {
        return band == nb - 1 || sy.progress[band + 1] >= std::min(s_ + 2, nx);
    }
    void step(Buf<double> &wz, Buf<double> &du, Buf<double> &x, int nx, int ny, Sync &sy)
    {
        const long N = (long)nx * ny;
        const int s = s_;
        if (s == 0 && band < nb - 1) {
            if (sy.progress[band + 1] < 1) fail("backward: unpublished prime", band, 0);
            nw[last] = wz.at((long)(band * W64 + last) * nx + nx + (nx - 1));
        }
        double dn[W64];
        for (int l = 0; l < W64; ++l) dn[l] = (l == W64 - 1) ? cur[W64 - 1] : cur[l + 1];
        for (int l = 0; l < W64; ++l) {
            ne[l] = n[l];
            n[l] = nw[l];
            nw[l] = dn[l];
        }
        if (band < nb - 1 && nx - 2 - s >= 0) {
            if (sy.progress[band + 1] < s + 2) fail("backward: unpublished read", band, s);
            const long rowbase = (long)(band * W64 + last) * nx;
            nw[last] = wz.at(rowbase + nx + (nx - 2 - s));
        } else {
            nw[last] = 0.0;
        }
        for (int l = 0; l < rows; ++l) {
            const int i = nx - 1 - s + 2 * (last - l), j = band * W64 + l;
            if (i < 0 || i >= nx) continue;
            const long p = (long)j * nx + i;
            const bool hn = j + 1 < ny, hw = i > 0, he = i + 1 < nx;
            const double v = backward_point<double>(
                wz.at(p), du.at(p), du.at(N + p), du.at(2 * N + p), du.at(3 * N + p),
                he ? cur[l] : 0.0, (hn && hw) ? nw[l] : 0.0, hn ? n[l] : 0.0,
                (hn && he) ? ne[l] : 0.0);
            wz.at(p) = v;
            x.at(p) = v + x.at(p);
            cur[l] = v;
            if (l == 0 && band > 0 && ((nx - i) % sy.K == 0 || i == 0))
                sy.progress[band] = nx - i;
        }
        ++s_;
    }
};

template <typename Band, typename StepFn>
bool schedule(std::vector<Band> &bands, Sync &sy, int nx, unsigned seed, StepFn fn)
{
    std::mt19937 rng(seed);
    for (;;) {
        bool all_done = true, moved = false;
        std::vector<int> order(bands.size());
        for (size_t q = 0; q < order.size(); ++q) order[q] = (int)q;
        std::shuffle(order.begin(), order.end(), rng);
        for (int q : order) {
            Band &b = bands[(size_t)q];
            if (b.done()) continue;
            all_done = false;
            const int burst = 1 + (int)(rng() % 7);
            for (int m = 0; m < burst && !b.done() && b.can_step(sy, nx); ++m) {
                fn(b);
                moved = true;
                if (g_failed) return false;
            }
        }
        if (all_done) return true;
        if (!moved) {
            fail("deadlock: no band can move");
            return false;
        }
    }
}

}  // namespace

extern "C" {

const char *emu_error(void) { return g_failed ? g_fail.msg : ""; }

// lu2d9 through the emulated wavefront kernel.  Returns 0 on success.
int emu_lu2d9(const double *a_in, double *dl_out, double *du_out, int nx, int ny, double alpha0,
              int K, unsigned seed)
{
    g_failed = false;
    const long N = (long)nx * ny;
    Buf<double> a(9 * N, "a"), dl(5 * N, "dl"), du(4 * N, "du");
    memcpy(a.v.data(), a_in, sizeof(double) * 9 * N);
    const int nb = (ny + W64 - 1) / W64;
    std::vector<FactorBand> bands;
    for (int b = 0; b < nb; ++b) bands.emplace_back(b, nb, nx, ny);
    Sync sy{std::vector<int>(nb, 0), K};
    schedule(bands, sy, nx, seed, [&](FactorBand &b) { b.step(a, dl, du, nx, ny, alpha0, sy); });
    memcpy(dl_out, dl.v.data(), sizeof(double) * 5 * N);
    memcpy(du_out, du.v.data(), sizeof(double) * 4 * N);
    return g_failed ? 1 : 0;
}

// One MSI iteration (forward + backward + update) through the emulated
// kernels, given r.  w_out receives z (in place over w, as the kernel).
int emu_msi_iteration(const double *r_in, const double *dl_in, const double *du_in,
                      double *x_io, double *w_out, int nx, int ny, int K, unsigned seed)
{
    g_failed = false;
    const long N = (long)nx * ny;
    Buf<double> r(N, "r"), dl(5 * N, "dl"), du(4 * N, "du"), w(N, "w"), x(N, "x");
    memcpy(r.v.data(), r_in, sizeof(double) * N);
    memcpy(dl.v.data(), dl_in, sizeof(double) * 5 * N);
    memcpy(du.v.data(), du_in, sizeof(double) * 4 * N);
    memcpy(x.v.data(), x_io, sizeof(double) * N);
    const int nb = (ny + W64 - 1) / W64;
    {
        std::vector<FwdBand> bands;
        for (int b = 0; b < nb; ++b) bands.emplace_back(b, nb, nx, ny);
        Sync sy{std::vector<int>(nb, 0), K};
        schedule(bands, sy, nx, seed, [&](FwdBand &b) { b.step(r, dl, w, nx, ny, sy); });
    }
    if (!g_failed) {
        std::vector<BwdBand> bands;
        for (int b = 0; b < nb; ++b) bands.emplace_back(b, nb, nx, ny);
        Sync sy{std::vector<int>(nb, 0), K};
        schedule(bands, sy, nx, seed + 1, [&](BwdBand &b) { b.step(w, du, x, nx, ny, sy); });
    }
    memcpy(x_io, x.v.data(), sizeof(double) * N);
    memcpy(w_out, w.v.data(), sizeof(double) * N);
    return g_failed ? 1 : 0;
}

// Residual of the 2-D aware kernel formula.
void emu_residual(const double *a, const double *x, const double *rhs, double *r, int nx, int ny)
{
    const long N = (long)nx * ny;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const long p = (long)j * nx + i;
            r[p] = residual_point<double, double, double>(a, N, x, rhs[p], i, j, nx, ny);
        }
}

}  // extern "C"
