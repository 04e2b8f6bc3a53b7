// wavefront_emu.cpp -- CPU emulation of the gfx950 wavefront kernels
// (mach-2d_amd/csrc/sweep.hip: k_factor, k_forward, k_backward, k_to_skew,
// k_from_skew, k_update).
//
// Test infrastructure.  The solver's own arrays live in the skewed layout of
// skew.h (same header as the kernels).  Each band (one wave of 64 lanes)
// runs the kernel's chunk loop: per-lane vector loads of F consecutive
// stored steps, F steps with the one-lane shift of the previous result and
// two registers of history, the neighbouring band's boundary values held by
// lanes 0..F (step s uses lane s+1), per-lane stores of the F results, and
// the generation-tagged publication of the edge lane's F results at the end
// of the chunk.  A scheduler interleaves bands chunk by chunk in a seeded
// random order; a band may start a chunk only when every boundary word it
// needs carries the current generation (the kernel's bounded wait), so the
// emulation checks that
//   * no value is consumed before its producer published it,
//   * the protocol cannot deadlock (some band can always move),
//   * every skewed index, boundary word and scratch word stays in range,
// and yields the numbers the kernels compute, for a bitwise comparison with
// the oracle.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <vector>

#include "msi_stencil.h"
#include "skew.h"

using namespace m2d;

namespace {

constexpr int W64 = 64;
constexpr int FS = 8;  // sweep steps per chunk (FP64)
constexpr int FF = 4;  // factor steps per chunk (FP64)

char g_msg[256];
bool g_failed = false;

void fail(const char *m, long a = 0, long b = 0)
{
    if (!g_failed) snprintf(g_msg, sizeof(g_msg), "%s (%ld, %ld)", m, a, b);
    g_failed = true;
}

struct Buf {
    std::vector<double> v;
    const char *name;
    Buf(size_t n, const char *nm) : v(n, 0.0), name(nm) {}
    double &at(long k)
    {
        static double dummy;
        if (k < 0 || k >= (long)v.size()) {
            fail(name, k, (long)v.size());
            return dummy;
        }
        return v[(size_t)k];
    }
};

// Boundary words (value + generation) per (slot, column, k) plus per-lane
// scratch, as the kernels' buffer.
struct Bnd {
    int nx, per, nb;
    std::vector<double> val;
    std::vector<unsigned> gen;
    long scratch_writes = 0;
    Bnd(int nb_, int nx_, int per_)
        : nx(nx_), per(per_), nb(nb_), val((size_t)nb_ * nx_ * per_, 0.0),
          gen((size_t)nb_ * nx_ * per_, 0u) {}
    long idx(int slot, int col, int k) const
    {
        if (slot < 0 || slot >= nb || col < 0 || col >= nx || k < 0 || k >= per) return -1;
        return ((long)slot * nx + col) * per + k;
    }
    void put(int slot, int col, int k, unsigned g, double v)
    {
        const long q = idx(slot, col, k);
        if (q < 0) return fail("boundary put out of range", slot, col);
        val[(size_t)q] = v;
        gen[(size_t)q] = g;
    }
    void put_scratch(int band, int lane)
    {
        if (band < 0 || band >= nb || lane < 0 || lane >= W64) fail("scratch out of range");
        ++scratch_writes;
    }
    bool ok(int slot, int col, int k, unsigned g) const
    {
        const long q = idx(slot, col, k);
        return q >= 0 && gen[(size_t)q] == g;
    }
    double get_clamped(int slot, int col, int k) const
    {
        col = std::min(std::max(col, 0), nx - 1);
        const long q = idx(slot, col, k);
        if (q < 0) {
            fail("boundary get out of range", slot, col);
            return 0.0;
        }
        return val[(size_t)q];
    }
};

// ---- forward ---------------------------------------------------------------
struct Fwd {
    const Skew &sk;
    int band, pad, nf, c = 0, phase = -1;
    bool has_up, has_down;
    double cur[W64], se[W64], sv[W64], sw[W64], bv[FS + 1], outv[W64][FS];
    Fwd(const Skew &k, int b) : sk(k), band(b)
    {
        pad = sk.pad(b);
        nf = sk.nst / FS;
        has_up = b > 0;
        has_down = b < sk.nb - 1;
        for (int l = 0; l < W64; ++l) cur[l] = se[l] = sv[l] = sw[l] = 0.0;
    }
    bool done() const { return c >= nf; }
    int bcol(int cc, int lane) const { return cc * FS - pad + lane; }
    bool bneed(int cc, int lane) const
    {
        const int col = bcol(cc, lane);
        return has_up && lane <= FS && col >= 0 && col < sk.nx;
    }
    bool valid(const Bnd &bd, unsigned g, int cc) const
    {
        if (cc >= nf) return true;
        for (int lane = 0; lane <= FS; ++lane)
            if (bneed(cc, lane) && !bd.ok(band - 1, bcol(cc, lane), 0, g)) return false;
        return true;
    }
    // phase -1: the prime waits for chunk 0's boundary; phase 0: loads and
    // steps of chunk c; phase 1: waits for chunk c+1's boundary, then
    // takes it, stores and publishes chunk c.
    bool can_run(const Bnd &bd, unsigned g) const
    {
        return phase == 0 || valid(bd, g, phase < 0 ? 0 : c + 1);
    }
    // columns outside the grid, and every column of the bottom band, read 0
    void take(const Bnd &bd, int cc)
    {
        for (int q = 0; q <= FS; ++q) {
            const double v = bd.get_clamped(has_up ? band - 1 : 0, bcol(cc, q), 0);
            bv[q] = bneed(cc, q) ? v : 0.0;
        }
    }
    void run(Buf &r, Buf &dl, Buf &w, Bnd &bd, unsigned g)
    {
        const long L = sk.size();
        if (phase < 0) {
            take(bd, 0);
            se[0] = bv[0];
            phase = 0;
            return;
        }
        if (phase == 1) return finish(w, bd, g);
        double cb[W64][7][FS];  // r, dl1..dl5, 1/dl1
        for (int l = 0; l < W64; ++l) {
            const long e = sk.chunk_base(band, std::min(c, nf - 1), l);
            for (int s = 0; s < FS; ++s) {
                cb[l][0][s] = r.at(e + s);
                for (int k = 0; k < 6; ++k) cb[l][1 + k][s] = dl.at(k * L + e + s);
            }
        }
        for (int s = 0; s < FS; ++s) {
            double up[W64];  // one-lane shift; lane 0 keeps the boundary value
            for (int l = 0; l < W64; ++l) up[l] = l == 0 ? bv[s + 1] : cur[l - 1];
            for (int l = 0; l < W64; ++l) {
                sw[l] = sv[l];
                sv[l] = se[l];
                se[l] = up[l];
            }
            for (int l = 0; l < W64; ++l) {  // no masks: padding values are finite
                const double v = forward_point_y<double>(
                    cb[l][0][s], cb[l][1][s], cb[l][6][s], cb[l][2][s], cb[l][3][s],
                    cb[l][4][s], cb[l][5][s], cur[l], se[l], sv[l], sw[l]);
                outv[l][s] = v;
                cur[l] = v;
            }
        }
        phase = 1;
    }
    void finish(Buf &w, Bnd &bd, unsigned g)
    {
        if (c + 1 < nf) take(bd, c + 1);  // chunk c+1's boundary, before storing chunk c
        const int cc = c++;
        phase = 0;
        for (int l = 0; l < W64; ++l) {
            const long e = sk.chunk_base(band, cc, l);
            for (int s = 0; s < FS; ++s) w.at(e + s) = outv[l][s];
        }
        for (int lane = 0; lane < W64; ++lane) {
            const int col = cc * FS + lane - pad - 2 * (W64 - 1);
            const bool ok = has_down && lane < FS && col >= 0 && col < sk.nx;
            if (ok)
                bd.put(band, col, 0, g, outv[W64 - 1][lane]);
            else
                bd.put_scratch(band, lane);
        }
    }
};

// ---- backward --------------------------------------------------------------
struct Bwd {
    const Skew &sk;
    int band, pad, nf, c = 0, phase = -1;
    bool has_above, has_below;
    double cur[W64], nw[W64], nv[W64], ne[W64], bv[FS + 1], outv[W64][FS];
    Bwd(const Skew &k, int b) : sk(k), band(b)
    {
        pad = sk.pad(b);
        nf = sk.nst / FS;
        has_above = b < sk.nb - 1;
        has_below = b > 0;
        for (int l = 0; l < W64; ++l) cur[l] = nw[l] = nv[l] = ne[l] = 0.0;
    }
    bool done() const { return c >= nf; }
    int bcol(int cc, int lane) const { return sk.nx - 1 - cc * FS - lane; }
    bool bneed(int cc, int lane) const
    {
        const int col = bcol(cc, lane);
        return has_above && lane <= FS && col >= 0 && col < sk.nx;
    }
    bool valid(const Bnd &bd, unsigned g, int cc) const
    {
        if (cc >= nf) return true;
        for (int lane = 0; lane <= FS; ++lane)
            if (bneed(cc, lane) && !bd.ok(band, bcol(cc, lane), 0, g)) return false;
        return true;
    }
    bool can_run(const Bnd &bd, unsigned g) const
    {
        return phase == 0 || valid(bd, g, phase < 0 ? 0 : c + 1);
    }
    void take(const Bnd &bd, int cc)
    {
        for (int q = 0; q <= FS; ++q) {
            const double v = bd.get_clamped(has_above ? band : 0, bcol(cc, q), 0);
            bv[q] = bneed(cc, q) ? v : 0.0;
        }
    }
    void run(Buf &wz, Buf &du, Bnd &bd, unsigned g)
    {
        const long L = sk.size();
        if (phase < 0) {
            take(bd, 0);
            nw[W64 - 1] = bv[0];
            phase = 0;
            return;
        }
        if (phase == 1) return finish(wz, bd, g);
        double cb[W64][5][FS];
        for (int l = 0; l < W64; ++l) {
            const long e = sk.chunk_base(band, std::max(nf - 1 - c, 0), l);
            for (int s = 0; s < FS; ++s) {
                cb[l][0][s] = wz.at(e + s);
                for (int k = 0; k < 4; ++k) cb[l][1 + k][s] = du.at(k * L + e + s);
            }
        }
        for (int s = 0; s < FS; ++s) {
            const int q = FS - 1 - s;
            double dn[W64];  // lane 63 keeps the boundary value; the top lane of a
                             // partial band reads the padding row above it
            for (int l = 0; l < W64; ++l) dn[l] = l == W64 - 1 ? bv[s + 1] : cur[l + 1];
            for (int l = 0; l < W64; ++l) {
                ne[l] = nv[l];
                nv[l] = nw[l];
                nw[l] = dn[l];
            }
            for (int l = 0; l < W64; ++l) {
                const double v = backward_point<double>(cb[l][0][q], cb[l][1][q], cb[l][2][q],
                                                        cb[l][3][q], cb[l][4][q], cur[l], nw[l],
                                                        nv[l], ne[l]);
                outv[l][q] = v;
                cur[l] = v;
            }
        }
        phase = 1;
    }
    void finish(Buf &wz, Bnd &bd, unsigned g)
    {
        const int sc = nf - 1 - c;
        if (c + 1 < nf) take(bd, c + 1);
        ++c;
        phase = 0;
        for (int l = 0; l < W64; ++l) {
            const long e = sk.chunk_base(band, sc, l);
            for (int s = 0; s < FS; ++s) wz.at(e + s) = outv[l][s];
        }
        for (int lane = 0; lane < W64; ++lane) {
            const int col = sc * FS + FS - 1 - lane - pad;
            const bool ok = has_below && lane < FS && col >= 0 && col < sk.nx;
            if (ok)
                bd.put(band - 1, col, 0, g, outv[0][FS - 1 - lane]);
            else
                bd.put_scratch(band, lane);
        }
    }
};

// ---- factor ----------------------------------------------------------------
struct Fac {
    const Skew &sk;
    int band, pad, nf, rows, c = 0, phase = -1;
    bool has_up, has_down;
    double cur[W64][4], se[W64][4], sv[W64][4], sw[W64][4], bv[4][FF + 1];
    double oL[W64][5][FF], oU[W64][4][FF];
    Fac(const Skew &k, int b) : sk(k), band(b)
    {
        pad = sk.pad(b);
        nf = sk.nst / FF;
        rows = sk.rows(b);
        has_up = b > 0;
        has_down = b < sk.nb - 1;
        memset(cur, 0, sizeof(cur));
        memset(se, 0, sizeof(se));
        memset(sv, 0, sizeof(sv));
        memset(sw, 0, sizeof(sw));
    }
    bool done() const { return c >= nf; }
    int bcol(int cc, int lane) const { return cc * FF - pad + lane; }
    bool bneed(int cc, int lane) const
    {
        const int col = bcol(cc, lane);
        return has_up && lane <= FF && col >= 0 && col < sk.nx;
    }
    bool valid(const Bnd &bd, unsigned g, int cc) const
    {
        if (cc >= nf) return true;
        for (int lane = 0; lane <= FF; ++lane)
            for (int k = 0; k < 4; ++k)
                if (bneed(cc, lane) && !bd.ok(band - 1, bcol(cc, lane), k, g)) return false;
        return true;
    }
    bool can_run(const Bnd &bd, unsigned g) const
    {
        return phase == 0 || valid(bd, g, phase < 0 ? 0 : c + 1);
    }
    void take(const Bnd &bd, int cc)
    {
        for (int k = 0; k < 4; ++k)
            for (int q = 0; q <= FF; ++q)
                bv[k][q] = bd.get_clamped(has_up ? band - 1 : 0, bcol(cc, q), k);
    }
    void run(Buf &a, Buf &dl, Buf &du, double alpha0, Bnd &bd, unsigned g)
    {
        const long L = sk.size();
        if (phase < 0) {
            take(bd, 0);
            for (int k = 0; k < 4; ++k) se[0][k] = has_up ? bv[k][0] : 0.0;
            phase = 0;
            return;
        }
        if (phase == 1) return finish(dl, du, bd, g);
        double cb[W64][9][FF];
        for (int l = 0; l < W64; ++l) {
            const long e = sk.index(band, l, std::min(c, nf - 1) * FF);
            for (int s = 0; s < FF; ++s)
                for (int k = 0; k < 9; ++k) cb[l][k][s] = a.at(k * L + e + s);
        }
        for (int s = 0; s < FF; ++s) {
            double up[W64][4];
            for (int l = 0; l < W64; ++l)
                for (int k = 0; k < 4; ++k) up[l][k] = l == 0 ? 0.0 : cur[l - 1][k];
            for (int l = 0; l < W64; ++l)
                for (int k = 0; k < 4; ++k) {
                    sw[l][k] = sv[l][k];
                    sv[l][k] = se[l][k];
                    se[l][k] = l == 0 ? (has_up ? bv[k][s + 1] : 0.0) : up[l][k];
                }
            for (int l = 0; l < W64; ++l) {
                const int j = band * W64 + l, i = c * FF + s - pad - 2 * l;
                const bool hs = j > 0, hw = i > 0, he = i + 1 < sk.nx;
                double b9[9], dw[4], ds[4], dse[4], dsw[4], Lv[5], Uv[4];
                for (int k = 0; k < 9; ++k) b9[k] = cb[l][k][s];
                for (int k = 0; k < 4; ++k) {
                    dw[k] = hw ? cur[l][k] : 0.0;
                    ds[k] = hs ? sv[l][k] : 0.0;
                    dse[k] = (hs && he) ? se[l][k] : 0.0;
                    dsw[k] = (hs && hw) ? sw[l][k] : 0.0;
                }
                factor_point<double>(b9, alpha_at<double>(i, j, sk.nx, sk.ny, alpha0), dw, ds,
                                     dse, dsw, Lv, Uv);
                for (int k = 0; k < 5; ++k) oL[l][k][s] = Lv[k];
                for (int k = 0; k < 4; ++k) {
                    oU[l][k][s] = Uv[k];
                    cur[l][k] = Uv[k];
                }
                (void)rows;
            }
        }
        phase = 1;
    }
    void finish(Buf &dl, Buf &du, Bnd &bd, unsigned g)
    {
        const long L = sk.size();
        if (c + 1 < nf) take(bd, c + 1);
        const int cc = c++;
        phase = 0;
        for (int l = 0; l < W64; ++l) {
            const long e = sk.index(band, l, cc * FF);
            for (int s = 0; s < FF; ++s) {
                for (int k = 0; k < 5; ++k) dl.at(k * L + e + s) = oL[l][k][s];
                for (int k = 0; k < 4; ++k) du.at(k * L + e + s) = oU[l][k][s];
            }
        }
        for (int lane = 0; lane < W64; ++lane) {
            const int col = cc * FF + lane - pad - 2 * (W64 - 1);
            const bool ok = has_down && lane < FF && col >= 0 && col < sk.nx;
            for (int k = 0; k < 4; ++k) {
                if (ok)
                    bd.put(band, col, k, g, oU[W64 - 1][k][lane]);
                else
                    bd.put_scratch(band, lane);
            }
        }
    }
};

template <typename Band, typename Run>
void schedule(std::vector<Band> &bands, const Bnd &bd, unsigned g, unsigned seed, Run run)
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
            const int burst = 1 + (int)(rng() % 5);
            for (int m = 0; m < burst && !b.done() && b.can_run(bd, g); ++m) {
                run(b);
                moved = true;
                if (g_failed) return;
            }
        }
        if (all_done) return;
        if (!moved) return fail("deadlock: no band can move");
    }
}

// k_prepare_factor: 1/dl1 at grid points; dl1 = 1/dl1 = 1 and zero
// couplings at padding entries.
void prepare_factor(Buf &dl, Buf &du, const Skew &sk)
{
    const long L = sk.size();
    for (long e = 0; e < L; ++e) {
        const int s = (int)(e % sk.S);
        const long t1 = e / sk.S;
        const int l = (int)(t1 % 64);
        const long t2 = t1 / 64;
        const int c = (int)(t2 % sk.nch), b = (int)(t2 / sk.nch);
        const int j = b * 64 + l, i = c * sk.S + s - sk.pad(b) - 2 * l;
        if (j < sk.ny && i >= 0 && i < sk.nx) {
            dl.at(5 * L + e) = 1.0 / dl.at(e);
            continue;
        }
        dl.at(e) = 1.0;
        dl.at(5 * L + e) = 1.0;
        for (int k = 1; k < 5; ++k) dl.at(k * L + e) = 0.0;
        for (int k = 0; k < 4; ++k) du.at(k * L + e) = 0.0;
    }
}

// reference layout <-> skewed layout (k_to_skew / k_from_skew)
void to_skew(const double *src, Buf &dst, int nk, const Skew &sk)
{
    const long L = sk.size(), N = (long)sk.nx * sk.ny;
    for (long e = 0; e < L; ++e) {
        const int s = (int)(e % sk.S);
        const long t1 = e / sk.S;
        const int l = (int)(t1 % 64);
        const long t2 = t1 / 64;
        const int c = (int)(t2 % sk.nch), b = (int)(t2 / sk.nch);
        const int u = c * sk.S + s, j = b * 64 + l, i = u - sk.pad(b) - 2 * l;
        const bool real = j < sk.ny && i >= 0 && i < sk.nx;
        for (int k = 0; k < nk; ++k)
            dst.at(k * L + e) = real ? src[k * N + (long)j * sk.nx + i] : 0.0;
    }
}
void from_skew(Buf &src, double *dst, int nk, const Skew &sk)
{
    const long L = sk.size(), N = (long)sk.nx * sk.ny;
    for (int j = 0; j < sk.ny; ++j)
        for (int i = 0; i < sk.nx; ++i) {
            const long e = sk.of_point(i, j);
            for (int k = 0; k < nk; ++k) dst[k * N + (long)j * sk.nx + i] = src.at(k * L + e);
        }
}

}  // namespace

extern "C" {

const char *emu_error(void) { return g_failed ? g_msg : ""; }

// lu2d9 through the emulated factor kernel (A and the factor round-trip
// through the skewed layout exactly as the handle does).
int emu_lu2d9(const double *a_in, double *dl_out, double *du_out, int nx, int ny, double alpha0,
              int unused, unsigned seed)
{
    (void)unused;
    g_failed = false;
    const Skew sk = Skew::make(nx, ny, 8);
    const long L = sk.size();
    Buf a(9 * L, "a_sk"), dl(5 * L, "dl_sk"), du(4 * L, "du_sk");
    to_skew(a_in, a, 9, sk);
    Bnd bd(sk.nb, nx, 4);
    std::vector<Fac> bands;
    for (int b = 0; b < sk.nb; ++b) bands.emplace_back(sk, b);
    schedule(bands, bd, 1u, seed, [&](Fac &b) { b.run(a, dl, du, alpha0, bd, 1u); });
    from_skew(dl, dl_out, 5, sk);
    from_skew(du, du_out, 4, sk);
    return g_failed ? 1 : 0;
}

// One MSI iteration: r (reference layout) -> skewed, forward, backward,
// var = z + var.  The boundary buffer persists across calls for one grid
// (stale words of earlier generations stay in it); generations 2*gen+1 and
// 2*gen+2.
int emu_msi_iteration(const double *r_in, const double *dl_in, const double *du_in,
                      double *x_io, double *w_out, int nx, int ny, int gen, unsigned seed)
{
    g_failed = false;
    const Skew sk = Skew::make(nx, ny, 8);
    const long L = sk.size(), N = (long)nx * ny;
    Buf r(L, "r_sk"), dl(6 * L, "dl_sk"), du(4 * L, "du_sk"), w(L, "w_sk");
    to_skew(r_in, r, 1, sk);
    to_skew(dl_in, dl, 5, sk);
    to_skew(du_in, du, 4, sk);
    prepare_factor(dl, du, sk);
    for (long e = 0; e < L; ++e)  // padding of w starts as garbage, never consumed
        w.at(e) = 1e300;
    static std::vector<Bnd> keep;
    if (keep.empty() || keep[0].nx != nx || keep[0].nb != sk.nb) keep.assign(1, Bnd(sk.nb, nx, 1));
    Bnd &bd = keep[0];
    const unsigned g1 = (unsigned)(2 * gen + 1), g2 = g1 + 1;
    {
        std::vector<Fwd> bands;
        for (int b = 0; b < sk.nb; ++b) bands.emplace_back(sk, b);
        schedule(bands, bd, g1, seed, [&](Fwd &b) { b.run(r, dl, w, bd, g1); });
    }
    if (!g_failed) {
        std::vector<Bwd> bands;
        for (int b = 0; b < sk.nb; ++b) bands.emplace_back(sk, b);
        schedule(bands, bd, g2, seed + 1, [&](Bwd &b) { b.run(w, du, bd, g2); });
    }
    // k_update: var = z + var
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const long p = (long)j * nx + i;
            x_io[p] = w.at(sk.of_point(i, j)) + x_io[p];
        }
    from_skew(w, w_out, 1, sk);
    (void)N;
    return g_failed ? 1 : 0;
}

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
