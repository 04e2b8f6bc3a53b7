// msi_stencil.h -- per-point arithmetic of the 9-diagonal MSI and TDMA paths,
// shared by the gfx950 kernels and by the CPU wavefront emulator in tests/.
//
// Every formula is the general 2-D stencil form: a neighbour outside the grid
// contributes a literal zero (its du / w / z value is 0), and alpha is alpha0
// on real volumes (2 <= i <= nx-1, 2 <= j <= ny-1, 1-based) and 0 on
// fictitious ones.  Term order follows the reference expressions, so with
// contraction off the results equal the reference's special-row formulas
// bit for bit whenever every coupling towards a neighbour outside the grid is
// exactly zero (the assumption stated at src/solvers/msi2d9-v03.f90:16-20).
#pragma once

#if defined(__HIPCC__)
#define M2D_HD __host__ __device__ __forceinline__
#else
#define M2D_HD inline
#endif

namespace m2d {

// Diagonal slots, 0-based: the reference's 1=SW 2=S 3=SE 4=W 5=P 6=E 7=NW 8=N 9=NE.
enum Slot { kSW = 0, kS = 1, kSE = 2, kW = 3, kP = 4, kE = 5, kNW = 6, kN = 7, kNE = 8 };

// Incomplete factorization of one point: src/solvers/msi2d9-v03.f90:356-369.
// b[9]: row of A at P.  du_w/du_s/du_se/du_sw: U entries (E, NW, N, NE) of
// the W, S, SE, SW neighbours (zeros when the neighbour is outside the grid).
// Out: dl[5] (P, W, SE, S, SW) and du[4] (E, NW, N, NE) at P.
template <typename T>
M2D_HD void factor_point(const T *b, T alpha, const T *du_w, const T *du_s,
                         const T *du_se, const T *du_sw, T *dl, T *du)
{
    const T one = T(1), two = T(2);
    T dl5 = b[kSW];
    T dl4 = (b[kS] - dl5 * du_sw[0] - alpha * b[kSE] * du_se[0]) /
            (one - alpha * du_s[0] * du_se[0]);
    T dl3 = b[kSE] - dl4 * du_s[0];
    T dl2 = (b[kW] - dl4 * du_s[1] - dl5 * du_sw[2] - two * alpha * dl5 * du_sw[1]) /
            (one + two * alpha * du_w[1]);
    T phi1 = dl3 * du_se[0];
    T phi2 = dl5 * du_sw[1];
    T phi3 = dl3 * du_se[3];
    T phi4 = dl2 * du_w[1];
    T dl1 = b[kP] - dl2 * du_w[0] - dl3 * du_se[1] - dl4 * du_s[2] - dl5 * du_sw[3] +
            alpha * (two * phi1 + phi2 + phi3 + two * phi4);
    du[0] = (b[kE] - dl3 * du_se[2] - dl4 * du_s[3] - two * alpha * (phi1 + phi3)) / dl1;
    du[1] = (b[kNW] - dl2 * du_w[2]) / dl1;
    du[2] = (b[kN] - dl2 * du_w[3] - alpha * phi4) / dl1;
    du[3] = b[kNE] / dl1;
    dl[0] = dl1;
    dl[1] = dl2;
    dl[2] = dl3;
    dl[3] = dl4;
    dl[4] = dl5;
}

// Forward substitution L w = -r at one point: msi2d9-v03.f90:155.
template <typename T>
M2D_HD T forward_point(T r, T dl1, T dl2, T dl3, T dl4, T dl5, T w_w, T w_se, T w_s,
                       T w_sw)
{
    return -(r + dl2 * w_w + dl3 * w_se + dl4 * w_s + dl5 * w_sw) / dl1;
}

M2D_HD double fma_t(double a, double b, double c) { return fma(a, b, c); }
M2D_HD float fma_t(float a, float b, float c) { return fmaf(a, b, c); }

// n / d from y = 1/d (correctly rounded): q0 = n*y, then two corrections
// q <- q + (n - q*d)*y with exact remainders.  After the first correction q
// is within one ulp of n/d, and Markstein's theorem (y within half an ulp of
// 1/d, q within one ulp of n/d) makes the second one return the correctly
// rounded quotient, i.e. exactly n / d, for normal operands and quotient.
// Only the sign of a zero quotient may differ (n = -0 gives +0).  Five
// dependent operations once y is known, where the quotient with its own
// scaling and reciprocal refinement needs about thirteen.
template <typename T>
M2D_HD T div_by_rcp(T n, T d, T y)
{
    T q = n * y;
    T e = fma_t(-q, d, n);
    q = fma_t(e, y, q);
    e = fma_t(-q, d, n);
    return fma_t(e, y, q);
}

// forward_point with the quotient taken through y = 1/dl1.
template <typename T>
M2D_HD T forward_point_y(T r, T dl1, T y, T dl2, T dl3, T dl4, T dl5, T w_w, T w_se, T w_s,
                         T w_sw)
{
    return div_by_rcp<T>(-(r + dl2 * w_w + dl3 * w_se + dl4 * w_s + dl5 * w_sw), dl1, y);
}

// factor_point with the four quotients by dl1 taken through y = 1/dl1
// (div_by_rcp: each equals the division, up to the sign of a zero quotient):
// one division where there were four.
template <typename T>
M2D_HD void factor_point_y(const T *b, T alpha, const T *du_w, const T *du_s,
                           const T *du_se, const T *du_sw, T *dl, T *du)
{
    const T one = T(1), two = T(2);
    T dl5 = b[kSW];
    T dl4 = (b[kS] - dl5 * du_sw[0] - alpha * b[kSE] * du_se[0]) /
            (one - alpha * du_s[0] * du_se[0]);
    T dl3 = b[kSE] - dl4 * du_s[0];
    T dl2 = (b[kW] - dl4 * du_s[1] - dl5 * du_sw[2] - two * alpha * dl5 * du_sw[1]) /
            (one + two * alpha * du_w[1]);
    T phi1 = dl3 * du_se[0];
    T phi2 = dl5 * du_sw[1];
    T phi3 = dl3 * du_se[3];
    T phi4 = dl2 * du_w[1];
    T dl1 = b[kP] - dl2 * du_w[0] - dl3 * du_se[1] - dl4 * du_s[2] - dl5 * du_sw[3] +
            alpha * (two * phi1 + phi2 + phi3 + two * phi4);
    const T y = one / dl1;
    du[0] = div_by_rcp<T>(b[kE] - dl3 * du_se[2] - dl4 * du_s[3] - two * alpha * (phi1 + phi3),
                          dl1, y);
    du[1] = div_by_rcp<T>(b[kNW] - dl2 * du_w[2], dl1, y);
    du[2] = div_by_rcp<T>(b[kN] - dl2 * du_w[3] - alpha * phi4, dl1, y);
    du[3] = div_by_rcp<T>(b[kNE], dl1, y);
    dl[0] = dl1;
    dl[1] = dl2;
    dl[2] = dl3;
    dl[3] = dl4;
    dl[4] = dl5;
}

// lu2d5's point (msi2d5-v03.f90:287-295) with the three quotients by dl1
// taken through y = 1/dl1.  b[5]: S, W, P, E, N; du_w/du_s/du_se: the U
// entries (E, NW, N) of the W, S and SE neighbours (zeros outside the grid).
// Out: dl[4] (P, W, SE, S) and du[3] (E, NW, N).
template <typename T>
M2D_HD void factor_point5_y(const T *b, T alpha, const T *du_w, const T *du_s,
                            const T *du_se, T *dl, T *du)
{
    const T one = T(1), two = T(2);
    T dl4 = b[0] / (one - alpha * du_s[0] * du_se[0]);
    T dl3 = -dl4 * du_s[0];
    T dl2 = (b[1] - dl4 * du_s[1]) / (one + two * alpha * du_w[1]);
    T phi1 = dl3 * du_se[0];
    T phi4 = dl2 * du_w[1];
    T dl1 = b[2] - dl2 * du_w[0] - dl3 * du_se[1] - dl4 * du_s[2] + two * alpha * (phi1 + phi4);
    const T y = one / dl1;
    du[0] = div_by_rcp<T>(b[3] - dl3 * du_se[2] - two * alpha * phi1, dl1, y);
    du[1] = -div_by_rcp<T>(dl2 * du_w[2], dl1, y);
    du[2] = div_by_rcp<T>(b[4] - alpha * phi4, dl1, y);
    dl[0] = dl1;
    dl[1] = dl2;
    dl[2] = dl3;
    dl[3] = dl4;
}

// Forward substitution of the FP32 sweeps (inner solver of the FP64
// refinement, where only the operator, not the rounding, must be the
// reference's): the row is pre-scaled by ny = -1/dl1 (ne_k = ny * dl_k), so
// w = ny*r + ne5*w_sw + ne4*w_s + ne2*w_w + ne3*w_se needs one multiply and
// four fused multiply-adds, and the two values of the previous step (w_w
// from this row, w_se from the row below) enter last.
template <typename T>
M2D_HD T forward_point_scaled(T r, T ny, T ne2, T ne3, T ne4, T ne5, T w_w, T w_se, T w_s,
                              T w_sw)
{
    T acc = r * ny;
    acc = fma_t(ne5, w_sw, acc);
    acc = fma_t(ne4, w_s, acc);
    acc = fma_t(ne2, w_w, acc);
    return fma_t(ne3, w_se, acc);
}

// Backward substitution U z = w at one point: msi2d9-v03.f90:175.
template <typename T>
M2D_HD T backward_point(T w, T du1, T du2, T du3, T du4, T z_e, T z_nw, T z_n, T z_ne)
{
    return w - du1 * z_e - du2 * z_nw - du3 * z_n - du4 * z_ne;
}

// Backward substitution of the FP32 sweeps: the same sum with fused
// multiply-adds, the values of the previous step (z_e, then z_nw from the
// row above) last.
template <typename T>
M2D_HD T backward_point_fused(T w, T du1, T du2, T du3, T du4, T z_e, T z_nw, T z_n, T z_ne)
{
    T acc = fma_t(-du4, z_ne, w);
    acc = fma_t(-du3, z_n, acc);
    acc = fma_t(-du1, z_e, acc);
    return fma_t(-du2, z_nw, acc);
}

// ---- 5-diagonal MSI (src/solvers/msi2d5-v03.f90) --------------------------
// L = dl1..dl4 (P, W, SE, S), U = I + du1..du3 (E, NW, N): the 9-diagonal
// factors of the zero-corner embedding of A, without the SW and NE entries,
// which are exactly zero there (dl5 = b1 = 0, du4 = b9 / dl1 = 0).  The sums
// are the 9-diagonal ones without their last term, in the same order.

// Forward substitution L w = -r at one point: msi2d5-v03.f90:139.
template <typename T>
M2D_HD T forward_point5_y(T r, T dl1, T y, T dl2, T dl3, T dl4, T w_w, T w_se, T w_s)
{
    return div_by_rcp<T>(-(r + dl2 * w_w + dl3 * w_se + dl4 * w_s), dl1, y);
}

// FP32 form on pre-scaled rows (see forward_point_scaled).
template <typename T>
M2D_HD T forward_point5_scaled(T r, T ny, T ne2, T ne3, T ne4, T w_w, T w_se, T w_s)
{
    T acc = r * ny;
    acc = fma_t(ne4, w_s, acc);
    acc = fma_t(ne2, w_w, acc);
    return fma_t(ne3, w_se, acc);
}

// Backward substitution U z = w at one point: msi2d5-v03.f90:158.
template <typename T>
M2D_HD T backward_point5(T w, T du1, T du2, T du3, T z_e, T z_nw, T z_n)
{
    return w - du1 * z_e - du2 * z_nw - du3 * z_n;
}

// FP32 form with fused multiply-adds, the previous step's values last.
template <typename T>
M2D_HD T backward_point5_fused(T w, T du1, T du2, T du3, T z_e, T z_nw, T z_n)
{
    T acc = fma_t(-du3, z_n, w);
    acc = fma_t(-du1, z_e, acc);
    return fma_t(-du2, z_nw, acc);
}

// 5-diagonal slots, 0-based: the reference's 1=S 2=W 3=P 4=E 5=N
// (src/solvers/tdma2d5.f90:47-51).
enum Slot5 { k5S = 0, k5W = 1, k5P = 2, k5E = 3, k5N = 4 };

// Residual of a 5-diagonal row, terms in slot order from -rhs, neighbours
// outside the grid omitted (the row classes of msi2d5-v03.f90:59-131).  The
// iterate is read through xv(i', j', flat index) of the neighbour.
template <typename TA, typename TR, typename XV>
M2D_HD TR residual_point5_f(const TA *a, long ld, XV xv, TR rhs, int i, int j, int nx, int ny)
{
    const long p = (long)j * nx + i;
    const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < ny - 1;
    // operands read unconditionally (clamped neighbours), absent terms skipped
    const int iw = hw ? i - 1 : i, ie = he ? i + 1 : i, js = hs ? j - 1 : j, jn = hn ? j + 1 : j;
    const TR x_s = (TR)xv(i, js, hs ? p - nx : p), x_w = (TR)xv(iw, j, hw ? p - 1 : p);
    const TR x_p = (TR)xv(i, j, p);
    const TR x_e = (TR)xv(ie, j, he ? p + 1 : p), x_n = (TR)xv(i, jn, hn ? p + nx : p);
    TR c[5];
    for (int k = 0; k < 5; ++k) c[k] = (TR)a[k * ld + p];
    TR acc = -rhs;
    if (hs) acc += c[k5S] * x_s;
    if (hw) acc += c[k5W] * x_w;
    acc += c[k5P] * x_p;
    if (he) acc += c[k5E] * x_e;
    if (hn) acc += c[k5N] * x_n;
    return acc;
}

// residual_point5_f on an iterate array x.
template <typename TA, typename TX, typename TR>
M2D_HD TR residual_point5(const TA *a, long ld, const TX *x, TR rhs, int i, int j, int nx, int ny)
{
    return residual_point5_f<TA, TR>(
        a, ld, [x](int, int, long q) { return x[q]; }, rhs, i, j, nx, ny);
}

// Residual r = A.x - rhs at (i, j) (0-based), 2-D aware: a term whose
// neighbour lies outside the grid is omitted, as in the seven row classes of
// msi2d9-v03.f90:59-131.  Terms are added in slot order starting from -rhs.
// `a` points at diagonal 0; `ld` is the diagonal stride (nx*ny).  The
// iterate is read through xv(i', j', flat index) of the neighbour.
template <typename TA, typename TR, typename XV>
M2D_HD TR residual_point_f(const TA *a, long ld, XV xv, TR rhs, int i, int j, int nx, int ny)
{
    const long p = (long)j * nx + i;
    const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < ny - 1;
    // Neighbour coordinates clamped into the grid: every operand is read
    // unconditionally (so all the loads can be in flight together) and the
    // term of an absent neighbour is skipped afterwards.
    const int iw = hw ? i - 1 : i, ie = he ? i + 1 : i;
    const int js = hs ? j - 1 : j, jn = hn ? j + 1 : j;
    const long rs = (long)js * nx, rp = (long)j * nx, rn = (long)jn * nx;
    const TR x_sw = (TR)xv(iw, js, rs + iw), x_s = (TR)xv(i, js, rs + i);
    const TR x_se = (TR)xv(ie, js, rs + ie), x_w = (TR)xv(iw, j, rp + iw);
    const TR x_p = (TR)xv(i, j, p), x_e = (TR)xv(ie, j, rp + ie);
    const TR x_nw = (TR)xv(iw, jn, rn + iw), x_n = (TR)xv(i, jn, rn + i);
    const TR x_ne = (TR)xv(ie, jn, rn + ie);
    TR c[9];
    for (int k = 0; k < 9; ++k) c[k] = (TR)a[k * ld + p];
    TR acc = -rhs;
    if (hs && hw) acc += c[kSW] * x_sw;
    if (hs) acc += c[kS] * x_s;
    if (hs && he) acc += c[kSE] * x_se;
    if (hw) acc += c[kW] * x_w;
    acc += c[kP] * x_p;
    if (he) acc += c[kE] * x_e;
    if (hn && hw) acc += c[kNW] * x_nw;
    if (hn) acc += c[kN] * x_n;
    if (hn && he) acc += c[kNE] * x_ne;
    return acc;
}

// residual_point_f on an iterate array x.
template <typename TA, typename TX, typename TR>
M2D_HD TR residual_point(const TA *a, long ld, const TX *x, TR rhs, int i, int j, int nx,
                         int ny)
{
    return residual_point_f<TA, TR>(
        a, ld, [x](int, int, long q) { return x[q]; }, rhs, i, j, nx, ny);
}

// alpha of msi2d9-v03.f90:287,347-351: alpha0 on real volumes, 0 on fictitious.
template <typename T>
M2D_HD T alpha_at(int i, int j, int nx, int ny, T alpha0)
{
    return (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2) ? alpha0 : T(0);
}

}  // namespace m2d
