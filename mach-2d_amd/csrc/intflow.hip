// intflow.hip -- the internal-flow routines of one time step that are not
// coefficient writers, on the GPU: intflow_boundary_simplec,
// intflow_extrapolate_u_v_to_fictitious, intflow_boundary_T_from_H_conservation
// and intflow_time_cycle_update (src/user/internal_flow/mod_intflow.f90
// :557-1134), norm_l1_5d (src/solvers/mod_solvers.f90:91-125) and
// get_average (src/util/mod_field_operations.f90:115-142).
//
// The boundary routines update fictitious rows and columns in place, the
// corners from values the same call writes.  One thread per line index t
// does row j = t+2 (west and east ends, then the corners next to it when j
// is 2 or ny-1) and column i = t+2 (south and north ends), reading before
// writing, so every value is the reference's (operation order kept, no
// contraction) and no thread reads what another writes.  The inflow
// pressure of intflow_time_cycle_update, a real power (x**y), is evaluated
// on the host with the C library's pow (the function the reference calls;
// the device pow may differ in the last bit), from u values read back, so
// it is the reference's too.
// get_average adds the real volumes in the reference's order on one lane
// (bit for bit); norm_l1_5d is a fixed-tree sum (as the 9-diagonal norms).
//
// Arrays are host or device memory (all of one kind); device arrays are
// processed on the assembly stream (enqueue only, except the two calls that
// return a value), host arrays are copied in and back.  Scalars by pointer
// (Fortran bind(C) callers pass them unchanged).  Status: m2d_last_status.
#include <math.h>

#include <algorithm>
#include <vector>

#include "m2d_internal.h"
#include "seqsum.h"

using namespace m2d;

namespace {

constexpr int kT = 256;

__device__ __forceinline__ long at(int i, int j, int nx) { return (long)nx * (j - 1) + i - 1; }

// intflow_boundary_simplec, mod_intflow.f90:791-877
__global__ __launch_bounds__(kT) void k_boundary_simplec(int nx, int ny, double *due,
                                                        double *dve, double *dun, double *dvn,
                                                        double *de, double *dn)
{
    const int t = blockIdx.x * kT + threadIdx.x;
    const int j = t + 2, i = t + 2;
    if (j <= ny - 1) {
        const long w = at(1, j, nx), e = at(nx - 1, j, nx);
        // west copies the volume 2 values; the last real volume copies the
        // one before it (with nx = 3 that is the new west value)
        const double ue2 = due[w + 1], de2 = de[w + 1];
        const double uew = nx == 3 ? ue2 : due[e - 1], vew = nx == 3 ? 0.0 : dve[e - 1],
                     dew = nx == 3 ? de2 : de[e - 1];
        due[w] = ue2;
        dve[w] = 0.0;
        de[w] = de2;
        due[e] = uew;
        dve[e] = vew;
        de[e] = dew;
    }
    if (i <= nx - 1) {
        const long s = at(i, 1, nx), n = at(i, ny - 1, nx);
        const double un2 = dun[s + nx];
        dun[s] = un2;
        dvn[s] = 0.0;
        dn[s] = 0.0;
        dun[n] = 0.0;
        dvn[n] = 0.0;
        dn[n] = 0.0;
    }
}

// tangential slip velocity on the north wall (get_u_v_north_slip,
// mod_intflow.f90:142-210); C = 0: u, 1: v
template <int C>
__device__ __forceinline__ double slip(const double *xk, const double *yk, const double *u,
                                       const double *v, long np)
{
    const double Up = sqrt(u[np] * u[np] + v[np] * v[np]);
    const double signal = copysign(1.0, xk[np] * u[np] + yk[np] * v[np]);
    const double t = sqrt(xk[np] * xk[np] + yk[np] * yk[np]);
    return signal * Up * ((C == 0 ? xk[np] : yk[np]) / t);
}

template <int C>
__device__ double slip_bn(const double *x, const double *xp, const double *xk, const double *yk,
                          const double *u, const double *v, int nx, int ny, int i)
{
    const long np = at(i, ny - 1, nx);
    if (i == 1) {
        const long npe = np + 1, npee = npe + 1;
        const double b2 = slip<C>(xk, yk, u, v, npe), b3 = slip<C>(xk, yk, u, v, npee);
        return b2 + (b2 - b3) * (xp[npe] - x[np]) / (xp[npee] - xp[npe]);
    }
    if (i == nx) {
        const long npw = np - 1, npww = npw - 1;
        const double b1 = slip<C>(xk, yk, u, v, npw), b2 = slip<C>(xk, yk, u, v, npww);
        return b1 + (b1 - b2) * (x[npw] - xp[npw]) / (xp[npw] - xp[npww]);
    }
    return slip<C>(xk, yk, u, v, np);
}

// intflow_extrapolate_u_v_to_fictitious, mod_intflow.f90:881-964 (nx >= 4)
__global__ __launch_bounds__(kT) void k_extrapolate_uv(int nx, int ny, int modvis,
                                                      const double *x, const double *xp,
                                                      const double *xk, const double *yk,
                                                      double *u, double *v)
{
    const int t = blockIdx.x * kT + threadIdx.x;
    const int j = t + 2, i = t + 2;
    if (j <= ny - 1) {
        const long w = at(1, j, nx), e = at(nx, j, nx);
        const long npe = w + 1, npee = w + 2, npw = e - 1, npww = e - 2;
        const double uw = u[npe] + 2.0 * xp[npe] / (xp[npee] - xp[npe]) * (u[npe] - u[npee]);
        const double vw = -v[npe];
        const double fe = 2.0 * (x[npw] - xp[npw]) / (xp[npw] - xp[npww]);
        const double ue = u[npw] + fe * (u[npw] - u[npww]);
        const double ve = v[npw] + fe * (v[npw] - v[npww]);
        u[w] = uw;
        v[w] = vw;
        u[e] = ue;
        v[e] = ve;
        // corners: the south row copies row 2, the north row reflects row ny-1
        if (j == 2) {
            u[at(1, 1, nx)] = uw;
            v[at(1, 1, nx)] = -vw;
            u[at(nx, 1, nx)] = ue;
            v[at(nx, 1, nx)] = -ve;
        }
        if (j == ny - 1) {
            double bu1 = 0.0, bv1 = 0.0, bun = 0.0, bvn = 0.0;
            if (modvis != 1) {
                bu1 = 2.0 * slip_bn<0>(x, xp, xk, yk, u, v, nx, ny, 1);
                bv1 = 2.0 * slip_bn<1>(x, xp, xk, yk, u, v, nx, ny, 1);
                bun = 2.0 * slip_bn<0>(x, xp, xk, yk, u, v, nx, ny, nx);
                bvn = 2.0 * slip_bn<1>(x, xp, xk, yk, u, v, nx, ny, nx);
            }
            const long c1 = at(1, ny, nx), cn = at(nx, ny, nx);
            if (modvis == 1) {
                u[c1] = -uw;
                v[c1] = -vw;
                u[cn] = -ue;
                v[cn] = -ve;
            } else {
                u[c1] = -uw + bu1;
                v[c1] = -vw + bv1;
                u[cn] = -ue + bun;
                v[cn] = -ve + bvn;
            }
        }
    }
    if (i <= nx - 1) {
        const long s = at(i, 1, nx), n = at(i, ny, nx);
        const double us = u[s + nx], vs = v[s + nx], un = u[n - nx], vn = v[n - nx];
        double bu = 0.0, bv = 0.0;
        if (modvis != 1) {
            bu = 2.0 * slip_bn<0>(x, xp, xk, yk, u, v, nx, ny, i);
            bv = 2.0 * slip_bn<1>(x, xp, xk, yk, u, v, nx, ny, i);
        }
        u[s] = us;
        v[s] = -vs;
        if (modvis == 1) {
            u[n] = -un;
            v[n] = -vn;
        } else {
            u[n] = -un + bu;
            v[n] = -vn + bv;
        }
    }
}

// intflow_boundary_T_from_H_conservation, mod_intflow.f90:557-688
__global__ __launch_bounds__(kT) void k_T_from_H(int nx, int ny, double Cpref, double Href,
                                                const double *Tin, const double *ue,
                                                const double *un, const double *ve,
                                                const double *vn, double *T, double *Tbe,
                                                double *Tbw, double *Tbn, double *Tbs)
{
    const int t = blockIdx.x * kT + threadIdx.x;
    const int j = t + 2, i = t + 2;
    if (j <= ny - 1) {
        const long e = at(nx - 1, j, nx), w = at(1, j, nx);
        const double tbe = (Href - (ue[e] * ue[e] + ve[e] * ve[e]) / 2.0) / Cpref;
        const double te = 2.0 * tbe - T[e];
        const double tbw = Tin[j - 1];
        const double tw = 2.0 * tbw - T[w + 1];
        Tbe[j - 1] = tbe;
        T[e + 1] = te;
        Tbw[j - 1] = tbw;
        T[w] = tw;
        if (j == 2) {
            T[at(1, 1, nx)] = tw;
            T[at(nx, 1, nx)] = te;
        }
        if (j == ny - 1) {
            T[at(1, ny, nx)] = tw;
            T[at(nx, ny, nx)] = te;
        }
    }
    if (i <= nx - 1) {
        const long n = at(i, ny - 1, nx), s = at(i, 1, nx);
        const double tbn = (Href - (un[n] * un[n] + vn[n] * vn[n]) / 2.0) / Cpref;
        const double tn = 2.0 * tbn - T[n];
        const double tbs = (Href - (un[s] * un[s] + vn[s] * vn[s]) / 2.0) / Cpref;
        const double ts = 2.0 * tbs - T[s + nx];
        Tbn[i - 1] = tbn;
        T[n + nx] = tn;
        Tbs[i - 1] = tbs;
        T[s] = ts;
    }
}

// The inflow pressure of get_uin_vin_pin_Tin_Mw (mod_intflow.f90:968-1006)
// for row j: a real power, evaluated on the host with the C library's pow,
// the function the reference's x**y calls, so the value is the reference's
// bit for bit (the device pow may differ in the last bit).  uw, ue: u of the
// west fictitious volume and its east neighbour.
static double inflow_pressure(double gamma, double Rg, double po, double T0, double uw,
                              double ue)
{
    const double ui = (uw + ue) / 2.0, vi = 0.0;
    const double Ti = T0 - (gamma - 1.0) / (2.0 * gamma * Rg) * (ui * ui + vi * vi);
    const double M = sqrt((ui * ui + vi * vi) / (gamma * Rg * Ti));
    return po * pow(1.0 + 0.5 * (gamma - 1.0) * (M * M), -gamma / (gamma - 1.0));
}

// u of the two west-most volumes of rows j = 2..ny-1 into uw[2(j-2)], [+1]
__global__ __launch_bounds__(kT) void k_inflow_u(int nx, int ny, const double *u,
                                                double *__restrict__ uw)
{
    const int j = blockIdx.x * kT + threadIdx.x + 2;
    if (j <= ny - 1) {
        const long w = at(1, j, nx);
        uw[2 * (j - 2)] = u[w];
        uw[2 * (j - 2) + 1] = u[w + 1];
    }
}

// intflow_time_cycle_update, mod_intflow.f90:1106-1134: pina = pin, then
// get_uin_vin_pin_Tin_Mw (:968-1006) and get_plin_and_p_fictitious
// (:1010-1101).  Thread t: pina of j = t+1; row j = t+2; column i = t+2.
// pnew[j-2]: the inflow pressure of row j (inflow_pressure).
__global__ __launch_bounds__(kT) void k_time_cycle(int nx, int ny, double gamma, double Rg,
                                                  double T0, const double *x,
                                                  const double *xp, const double *u,
                                                  const double *pnew, double *uin,
                                                  double *vin, double *pin, double *Tin,
                                                  double *Mw, double *plin, double *pina,
                                                  double *p)
{
    const int t = blockIdx.x * kT + threadIdx.x;
    if (t + 1 == 1 || t + 1 == ny) pina[t] = pin[t];  // rows the loops below leave
    const int j = t + 2, i = t + 2;
    if (j <= ny - 1) {
        const int k = j - 1;
        const long w = at(1, j, nx), e = at(nx, j, nx);
        const double pa = pin[k];
        const double ui = (u[w] + u[w + 1]) / 2.0, vi = 0.0;
        const double Ti = T0 - (gamma - 1.0) / (2.0 * gamma * Rg) * (ui * ui + vi * vi);
        const double M = sqrt((ui * ui + vi * vi) / (gamma * Rg * Ti));
        const double pi = pnew[j - 2];
        pina[k] = pa;
        uin[k] = ui;
        vin[k] = vi;
        Tin[k] = Ti;
        Mw[k] = M;
        pin[k] = pi;
        plin[k] = pi - pa;
        const double pw = 2.0 * pi - p[w + 1];
        const long npw = e - 1, npww = npw - 1;
        const double pww = nx == 3 ? pw : p[npww];
        const double pe = p[npw] + 2.0 * (x[npw] - xp[npw]) / (xp[npw] - xp[npww]) * (p[npw] - pww);
        p[w] = pw;
        p[e] = pe;
        if (j == 2) {
            p[at(1, 1, nx)] = pw;
            p[at(nx, 1, nx)] = pe;
        }
        if (j == ny - 1) {
            p[at(1, ny, nx)] = pw;
            p[at(nx, ny, nx)] = pe;
        }
    }
    if (i <= nx - 1) {
        const long s = at(i, 1, nx), n = at(i, ny, nx);
        const double ps = p[s + nx], pn = p[n - nx];
        p[s] = ps;
        p[n] = pn;
    }
}

// per-block partial sums of |A var - b| over real volumes (5 diagonals)
__global__ __launch_bounds__(kT) void k_norm5_partial(const double *__restrict__ a,
                                                     const double *__restrict__ b,
                                                     const double *__restrict__ var, int nx,
                                                     int ny, double *__restrict__ part)
{
    const long N = (long)nx * ny;
    const int i = blockIdx.x * kT + threadIdx.x + 2, j = blockIdx.y + 2;
    double r = 0.0;
    if (i <= nx - 1) {
        const long np = at(i, j, nx);
        r = fabs(+a[np] * var[np - nx] + a[N + np] * var[np - 1] + a[2 * N + np] * var[np] +
                 a[3 * N + np] * var[np + 1] + a[4 * N + np] * var[np + nx] - b[np]);
    }
    __shared__ double red[kT];
    red[threadIdx.x] = r;
    __syncthreads();
    for (int o = kT / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kT) void k_sum(const double *__restrict__ part, long n,
                                           double *__restrict__ out)
{
    __shared__ double red[kT];
    double s = 0.0;
    for (long q = threadIdx.x; q < n; q += kT) s += part[q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kT / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// get_average: F over real volumes in the reference's order (j outer, i
// inner) gathered into t (contiguous, in that order), then added in order
// by one wave and divided by (nx-2)(ny-2)
__global__ __launch_bounds__(kT) void k_average_gather(const double *__restrict__ F, int nx,
                                                      int ny, double *__restrict__ t)
{
    const int i = blockIdx.x * kT + threadIdx.x + 1, j = blockIdx.y + 1;
    if (i <= nx - 2) t[(long)(j - 1) * (nx - 2) + (i - 1)] = F[(long)nx * j + i];
}
__global__ __launch_bounds__(kSeqThreads) void k_average(const double *__restrict__ t, int nx,
                                                        int ny, double *__restrict__ out)
{
    const long n = (long)(nx - 2) * (ny - 2);
    __shared__ double lds[kSeqSlots * kSeqChunk];
    const double s = seq_sum(t, n, lds);
    if (threadIdx.x == 0) out[0] = s / ((double)(nx - 2) * (double)(ny - 2));
}

// arrays of one call: pointer, length, written by the call
struct Arr {
    const double *p;
    long n;
    bool out;
};

struct Buf {
    double *d = nullptr;
    ~Buf()
    {
        if (d) (void)hipFree(d);
    }
};

// Checks the arrays, stages host arrays, runs launch(device pointers,
// stream, scratch) and copies written host arrays back.  scratch: doubles of
// device scratch the launch needs.  sync: wait for the stream also with
// device arrays (results read on the host).
template <typename Launch>
int run(int nx, int ny, std::vector<Arr> arrs, long scratch, bool sync, Launch launch)
{
    if (nx < 3 || ny < 3) return M2D_ERR_ARG;
    for (const Arr &a : arrs)
        if (!a.p) return M2D_ERR_ARG;
    const bool dev = is_device_array(arrs[0].p);
    for (const Arr &a : arrs)
        if (is_device_array(a.p) != dev) return M2D_ERR_ARG;
    const hipStream_t s = assembly_stream_current();
    long tot = scratch;
    if (!dev)
        for (const Arr &a : arrs) tot += a.n;
    Buf B;
    if (tot > 0 && hipMalloc((void **)&B.d, sizeof(double) * tot) != hipSuccess)
        return M2D_ERR_HIP;
    std::vector<double *> d(arrs.size());
    long off = 0;
    for (size_t k = 0; k < arrs.size(); ++k) {
        if (dev) {
            d[k] = const_cast<double *>(arrs[k].p);
            continue;
        }
        d[k] = B.d + off;
        off += arrs[k].n;
        if (hipMemcpyAsync(d[k], arrs[k].p, sizeof(double) * arrs[k].n, hipMemcpyHostToDevice,
                           s) != hipSuccess)
            return M2D_ERR_HIP;
    }
    double *scr = B.d ? B.d + off : nullptr;
    const int L = nx > ny ? nx : ny;
    launch(d, s, scr, dim3((unsigned)((L + kT - 1) / kT)));
    if (hipGetLastError() != hipSuccess) return M2D_ERR_HIP;
    if (!dev)
        for (size_t k = 0; k < arrs.size(); ++k)
            if (arrs[k].out && hipMemcpyAsync(const_cast<double *>(arrs[k].p), d[k],
                                              sizeof(double) * arrs[k].n,
                                              hipMemcpyDeviceToHost, s) != hipSuccess)
                return M2D_ERR_HIP;
    // host arrays, a value to read, or a scratch buffer that the free below
    // would release under a running kernel
    if ((!dev || sync || B.d) && hipStreamSynchronize(s) != hipSuccess) return M2D_ERR_HIP;
    return M2D_OK;
}

}  // namespace

// scratch: the mean, the n gathered terms, one more (seq_sum reads whole
// 16-byte pairs)
long m2d::average_scratch(int nx, int ny) { return (long)(nx - 2) * (ny - 2) + 2; }

void m2d::average_enqueue(int nx, int ny, const double *F, double *scr, hipStream_t s)
{
    const dim3 g((unsigned)((nx - 2 + kT - 1) / kT), (unsigned)(ny - 2));
    hipLaunchKernelGGL(k_average_gather, g, dim3(kT), 0, s, F, nx, ny, scr + 1);
    hipLaunchKernelGGL(k_average, dim3(1), dim3(kSeqThreads), 0, s, scr + 1, nx, ny, scr);
}

void m2d_set_last_status_code(int status);

#define NEED(...)                                               \
    do {                                                        \
        const void *s_[] = {__VA_ARGS__};                       \
        for (const void *p_ : s_)                               \
            if (!p_) {                                          \
                m2d_set_last_status_code(M2D_ERR_ARG);          \
                return;                                         \
            }                                                   \
    } while (0)

// intflow_boundary_simplec, mod_intflow.f90:791-877
extern "C" void m2d_intflow_boundary_simplec(const int *nx, const int *ny, const int *modvis,
                                             double *due, double *dve, double *dun, double *dvn,
                                             double *de, double *dn)
{
    NEED(nx, ny, modvis);
    const int NX = *nx, NY = *ny;
    const long N = (long)NX * NY;
    const int rc = run(NX, NY,
                       {{due, N, true}, {dve, N, true}, {dun, N, true}, {dvn, N, true},
                        {de, N, true}, {dn, N, true}},
                       0, false, [&](std::vector<double *> &d, hipStream_t s, double *, dim3 g) {
                           hipLaunchKernelGGL(k_boundary_simplec, g, dim3(kT), 0, s, NX, NY,
                                              d[0], d[1], d[2], d[3], d[4], d[5]);
                       });
    m2d_set_last_status_code(rc);
}

// intflow_extrapolate_u_v_to_fictitious, mod_intflow.f90:881-964 (nx >= 4:
// the north slip extrapolation reads the third column)
extern "C" void m2d_intflow_extrapolate_u_v_to_fictitious(const int *nx, const int *ny,
                                                          const int *modvis, const double *x,
                                                          const double *xp, const double *xk,
                                                          const double *yk, double *u,
                                                          double *v)
{
    NEED(nx, ny, modvis);
    const int NX = *nx, NY = *ny, mv = *modvis;
    if (NX < 4) {
        m2d_set_last_status_code(M2D_ERR_ARG);
        return;
    }
    const long N = (long)NX * NY;
    const int rc = run(NX, NY,
                       {{x, N, false}, {xp, N, false}, {xk, N, false}, {yk, N, false},
                        {u, N, true}, {v, N, true}},
                       0, false, [&](std::vector<double *> &d, hipStream_t s, double *, dim3 g) {
                           hipLaunchKernelGGL(k_extrapolate_uv, g, dim3(kT), 0, s, NX, NY, mv,
                                              d[0], d[1], d[2], d[3], d[4], d[5]);
                       });
    m2d_set_last_status_code(rc);
}

// intflow_boundary_T_from_H_conservation, mod_intflow.f90:557-688 (iflow
// contributes Tin, ny values; Tbe, Tbw: ny values, Tbn, Tbs: nx values)
extern "C" void m2d_intflow_boundary_T_from_H_conservation(
    const int *nx, const int *ny, const double *Cpref, const double *Href, const double *Tin,
    const double *u, const double *ue, const double *un, const double *v, const double *ve,
    const double *vn, double *T, double *Tbe, double *Tbw, double *Tbn, double *Tbs)
{
    NEED(nx, ny, Cpref, Href, u, v);
    const int NX = *nx, NY = *ny;
    const double cp = *Cpref, h = *Href;
    const long N = (long)NX * NY;
    const int rc = run(NX, NY,
                       {{Tin, NY, false}, {ue, N, false}, {un, N, false}, {ve, N, false},
                        {vn, N, false}, {T, N, true}, {Tbe, NY, true}, {Tbw, NY, true},
                        {Tbn, NX, true}, {Tbs, NX, true}},
                       0, false, [&](std::vector<double *> &d, hipStream_t s, double *, dim3 g) {
                           hipLaunchKernelGGL(k_T_from_H, g, dim3(kT), 0, s, NX, NY, cp, h,
                                              d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7],
                                              d[8], d[9]);
                       });
    m2d_set_last_status_code(rc);
}

// intflow_time_cycle_update, mod_intflow.f90:1106-1134; gamma is
// thermomodel%gm(T0), Rg thermomodel%Rg; the inflow vectors (ny values) are
// the iflow fields of the same names.
extern "C" void m2d_intflow_time_cycle_update(const int *nx, const int *ny, const double *gamma,
                                              const double *Rg, const double *P0,
                                              const double *T0, const double *x,
                                              const double *xp, const double *u, double *uin,
                                              double *vin, double *pin, double *Tin, double *Mw,
                                              double *plin, double *pina, double *p)
{
    NEED(nx, ny, gamma, Rg, P0, T0);
    const int NX = *nx, NY = *ny;
    const double g = *gamma, rg = *Rg, po = *P0, t0 = *T0;
    const long N = (long)NX * NY;
    const int nr = NY - 2;  // inflow rows j = 2..ny-1
    // host copies of the west u values and the inflow pressures; they stay
    // alive until run() has synchronised the stream
    std::vector<double> hu(2 * (size_t)std::max(nr, 1)), hp((size_t)std::max(nr, 1));
    int inner = M2D_OK;
    const int rc = run(NX, NY,
                       {{x, N, false}, {xp, N, false}, {u, N, false}, {uin, NY, true},
                        {vin, NY, true}, {pin, NY, true}, {Tin, NY, true}, {Mw, NY, true},
                        {plin, NY, true}, {pina, NY, true}, {p, N, true}},
                       3 * (long)std::max(nr, 1), false,
                       [&](std::vector<double *> &d, hipStream_t s, double *scr, dim3 gr) {
                           // the pressures need the C library's pow: one read of 2(ny-2)
                           // values, the pressures computed here, then written back
                           double *duw = scr, *dp = scr + 2 * std::max(nr, 1);
                           hipLaunchKernelGGL(k_inflow_u, dim3((unsigned)((nr + kT - 1) / kT)),
                                              dim3(kT), 0, s, NX, NY, d[2], duw);
                           if (hipGetLastError() != hipSuccess ||
                               hipMemcpyAsync(hu.data(), duw, sizeof(double) * 2 * nr,
                                              hipMemcpyDeviceToHost, s) != hipSuccess ||
                               hipStreamSynchronize(s) != hipSuccess) {
                               inner = M2D_ERR_HIP;
                               return;
                           }
                           for (int q = 0; q < nr; ++q)
                               hp[q] = inflow_pressure(g, rg, po, t0, hu[2 * q], hu[2 * q + 1]);
                           if (hipMemcpyAsync(dp, hp.data(), sizeof(double) * nr,
                                              hipMemcpyHostToDevice, s) != hipSuccess) {
                               inner = M2D_ERR_HIP;
                               return;
                           }
                           hipLaunchKernelGGL(k_time_cycle, gr, dim3(kT), 0, s, NX, NY, g, rg,
                                              t0, d[0], d[1], d[2], dp, d[3], d[4], d[5], d[6],
                                              d[7], d[8], d[9], d[10]);
                       });
    m2d_set_last_status_code(rc ? rc : inner);
}

// norm_l1_5d, mod_solvers.f90:91-125 (norm += sum |A var - b|, a: 5 N)
extern "C" void m2d_norm_l1_5d(const int *nx, const int *ny, const double *var, const double *b,
                               const double *a, double *norm)
{
    NEED(nx, ny, norm);
    const int NX = *nx, NY = *ny;
    const long N = (long)NX * NY;
    const dim3 grid((unsigned)((NX - 2 + kT - 1) / kT), (unsigned)(NY > 2 ? NY - 2 : 1));
    const long np = (long)grid.x * grid.y;
    double out = 0.0;
    int rc = run(NX, NY, {{var, N, false}, {b, N, false}, {a, 5 * N, false}}, np + 1, true,
                 [&](std::vector<double *> &d, hipStream_t s, double *scr, dim3) {
                     hipLaunchKernelGGL(k_norm5_partial, grid, dim3(kT), 0, s, d[2], d[1], d[0],
                                        NX, NY, scr);
                     hipLaunchKernelGGL(k_sum, dim3(1), dim3(kT), 0, s, scr, np, scr + np);
                     (void)hipMemcpyAsync(&out, scr + np, sizeof(double), hipMemcpyDeviceToHost,
                                          s);
                 });
    if (!rc) *norm = *norm + out;
    m2d_set_last_status_code(rc);
}

namespace m2d {
// intflow_time_cycle_update on device arrays for the time step, with the
// caller's scratch: d_scr and the pinned h_buf, 3 max(ny-2, 1) doubles
// each.  One wait, for the west u values the C library's pow needs; the
// pressures go back from h_buf, which the caller leaves alone until its
// stream has passed this point.
int time_cycle_enqueue(int NX, int NY, double g, double rg, double po, double t0,
                       const double *x, const double *xp, const double *u, double *uin,
                       double *vin, double *pin, double *Tin, double *Mw, double *plin,
                       double *pina, double *p, double *d_scr, double *h_buf, hipStream_t s)
{
    if (NX < 3 || NY < 3) return M2D_ERR_ARG;
    const int nr = NY - 2;  // inflow rows j = 2..ny-1
    double *duw = d_scr, *dp = d_scr + 2 * nr, *hu = h_buf, *hp = h_buf + 2 * nr;
    hipLaunchKernelGGL(k_inflow_u, dim3((unsigned)((nr + kT - 1) / kT)), dim3(kT), 0, s, NX, NY,
                       u, duw);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(hu, duw, sizeof(double) * 2 * nr, hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return M2D_ERR_HIP;
    for (int q = 0; q < nr; ++q) hp[q] = inflow_pressure(g, rg, po, t0, hu[2 * q], hu[2 * q + 1]);
    if (hipMemcpyAsync(dp, hp, sizeof(double) * nr, hipMemcpyHostToDevice, s) != hipSuccess)
        return M2D_ERR_HIP;
    const int L = NX > NY ? NX : NY;
    hipLaunchKernelGGL(k_time_cycle, dim3((unsigned)((L + kT - 1) / kT)), dim3(kT), 0, s, NX, NY,
                       g, rg, t0, x, xp, u, dp, uin, vin, pin, Tin, Mw, plin, pina, p);
    return hipGetLastError() == hipSuccess ? M2D_OK : M2D_ERR_HIP;
}

// norm_l1_5d's sum enqueued on `s` into d_res[0] (partials in d_scr, one per
// block of 256 columns of a real row)
void norm5_enqueue(const double *var, const double *b, const double *a, int nx, int ny,
                   double *d_scr, double *d_res, hipStream_t s)
{
    const dim3 grid((unsigned)((nx - 2 + kT - 1) / kT), (unsigned)(ny > 2 ? ny - 2 : 1));
    const long np = (long)grid.x * grid.y;
    hipLaunchKernelGGL(k_norm5_partial, grid, dim3(kT), 0, s, a, b, var, nx, ny, d_scr);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(kT), 0, s, d_scr, np, d_res);
}
}  // namespace m2d

// get_average, mod_field_operations.f90:115-142
extern "C" double m2d_get_average(const int *nx, const int *ny, const double *F)
{
    if (!nx || !ny) {
        m2d_set_last_status_code(M2D_ERR_ARG);
        return 0.0;
    }
    const int NX = *nx, NY = *ny;
    const long n = (long)(NX - 2) * (NY - 2);
    double out = 0.0;
    const int rc = run(NX, NY, {{F, (long)NX * NY, false}}, average_scratch(NX, NY), true,
                       [&](std::vector<double *> &d, hipStream_t s, double *scr, dim3) {
                           average_enqueue(NX, NY, d[0], scr, s);
                           (void)hipMemcpyAsync(&out, scr, sizeof(double),
                                                hipMemcpyDeviceToHost, s);
                       });
    m2d_set_last_status_code(rc);
    return out;
}
