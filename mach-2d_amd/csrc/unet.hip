// unet.hip -- the learned initial guess of the p' solve: a 4-level U-Net
// (16-32-64-128 channels, 3x3 convolutions) with BF16 weights and
// activations and FP32 accumulation on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16), for inference and for training.  No reference
// counterpart: the reference always starts MSI from the caller's var.
//
// Tensors are NHWC (channels fastest), bf16, channel counts multiples of 8,
// so one lane moves 8 channels (16 bytes) at a time.  A convolution is an
// implicit GEMM: M = pixels, N = output channels, K = 9 taps x input
// channels, k = tap * Cin + ci (tap = 3 ky + kx); weights are stored
// [Cout_pad][Kpad] (Kpad = 9 Cin rounded up to 32, zeros beyond).  Its input
// is up to two tensors whose channels are concatenated: the first may be read
// directly, as the 2x2 average of a tensor of twice the resolution (the
// down path's pooling, fused), or as the nearest-neighbour 2x upsampling of a
// tensor of half the resolution (the up path, fused); the second (the skip
// connection) is read directly.  The backward pass reuses the same kernel for
// input gradients (weights rotated by 180 degrees, channels transposed) and
// has its own MFMA kernel for weight gradients.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "m2d_internal.h"
#include "msi_stencil.h"

namespace m2d {
namespace unet {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int kMode0 = 0;     // direct
constexpr int kModePool = 1;  // 2x2 average of a 2H x 2W tensor
constexpr int kModeUp = 2;    // nearest 2x upsampling of an H/2 x W/2 tensor

struct In {
    const __bf16 *x;
    int c, mode;
};

struct ConvArgs {
    In in1, in2;          // in2.c == 0: no second input
    const __bf16 *w;      // [cout_pad][kpad]
    const float *bias;    // cout values, or null
    __bf16 *out;          // NHWC, cout_stride channels, or null
    float *out32;         // NHWC fp32, cout_stride channels, or null
    int cout, cout_stride, kpad;
    int n, h, w_;         // output (and direct-input) resolution
    int relu;
    int accumulate;       // add into out32 instead of overwriting
};

__device__ __forceinline__ bf16x8 zero8()
{
    bf16x8 z;
#pragma unroll
    for (int i = 0; i < 8; ++i) z[i] = (__bf16)0.0f;
    return z;
}

__device__ __forceinline__ bf16x8 ld8(const __bf16 *p)
{
    return *reinterpret_cast<const bf16x8 *>(p);
}

// 8 channels ci0.. of the (virtual) input at pixel (n, y, x) of the output
// resolution h x w, for input `in` (tensor resolution per its mode).
__device__ __forceinline__ bf16x8 load_in(const In &in, int ci0, int n, int y, int x, int h,
                                          int w)
{
    if (in.mode == kMode0) return ld8(in.x + (((long)n * h + y) * w + x) * in.c + ci0);
    if (in.mode == kModeUp) {
        const int hc = h >> 1, wc = w >> 1;
        return ld8(in.x + (((long)n * hc + (y >> 1)) * wc + (x >> 1)) * in.c + ci0);
    }
    const int hf = h << 1, wf = w << 1;  // pooled: mean of the 2x2 block
    const __bf16 *p = in.x + (((long)n * hf + 2 * y) * wf + 2 * x) * in.c + ci0;
    const bf16x8 a = ld8(p), b = ld8(p + in.c), c = ld8(p + (long)wf * in.c),
                 d = ld8(p + (long)wf * in.c + in.c);
    bf16x8 r;
#pragma unroll
    for (int i = 0; i < 8; ++i)
        r[i] = (__bf16)((((float)a[i] + (float)b[i]) + ((float)c[i] + (float)d[i])) * 0.25f);
    return r;
}

// im2col element group: 8 consecutive k = k0.. (k0 multiple of 8) of pixel
// p (flat n*h*w index); zeros outside the image and beyond 9 * cin.
__device__ __forceinline__ bf16x8 im2col8(const ConvArgs &A, int p, int k0)
{
    const int cin = A.in1.c + A.in2.c;
    if (k0 >= 9 * cin) return zero8();
    const int tap = k0 / cin, ci0 = k0 - tap * cin;
    const int hw = A.h * A.w_;
    const int n = p / hw, rem = p - n * hw, y0 = rem / A.w_, x0 = rem - y0 * A.w_;
    const int y = y0 + tap / 3 - 1, x = x0 + tap % 3 - 1;
    if (y < 0 || y >= A.h || x < 0 || x >= A.w_) return zero8();
    return ci0 < A.in1.c ? load_in(A.in1, ci0, n, y, x, A.h, A.w_)
                         : load_in(A.in2, ci0 - A.in1.c, n, y, x, A.h, A.w_);
}

// Forward convolution (also the input gradient, with rotated weights).
// Workgroup: 4 waves x 16 pixels; wave: 16 pixels x 16*NT output channels.
template <int NT>
__global__ __launch_bounds__(256) void k_conv3x3(ConvArgs A)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int npix = A.n * A.h * A.w_;
    const int pix0 = (blockIdx.x * 4 + wave) * 16;
    const int co0 = blockIdx.y * 16 * NT;
    if (pix0 >= npix) return;
    const int pa = min(pix0 + (lane & 15), npix - 1);  // A row of this lane
    const int kl = 8 * (lane >> 4);                    // k offset of this lane
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const __bf16 *wrow = A.w + (long)(co0 + (lane & 15)) * A.kpad + kl;
    for (int kb = 0; kb < A.kpad; kb += 32) {
        const bf16x8 a = im2col8(A, pa, kb + kl);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bf16x8 b = ld8(wrow + (long)t * 16 * A.kpad + kb);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = co0 + 16 * t + (lane & 15);
        if (co >= A.cout) continue;
        const float bias = A.bias ? A.bias[co] : 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int p = pix0 + 4 * (lane >> 4) + i;
            if (p >= npix) continue;
            float v = acc[t][i] + bias;
            if (A.relu) v = fmaxf(v, 0.0f);
            const long o = (long)p * A.cout_stride + co;
            if (A.out) A.out[o] = (__bf16)v;
            if (A.out32) A.out32[o] = A.accumulate ? A.out32[o] + v : v;
        }
    }
}

// Weight gradient dW[co][k] += sum_p G[p][co] * im2col(p, k): M = output
// channels, N = k, K = pixels.  One wave per (16 co, 16 k, pixel range);
// partial sums leave with FP32 atomics.
struct WgradArgs {
    ConvArgs c;          // inputs and geometry of the forward convolution
    const __bf16 *g;     // gradient at the convolution's output (pre-activation), cout channels
    float *dw;           // [cout_pad][kpad]
    float *db;           // cout values, or null
    int pix_per_block;
};

__global__ __launch_bounds__(64) void k_conv3x3_wgrad(WgradArgs W)
{
    const ConvArgs &A = W.c;
    const int lane = threadIdx.x;
    const int k0 = blockIdx.x * 16, co0 = blockIdx.y * 16;
    const int npix = A.n * A.h * A.w_;
    const int p_begin = blockIdx.z * W.pix_per_block;
    const int p_end = min(npix, p_begin + W.pix_per_block);
    const int cin = A.in1.c + A.in2.c;
    const int co = co0 + (lane & 15), kk = k0 + (lane & 15);
    const int pl = 8 * (lane >> 4);
    const bool co_ok = co < A.cout;
    // the k column of this lane as (tap, channel group); 8 pixels at a time
    const bool k_ok = kk < 9 * cin;
    const int tap = k_ok ? kk / cin : 0, ci = k_ok ? kk - tap * cin : 0;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    float bsum = 0.0f;
    for (int p0 = p_begin; p0 < p_end; p0 += 32) {
        bf16x8 a, b;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int p = p0 + pl + i;
            const bool ok = p < p_end;
            a[i] = (ok && co_ok) ? W.g[(long)p * A.cout + co] : (__bf16)0.0f;
            __bf16 x = (__bf16)0.0f;
            if (ok && k_ok) {
                const int hw = A.h * A.w_;
                const int n = p / hw, rem = p - n * hw, y0 = rem / A.w_, x0 = rem - y0 * A.w_;
                const int y = y0 + tap / 3 - 1, xx = x0 + tap % 3 - 1;
                if (y >= 0 && y < A.h && xx >= 0 && xx < A.w_) {
                    const int cg = ci & ~7;
                    const bf16x8 v = ci < A.in1.c ? load_in(A.in1, cg, n, y, xx, A.h, A.w_)
                                                  : load_in(A.in2, cg - A.in1.c, n, y, xx,
                                                            A.h, A.w_);
                    x = v[ci & 7];
                }
            }
            b[i] = x;
            if (W.db && blockIdx.x == 0) bsum += (float)a[i];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    // D[m][n]: lane holds n = lane%16 (k), m = 4*(lane/16) + i (co)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int com = co0 + 4 * (lane >> 4) + i, kn = k0 + (lane & 15);
        if (com < A.cout && kn < A.kpad) atomicAdd(&W.dw[(long)com * A.kpad + kn], acc[i]);
    }
    if (W.db && blockIdx.x == 0) {
        // lanes l, l+16, l+32, l+48 hold partial sums of channel co0 + l%16
        bsum += __shfl_xor(bsum, 16);
        bsum += __shfl_xor(bsum, 32);
        if (lane < 16 && co_ok) atomicAdd(&W.db[co], bsum);
    }
}

// Elementwise pieces of the backward pass.
// g <- g * (y > 0): gradient through the ReLU, from the activation y.
__global__ __launch_bounds__(256) void k_relu_bwd(__bf16 *g, const __bf16 *y, long n)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n && !((float)y[i] > 0.0f)) g[i] = (__bf16)0.0f;
}

// Gradient of a fused input mode, from the gradient `gi` of the effective
// input (output resolution h x w, `cs` channels, first `c` of them belong to
// this input) into `dst` (the tensor's own resolution, c channels), added in
// FP32 into dst32.  Pool: each fine pixel gets a quarter; upsampling: a
// coarse pixel sums its four.
__global__ __launch_bounds__(256) void k_mode_bwd(const float *gi, int cs, int c0, int c,
                                                  int mode, int n, int h, int w, float *dst32)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (mode == kModePool) {
        const int hf = 2 * h, wf = 2 * w;
        const long tot = (long)n * hf * wf * c;
        if (i >= tot) return;
        const int ch = (int)(i % c);
        const long px = i / c;
        const int x = (int)(px % wf), y = (int)((px / wf) % hf), b = (int)(px / ((long)wf * hf));
        dst32[i] += 0.25f * gi[(((long)b * h + y / 2) * w + x / 2) * cs + c0 + ch];
    } else if (mode == kModeUp) {
        const int hc = h / 2, wc = w / 2;
        const long tot = (long)n * hc * wc * c;
        if (i >= tot) return;
        const int ch = (int)(i % c);
        const long px = i / c;
        const int x = (int)(px % wc), y = (int)((px / wc) % hc), b = (int)(px / ((long)wc * hc));
        float s = 0.0f;
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx)
                s += gi[(((long)b * h + 2 * y + dy) * w + 2 * x + dx) * cs + c0 + ch];
        dst32[i] += s;
    } else {
        const long tot = (long)n * h * w * c;
        if (i >= tot) return;
        const int ch = (int)(i % c);
        dst32[i] += gi[(i / c) * cs + c0 + ch];
    }
}

// bf16 copy of an FP32 tensor.
__global__ __launch_bounds__(256) void k_to_bf16(const float *src, __bf16 *dst, long n)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = (__bf16)src[i];
}

// Adam on FP32 master weights w[cout_pad][kpad] (and bias), then the bf16
// copies the kernels read: wb (same layout) and wr, the input-gradient
// weights [cin_pad][kpad_r] with wr[ci][tap' * cout + co] = w[co][(8 - tap') * cin + ci].
struct AdamArgs {
    float *w, *m, *v;
    const float *g;
    long n;
    float lr, b1, b2, eps, c1, c2;  // c1 = 1 - b1^t, c2 = 1 - b2^t
};

__global__ __launch_bounds__(256) void k_adam(AdamArgs a)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const float g = a.g[i];
    const float m = a.b1 * a.m[i] + (1.0f - a.b1) * g;
    const float v = a.b2 * a.v[i] + (1.0f - a.b2) * g * g;
    a.m[i] = m;
    a.v[i] = v;
    a.w[i] -= a.lr * (m / a.c1) / (sqrtf(v / a.c2) + a.eps);
}

__global__ __launch_bounds__(256) void k_pack_weights(const float *w, int cout, int cin, int kpad,
                                                      __bf16 *wb, __bf16 *wr, int kpad_r,
                                                      int cin_pad)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const long nw = (long)cout * kpad;
    if (i < nw) wb[i] = (__bf16)w[i];
    if (wr && i < (long)cin_pad * kpad_r) {
        const int ci = (int)(i / kpad_r), k = (int)(i % kpad_r);
        float v = 0.0f;
        if (ci < cin && k < 9 * cout) {
            const int tr = k / cout, co = k - tr * cout;
            v = w[(long)co * kpad + (8 - tr) * cin + ci];
        }
        wr[i] = (__bf16)v;
    }
}

// Loss 0.5 * sum over masked pixels of (pred - target)^2 / count, its
// gradient (bf16, `cs` channels with the value in channel 0) and per-block
// partial sums of the loss.
__global__ __launch_bounds__(256) void k_mse(const float *pred, int ps, const float *target,
                                             const float *mask, long npix, float inv_count,
                                             __bf16 *grad, int gs, float *partial)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    float l = 0.0f;
    if (i < npix) {
        const float d = (pred[i * ps] - target[i]) * mask[i];
        l = 0.5f * d * d * inv_count;
        grad[i * gs] = (__bf16)(d * inv_count);
        for (int c = 1; c < gs; ++c) grad[i * gs + c] = (__bf16)0.0f;
    }
    __shared__ float red[256];
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Network input from a 9-diagonal system (FP64, reference layout) and its
// right-hand side: the Jacobi-scaled form  p - sum (a_nb/a_P) p_nb = b/a_P.
// Channels: 0 b/(a_P s), 1..4 a_W, a_E, a_S, a_N over a_P, 5 the corner sum
// over a_P, 6..7 zero; s = RMS of b/a_P over the grid.  Pixels beyond the
// grid (the tensor is h x w >= ny x nx) are zero, as are fictitious rows
// with a_P = 0.
__global__ __launch_bounds__(256) void k_pack_input(const double *a, const double *b, int nx,
                                                    int ny, int h, int w, double inv_s,
                                                    __bf16 *x, int nb)
{
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    const long np_ = (long)h * w;
    if (q >= np_ * nb) return;
    const int bi = (int)(q / np_);
    const long pix = q - (long)bi * np_;
    const int xx = (int)(pix % w), yy = (int)(pix / w);
    float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (xx < nx && yy < ny) {
        const long N = (long)nx * ny, p = (long)yy * nx + xx;
        const double *as = a + (long)bi * 9 * N, *bs = b + (long)bi * N;
        const double ap = as[kP * N + p];
        if (ap != 0.0) {
            const double r = 1.0 / ap;
            f[0] = (float)(bs[p] * r * inv_s);
            f[1] = (float)(as[kW * N + p] * r);
            f[2] = (float)(as[kE * N + p] * r);
            f[3] = (float)(as[kS * N + p] * r);
            f[4] = (float)(as[kN * N + p] * r);
            f[5] = (float)((as[kSW * N + p] + as[kSE * N + p] + as[kNW * N + p] +
                            as[kNE * N + p]) * r);
        }
    }
    bf16x8 v;
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = (__bf16)f[i];
    *reinterpret_cast<bf16x8 *>(x + q * 8) = v;
}

// Guess var = s * out (channel 0 of the FP32 output) on the grid.
__global__ __launch_bounds__(256) void k_unpack_output(const float *out, int os, int nx, int ny,
                                                       int w, double s, double *var)
{
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q >= (long)nx * ny) return;
    const int xx = (int)(q % nx), yy = (int)(q / nx);
    var[q] = s * (double)out[((long)yy * w + xx) * os];
}

}  // namespace unet
}  // namespace m2d

using namespace m2d;
using namespace m2d::unet;

namespace {
unsigned blocks256(long n) { return (unsigned)((n + 255) / 256); }
int rc_of(hipError_t e) { return e == hipSuccess ? M2D_OK : M2D_ERR_HIP; }
bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }
}  // namespace

// ---------------------------------------------------------------------------
// C ABI (device pointers, stream as void*).  No reference counterpart.
// ---------------------------------------------------------------------------

extern "C" int m2d_unet_conv3x3(const void *in1, int c1, int mode1, const void *in2, int c2,
                                const void *w, const float *bias, int cout, int cout_stride,
                                int kpad, int n, int h, int w_, int relu, void *out,
                                float *out32, int accumulate, void *stream)
{
    const int cin = c1 + c2;
    if (!in1 || c1 <= 0 || c1 % 8 || c2 < 0 || c2 % 8 || (c2 && !in2) || !w || cout <= 0 ||
        cout_stride < cout || kpad < 9 * cin || kpad % 32 || n <= 0 || h <= 0 || w_ <= 0 ||
        (!out && !out32) || mode1 < 0 || mode1 > 2 || (mode1 == kModeUp && (h % 2 || w_ % 2)))
        return M2D_ERR_ARG;
    if (!aligned16(in1) || (in2 && !aligned16(in2)) || !aligned16(w)) return M2D_ERR_ARG;
    ConvArgs A;
    A.in1 = In{(const __bf16 *)in1, c1, mode1};
    A.in2 = In{(const __bf16 *)in2, c2, kMode0};
    A.w = (const __bf16 *)w;
    A.bias = bias;
    A.out = (__bf16 *)out;
    A.out32 = out32;
    A.cout = cout;
    A.cout_stride = cout_stride;
    A.kpad = kpad;
    A.n = n;
    A.h = h;
    A.w_ = w_;
    A.relu = relu;
    A.accumulate = accumulate;
    const long npix = (long)n * h * w_;
    if (npix > 0x7fffffffL / 2) return M2D_ERR_ARG;
    const int cpad = (cout + 15) / 16 * 16;  // weight rows available: cout rounded to 16
    const int nt = cpad >= 64 && cpad % 64 == 0 ? 4 : cpad % 32 == 0 ? 2 : 1;
    dim3 grid((unsigned)((npix + 63) / 64), (unsigned)(cpad / (16 * nt)));
    hipStream_t s = (hipStream_t)stream;
    if (nt == 4)
        hipLaunchKernelGGL(k_conv3x3<4>, grid, dim3(256), 0, s, A);
    else if (nt == 2)
        hipLaunchKernelGGL(k_conv3x3<2>, grid, dim3(256), 0, s, A);
    else
        hipLaunchKernelGGL(k_conv3x3<1>, grid, dim3(256), 0, s, A);
    return rc_of(hipGetLastError());
}

extern "C" int m2d_unet_conv3x3_wgrad(const void *in1, int c1, int mode1, const void *in2, int c2,
                                      const void *g, int cout, int kpad, int n, int h, int w_,
                                      float *dw, float *db, void *stream)
{
    const int cin = c1 + c2;
    if (!in1 || c1 <= 0 || c1 % 8 || c2 < 0 || c2 % 8 || (c2 && !in2) || !g || !dw ||
        cout <= 0 || kpad < 9 * cin || kpad % 32 || n <= 0 || h <= 0 || w_ <= 0 ||
        mode1 < 0 || mode1 > 2)
        return M2D_ERR_ARG;
    WgradArgs W;
    W.c.in1 = In{(const __bf16 *)in1, c1, mode1};
    W.c.in2 = In{(const __bf16 *)in2, c2, kMode0};
    W.c.cout = cout;
    W.c.kpad = kpad;
    W.c.n = n;
    W.c.h = h;
    W.c.w_ = w_;
    W.g = (const __bf16 *)g;
    W.dw = dw;
    W.db = db;
    const long npix = (long)n * h * w_;
    W.pix_per_block = 2048;
    dim3 grid((unsigned)((9 * cin + 15) / 16), (unsigned)((cout + 15) / 16),
              (unsigned)((npix + W.pix_per_block - 1) / W.pix_per_block));
    hipLaunchKernelGGL(k_conv3x3_wgrad, grid, dim3(64), 0, (hipStream_t)stream, W);
    return rc_of(hipGetLastError());
}

extern "C" int m2d_unet_relu_bwd(void *g, const void *y, long count, void *stream)
{
    if (!g || !y || count < 0) return M2D_ERR_ARG;
    if (!count) return M2D_OK;
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks256(count)), dim3(256), 0, (hipStream_t)stream,
                       (__bf16 *)g, (const __bf16 *)y, count);
    return rc_of(hipGetLastError());
}

extern "C" int m2d_unet_mode_bwd(const float *gi, int cs, int c0, int c, int mode, int n, int h,
                                 int w, float *dst32, void *stream)
{
    if (!gi || !dst32 || c <= 0 || c0 < 0 || c0 + c > cs || mode < 0 || mode > 2) return M2D_ERR_ARG;
    const long tot = mode == kModePool  ? (long)n * 4 * h * w * c
                     : mode == kModeUp ? (long)n * (h / 2) * (w / 2) * c
                                       : (long)n * h * w * c;
    hipLaunchKernelGGL(k_mode_bwd, dim3(blocks256(tot)), dim3(256), 0, (hipStream_t)stream, gi,
                       cs, c0, c, mode, n, h, w, dst32);
    return rc_of(hipGetLastError());
}

extern "C" int m2d_unet_to_bf16(const float *src, void *dst, long count, void *stream)
{
    if (!src || !dst || count < 0) return M2D_ERR_ARG;
    if (!count) return M2D_OK;
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks256(count)), dim3(256), 0, (hipStream_t)stream, src,
                       (__bf16 *)dst, count);
    return rc_of(hipGetLastError());
}

extern "C" int m2d_unet_adam(float *w, float *m, float *v, const float *g, long count, float lr,
                             float b1, float b2, float eps, int step, void *stream)
{
    if (!w || !m || !v || !g || count < 0 || step < 1) return M2D_ERR_ARG;
    if (!count) return M2D_OK;
    AdamArgs a{w, m, v, g, count, lr, b1, b2, eps, 1.0f - powf(b1, (float)step),
               1.0f - powf(b2, (float)step)};
    hipLaunchKernelGGL(k_adam, dim3(blocks256(count)), dim3(256), 0, (hipStream_t)stream, a);
    return rc_of(hipGetLastError());
}

extern "C" int m2d_unet_pack_weights(const float *w, int cout, int cin, int kpad, void *wb,
                                     void *wr, int kpad_r, int cin_pad, void *stream)
{
    if (!w || !wb || cout <= 0 || cin <= 0 || kpad < 9 * cin ||
        (wr && (kpad_r < 9 * cout || cin_pad < cin)))
        return M2D_ERR_ARG;
    const long nw = (long)cout * kpad, nr = wr ? (long)cin_pad * kpad_r : 0;
    const long tot = nw > nr ? nw : nr;
    hipLaunchKernelGGL(k_pack_weights, dim3(blocks256(tot)), dim3(256), 0, (hipStream_t)stream, w,
                       cout, cin, kpad, (__bf16 *)wb, (__bf16 *)wr, kpad_r, cin_pad);
    return rc_of(hipGetLastError());
}

extern "C" int m2d_unet_mse(const float *pred, int ps, const float *target, const float *mask,
                            long npix, float inv_count, void *grad, int gs, float *partial,
                            void *stream)
{
    if (!pred || !target || !mask || !grad || !partial || npix <= 0 || ps < 1 || gs < 1)
        return M2D_ERR_ARG;
    hipLaunchKernelGGL(k_mse, dim3(blocks256(npix)), dim3(256), 0, (hipStream_t)stream, pred, ps,
                       target, mask, npix, inv_count, (__bf16 *)grad, gs, partial);
    return rc_of(hipGetLastError());
}

extern "C" int m2d_unet_pack_input(const double *a, const double *b, int nx, int ny, int h, int w,
                                   double inv_s, void *x, int nb, void *stream)
{
    if (!a || !b || !x || nx <= 0 || ny <= 0 || h < ny || w < nx || nb <= 0) return M2D_ERR_ARG;
    hipLaunchKernelGGL(k_pack_input, dim3(blocks256((long)h * w * nb)), dim3(256), 0,
                       (hipStream_t)stream, a, b, nx, ny, h, w, inv_s, (__bf16 *)x, nb);
    return rc_of(hipGetLastError());
}

extern "C" int m2d_unet_unpack_output(const float *out, int os, int nx, int ny, int w, double s,
                                      double *var, void *stream)
{
    if (!out || !var || nx <= 0 || ny <= 0 || w < nx || os < 1) return M2D_ERR_ARG;
    hipLaunchKernelGGL(k_unpack_output, dim3(blocks256((long)nx * ny)), dim3(256), 0,
                       (hipStream_t)stream, out, os, nx, ny, w, s, var);
    return rc_of(hipGetLastError());
}
