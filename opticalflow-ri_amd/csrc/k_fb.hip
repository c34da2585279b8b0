// Farneback dense flow, one internal pyramid level (reference:
// Farneback_PyCL.py:461-603 with optical_flow_farneback.cl:72-429).
//
//   pre-blur (reflect-101 separable Gaussian) -> polynomial expansion of both
//   frames (5 coefficient planes each) -> matrices from the initial flow ->
//   iters x [window blur (Gaussian or box) of the 5 planes -> flow update ->
//   matrices update].
//
// The flow update and the matrices update touch the same pixel only, so they
// run as one fused kernel (the blurred planes are read once, the flow and the
// new planes written once).  Blurs are separable: a vertical pass into a
// scratch buffer and a horizontal pass staged through LDS.  Arithmetic is
// float32 as the OpenCL source is written, with its `acc += a*b`
// accumulations as explicit FMAs (OpenCL contracts them by default).
#include "ofri_internal.h"

#pragma clang fp contract(off)

namespace ofri {
namespace {

constexpr int kMaxHalf = 192;
constexpr int kMaxPoly = 8;

struct HalfKernel {
    float g[kMaxHalf + 1];
    int kh;
};

struct PolyConsts {
    float g[kMaxPoly + 1], xg[kMaxPoly + 1], xxg[kMaxPoly + 1];
    float ig[4];
    int n;
};

__device__ __forceinline__ int r101_low(int y, int last) { return abs(y) % (last + 1); }
__device__ __forceinline__ int r101_high(int y, int last) { return abs(last - abs(last - y)) % (last + 1); }
__device__ __forceinline__ int r101_col(int x, int last) { return r101_low(r101_high(x, last), last); }

// Vertical pass of the reflect-101 blur over `planes` stacked planes.
__global__ void fb_blur_v(const float* __restrict__ src, float* __restrict__ tmp, int H, int W,
                          HalfKernel k) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    int p = blockIdx.z;
    if (x >= W) return;
    const float* s = src + (size_t)p * H * W;
    float v = s[(size_t)y * W + x] * k.g[0];
    for (int j = 1; j <= k.kh; ++j) {
        float a = s[(size_t)r101_low(y - j, H - 1) * W + x] + s[(size_t)r101_high(y + j, H - 1) * W + x];
        v = fmaf(a, k.g[j], v);
    }
    tmp[(size_t)p * H * W + (size_t)y * W + x] = v;
}

// Horizontal pass: a 256-wide output segment with its reflect-101 halo in LDS.
__global__ __launch_bounds__(256) void fb_blur_h(const float* __restrict__ tmp, float* __restrict__ dst,
                                                 int H, int W, HalfKernel k) {
    __shared__ float row[256 + 2 * kMaxHalf];
    const int y = blockIdx.y, p = blockIdx.z;
    const int x0 = blockIdx.x * 256;
    const float* t = tmp + (size_t)p * H * W + (size_t)y * W;
    for (int i = threadIdx.x; i < 256 + 2 * k.kh; i += 256) {
        int xe = x0 + i - k.kh;
        row[i] = t[r101_col(xe, W - 1)];
    }
    __syncthreads();
    int x = x0 + threadIdx.x;
    if (x >= W) return;
    const float* r = row + threadIdx.x + k.kh;
    float res = r[0] * k.g[0];
    for (int i = 1; i <= k.kh; ++i) res = fmaf(r[-i] + r[i], k.g[i], res);
    dst[(size_t)p * H * W + (size_t)y * W + x] = res;
}

// The same two passes for a half size KH fixed at compile time (the
// Farnebäck window and pre-filter sizes), each output formed with the same
// multiply and FMA sequence as above.  Vertical: a thread takes kTY
// consecutive rows of one column and holds the kTY + 2 KH source values of
// its window in registers, each read once (reflect-101 index per source row:
// rows above the image only arise as y - j, rows below only as y + j).
// Horizontal: a thread forms 4 consecutive outputs of a 1024-wide segment
// from 16-byte LDS reads of its 4 + 2 KH window.
constexpr int kTY = 16, kHX = 1024;

template <int KH>
__global__ __launch_bounds__(256) void fb_blur_v_reg(const float* __restrict__ src, float* __restrict__ tmp, int H,
                                                     int W, HalfKernel k) {
    // XCD-aware order: vertically adjacent row blocks, which read 2 KH of
    // the same rows, run on one XCD (row block fastest within a run)
    const int gx = gridDim.x, gy = gridDim.y;
    const int lin = xcd_run((blockIdx.z * gy + blockIdx.y) * gx + blockIdx.x, gx * gy * gridDim.z);
    const int by = lin % gy, bx = (lin / gy) % gx, p = lin / (gy * gx);
    const int x = bx * 256 + threadIdx.x;
    const int y0 = by * kTY;
    if (x >= W) return;
    const float* s = src + (size_t)p * H * W + x;
    float w[kTY + 2 * KH];
    if (y0 >= KH && y0 + kTY + KH <= H) {
#pragma unroll
        for (int m = 0; m < kTY + 2 * KH; ++m) w[m] = s[(size_t)(y0 - KH + m) * W];
    } else {
#pragma unroll
        for (int m = 0; m < kTY + 2 * KH; ++m) {
            const int r = y0 - KH + m;
            w[m] = s[(size_t)(r < 0 ? r101_low(r, H - 1) : r101_high(r, H - 1)) * W];
        }
    }
    float* d = tmp + (size_t)p * H * W + x;
#pragma unroll
    for (int i = 0; i < kTY; ++i) {
        if (y0 + i >= H) break;
        float v = w[i + KH] * k.g[0];
#pragma unroll
        for (int j = 1; j <= KH; ++j) v = fmaf(w[i + KH - j] + w[i + KH + j], k.g[j], v);
        d[(size_t)(y0 + i) * W] = v;
    }
}

template <int KH>
__global__ __launch_bounds__(256) void fb_blur_h_reg(const float* __restrict__ tmp, float* __restrict__ dst, int H,
                                                     int W, HalfKernel k) {
    constexpr int WN = (4 + 2 * KH + 3) / 4;   // float4 reads per thread
    __shared__ __attribute__((aligned(16))) float row[kHX + 4 * WN];
    const int y = blockIdx.y, p = blockIdx.z;
    const int x0 = blockIdx.x * kHX;
    const float* t = tmp + (size_t)p * H * W + (size_t)y * W;
    const int n = kHX + 2 * KH < W - x0 + 2 * KH ? kHX + 2 * KH : W - x0 + 2 * KH;
    for (int i = threadIdx.x; i < n; i += 256) row[i] = t[r101_col(x0 + i - KH, W - 1)];
    __syncthreads();
    const int xl = 4 * threadIdx.x;
    if (x0 + xl >= W) return;
    float r[4 * WN];
#pragma unroll
    for (int q = 0; q < WN; ++q) {
        const float4 f = *(const float4*)(row + xl + 4 * q);
        r[4 * q] = f.x;
        r[4 * q + 1] = f.y;
        r[4 * q + 2] = f.z;
        r[4 * q + 3] = f.w;
    }
    float o[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        float res = r[e + KH] * k.g[0];
#pragma unroll
        for (int i = 1; i <= KH; ++i) res = fmaf(r[e + KH - i] + r[e + KH + i], k.g[i], res);
        o[e] = res;
    }
    float* d = dst + (size_t)p * H * W + (size_t)y * W + x0 + xl;
    if (x0 + xl + 4 <= W && ((uintptr_t)d & 15) == 0) {
        *(float4*)d = float4{o[0], o[1], o[2], o[3]};
    } else {
        for (int e = 0; e < 4 && x0 + xl + e < W; ++e) d[e] = o[e];
    }
}

// Box window (updateFlowBoxFilter): replicate borders, running sums
// v + (a + b) in float32 outward from the centre, vertical then horizontal,
// times 1 / (2 kh + 1)^2.  The vertical sums of clamped columns equal those of
// the edge columns, so the horizontal pass clamps its reads into the
// vertical results.
__global__ void fb_box_v(const float* __restrict__ src, float* __restrict__ tmp, int H, int W, int kh) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    int p = blockIdx.z;
    if (x >= W) return;
    const float* s = src + (size_t)p * H * W;
    float v = s[(size_t)y * W + x];
    for (int j = 1; j <= kh; ++j) {
        int ylo = y - j < 0 ? 0 : y - j, yhi = y + j > H - 1 ? H - 1 : y + j;
        v = v + (s[(size_t)ylo * W + x] + s[(size_t)yhi * W + x]);
    }
    tmp[(size_t)p * H * W + (size_t)y * W + x] = v;
}

__global__ __launch_bounds__(256) void fb_box_h(const float* __restrict__ tmp, float* __restrict__ dst, int H,
                                                int W, int kh, float inv) {
    __shared__ float row[256 + 2 * kMaxHalf];
    const int y = blockIdx.y, p = blockIdx.z;
    const int x0 = blockIdx.x * 256;
    const float* t = tmp + (size_t)p * H * W + (size_t)y * W;
    for (int i = threadIdx.x; i < 256 + 2 * kh; i += 256) {
        int xe = x0 + i - kh;
        row[i] = t[xe < 0 ? 0 : (xe > W - 1 ? W - 1 : xe)];
    }
    __syncthreads();
    int x = x0 + threadIdx.x;
    if (x >= W) return;
    const float* r = row + threadIdx.x + kh;
    float res = r[0];
    for (int i = 1; i <= kh; ++i) res = res + (r[-i] + r[i]);
    dst[(size_t)p * H * W + (size_t)y * W + x] = res * inv;
}

// Polynomial expansion: vertical 3-channel pass into LDS (clamped rows and
// columns), horizontal pass to the 5 planes.
__global__ __launch_bounds__(256) void fb_polyexp_k(const float* __restrict__ src, int H, int W,
                                                    PolyConsts c, float* __restrict__ dst) {
    __shared__ float r0[256 + 2 * kMaxPoly], r1[256 + 2 * kMaxPoly], r2[256 + 2 * kMaxPoly];
    const int y = blockIdx.y, n = c.n;
    const int x0 = blockIdx.x * 256;
    for (int i = threadIdx.x; i < 256 + 2 * n; i += 256) {
        int xe = x0 + i - n;
        xe = xe < 0 ? 0 : (xe > W - 1 ? W - 1 : xe);
        float a = src[(size_t)y * W + xe] * c.g[0], b = 0.0f, d = 0.0f;
        for (int k = 1; k <= n; ++k) {
            float t0 = src[(size_t)(y - k < 0 ? 0 : y - k) * W + xe];
            float t1 = src[(size_t)(y + k > H - 1 ? H - 1 : y + k) * W + xe];
            a = fmaf(c.g[k], t0 + t1, a);
            b = fmaf(c.xg[k], t1 - t0, b);
            d = fmaf(c.xxg[k], t0 + t1, d);
        }
        r0[i] = a;
        r1[i] = b;
        r2[i] = d;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x;
    if (x >= W) return;
    const int ci = threadIdx.x + n;
    float b1 = c.g[0] * r0[ci], b3 = c.g[0] * r1[ci], b5 = c.g[0] * r2[ci];
    float b2 = 0.0f, b4 = 0.0f, b6 = 0.0f;
    for (int k = 1; k <= n; ++k) {
        // columns past the image edge hold the clamped column's sums
        float p0 = r0[ci + k], m0 = r0[ci - k];
        float p1 = r1[ci + k], m1 = r1[ci - k];
        float p2 = r2[ci + k], m2 = r2[ci - k];
        b1 = fmaf(p0 + m0, c.g[k], b1);
        b4 = fmaf(p0 + m0, c.xxg[k], b4);
        b2 = fmaf(p0 - m0, c.xg[k], b2);
        b3 = fmaf(p1 + m1, c.g[k], b3);
        b6 = fmaf(p1 - m1, c.xg[k], b6);
        b5 = fmaf(p2 + m2, c.g[k], b5);
    }
    const size_t N = (size_t)H * W, q = (size_t)y * W + x;
    dst[q] = b3 * c.ig[0];
    dst[N + q] = b2 * c.ig[0];
    dst[2 * N + q] = fmaf(b1, c.ig[1], b5 * c.ig[2]);
    dst[3 * N + q] = fmaf(b1, c.ig[1], b4 * c.ig[2]);
    dst[4 * N + q] = b6 * c.ig[3];
}

__constant__ float c_border[6] = {0.14f, 0.14f, 0.4472f, 0.4472f, 0.4472f, 1.f};

// updateMatrices at one pixel from its flow (dx, dy).
__device__ __forceinline__ void fb_matrices_at(int x, int y, int H, int W, float dx, float dy,
                                               const float* __restrict__ R0,
                                               const float* __restrict__ R1,
                                               float* __restrict__ M) {
    const size_t N = (size_t)H * W, q = (size_t)y * W + x;
    float fx = (float)x + dx, fy = (float)y + dy;
    int x1 = (int)floorf(fx), y1 = (int)floorf(fy);
    fx -= (float)x1;
    fy -= (float)y1;
    float r2, r3, r4, r5, r6;
    if (x1 >= 0 && y1 >= 0 && x1 < W - 1 && y1 < H - 1) {
        float a00 = (1.f - fx) * (1.f - fy), a01 = fx * (1.f - fy), a10 = (1.f - fx) * fy, a11 = fx * fy;
        float r[5];
        const size_t b = (size_t)y1 * W + x1;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const float* P = R1 + (size_t)k * N;
            float s = a00 * P[b];
            s = fmaf(a01, P[b + 1], s);
            s = fmaf(a10, P[b + W], s);
            s = fmaf(a11, P[b + W + 1], s);
            r[k] = s;
        }
        r2 = r[0];
        r3 = r[1];
        r4 = (R0[2 * N + q] + r[2]) * 0.5f;
        r5 = (R0[3 * N + q] + r[3]) * 0.5f;
        r6 = (R0[4 * N + q] + r[4]) * 0.25f;
    } else {
        r2 = r3 = 0.f;
        r4 = R0[2 * N + q];
        r5 = R0[3 * N + q];
        r6 = R0[4 * N + q] * 0.5f;
    }
    r2 = (R0[q] - r2) * 0.5f;
    r3 = (R0[N + q] - r3) * 0.5f;
    r2 = r2 + fmaf(r4, dy, r6 * dx);
    r3 = r3 + fmaf(r6, dy, r5 * dx);
    float sc = c_border[x < 5 ? x : 5] * c_border[y < 5 ? y : 5];
    sc = sc * c_border[W - x - 1 < 5 ? W - x - 1 : 5];
    sc = sc * c_border[H - y - 1 < 5 ? H - y - 1 : 5];
    r2 *= sc; r3 *= sc; r4 *= sc; r5 *= sc; r6 *= sc;
    M[q] = fmaf(r4, r4, r6 * r6);
    M[N + q] = (r4 + r5) * r6;
    M[2 * N + q] = fmaf(r5, r5, r6 * r6);
    M[3 * N + q] = fmaf(r4, r2, r6 * r3);
    M[4 * N + q] = fmaf(r6, r2, r5 * r3);
}

__global__ void fb_matrices_k(const float* __restrict__ flx, const float* __restrict__ fly, int H, int W,
                              const float* __restrict__ R0, const float* __restrict__ R1,
                              float* __restrict__ M) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= W) return;
    size_t q = (size_t)y * W + x;
    fb_matrices_at(x, y, H, W, flx[q], fly[q], R0, R1, M);
}

// updateFlow from the blurred planes Mb, then (when `matrices`) the
// matrices of the new flow into M.
__global__ void fb_flow_k(const float* __restrict__ Mb, float* __restrict__ flx, float* __restrict__ fly,
                          int H, int W, const float* __restrict__ R0, const float* __restrict__ R1,
                          float* __restrict__ M, int matrices) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= W) return;
    const size_t N = (size_t)H * W, q = (size_t)y * W + x;
    float g11 = Mb[q], g12 = Mb[N + q], g22 = Mb[2 * N + q], h1 = Mb[3 * N + q], h2 = Mb[4 * N + q];
    float det = fmaf(g11, g22, -(g12 * g12)) + 1e-3f;
    float inv = 1.0f / det;
    float dx = fmaf(g11, h2, -(g12 * h1)) * inv;
    float dy = fmaf(g22, h1, -(g12 * h2)) * inv;
    flx[q] = dx;
    fly[q] = dy;
    if (matrices) fb_matrices_at(x, y, H, W, dx, dy, R0, R1, M);
}

// curFlow *= scale of the level loop, float32 products.
__global__ void fb_scale_k(float* __restrict__ fx, float* __restrict__ fy, size_t n, float s) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        fx[i] = fx[i] * s;
        fy[i] = fy[i] * s;
    }
}

template <int KH>
void blur_reg(ofri_ctx* ctx, const float* src, float* tmp, float* dst, int H, int W, int planes,
              const HalfKernel& k) {
    fb_blur_v_reg<KH><<<dim3((W + 255) / 256, (H + kTY - 1) / kTY, planes), 256, 0, ctx->stream>>>(src, tmp, H, W,
                                                                                                  k);
    fb_blur_h_reg<KH><<<dim3((W + kHX - 1) / kHX, H, planes), 256, 0, ctx->stream>>>(tmp, dst, H, W, k);
}

int blur(ofri_ctx* ctx, const float* src, float* dst, int H, int W, int planes, const HalfKernel& k) {
    void* tmp;
    OFRI_TRY(scratch(ctx, "fb_blur_tmp", (size_t)planes * H * W * sizeof(float), &tmp));
    // the pre-filter (3 taps) and the Farnebäck windows of 21 and 33 taps
    // have register-window passes; other sizes take the general ones
    if (k.kh == 1) {
        blur_reg<1>(ctx, src, (float*)tmp, dst, H, W, planes, k);
    } else if (k.kh == 10) {
        blur_reg<10>(ctx, src, (float*)tmp, dst, H, W, planes, k);
    } else if (k.kh == 16) {
        blur_reg<16>(ctx, src, (float*)tmp, dst, H, W, planes, k);
    } else {
        fb_blur_v<<<dim3((W + 255) / 256, H, planes), 256, 0, ctx->stream>>>(src, (float*)tmp, H, W, k);
        fb_blur_h<<<dim3((W + 255) / 256, H, planes), 256, 0, ctx->stream>>>((const float*)tmp, dst, H, W, k);
    }
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

int box(ofri_ctx* ctx, const float* src, float* dst, int H, int W, int planes, int kh) {
    void* tmp;
    OFRI_TRY(scratch(ctx, "fb_blur_tmp", (size_t)planes * H * W * sizeof(float), &tmp));
    const float inv = 1.0f / (float)((1 + 2 * kh) * (1 + 2 * kh));
    fb_box_v<<<dim3((W + 255) / 256, H, planes), 256, 0, ctx->stream>>>(src, (float*)tmp, H, W, kh);
    fb_box_h<<<dim3((W + 255) / 256, H, planes), 256, 0, ctx->stream>>>((const float*)tmp, dst, H, W, kh, inv);
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

}  // namespace

// The reflect-101 index maps reduce modulo the side, so any half size up to
// kMaxHalf stays in bounds, also when it exceeds the image (the reference's
// index helpers behave the same way).
int fb_blur(ofri_ctx* ctx, const float* src, float* dst, int H, int W, const float* half, int kh) {
    if (kh < 0 || kh > kMaxHalf) return fail(OFRI_E_ARG, "fb_blur: kernel half size %d outside 0..%d", kh, kMaxHalf);
    if (H < 1 || W < 1) return fail(OFRI_E_ARG, "fb_blur: image %dx%d", H, W);
    HalfKernel k{};
    k.kh = kh;
    for (int i = 0; i <= kh; ++i) k.g[i] = half[i];
    return blur(ctx, src, dst, H, W, 1, k);
}

int farneback(ofri_ctx* ctx, const float* im1, const float* im2, int H, int W, float* flx, float* fly,
              const ofri_fb_params* p) {
    if (p->poly_n < 1 || p->poly_n > kMaxPoly) return fail(OFRI_E_ARG, "farneback: polyN %d", p->poly_n);
    if (p->pre_kh < 0 || p->pre_kh > kMaxHalf || p->win_kh < 0 || p->win_kh > kMaxHalf)
        return fail(OFRI_E_ARG, "farneback: kernel half sizes %d/%d above %d", p->pre_kh, p->win_kh, kMaxHalf);
    if (H < 2 || W < 2) return fail(OFRI_E_ARG, "farneback: image %dx%d too small", H, W);
    const size_t N = (size_t)H * W;
    void* ws;
    OFRI_TRY(scratch(ctx, "fb_ws", 21 * N * sizeof(float), &ws));
    float* bl = (float*)ws;
    float* RA = bl + N;
    float* RB = RA + 5 * N;
    float* M = RB + 5 * N;
    float* Mb = M + 5 * N;
    HalfKernel pk{}, wk{};
    pk.kh = p->pre_kh;
    wk.kh = p->win_kh;
    for (int i = 0; i <= pk.kh; ++i) pk.g[i] = p->pre_kernel[i];
    if (!p->box_filter)
        for (int i = 0; i <= wk.kh; ++i) wk.g[i] = p->win_kernel[i];
    PolyConsts c{};
    c.n = p->poly_n;
    for (int i = 0; i <= c.n; ++i) {
        c.g[i] = p->g[i];
        c.xg[i] = p->xg[i];
        c.xxg[i] = p->xxg[i];
    }
    for (int i = 0; i < 4; ++i) c.ig[i] = p->ig[i];
    dim3 g2((W + 255) / 256, H);
    if (p->flow_scale != 1.0f)
        fb_scale_k<<<(unsigned)((N + 255) / 256 < 4096 ? (N + 255) / 256 : 4096), 256, 0, ctx->stream>>>(flx, fly, N, p->flow_scale);
    OFRI_TRY(blur(ctx, im1, bl, H, W, 1, pk));
    fb_polyexp_k<<<g2, 256, 0, ctx->stream>>>(bl, H, W, c, RA);
    OFRI_TRY(blur(ctx, im2, bl, H, W, 1, pk));
    fb_polyexp_k<<<g2, 256, 0, ctx->stream>>>(bl, H, W, c, RB);
    fb_matrices_k<<<g2, 256, 0, ctx->stream>>>(flx, fly, H, W, RA, RB, M);
    OFRI_HIP(hipGetLastError());
    for (int i = 0; i < p->iters; ++i) {
        if (p->box_filter)
            OFRI_TRY(box(ctx, M, Mb, H, W, 5, p->win_kh));
        else
            OFRI_TRY(blur(ctx, M, Mb, H, W, 5, wk));
        fb_flow_k<<<g2, 256, 0, ctx->stream>>>(Mb, flx, fly, H, W, RA, RB, M, i < p->iters - 1);
        OFRI_HIP(hipGetLastError());
    }
    return OFRI_OK;
}

}  // namespace ofri
