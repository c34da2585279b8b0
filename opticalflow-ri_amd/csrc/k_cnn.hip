// Warm-start network: a 4-level convolutional encoder-decoder on MFMA.
//
// Tensors are NHWC bfloat16 (stored as raw 16-bit words) with a per-pixel
// stride, so a skip connection is a channel slice of its concatenation
// buffer and concatenation costs nothing.  Down-sampling is space-to-depth
// (2x2 pixels -> 4x channels) folded into the next convolution's operand
// load; up-sampling is depth-to-space (pixel shuffle) folded into the
// producing convolution's store.
//
// conv3_k: 3x3 'same' convolution as an implicit GEMM on
// v_mfma_f32_16x16x32_bf16.  M = output pixels, N = output channels,
// K = (tap, input channel) with the channel fastest, padded to 32.  A wave
// owns MT x 16 consecutive pixels and NT x 16 channels; every lane of a
// 16-lane group supplies 8 consecutive K values of one pixel (one 16-byte
// load, channels of one tap) and of one weight row.  Epilogue: bias,
// leaky ReLU (slope 0.1), optional accumulation into the destination, store
// as bfloat16 or float32.  The same kernel computes the data gradient of a
// convolution: the gradient is a convolution of the output gradient (gated
// by the activation's derivative, read from the forward output) with the
// weights flipped in space and transposed in channels.
//
// dw_k: weight and bias gradients, sum over pixels of gated output gradient
// times shifted input, in float32 on the vector ALUs from LDS-staged tiles,
// partial sums added atomically (training only).
#include <cmath>
#include <random>

#include "ofri_internal.h"

#pragma clang fp contract(off)

namespace ofri {

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

constexpr float kSlope = 0.1f;   // leaky ReLU

__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((unsigned)h << 16); }
// float -> bfloat16, round to nearest even (finite inputs)
__device__ __forceinline__ u16 f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}

// A tensor as one convolution sees it: conv-space size B x H x W x C.
// mode 1 reads it through space-to-depth from a (2H) x (2W) x (C/4) source
// (channel c' = sub * C/4 + c, sub = 2 * dy + dx) or, as a destination,
// writes depth-to-space into one.
struct View {
    void* p;
    long long img;   // elements per image of the stored tensor
    int ps;          // elements per stored pixel
    int f32;         // 1: float32 elements, 0: bfloat16
    int mode;        // 0 direct, 1 space-to-depth read / depth-to-space write
};

__device__ __forceinline__ long long view_off(const View& v, int b, int y, int x, int c, int H, int W,
                                              int C) {
    if (v.mode == 0) return (long long)b * v.img + ((long long)y * W + x) * v.ps + c;
    const int Cs = C >> 2;
    const int sub = c / Cs, cc = c - sub * Cs;
    const int sy = 2 * y + (sub >> 1), sx = 2 * x + (sub & 1);
    return (long long)b * v.img + ((long long)sy * (2 * W) + sx) * v.ps + cc;
}

// 8 consecutive conv-space channels c0..c0+7 (c0 % 8 == 0, and 8 channels
// never straddle a space-to-depth group) as floats
__device__ __forceinline__ void load8f(const View& v, long long off, float* o) {
    if (v.f32) {
        const float4* q = (const float4*)((const float*)v.p + off);
        float4 a = q[0], b = q[1];
        o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
        o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
    } else {
        uint4 r = *(const uint4*)((const u16*)v.p + off);
        unsigned w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            o[2 * i] = bf2f((u16)(w[i] & 0xffffu));
            o[2 * i + 1] = bf2f((u16)(w[i] >> 16));
        }
    }
}

struct ConvArgs {
    View x;            // input (conv space B x H x W x Cin)
    View gate;         // optional: x is multiplied by the leaky-ReLU derivative at gate (p == null: none)
    const u16* w;      // weights [rows][Kp], row = output channel, k = tap * Cin + ci
    int Kp;
    const float* bias; // [Cout] or null
    View y;            // output (conv space B x H x W x Cout)
    int B, H, W, Cin, Cout;
    int act;           // 1: leaky ReLU
    int beta;          // 1: add into the destination
    // non-null (16 output channels, channel = sub * 4 + field): instead of
    // storing y, fields 0 and 1 are added depth-to-space into these two
    // float32 planes of size 2H x 2W (fields 2, 3 are dropped)
    float* add_u;
    float* add_v;
};

template <int MT, int NT>
__global__ __launch_bounds__(256) void conv3_k(ConvArgs a) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int H = a.H, W = a.W, HW = H * W;
    const long long P = (long long)a.B * HW;
    const long long pbase = ((long long)blockIdx.x * 4 + wv) * (MT * 16);
    const int cobase = blockIdx.y * (NT * 16);
    int pb[MT], py[MT], px[MT];
    bool pv[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const long long p = pbase + mt * 16 + r16;
        pv[mt] = p < P;
        const long long pc = pv[mt] ? p : 0;
        pb[mt] = (int)(pc / HW);
        const int rem = (int)(pc - (long long)pb[mt] * HW);
        py[mt] = rem / W;
        px[mt] = rem - py[mt] * W;
    }
    f4 acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f4{0.0f, 0.0f, 0.0f, 0.0f};

    const bool plain_x = !a.x.f32 && !a.gate.p;
    const int nk = a.Kp >> 5;
    for (int kc = 0; kc < nk; ++kc) {
        const int k0 = kc * 32 + g * 8;
        const int tap = k0 / a.Cin, ci0 = k0 - tap * a.Cin;
        const int dy = tap / 3 - 1, dx = tap - 3 * (tap / 3) - 1;
        bf8 bfr[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int co = cobase + nt * 16 + r16;
            bfr[nt] = __builtin_bit_cast(bf8, *(const uint4*)(a.w + (size_t)co * a.Kp + k0));
        }
        bf8 afr[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            uint4 raw = {0u, 0u, 0u, 0u};
            const int yy = py[mt] + dy, xx = px[mt] + dx;
            if (tap < 9 && pv[mt] && yy >= 0 && yy < H && xx >= 0 && xx < W) {
                const long long off = view_off(a.x, pb[mt], yy, xx, ci0, H, W, a.Cin);
                if (plain_x) {
                    raw = *(const uint4*)((const u16*)a.x.p + off);
                } else {
                    float f[8];
                    load8f(a.x, off, f);
                    if (a.gate.p) {
                        float gt[8];
                        load8f(a.gate, view_off(a.gate, pb[mt], yy, xx, ci0, H, W, a.Cin), gt);
#pragma unroll
                        for (int i = 0; i < 8; ++i) f[i] = gt[i] > 0.0f ? f[i] : f[i] * kSlope;
                    }
                    raw.x = (unsigned)f2bf(f[0]) | ((unsigned)f2bf(f[1]) << 16);
                    raw.y = (unsigned)f2bf(f[2]) | ((unsigned)f2bf(f[3]) << 16);
                    raw.z = (unsigned)f2bf(f[4]) | ((unsigned)f2bf(f[5]) << 16);
                    raw.w = (unsigned)f2bf(f[6]) | ((unsigned)f2bf(f[7]) << 16);
                }
            }
            afr[mt] = __builtin_bit_cast(bf8, raw);
        }
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[mt], bfr[nt], acc[mt][nt], 0, 0, 0);
    }

    // D layout: lane holds rows 4g..4g+3 (pixels) of column r16 (channel)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int co = cobase + nt * 16 + r16;
        if (co >= a.Cout) continue;
        const float bv = a.bias ? a.bias[co] : 0.0f;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const long long p = pbase + mt * 16 + 4 * g + i;
                if (p >= P) continue;
                const int b = (int)(p / HW);
                const int rem = (int)(p - (long long)b * HW);
                const int y = rem / W, x = rem - (rem / W) * W;
                float v = acc[mt][nt][i] + bv;
                if (a.act) v = v > 0.0f ? v : v * kSlope;
                const long long off = view_off(a.y, b, y, x, co, H, W, a.Cout);
                if (a.y.f32) {
                    float* d = (float*)a.y.p + off;
                    if (a.beta) v = *d + v;
                    *d = v;
                } else {
                    u16* d = (u16*)a.y.p + off;
                    if (a.beta) v = bf2f(*d) + v;
                    *d = f2bf(v);
                }
            }
    }
}

// LDS-tiled form of conv3_k.  A workgroup (4 wavefronts) owns a TS x TS
// output tile of one image (TS = 16, or 8 for the small low-resolution
// levels, which would otherwise leave most CUs idle) and NT x 16 output
// channels; wavefront w computes pixels [16 MT w, 16 MT (w + 1)) of the tile
// in row-major order, one 16-pixel m-tile each.  Input channels go in chunks
// of 32: the (TS + 2)^2 input patch of the chunk (halo included, zero outside
// the image, gated / converted / space-to-depth-read on the way in) and the
// chunk's weights for all 9 taps are staged in LDS once, then every tap's
// A and B fragments are 16-byte LDS reads and 9 x MT x NT MFMAs run per
// wavefront per chunk.  Pixel rows are padded to 40 bf16 (80 bytes) and
// weight rows to 296 bf16 so that the 16 lanes of a fragment read spread
// over the LDS banks.
constexpr int kPixStride = 40;   // bf16 per staged pixel (32 used)
constexpr int kCoStride = 296;   // bf16 per staged weight row (9 * 32 used)

template <int TS, int NT>
__global__ __launch_bounds__(256) void conv3_lds_k(ConvArgs a) {
    constexpr int PS = TS + 2;                 // staged patch side
    constexpr int MT = TS * TS / 64;           // m-tiles per wavefront
    __shared__ __attribute__((aligned(16))) u16 sx[PS * PS * kPixStride];
    __shared__ __attribute__((aligned(16))) u16 sw[NT * 16 * kCoStride];
    const int t = threadIdx.x;
    const int lane = t & 63, wv = t >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int H = a.H, W = a.W;
    const int tiles_x = (W + TS - 1) / TS, tiles_y = (H + TS - 1) / TS;
    // XCD-aware order: a workgroup's run of (tile, channel group) pairs is
    // contiguous with the channel group fastest, so the groups that stage the
    // same patch, and neighbouring tiles, share one XCD's L2
    const int lin = xcd_run(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    int tile = lin / gridDim.y;
    const int b = tile / (tiles_x * tiles_y);
    tile -= b * tiles_x * tiles_y;
    const int y0 = (tile / tiles_x) * TS, x0 = (tile % tiles_x) * TS;
    const int cobase = (lin - (lin / gridDim.y) * gridDim.y) * (NT * 16);

    f4 acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f4{0.0f, 0.0f, 0.0f, 0.0f};

    // the MFMAs of one staged 32-channel chunk: 9 taps
    auto chunk_mfma = [&]() {
#pragma unroll 3
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap - 3 * (tap / 3);   // patch offsets 0..2
            bf8 bfr[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                bfr[nt] = __builtin_bit_cast(bf8, *(const uint4*)(sw + (nt * 16 + r16) * kCoStride + tap * 32 + g * 8));
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const int p = (wv * MT + mt) * 16 + r16;
                const int py = p / TS + dy, px = p % TS + dx;
                const bf8 afr = __builtin_bit_cast(bf8, *(const uint4*)(sx + (py * PS + px) * kPixStride + g * 8));
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr[nt], acc[mt][nt], 0, 0, 0);
            }
        }
    };

    const bool plain_x = !a.x.f32 && !a.gate.p;
    if (plain_x) {
        // bfloat16 input: the next chunk's patch and weights are fetched into
        // registers while the current chunk's MFMAs run
        constexpr int XN = (PS * PS * 4 + 255) / 256, WN = (NT * 16 * 36 + 255) / 256;
        uint4 rx[XN], rw[WN];
        // per-thread load offsets at channel 0, formed once: a direct input
        // and the weights then only add the chunk's first channel (-1: the
        // patch pixel lies outside the image or the slot is unused)
        const bool direct = a.x.mode == 0;
        long long xo[XN];
        int xpart[XN];
#pragma unroll
        for (int k = 0; k < XN; ++k) {
            const int idx = t + k * 256;
            xo[k] = -1;
            xpart[k] = (idx & 3) * 8;
            if (idx < PS * PS * 4) {
                const int pix = idx >> 2;
                const int py = pix / PS, px = pix - py * PS;
                const int gy = y0 - 1 + py, gx = x0 - 1 + px;
                if (gy >= 0 && gy < H && gx >= 0 && gx < W)
                    xo[k] = direct ? view_off(a.x, b, gy, gx, xpart[k], H, W, a.Cin)
                                   : (((long long)gy << 32) | (unsigned)gx);
            }
        }
        long long wo[WN];
        int wpart[WN];
#pragma unroll
        for (int k = 0; k < WN; ++k) {
            const int idx = t + k * 256;
            wo[k] = -1;
            wpart[k] = 0;
            if (idx < NT * 16 * 36) {
                const int co = idx / 36, rem = idx - co * 36;
                const int tap = rem >> 2, part = rem & 3;
                wpart[k] = part * 8;
                wo[k] = (long long)(cobase + co) * a.Kp + tap * a.Cin + part * 8;
            }
        }
        auto fetch = [&](int c0) {
#pragma unroll
            for (int k = 0; k < XN; ++k) {
                rx[k] = uint4{0u, 0u, 0u, 0u};
                if (xo[k] >= 0 && c0 + xpart[k] < a.Cin) {
                    long long off;
                    if (direct) off = xo[k] + c0;
                    else off = view_off(a.x, b, (int)(xo[k] >> 32), (int)(xo[k] & 0xffffffff), c0 + xpart[k], H, W,
                                        a.Cin);
                    rx[k] = *(const uint4*)((const u16*)a.x.p + off);
                }
            }
#pragma unroll
            for (int k = 0; k < WN; ++k) {
                rw[k] = uint4{0u, 0u, 0u, 0u};
                if (wo[k] >= 0 && c0 + wpart[k] < a.Cin) rw[k] = *(const uint4*)(a.w + wo[k] + c0);
            }
        };
        auto put = [&]() {
#pragma unroll
            for (int k = 0; k < XN; ++k) {
                const int idx = t + k * 256;
                if (idx < PS * PS * 4) *(uint4*)(sx + (idx >> 2) * kPixStride + (idx & 3) * 8) = rx[k];
            }
#pragma unroll
            for (int k = 0; k < WN; ++k) {
                const int idx = t + k * 256;
                if (idx < NT * 16 * 36) {
                    const int co = idx / 36, rem = idx - co * 36;
                    *(uint4*)(sw + co * kCoStride + (rem >> 2) * 32 + (rem & 3) * 8) = rw[k];
                }
            }
        };
        fetch(0);
        for (int c0 = 0; c0 < a.Cin; c0 += 32) {
            put();
            __syncthreads();
            if (c0 + 32 < a.Cin) fetch(c0 + 32);
            chunk_mfma();
            __syncthreads();
        }
    } else {
        for (int c0 = 0; c0 < a.Cin; c0 += 32) {
            // stage the input patch: PS x PS pixels x 4 groups of 8 channels
            for (int idx = t; idx < PS * PS * 4; idx += 256) {
                const int pix = idx >> 2, part = idx & 3;
                const int py = pix / PS, px = pix - py * PS;
                const int gy = y0 - 1 + py, gx = x0 - 1 + px;
                const int ch = c0 + part * 8;
                uint4 raw = {0u, 0u, 0u, 0u};
                if (gy >= 0 && gy < H && gx >= 0 && gx < W && ch < a.Cin) {
                    const long long off = view_off(a.x, b, gy, gx, ch, H, W, a.Cin);
                    float f[8];
                    load8f(a.x, off, f);
                    if (a.gate.p) {
                        float gt[8];
                        load8f(a.gate, view_off(a.gate, b, gy, gx, ch, H, W, a.Cin), gt);
#pragma unroll
                        for (int i = 0; i < 8; ++i) f[i] = gt[i] > 0.0f ? f[i] : f[i] * kSlope;
                    }
                    raw.x = (unsigned)f2bf(f[0]) | ((unsigned)f2bf(f[1]) << 16);
                    raw.y = (unsigned)f2bf(f[2]) | ((unsigned)f2bf(f[3]) << 16);
                    raw.z = (unsigned)f2bf(f[4]) | ((unsigned)f2bf(f[5]) << 16);
                    raw.w = (unsigned)f2bf(f[6]) | ((unsigned)f2bf(f[7]) << 16);
                }
                *(uint4*)(sx + pix * kPixStride + part * 8) = raw;
            }
            // stage the weights: NT*16 rows x 9 taps x 4 groups of 8 channels
            for (int idx = t; idx < NT * 16 * 36; idx += 256) {
                const int co = idx / 36, rem = idx - co * 36;
                const int tap = rem >> 2, part = rem & 3;
                const int ch = c0 + part * 8;
                uint4 raw = {0u, 0u, 0u, 0u};
                if (ch < a.Cin) raw = *(const uint4*)(a.w + (size_t)(cobase + co) * a.Kp + tap * a.Cin + ch);
                *(uint4*)(sw + co * kCoStride + tap * 32 + part * 8) = raw;
            }
            __syncthreads();
            chunk_mfma();
            __syncthreads();
        }
    }

    // D layout: lane holds m-tile pixels 4g..4g+3, channel r16.  The
    // channel's part of the destination offset (and, through depth-to-space,
    // its sub-pixel) is formed once per channel; pixels add 32-bit offsets.
    // A correction head (add_u set) adds its fields 0, 1 into the flow
    // planes; a bfloat16 output written (not accumulated) from a tile inside
    // the image takes a straight-line path.
    if (a.add_u) {
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int co = cobase + nt * 16 + r16;
            if (co >= a.Cout || (co & 3) >= 2) continue;
            const float bv = a.bias ? a.bias[co] : 0.0f;
            const int sub = co >> 2, dy = sub >> 1, dx = sub & 1;
            float* pl = (co & 1) ? a.add_v : a.add_u;
            const int Wd = W << 1;
            // every flow value this lane updates is read before any is
            // written (distinct pixels), so the reads' latencies overlap
            float old[MT][4];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int p = (wv * MT + mt) * 16 + 4 * g + i;
                    const int y = y0 + p / TS, x = x0 + p % TS;
                    old[mt][i] = 0.0f;
                    if (y < H && x < W)
                        old[mt][i] = pl[(long long)b * 4 * H * W + ((y << 1) + dy) * Wd + ((x << 1) + dx)];
                }
            }
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int p = (wv * MT + mt) * 16 + 4 * g + i;
                    const int y = y0 + p / TS, x = x0 + p % TS;
                    if (y >= H || x >= W) continue;
                    float v = acc[mt][nt][i] + bv;
                    if (a.act) v = v > 0.0f ? v : v * kSlope;
                    pl[(long long)b * 4 * H * W + ((y << 1) + dy) * Wd + ((x << 1) + dx)] = old[mt][i] + v;
                }
            }
        }
        return;
    }
    if (!a.y.f32 && !a.beta && y0 + TS <= H && x0 + TS <= W) {
        const bool act = a.act;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int co = cobase + nt * 16 + r16;
            if (co >= a.Cout) continue;
            const float bv = a.bias ? a.bias[co] : 0.0f;
            int cc = co, dy = 0, dx = 0, sh = 0;
            if (a.y.mode) {
                const int Cs = a.Cout >> 2;
                const int sub = co / Cs;
                cc = co - sub * Cs;
                dy = sub >> 1;
                dx = sub & 1;
                sh = 1;
            }
            const int Wd = W << sh;
            u16* dbase = (u16*)a.y.p + ((long long)b * a.y.img + cc);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int p = (wv * MT + mt) * 16 + 4 * g + i;
                    const int y = y0 + p / TS, x = x0 + p % TS;
                    float v = acc[mt][nt][i] + bv;
                    v = act ? (v > 0.0f ? v : v * kSlope) : v;
                    dbase[(((y << sh) + dy) * Wd + ((x << sh) + dx)) * a.y.ps] = f2bf(v);
                }
            }
        }
        return;
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int co = cobase + nt * 16 + r16;
        if (co >= a.Cout) continue;
        const float bv = a.bias ? a.bias[co] : 0.0f;
        int cc = co, dy = 0, dx = 0, sh = 0;
        if (a.y.mode) {
            const int Cs = a.Cout >> 2;
            const int sub = co / Cs;
            cc = co - sub * Cs;
            dy = sub >> 1;
            dx = sub & 1;
            sh = 1;
        }
        const int Wd = W << sh;
        const long long base_off = (long long)b * a.y.img + cc;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int p = (wv * MT + mt) * 16 + 4 * g + i;
                const int y = y0 + p / TS, x = x0 + p % TS;
                if (y >= H || x >= W) continue;
                float v = acc[mt][nt][i] + bv;
                if (a.act) v = v > 0.0f ? v : v * kSlope;
                const int po = (((y << sh) + dy) * Wd + ((x << sh) + dx)) * a.y.ps;
                if (a.y.f32) {
                    float* d = (float*)a.y.p + base_off + po;
                    if (a.beta) v = *d + v;
                    *d = v;
                } else {
                    u16* d = (u16*)a.y.p + base_off + po;
                    if (a.beta) v = bf2f(*d) + v;
                    *d = f2bf(v);
                }
            }
        }
    }
}

// Weight / bias gradients of one convolution.  Block (kt, ct, chunk): a
// 64-wide slice of K, a 32-wide slice of output channels, and `chunk` pixels
// in stages of 64.  dZ = dY gated by the activation derivative (from the
// forward output) is staged as float in LDS, the shifted input likewise;
// every thread accumulates 2 channels x 4 K values.
struct DwArgs {
    View x;      // forward input (conv space B x H x W x Cin)
    View dy;     // output gradient (conv space B x H x W x Cout)
    View gate;   // forward output (null: no activation)
    float* dw;   // [rows][Kp] (float32, accumulated)
    float* db;   // [Cout] (accumulated)
    int Kp;
    int B, H, W, Cin, Cout;
    int chunk;   // pixels per block (multiple of 64)
};

__global__ __launch_bounds__(256) void dw_k(DwArgs a) {
    __shared__ float sz[64][33];
    __shared__ float sx[64][65];
    const int t = threadIdx.x;
    const int k0 = blockIdx.x * 64, c0 = blockIdx.y * 32;
    const int H = a.H, W = a.W, HW = H * W;
    const long long P = (long long)a.B * HW;
    const long long p0 = (long long)blockIdx.z * a.chunk;
    const int K = 9 * a.Cin;
    const int tc = (t & 7) * 4;        // 4 K values
    const int tr = (t >> 3);           // 0..31: channel pair index -> channels 2*(tr%16)... see below
    const int ch0 = (tr & 15) * 2;     // 2 channels
    const int kq = tc + (tr >> 4) * 32;  // K offset (0..60)
    float acc[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    float bacc[2] = {0, 0};
    for (long long ps = p0; ps < p0 + a.chunk && ps < P; ps += 64) {
        // stage dZ: 64 pixels x 32 channels, 8-channel groups
        {
            const int pp = t >> 2, cg = (t & 3) * 8;
            const long long p = ps + pp;
            float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (p < P && c0 + cg < a.Cout) {
                const int b = (int)(p / HW), rem = (int)(p - (long long)b * HW);
                const int y = rem / W, x = rem - (rem / W) * W;
                load8f(a.dy, view_off(a.dy, b, y, x, c0 + cg, H, W, a.Cout), f);
                if (a.gate.p) {
                    float gt[8];
                    load8f(a.gate, view_off(a.gate, b, y, x, c0 + cg, H, W, a.Cout), gt);
#pragma unroll
                    for (int i = 0; i < 8; ++i) f[i] = gt[i] > 0.0f ? f[i] : f[i] * kSlope;
                }
                // channels past Cout inside the group stay as loaded: their rows are never written
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) sz[pp][cg + i] = f[i];
        }
        // stage shifted input: 64 pixels x 64 K values, 8-value groups
#pragma unroll
        for (int rep = 0; rep < 2; ++rep) {
            const int item = t + rep * 256;
            const int pp = item >> 3, kg = (item & 7) * 8;
            const long long p = ps + pp;
            const int k = k0 + kg;
            float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (p < P && k < K) {
                const int tap = k / a.Cin, ci = k - tap * a.Cin;
                const int dy = tap / 3 - 1, dx = tap - 3 * (tap / 3) - 1;
                const int b = (int)(p / HW), rem = (int)(p - (long long)b * HW);
                const int y = rem / W + dy, x = rem - (rem / W) * W + dx;
                if (y >= 0 && y < H && x >= 0 && x < W) load8f(a.x, view_off(a.x, b, y, x, ci, H, W, a.Cin), f);
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) sx[pp][kg + i] = f[i];
        }
        __syncthreads();
#pragma unroll 8
        for (int pp = 0; pp < 64; ++pp) {
            const float z0 = sz[pp][ch0], z1 = sz[pp][ch0 + 1];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float xv = sx[pp][kq + j];
                acc[0][j] = fmaf(z0, xv, acc[0][j]);
                acc[1][j] = fmaf(z1, xv, acc[1][j]);
            }
            bacc[0] += z0;
            bacc[1] += z1;
        }
        __syncthreads();
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int co = c0 + ch0 + c;
        if (co >= a.Cout) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + kq + j;
            if (k < K) atomicAdd(a.dw + (size_t)co * a.Kp + k, acc[c][j]);
        }
        if (blockIdx.x == 0 && (tr >> 4) == 0 && (t & 7) == 0) atomicAdd(a.db + co, bacc[c]);
    }
}

// Wt[ci][tap' * Cout + co] = W[co][(8 - tap') * Cin + ci]: the weights of the
// data-gradient convolution (bfloat16, padding zero).
__global__ void wflip_k(const u16* w, int Kp, int Cin, int Cout, u16* wt, int rows_t, int Kpt) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)rows_t * Kpt) return;
    const int ci = (int)(i / Kpt), k = (int)(i - (long long)ci * Kpt);
    u16 v = 0;
    if (ci < Cin && k < 9 * Cout) {
        const int tp = k / Cout, co = k - tp * Cout;
        v = w[(size_t)co * Kp + (8 - tp) * Cin + ci];
    }
    wt[i] = v;
}

// Adam on every parameter (float32 master copy), bfloat16 copy of the
// weights refreshed; entries outside a layer's real shape have zero gradient
// and stay zero.
__global__ void adam_k(float* p, const float* gr, float* m, float* v, u16* pb, long long nb, long long n,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = gr[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    const float pn = p[i] - upd;
    p[i] = pn;
    if (i < nb) pb[i] = f2bf(pn);
}

__global__ void f2bf_k(const float* s, u16* d, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = f2bf(s[i]);
}

__global__ void zero_f_k(float* p, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// The network's input planes: plane p of batch item b starts at p[p] + b *
// bstride (a stacked [B][P][H][W] tensor, or separate planes when B = 1).
struct InPlanes {
    const float* p[4];
    long long bstride;
};

// Input planes -> space-to-depth bf16 NHWC at H/2 x W/2 (4P channels,
// channel = sub * P + plane, sub = 2 * dy + dx); planes 0, 1 (frames) scaled
// to (I - 128) / 64, the others (flow) as they are.
template <int P>
__global__ void net_input_k(InPlanes in, int H, int W, u16* x) {
    const int h = H / 2, w = W / 2;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= (long long)h * w) return;
    const int y = (int)(i / w), xx = (int)(i - (long long)y * w);
    u16 out[4 * P];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const long long q = (long long)(2 * y + (s >> 1)) * W + 2 * xx + (s & 1);
#pragma unroll
        for (int pl = 0; pl < P; ++pl) {
            const float v = in.p[pl][(long long)b * in.bstride + q];
            out[s * P + pl] = f2bf(pl < 2 ? (v - 128.0f) * (1.0f / 64.0f) : v);
        }
    }
    u16* dst = x + ((long long)b * h * w + i) * 4 * P;
#pragma unroll
    for (int c = 0; c < 4 * P; c += 8) {
        uint4 r;
        r.x = out[c] | ((unsigned)out[c + 1] << 16);
        r.y = out[c + 2] | ((unsigned)out[c + 3] << 16);
        r.z = out[c + 4] | ((unsigned)out[c + 5] << 16);
        r.w = out[c + 6] | ((unsigned)out[c + 7] << 16);
        *(uint4*)(dst + c) = r;
    }
}

// Head outputs (float32 NHWC at H/2) -> planar fields: coarse [B][4][h][w]
// from the first 4 of the 8-channel head, fine [B][4][H][W] from the
// 16-channel head
// (channel = sub * 4 + field, sub = 2 * dy + dx).
__global__ void heads_to_planes_k(const float* hc, const float* hf, int B, int h, int w, float* coarse,
                                  float* fine) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)B * h * w;
    if (i >= n) return;
    const int b = (int)(i / ((long long)h * w));
    const long long r = i - (long long)b * h * w;
    const int y = (int)(r / w), x = (int)(r - (long long)y * w);
    const long long hw = (long long)h * w, HW = 4 * hw;
#pragma unroll
    for (int f = 0; f < 4; ++f) coarse[(long long)b * 4 * hw + f * hw + r] = hc[i * 8 + f];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int f = 0; f < 4; ++f)
            fine[(long long)b * 4 * HW + f * HW + (long long)(2 * y + (s >> 1)) * (2 * w) + 2 * x + (s & 1)] =
                hf[i * 16 + s * 4 + f];
}

// Training loss: per-field weighted Charbonnier penalty sqrt(d^2 + e^2) of
// the head outputs against planar labels, as a mean over each level's
// pixels; writes the gradients with respect to the heads and adds the loss.
__global__ void loss_k(const float* hc, const float* hf, const float* lc, const float* lf, int B, int h, int w,
                       const float* fw, float* ghc, float* ghf, float* loss) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)B * h * w;
    const float e2 = 1e-6f;
    float s = 0.0f;
    if (i < n) {
        const int b = (int)(i / ((long long)h * w));
        const long long r = i - (long long)b * h * w;
        const int y = (int)(r / w), x = (int)(r - (long long)y * w);
        const long long hw = (long long)h * w, HW = 4 * hw;
        const float nc = 1.0f / (float)n, nf = 1.0f / (float)(4 * n);
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const float d = hc[i * 8 + f] - lc[(long long)b * 4 * hw + f * hw + r];
            const float q = sqrtf(d * d + e2);
            s += fw[f] * q * nc;
            ghc[i * 8 + f] = fw[f] * nc * d / q;
            ghc[i * 8 + 4 + f] = 0.0f;
        }
#pragma unroll
        for (int sb = 0; sb < 4; ++sb)
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                const long long q0 = (long long)(2 * y + (sb >> 1)) * (2 * w) + 2 * x + (sb & 1);
                const float d = hf[i * 16 + sb * 4 + f] - lf[(long long)b * 4 * HW + f * HW + q0];
                const float q = sqrtf(d * d + e2);
                s += fw[4 + f] * q * nf;
                ghf[i * 16 + sb * 4 + f] = fw[4 + f] * nf * d / q;
            }
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, s);
}

// ---------------------------------------------------------------- network

// Layer table.  Activations of a B x H x W batch (H, W multiples of 16)
// live at four resolutions H/2 .. H/16; `src`/`dst` name tensors, modes are
// space-to-depth reads and depth-to-space writes.
enum Tensor { T_X0, T_T0, T_CAT0, T_T1, T_CAT1, T_T2, T_CAT2, T_T3, T_T3B, T_R2, T_R1, T_R0, T_HC, T_HF, T_N };
struct TensorDef {
    int level;   // resolution H / 2^(level+1)
    int C;       // stored channels per pixel
    int f32;
};
const TensorDef kTensors[T_N] = {
    {0, 8, 0},  {0, 32, 0}, {0, 64, 0}, {1, 32, 0}, {1, 64, 0}, {2, 48, 0}, {2, 96, 0},
    {3, 64, 0}, {3, 64, 0}, {2, 48, 0}, {1, 32, 0}, {0, 32, 0}, {0, 8, 1},  {0, 16, 1},
};
struct LayerDef {
    const char* name;
    int src, src_c0, src_mode;   // input tensor, channel offset, 1 = space-to-depth
    int dst, dst_c0, dst_mode;   // output tensor, channel offset, 1 = depth-to-space
    int cin, cout;               // conv-space channels
    int act;
    int level;                   // conv resolution
};
// enc0a's input channels are 4 x the network's input planes (set per network)
const LayerDef kLayers[] = {
    {"enc0a", T_X0, 0, 0, T_T0, 0, 0, 8, 32, 1, 0},
    {"enc0b", T_T0, 0, 0, T_CAT0, 32, 0, 32, 32, 1, 0},
    {"enc1a", T_CAT0, 32, 1, T_T1, 0, 0, 128, 32, 1, 1},
    {"enc1b", T_T1, 0, 0, T_CAT1, 32, 0, 32, 32, 1, 1},
    {"enc2a", T_CAT1, 32, 1, T_T2, 0, 0, 128, 48, 1, 2},
    {"enc2b", T_T2, 0, 0, T_CAT2, 48, 0, 48, 48, 1, 2},
    {"enc3a", T_CAT2, 48, 1, T_T3, 0, 0, 192, 64, 1, 3},
    {"enc3b", T_T3, 0, 0, T_T3B, 0, 0, 64, 64, 1, 3},
    {"dec2u", T_T3B, 0, 0, T_CAT2, 0, 1, 64, 192, 1, 3},
    {"dec2", T_CAT2, 0, 0, T_R2, 0, 0, 96, 48, 1, 2},
    {"dec1u", T_R2, 0, 0, T_CAT1, 0, 1, 48, 128, 1, 2},
    {"dec1", T_CAT1, 0, 0, T_R1, 0, 0, 64, 32, 1, 1},
    {"dec0u", T_R1, 0, 0, T_CAT0, 0, 1, 32, 128, 1, 1},
    {"dec0", T_CAT0, 0, 0, T_R0, 0, 0, 64, 32, 1, 0},
    {"head_c", T_R0, 0, 0, T_HC, 0, 0, 32, 8, 0, 0},   // channels 4..7 unused
    {"head_f", T_R0, 0, 0, T_HF, 0, 0, 32, 16, 0, 0},
};
constexpr int kNumLayers = sizeof(kLayers) / sizeof(kLayers[0]);

inline int rup(int v, int m) { return (v + m - 1) / m * m; }

struct LayerParams {
    int rows, Kp;          // forward weights [rows][Kp]
    int rows_t, Kp_t;      // data-gradient weights
    size_t w_off, b_off;   // float offsets in the parameter vector
    size_t wb_off;         // bf16 offsets (forward weights)
    size_t wt_off;         // bf16 offsets (flipped weights)
};

}  // namespace

// Offsets of every tensor of a (B, H, W) batch in one activation buffer.
struct ActLayout {
    int B = 0, H = 0, W = 0;
    size_t t_off[T_N] = {};
    size_t bytes = 0;
};

struct Net {
    int planes = 2;            // input planes (2: frame pair; 4: pair + flow)
    LayerDef def[kNumLayers];  // the layer table with enc0a's input width
    LayerParams L[kNumLayers];
    size_t n_par = 0;      // floats in the padded parameter vector
    size_t n_wb = 0;       // bf16 words of forward weights (prefix of the vector's layout)
    size_t n_wt = 0;
    size_t n_valid = 0;    // real parameters
    float *par = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    u16 *wb = nullptr, *wt = nullptr;
    long long step = 0;
    bool wt_fresh = false;
    // training activations / gradients for the current batch shape (the
    // inference entry takes its activations from the calling context)
    ActLayout lay;
    void* act = nullptr;
    void* gact = nullptr;
    size_t act_bytes = 0;
};

namespace {

void net_layout(Net* n) {
    for (int l = 0; l < kNumLayers; ++l) n->def[l] = kLayers[l];
    n->def[0].cin = 4 * n->planes;
    size_t off = 0, wb = 0, wt = 0, valid = 0;
    // forward weights of every layer first (their bf16 copies mirror them), then biases
    for (int l = 0; l < kNumLayers; ++l) {
        LayerParams& P = n->L[l];
        const LayerDef& d = n->def[l];
        P.rows = rup(d.cout, 64);
        P.Kp = rup(9 * d.cin, 32);
        P.rows_t = rup(d.cin, 64);
        P.Kp_t = rup(9 * d.cout, 32);
        P.w_off = off;
        P.wb_off = wb;
        off += (size_t)P.rows * P.Kp;
        wb += (size_t)P.rows * P.Kp;
        P.wt_off = wt;
        wt += (size_t)P.rows_t * P.Kp_t;
        valid += (size_t)9 * d.cin * d.cout + d.cout;
    }
    n->n_wb = wb;
    for (int l = 0; l < kNumLayers; ++l) {
        n->L[l].b_off = off;
        off += (size_t)rup(n->def[l].cout, 64);
    }
    n->n_par = off;
    n->n_wt = wt;
    n->n_valid = valid;
}

inline dim3 grid1(long long n) { return dim3((unsigned)((n + 255) / 256)); }

// stored channels of a tensor (the input tensor's follow the input planes)
inline int tchan(const Net* n, int t) { return t == T_X0 ? 4 * n->planes : kTensors[t].C; }

int check_shape(int B, int H, int W) {
    if (B < 1 || H < 16 || W < 16 || H % 16 || W % 16)
        return fail(OFRI_E_ARG, "net: batch %d of %dx%d (H, W must be multiples of 16)", B, H, W);
    return OFRI_OK;
}

ActLayout act_layout(const Net* n, int B, int H, int W) {
    ActLayout L;
    L.B = B;
    L.H = H;
    L.W = W;
    size_t off = 0;
    for (int t = 0; t < T_N; ++t) {
        const int lh = H >> (kTensors[t].level + 1), lw = W >> (kTensors[t].level + 1);
        L.t_off[t] = off;
        off += align256((size_t)B * lh * lw * tchan(n, t) * (kTensors[t].f32 ? 4 : 2));
    }
    L.bytes = off;
    return L;
}

// The training buffers (act, gact) laid out for a (B, H, W) batch.
int net_shape(ofri_ctx* ctx, Net* n, int B, int H, int W) {
    OFRI_TRY(check_shape(B, H, W));
    if (n->act && n->lay.B == B && n->lay.H == H && n->lay.W == W) return OFRI_OK;
    const ActLayout L = act_layout(n, B, H, W);
    // the buffers only grow: a smaller batch reuses them with new offsets
    if (L.bytes > n->act_bytes) {
        OFRI_HIP(hipStreamSynchronize(ctx->stream));
        if (n->act) OFRI_HIP(hipFree(n->act));
        if (n->gact) OFRI_HIP(hipFree(n->gact));
        n->act = n->gact = nullptr;
        n->act_bytes = 0;
        OFRI_HIP(hipMalloc(&n->act, L.bytes));
        OFRI_HIP(hipMalloc(&n->gact, L.bytes));
        n->act_bytes = L.bytes;
        ctx->ws_gen += 1;   // captured graphs holding the old buffers are stale
    }
    n->lay = L;
    return OFRI_OK;
}

View tview(const Net* n, const ActLayout& L, void* base, int t, int c0, int mode) {
    const TensorDef& d = kTensors[t];
    const int lh = L.H >> (d.level + 1), lw = L.W >> (d.level + 1);
    const int C = tchan(n, t);
    View v;
    const size_t esz = d.f32 ? 4 : 2;
    v.p = (char*)base + L.t_off[t] + (size_t)c0 * esz;
    v.img = (long long)lh * lw * C;
    v.ps = C;
    v.f32 = d.f32;
    v.mode = mode;
    return v;
}

// 16 x 16 tiles while they give a few hundred workgroups; from 128 to 511
// tiles, 16 x 16 tiles with 16-channel groups (the weights staged per
// workgroup shrink, the grid keeps enough workgroups: measured 3 % faster
// over the whole network than 8 x 8 tiles there); below, 8 x 8 tiles with
// 32-channel groups (low-resolution levels).
int launch_conv(ofri_ctx* ctx, const ConvArgs& a) {
    const unsigned t16 = (unsigned)(a.B * ((a.H + 15) / 16) * ((a.W + 15) / 16));
    const unsigned t8 = (unsigned)(a.B * ((a.H + 7) / 8) * ((a.W + 7) / 8));
    if (t16 >= 512) {
        if (a.Cout > 32) conv3_lds_k<16, 4><<<dim3(t16, (unsigned)((a.Cout + 63) / 64)), 256, 0, ctx->stream>>>(a);
        else if (a.Cout > 16) conv3_lds_k<16, 2><<<dim3(t16, 1), 256, 0, ctx->stream>>>(a);
        else conv3_lds_k<16, 1><<<dim3(t16, 1), 256, 0, ctx->stream>>>(a);
    } else if (t16 >= 128) {
        conv3_lds_k<16, 1><<<dim3(t16, (unsigned)((a.Cout + 15) / 16)), 256, 0, ctx->stream>>>(a);
    } else {
        if (a.Cout > 16) conv3_lds_k<8, 2><<<dim3(t8, (unsigned)((a.Cout + 31) / 32)), 256, 0, ctx->stream>>>(a);
        else conv3_lds_k<8, 1><<<dim3(t8, 1), 256, 0, ctx->stream>>>(a);
    }
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

// coarse_head false: the coarse head (layer head_c) is skipped.  add_u,
// add_v non-null (B = 1): the fine head's fields 0, 1 are added into these
// full-size planes instead of being stored.  own_act: the activations go to
// the network's training buffers (kept for the backward pass), else to the
// context's scratch, so one network can run on several streams at once.
int net_forward(ofri_ctx* ctx, Net* n, const InPlanes& in, int B, int H, int W, bool coarse_head = true,
                float* add_u = nullptr, float* add_v = nullptr, bool own_act = true) {
    ActLayout L;
    void* base;
    if (own_act) {
        OFRI_TRY(net_shape(ctx, n, B, H, W));
        L = n->lay;
        base = n->act;
    } else {
        OFRI_TRY(check_shape(B, H, W));
        L = act_layout(n, B, H, W);
        OFRI_TRY(scratch(ctx, "net_act", L.bytes, &base));
    }
    const long long hw = (long long)(H / 2) * (W / 2);
    const dim3 g((unsigned)((hw + 255) / 256), B);
    u16* x0 = (u16*)((char*)base + L.t_off[T_X0]);
    if (n->planes == 4) net_input_k<4><<<g, 256, 0, ctx->stream>>>(in, H, W, x0);
    else if (n->planes == 2) net_input_k<2><<<g, 256, 0, ctx->stream>>>(in, H, W, x0);
    else return fail(OFRI_E_ARG, "net: %d input planes not supported", n->planes);
    OFRI_HIP(hipGetLastError());
    for (int l = 0; l < kNumLayers; ++l) {
        const LayerDef& d = n->def[l];
        if (!coarse_head && d.dst == T_HC) continue;
        const LayerParams& P = n->L[l];
        ConvArgs a{};
        a.x = tview(n, L, base, d.src, d.src_c0, d.src_mode);
        a.gate.p = nullptr;
        a.w = n->wb + P.wb_off;
        a.Kp = P.Kp;
        a.bias = n->par + P.b_off;
        a.y = tview(n, L, base, d.dst, d.dst_c0, d.dst_mode);
        a.B = B;
        a.H = H >> (d.level + 1);
        a.W = W >> (d.level + 1);
        a.Cin = d.cin;
        a.Cout = d.cout;
        a.act = d.act;
        a.beta = 0;
        if (add_u && d.dst == T_HF) {
            a.add_u = add_u;
            a.add_v = add_v;
        }
        OFRI_TRY(launch_conv(ctx, a));
    }
    return OFRI_OK;
}

// Backward through every layer from the head gradients in gact's T_HC and
// T_HF slots: parameter gradients into n->grad (zeroed here), activation
// gradients through gact.
int net_backward(ofri_ctx* ctx, Net* n) {
    const int B = n->lay.B, H = n->lay.H, W = n->lay.W;
    zero_f_k<<<grid1((long long)n->n_par), 256, 0, ctx->stream>>>(n->grad, (long long)n->n_par);
    OFRI_HIP(hipGetLastError());
    if (!n->wt_fresh) {
        for (int l = 0; l < kNumLayers; ++l) {
            const LayerParams& P = n->L[l];
            const long long cnt = (long long)P.rows_t * P.Kp_t;
            wflip_k<<<grid1(cnt), 256, 0, ctx->stream>>>(n->wb + P.wb_off, P.Kp, n->def[l].cin, n->def[l].cout,
                                                         n->wt + P.wt_off, P.rows_t, P.Kp_t);
        }
        OFRI_HIP(hipGetLastError());
        n->wt_fresh = true;
    }
    // which tensors already hold a gradient (a second consumer adds into it)
    bool have[T_N] = {};
    have[T_HC] = have[T_HF] = true;
    for (int l = kNumLayers - 1; l >= 0; --l) {
        const LayerDef& d = n->def[l];
        const LayerParams& P = n->L[l];
        const int lh = H >> (d.level + 1), lw = W >> (d.level + 1);
        DwArgs g{};
        g.x = tview(n, n->lay, n->act, d.src, d.src_c0, d.src_mode);
        g.dy = tview(n, n->lay, n->gact, d.dst, d.dst_c0, d.dst_mode);
        g.gate = d.act ? tview(n, n->lay, n->act, d.dst, d.dst_c0, d.dst_mode) : View{nullptr, 0, 0, 0, 0};
        g.dw = n->grad + P.w_off;
        g.db = n->grad + P.b_off;
        g.Kp = P.Kp;
        g.B = B;
        g.H = lh;
        g.W = lw;
        g.Cin = d.cin;
        g.Cout = d.cout;
        const long long Pn = (long long)B * lh * lw;
        g.chunk = 4096;
        dim3 gg((unsigned)((9 * d.cin + 63) / 64), (unsigned)((d.cout + 31) / 32),
                (unsigned)((Pn + g.chunk - 1) / g.chunk));
        dw_k<<<gg, 256, 0, ctx->stream>>>(g);
        OFRI_HIP(hipGetLastError());
        if (l == 0) break;  // no gradient for the network input
        ConvArgs a{};
        a.x = g.dy;
        a.gate = g.gate;
        a.w = n->wt + P.wt_off;
        a.Kp = P.Kp_t;
        a.bias = nullptr;
        a.y = tview(n, n->lay, n->gact, d.src, d.src_c0, d.src_mode);
        a.B = B;
        a.H = lh;
        a.W = lw;
        a.Cin = d.cout;
        a.Cout = d.cin;
        a.act = 0;
        // a tensor with two consumers (a skip connection's concatenation buffer,
        // the two heads' input) takes the second gradient on top of the first;
        // in reverse order the consumer reading the whole tensor comes first
        a.beta = have[d.src] ? 1 : 0;
        OFRI_TRY(launch_conv(ctx, a));
        have[d.src] = true;
    }
    return OFRI_OK;
}

}  // namespace

// ---- interface used by the C ABI and the learned pipeline ----

int net_create(ofri_ctx* ctx, int planes, Net** out) {
    if (planes != 2 && planes != 4) return fail(OFRI_E_ARG, "net: %d input planes (2 or 4)", planes);
    Net* n = new Net();
    n->planes = planes;
    net_layout(n);
    hipError_t e = hipSuccess;
    auto al = [&](void** p, size_t b) {
        if (e == hipSuccess) e = hipMalloc(p, b);
    };
    al((void**)&n->par, n->n_par * 4);
    al((void**)&n->grad, n->n_par * 4);
    al((void**)&n->m, n->n_par * 4);
    al((void**)&n->v, n->n_par * 4);
    al((void**)&n->wb, n->n_wb * 2);
    al((void**)&n->wt, n->n_wt * 2);
    if (e == hipSuccess) e = hipMemsetAsync(n->par, 0, n->n_par * 4, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(n->m, 0, n->n_par * 4, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(n->v, 0, n->n_par * 4, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(n->wb, 0, n->n_wb * 2, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        net_destroy(n);
        return fail(OFRI_E_HIP, "net_create: %s", hipGetErrorString(e));
    }
    *out = n;
    return OFRI_OK;
}

void net_destroy(Net* n) {
    if (!n) return;
    for (void* p : {(void*)n->par, (void*)n->grad, (void*)n->m, (void*)n->v, (void*)n->wb, (void*)n->wt, n->act,
                    n->gact})
        if (p) (void)hipFree(p);
    delete n;
}

size_t net_num_params(const Net* n) { return n->n_valid; }
int net_planes(const Net* n) { return n->planes; }

// Multiply-adds of one forward pass over one H x W input, counted on the
// real (unpadded) channels: 9 * cin * cout per output pixel of each layer.
double net_forward_flops(const Net* n, int H, int W) {
    double macs = 0.0;
    for (int l = 0; l < kNumLayers; ++l) {
        const LayerDef& d = n->def[l];
        macs += 9.0 * d.cin * d.cout * (double)(H >> (d.level + 1)) * (double)(W >> (d.level + 1));
    }
    return 2.0 * macs;
}

// Packed order: layer by layer, weights [cout][tap][cin] then bias [cout].
int net_set_params(ofri_ctx* ctx, Net* n, const float* h, size_t count) {
    if (count != n->n_valid) return fail(OFRI_E_ARG, "net: %zu parameters given, %zu expected", count, n->n_valid);
    std::vector<float> pad(n->n_par, 0.0f);
    size_t q = 0;
    for (int l = 0; l < kNumLayers; ++l) {
        const LayerDef& d = n->def[l];
        const LayerParams& P = n->L[l];
        for (int co = 0; co < d.cout; ++co)
            for (int k = 0; k < 9 * d.cin; ++k) pad[P.w_off + (size_t)co * P.Kp + k] = h[q++];
        for (int co = 0; co < d.cout; ++co) pad[P.b_off + co] = h[q++];
    }
    OFRI_HIP(hipMemcpyAsync(n->par, pad.data(), n->n_par * 4, hipMemcpyHostToDevice, ctx->stream));
    OFRI_HIP(hipMemsetAsync(n->m, 0, n->n_par * 4, ctx->stream));
    OFRI_HIP(hipMemsetAsync(n->v, 0, n->n_par * 4, ctx->stream));
    f2bf_k<<<grid1((long long)n->n_wb), 256, 0, ctx->stream>>>(n->par, n->wb, (long long)n->n_wb);
    OFRI_HIP(hipGetLastError());
    OFRI_HIP(hipStreamSynchronize(ctx->stream));
    n->step = 0;
    n->wt_fresh = false;
    return OFRI_OK;
}

// He initialisation for the leaky-ReLU layers (std sqrt(2 / fan_in)), heads
// 10x smaller (std 0.1 sqrt(1 / fan_in)), biases zero; Adam state reset.
int net_init_params(ofri_ctx* ctx, Net* n, unsigned long long seed) {
    std::mt19937_64 gen(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<float> h;
    h.reserve(n->n_valid);
    for (int l = 0; l < kNumLayers; ++l) {
        const LayerDef& d = n->def[l];
        const double fan = 9.0 * d.cin;
        const double sd = d.act ? std::sqrt(2.0 / fan) : 0.1 * std::sqrt(1.0 / fan);
        for (int i = 0; i < 9 * d.cin * d.cout; ++i) h.push_back((float)(nd(gen) * sd));
        for (int co = 0; co < d.cout; ++co) h.push_back(0.0f);
    }
    return net_set_params(ctx, n, h.data(), h.size());
}

static int unpack(ofri_ctx* ctx, Net* n, const float* d, float* h, size_t count) {
    if (count != n->n_valid) return fail(OFRI_E_ARG, "net: %zu parameters asked, %zu held", count, n->n_valid);
    std::vector<float> pad(n->n_par);
    OFRI_HIP(hipMemcpyAsync(pad.data(), d, n->n_par * 4, hipMemcpyDeviceToHost, ctx->stream));
    OFRI_HIP(hipStreamSynchronize(ctx->stream));
    size_t q = 0;
    for (int l = 0; l < kNumLayers; ++l) {
        const LayerDef& dd = n->def[l];
        const LayerParams& P = n->L[l];
        for (int co = 0; co < dd.cout; ++co)
            for (int k = 0; k < 9 * dd.cin; ++k) h[q++] = pad[P.w_off + (size_t)co * P.Kp + k];
        for (int co = 0; co < dd.cout; ++co) h[q++] = pad[P.b_off + co];
    }
    return OFRI_OK;
}

int net_get_params(ofri_ctx* ctx, Net* n, float* h, size_t count) { return unpack(ctx, n, n->par, h, count); }
int net_get_grads(ofri_ctx* ctx, Net* n, float* h, size_t count) { return unpack(ctx, n, n->grad, h, count); }

InPlanes stacked(const Net* n, const float* in, int H, int W) {
    InPlanes r{};
    const long long HW = (long long)H * W;
    for (int i = 0; i < n->planes && i < 4; ++i) r.p[i] = in + i * HW;
    r.bstride = n->planes * HW;
    return r;
}

// Forward pass; planar outputs: coarse [B][4][H/2][W/2], fine [B][4][H][W].
int net_predict(ofri_ctx* ctx, Net* n, const float* in, int B, int H, int W, float* coarse, float* fine) {
    OFRI_TRY(net_forward(ctx, n, stacked(n, in, H, W), B, H, W));
    const int h = H / 2, w = W / 2;
    heads_to_planes_k<<<grid1((long long)B * h * w), 256, 0, ctx->stream>>>(
        (const float*)((char*)n->act + n->lay.t_off[T_HC]), (const float*)((char*)n->act + n->lay.t_off[T_HF]), B, h, w,
        coarse, fine);
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

// One H x W input given as separate planes (planes[0 .. planes-1]); the fine
// head's fields 0, 1 are added into add_u, add_v (full size).  The coarse
// head is not computed.
int net_correct(ofri_ctx* ctx, Net* n, const float* const* planes, int H, int W, float* add_u, float* add_v) {
    InPlanes in{};
    for (int i = 0; i < n->planes && i < 4; ++i) in.p[i] = planes[i];
    in.bstride = 0;
    if (!add_u || !add_v) return fail(OFRI_E_ARG, "net_correct: null flow planes");
    // the head's last kernel adds the correction into the flow; activations
    // in the context's scratch
    return net_forward(ctx, n, in, 1, H, W, false, add_u, add_v, false);
}

// Forward, loss against planar labels, backward; with lr > 0 an Adam step.
// d_loss: one float on the device, overwritten.
int net_train(ofri_ctx* ctx, Net* n, const float* in, int B, int H, int W, const float* lab_c, const float* lab_f,
              const float* h_field_w, float lr, float* d_loss) {
    OFRI_TRY(net_forward(ctx, n, stacked(n, in, H, W), B, H, W));
    const int h = H / 2, w = W / 2;
    void* fwb;
    OFRI_TRY(scratch(ctx, "net_fw", 8 * sizeof(float), &fwb));
    OFRI_HIP(hipMemcpyAsync(fwb, h_field_w, 8 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    zero_f_k<<<1, 64, 0, ctx->stream>>>(d_loss, 1);
    loss_k<<<grid1((long long)B * h * w), 256, 0, ctx->stream>>>(
        (const float*)((char*)n->act + n->lay.t_off[T_HC]), (const float*)((char*)n->act + n->lay.t_off[T_HF]), lab_c,
        lab_f, B, h, w, (const float*)fwb, (float*)((char*)n->gact + n->lay.t_off[T_HC]),
        (float*)((char*)n->gact + n->lay.t_off[T_HF]), d_loss);
    OFRI_HIP(hipGetLastError());
    OFRI_TRY(net_backward(ctx, n));
    if (lr > 0.0f) {
        n->step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - powf(b1, (float)n->step), c2 = 1.0f - powf(b2, (float)n->step);
        adam_k<<<grid1((long long)n->n_par), 256, 0, ctx->stream>>>(n->par, n->grad, n->m, n->v, n->wb,
                                                                    (long long)n->n_wb, (long long)n->n_par, lr, b1,
                                                                    b2, 1e-8f, c1, c2);
        OFRI_HIP(hipGetLastError());
        n->wt_fresh = false;
    }
    return OFRI_OK;
}

}  // namespace ofri
