// Pre-filter, Pillow-compatible resize, not-a-knot spline upsample, symmetric
// bilinear warp, accumulation and small reductions.  All are HBM-bound
// per-pixel kernels: one thread per output pixel, rows mapped to x so every
// wavefront reads and writes 64 consecutive floats.
#include <cmath>
#include <cstring>

#include "ofri_internal.h"

#pragma clang fp contract(off)

namespace ofri {
namespace {

// Row pass: out[m] = sum_j w[j] * line(m + h - j), float32, tap order, from 0.
// Up to two planes per launch (blockIdx.z): source s[z] -> destination d[z].
struct Planes2 {
    const float* s[2];
    float* d[2];
};

__global__ void gauss_rows(Planes2 p, int H, int W, Taps t) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= W) return;
    float* __restrict__ dst = p.d[blockIdx.z];
    const float* row = p.s[blockIdx.z] + (size_t)y * W;
    int h = t.n / 2;
    float acc = 0.0f;
    for (int j = 0; j < t.n; ++j) {
        float prod = t.w[j] * row[pf_index(x + h - j, W, h)];
        acc = acc + prod;
    }
    dst[(size_t)y * W + x] = acc;
}

// Column pass on the row-filtered image.
__global__ void gauss_cols(Planes2 p, int H, int W, Taps t) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= W) return;
    const float* __restrict__ src = p.s[blockIdx.z];
    float* __restrict__ dst = p.d[blockIdx.z];
    int h = t.n / 2;
    float acc = 0.0f;
    for (int j = 0; j < t.n; ++j) {
        float prod = t.w[j] * src[(size_t)pf_index(y + h - j, H, h) * W + x];
        acc = acc + prod;
    }
    dst[(size_t)y * W + x] = acc;
}

// Both passes in one launch for up to 9 taps: a workgroup stages a source
// tile (padded positions mapped through pf_index), filters its rows into LDS
// and then its columns, with the two-pass kernels' operations and order.
constexpr int kGTW = 64, kGTH = 16, kGMaxH = 4;
__global__ __launch_bounds__(256) void gauss_fused(Planes2 p, int H, int W, Taps t) {
    __shared__ float sin_[kGTH + 2 * kGMaxH][kGTW + 2 * kGMaxH];
    __shared__ float srow[kGTH + 2 * kGMaxH][kGTW];
    const int h = t.n / 2;
    const int x0 = blockIdx.x * kGTW, y0 = blockIdx.y * kGTH;
    const float* __restrict__ src = p.s[blockIdx.z];
    float* __restrict__ dst = p.d[blockIdx.z];
    const int R = kGTH + 2 * h, Cw = kGTW + 2 * h;
    // source tile: positions [y0 - h, y0 + kGTH + h) x [x0 - h, x0 + kGTW + h);
    // positions past the padding domain are never read
    for (int i = threadIdx.x; i < R * Cw; i += 256) {
        const int rr = i / Cw, cc = i - rr * Cw;
        const int pr = y0 - h + rr, pc = x0 - h + cc;
        float v = 0.0f;
        if (pr < H + h && pc < W + h) v = src[(size_t)pf_index(pr, H, h) * W + pf_index(pc, W, h)];
        sin_[rr][cc] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < R * kGTW; i += 256) {
        const int rr = i / kGTW, cc = i - rr * kGTW;
        float acc = 0.0f;
        for (int j = 0; j < t.n; ++j) {
            float prod = t.w[j] * sin_[rr][cc + 2 * h - j];
            acc = acc + prod;
        }
        srow[rr][cc] = acc;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kGTH * kGTW; i += 256) {
        const int ry = i / kGTW, cc = i - ry * kGTW;
        const int y = y0 + ry, x = x0 + cc;
        if (y >= H || x >= W) continue;
        float acc = 0.0f;
        for (int j = 0; j < t.n; ++j) {
            float prod = t.w[j] * srow[ry + 2 * h - j][cc];
            acc = acc + prod;
        }
        dst[(size_t)y * W + x] = acc;
    }
}

// Pillow resample, horizontal pass over the rows [first, first+Ht).
__global__ void resize_h(Planes2 p, int W, int W2, int Ht, int first, const double* __restrict__ k,
                         const int* __restrict__ b, int ks) {
    int xx = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (xx >= W2 || y >= Ht) return;
    const float* __restrict__ src = p.s[blockIdx.z];
    float* __restrict__ tmp = p.d[blockIdx.z];
    int xmin = b[2 * xx], xn = b[2 * xx + 1];
    const double* kk = k + (size_t)xx * ks;
    const float* row = src + (size_t)(y + first) * W + xmin;
    double ss = 0.0;
    // 8 taps per step, loaded before use (taps past xn read tap xn - 1 and
    // are not added)
    for (int x0 = 0; x0 < xn; x0 += 8) {
        float v[8];
        double kv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int x = min(x0 + u, xn - 1);
            v[u] = row[x];
            kv[u] = kk[x];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (x0 + u < xn) ss = ss + (double)v[u] * kv[u];
    }
    tmp[(size_t)y * W2 + xx] = (float)ss;
}

__global__ void resize_v(Planes2 p, int W2, int H2, int first, const double* __restrict__ k,
                         const int* __restrict__ b, int ks) {
    int xx = blockIdx.x * blockDim.x + threadIdx.x;
    int yy = blockIdx.y;
    if (xx >= W2 || yy >= H2) return;
    const float* __restrict__ tmp = p.s[blockIdx.z];
    float* __restrict__ dst = p.d[blockIdx.z];
    int ymin = b[2 * yy] - first, yn = b[2 * yy + 1];
    const double* kk = k + (size_t)yy * ks;
    double ss = 0.0;
    // 8 taps per step, loaded before use (taps past yn read tap yn - 1 and
    // are not added)
    for (int y0 = 0; y0 < yn; y0 += 8) {
        float v[8];
        double kv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int y = min(y0 + u, yn - 1);
            v[u] = tmp[(size_t)(y + ymin) * W2 + xx];
            kv[u] = kk[y];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (y0 + u < yn) ss = ss + (double)v[u] * kv[u];
    }
    dst[(size_t)yy * W2 + xx] = (float)ss;
}

// Not-a-knot spline upsample, two passes in double over a banded weight
// table: tmp[r][x] = sum_t w[x][t] * src[r][st[x] + t], then
// dst[y][x] = float(sum_t w'[y][t] * tmp[st'[y] + t][x]) * scale.  Every sum
// runs t = 0, 1, ... from 0.0.  blockIdx.z selects one of up to two planes
// (U and V of a flow share the tables).
struct SplinePlanes {
    const float* src[2];
    double* tmp[2];
    float* dst[2];
    float scale[2];
};

constexpr int kSplineRowsX = 2;    // coarse rows per workgroup of the x pass
constexpr int kSplineRowsY = 4;    // output rows per thread of the y pass
constexpr int kSplineChunk = 22;   // weights in flight per thread of the x pass
constexpr int kSplineMaxWidth = 66;  // band width of the weight tables (2 * 32 + 2)

// x pass: kSplineRowsX coarse rows staged in LDS (nrows * Wc floats); the
// weights are read transposed (wT[t][x]) so a wavefront's loads are
// contiguous, and each weight serves every staged row.
__global__ __launch_bounds__(256) void spline_x(SplinePlanes p, int Wc, int W, int Hc, const double* __restrict__ wT,
                                                const int* __restrict__ st, int width, int nrows) {
    extern __shared__ float rows[];
    const int z = blockIdx.z;
    const int r0 = blockIdx.y * nrows;
    const int nr = min(nrows, Hc - r0);
    const float* src = p.src[z] + (size_t)r0 * Wc;
    for (int i = threadIdx.x; i < nr * Wc; i += 256) rows[i] = src[i];
    __syncthreads();
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= W) return;
    const float* b = rows + st[x];
    double s[kSplineRowsX];
#pragma unroll
    for (int i = 0; i < kSplineRowsX; ++i) s[i] = 0.0;
    // weights fetched kSplineChunk at a time so their loads overlap
    for (int t0 = 0; t0 < width; t0 += kSplineChunk) {
        double wc[kSplineChunk];
#pragma unroll
        for (int k = 0; k < kSplineChunk; ++k) wc[k] = t0 + k < width ? wT[(size_t)(t0 + k) * W + x] : 0.0;
#pragma unroll
        for (int k = 0; k < kSplineChunk; ++k) {
            if (t0 + k < width) {
#pragma unroll
                for (int i = 0; i < kSplineRowsX; ++i)
                    if (i < nr) s[i] = s[i] + wc[k] * (double)b[i * Wc + t0 + k];
            }
        }
    }
    double* tmp = p.tmp[z] + (size_t)r0 * W + x;
#pragma unroll
    for (int i = 0; i < kSplineRowsX; ++i)
        if (i < nr) tmp[(size_t)i * W] = s[i];
}

// Exact factor 2 (W = 2 Wc): output 2i sits on sample i (its weight row is
// the unit vector, so the band sum is the sample itself) and the last output
// is clamped onto the last sample; only the odd outputs 2i+1 < W-1 need the
// band.  x pass: thread i writes outputs 2i and 2i+1 of its staged rows.
__global__ __launch_bounds__(256) void spline_x2(SplinePlanes p, int Wc, int W, int Hc, const double* __restrict__ wT,
                                                 const int* __restrict__ st, int width, int nrows) {
    extern __shared__ float rows[];
    const int z = blockIdx.z;
    const int r0 = blockIdx.y * nrows;
    const int nr = min(nrows, Hc - r0);
    const float* src = p.src[z] + (size_t)r0 * Wc;
    for (int i = threadIdx.x; i < nr * Wc; i += 256) rows[i] = src[i];
    __syncthreads();
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= Wc) return;
    const int x = 2 * i + 1;
    double* tmp = p.tmp[z] + (size_t)r0 * W;
    double s[kSplineRowsX];
#pragma unroll
    for (int q = 0; q < kSplineRowsX; ++q) s[q] = q < nr ? (double)rows[q * Wc + i] : 0.0;
#pragma unroll
    for (int q = 0; q < kSplineRowsX; ++q)
        if (q < nr) tmp[(size_t)q * W + 2 * i] = s[q];
    if (x == W - 1) {
#pragma unroll
        for (int q = 0; q < kSplineRowsX; ++q)
            if (q < nr) tmp[(size_t)q * W + x] = s[q];
        return;
    }
    const float* b = rows + st[x];
#pragma unroll
    for (int q = 0; q < kSplineRowsX; ++q) s[q] = 0.0;
    for (int t0 = 0; t0 < width; t0 += kSplineChunk) {
        double wc[kSplineChunk];
#pragma unroll
        for (int k = 0; k < kSplineChunk; ++k) wc[k] = t0 + k < width ? wT[(size_t)(t0 + k) * W + x] : 0.0;
#pragma unroll
        for (int k = 0; k < kSplineChunk; ++k) {
            if (t0 + k < width) {
#pragma unroll
                for (int q = 0; q < kSplineRowsX; ++q)
                    if (q < nr) s[q] = s[q] + wc[k] * (double)b[q * Wc + t0 + k];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kSplineRowsX; ++q)
        if (q < nr) tmp[(size_t)q * W + x] = s[q];
}

// Exact factor 2 (H = 2 Hc), y pass: a thread owns column x of output rows
// 2k and 2k+1 for kSplineRowsY consecutive coarse rows k; even rows are the
// x-pass rows, odd rows walk the coarse rows their bands cover once.
__global__ __launch_bounds__(256) void spline_y2(SplinePlanes p, int W, int H, int Hc, const double* __restrict__ w,
                                                 const int* __restrict__ st, int width) {
    __shared__ double ws[kSplineRowsY][kSplineMaxWidth];
    const int z = blockIdx.z;
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int k0 = blockIdx.y * kSplineRowsY;
    const int nk = min(kSplineRowsY, Hc - k0);
    // weights of the odd rows 2(k0+q)+1 (the last output row has none)
    for (int i = threadIdx.x; i < nk * width; i += 256) {
        const int q = i / width, y = 2 * (k0 + q) + 1;
        ws[q][i % width] = y < H - 1 ? w[(size_t)y * width + i % width] : 0.0;
    }
    __syncthreads();
    if (x >= W) return;
    const double* tmp = p.tmp[z] + x;
    float* dst = p.dst[z] + x;
    const float sc = p.scale[z];
    // even rows: the x-pass rows themselves (all read before any store); the
    // last coarse row is read up front for the clamped last output row
    double ev[kSplineRowsY];
#pragma unroll
    for (int q = 0; q < kSplineRowsY; ++q) ev[q] = tmp[(size_t)min(k0 + q, Hc - 1) * W];
    const double last = tmp[(size_t)(Hc - 1) * W];
#pragma unroll
    for (int q = 0; q < kSplineRowsY; ++q)
        if (q < nk) dst[(size_t)(2 * (k0 + q)) * W] = (float)ev[q] * sc;
    double s[kSplineRowsY];
    int sj[kSplineRowsY];
    int j0 = 0x3fffffff, j1 = 0;
#pragma unroll
    for (int q = 0; q < kSplineRowsY; ++q) {
        const int y = 2 * (k0 + q) + 1;
        const bool band = q < nk && y < H - 1;
        s[q] = 0.0;
        sj[q] = band ? st[y] : 0x3fffffff;
        if (band) {
            j0 = min(j0, sj[q]);
            j1 = max(j1, sj[q] + width);
        }
    }
    // 8 coarse rows per step, all loaded before any is used (rows past j1
    // read row j1 - 1 again); a row outside an output's band leaves its sum
    // as it was (a select, not a branch)
    for (int j = j0; j < j1; j += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = tmp[(size_t)min(j + u, j1 - 1) * W];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
#pragma unroll
            for (int q = 0; q < kSplineRowsY; ++q) {
                const int t = j + u - sj[q];
                const bool in = t >= 0 && t < width;
                const double sn = s[q] + ws[q][in ? t : 0] * v[u];
                s[q] = in ? sn : s[q];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kSplineRowsY; ++q) {
        const int y = 2 * (k0 + q) + 1;
        if (q < nk) {
            // the last output row is clamped onto the last coarse row
            float o = y < H - 1 ? (float)s[q] : (float)last;
            dst[(size_t)y * W] = o * sc;
        }
    }
}

// y pass: a thread owns column x of kSplineRowsY consecutive output rows and
// walks the coarse rows their bands cover once, adding each row's value to
// the outputs whose band holds it (in band order).
__global__ __launch_bounds__(256) void spline_y(SplinePlanes p, int W, int H, const double* __restrict__ w,
                                                const int* __restrict__ st, int width) {
    // the weights of this workgroup's output rows, staged in LDS
    __shared__ double ws[kSplineRowsY][kSplineMaxWidth];
    const int z = blockIdx.z;
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y0 = blockIdx.y * kSplineRowsY;
    const int ny = min(kSplineRowsY, H - y0);
    for (int i = threadIdx.x; i < ny * width; i += 256) ws[i / width][i % width] = w[(size_t)y0 * width + i];
    __syncthreads();
    if (x >= W) return;
    const int j0 = st[y0], j1 = st[y0 + ny - 1] + width;
    const double* tmp = p.tmp[z] + x;
    double s[kSplineRowsY];
    int sj[kSplineRowsY];
#pragma unroll
    for (int i = 0; i < kSplineRowsY; ++i) {
        s[i] = 0.0;
        sj[i] = i < ny ? st[y0 + i] : 0x3fffffff;
    }
    // 8 coarse rows per step, all loaded before any is used (rows past j1
    // read row j1 - 1 again); a row outside an output's band leaves its sum
    // as it was (a select, not a branch)
    for (int j = j0; j < j1; j += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = tmp[(size_t)min(j + u, j1 - 1) * W];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
#pragma unroll
            for (int i = 0; i < kSplineRowsY; ++i) {
                const int t = j + u - sj[i];
                const bool in = t >= 0 && t < width;
                const double sn = s[i] + ws[i][in ? t : 0] * v[u];
                s[i] = in ? sn : s[i];
            }
        }
    }
    float* dst = p.dst[z] + x;
    const float sc = p.scale[z];
#pragma unroll
    for (int i = 0; i < kSplineRowsY; ++i)
        if (i < ny) {
            float o = (float)s[i];
            dst[(size_t)(y0 + i) * W] = o * sc;
        }
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Bilinear sample with the reference's rounding (GenericPyramidalOpticalFlow.py:69-110).
__device__ __forceinline__ float warp_px(const float* __restrict__ img, int H, int W, float cy,
                                         float cx) {
    double ry = rint((double)cy), rx = rint((double)cx);
    int iy = (int)ry, ix = (int)rx;
    double dy = (double)cy - ry, dx = (double)cx - rx;
    int iyn = dy < 0 ? iy - 1 : iy + 1;
    int ixn = dx < 0 ? ix - 1 : ix + 1;
    dy = fabs(dy);
    dx = fabs(dx);
    iy = clampi(iy, 0, H - 1);
    iyn = clampi(iyn, 0, H - 1);
    ix = clampi(ix, 0, W - 1);
    ixn = clampi(ixn, 0, W - 1);
    double r = (1.0 - dy) * (1.0 - dx) * (double)img[(size_t)iy * W + ix];
    r = r + (1.0 - dy) * dx * (double)img[(size_t)iy * W + ixn];
    r = r + dy * (1.0 - dx) * (double)img[(size_t)iyn * W + ix];
    r = r + dy * dx * (double)img[(size_t)iyn * W + ixn];
    return (float)r;
}

__global__ void warp_sym(const float* __restrict__ im1, const float* __restrict__ im2,
                         const float* __restrict__ U, const float* __restrict__ V, int H, int W,
                         float* __restrict__ o1, float* __restrict__ o2) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= W) return;
    size_t i = (size_t)y * W + x;
    float hu = U[i] / 2.0f, hv = V[i] / 2.0f;
    float cy1 = (float)((double)y - (double)hv), cx1 = (float)((double)x - (double)hu);
    float cy2 = (float)((double)y + (double)hv), cx2 = (float)((double)x + (double)hu);
    o1[i] = warp_px(im1, H, W, cy1, cx1);
    o2[i] = warp_px(im2, H, W, cy2, cx2);
}

// Bilinear sample at given float32 coordinates (doBiLinearWarping).
__global__ void warp_coords_k(const float* __restrict__ img, const float* __restrict__ cy,
                              const float* __restrict__ cx, int H, int W, float* __restrict__ out) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= W) return;
    size_t i = (size_t)y * W + x;
    out[i] = warp_px(img, H, W, cy[i], cx[i]);
}

// d[z] += s[z]
__global__ void accumulate_k(Planes2 p, size_t n) {
    float* __restrict__ acc = p.d[blockIdx.z];
    const float* __restrict__ x = p.s[blockIdx.z];
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) acc[i] = acc[i] + x[i];
}

__global__ void sub_k(float* __restrict__ d, const float* __restrict__ a, const float* __restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) d[i] = a[i] - b[i];
}

__global__ void fill_zero_k(Planes2 q, size_t n) {
    float* __restrict__ p = q.d[blockIdx.z];
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.0f;
}

// d = c->final_buf ? s1 : s0 for plane blockIdx.z; s0 may alias d (the
// result may already be in place).
struct Select2 {
    const float* s0[2];
    const float* s1[2];
    float* d[2];
};
__global__ void select_copy_k(const SolveCtrl* c, Select2 p, size_t n, int add) {
    const int z = blockIdx.z;
    const float* s = c->final_buf ? p.s1[z] : p.s0[z];
    float* d = p.d[z];
    if (add) {
        // d += the selected plane (the same float add as accumulate_k)
        size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
        size_t stride = (size_t)gridDim.x * blockDim.x;
        for (; i < n; i += stride) d[i] = d[i] + s[i];
        return;
    }
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) d[i] = s[i];
}

constexpr int kRedBlocks = 512;
constexpr int kRedThreads = 256;

// Stage 1 of a deterministic max: per-block maxima.
__global__ void max_stage1(Planes2 p, size_t n, float* __restrict__ part_all) {
    __shared__ float sm[kRedThreads];
    const float* __restrict__ x = p.s[blockIdx.z];
    float* __restrict__ part = part_all + (size_t)blockIdx.z * gridDim.x;
    float m = -INFINITY;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        m = fmaxf(m, x[i]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] = fmaxf(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ void max_stage2(const float* __restrict__ part_all, int np, float* __restrict__ out_all) {
    __shared__ float sm[kRedThreads];
    const float* __restrict__ part = part_all + (size_t)blockIdx.x * np;
    float* __restrict__ out = out_all + blockIdx.x;
    float m = -INFINITY;
    for (int i = threadIdx.x; i < np; i += blockDim.x) m = fmaxf(m, part[i]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] = fmaxf(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sm[0];
}

// Stage 1 of a deterministic pair of squared-difference sums.
__global__ void dnorm_stage1(const float* __restrict__ A, const float* __restrict__ A0,
                             const float* __restrict__ B, const float* __restrict__ B0, size_t n,
                             double* __restrict__ part) {
    __shared__ double sa[kRedThreads], sb[kRedThreads];
    double a = 0.0, b = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        float da = A[i] - A0[i], db = B[i] - B0[i];
        a = a + (double)da * (double)da;
        b = b + (double)db * (double)db;
    }
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sa[threadIdx.x] = sa[threadIdx.x] + sa[threadIdx.x + s];
            sb[threadIdx.x] = sb[threadIdx.x] + sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sa[0];
        part[2 * blockIdx.x + 1] = sb[0];
    }
}

// Stage 1 of the mean endpoint error of two flows: per-block double sums
// of sqrt(du^2 + dv^2) (the differences and the root in float32).
__global__ void epe_stage1(const float* __restrict__ U1, const float* __restrict__ V1,
                           const float* __restrict__ U2, const float* __restrict__ V2, size_t n,
                           double* __restrict__ part) {
    __shared__ double sa[kRedThreads];
    double a = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const float du = U1[i] - U2[i], dv = V1[i] - V2[i];
        a = a + (double)sqrtf(du * du + dv * dv);
    }
    sa[threadIdx.x] = a;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sa[threadIdx.x] = sa[threadIdx.x] + sa[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sa[0];
}

__global__ void epe_stage2(const double* __restrict__ part, int np, size_t n, float* __restrict__ out) {
    __shared__ double sa[kRedThreads];
    double a = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) a = a + part[i];
    sa[threadIdx.x] = a;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sa[threadIdx.x] = sa[threadIdx.x] + sa[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = (float)(sa[0] / (double)n);
}

__global__ void dnorm_stage2(const double* __restrict__ part, int np, size_t n,
                             float* __restrict__ out) {
    __shared__ double sa[kRedThreads], sb[kRedThreads];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) {
        a = a + part[2 * i];
        b = b + part[2 * i + 1];
    }
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sa[threadIdx.x] = sa[threadIdx.x] + sa[threadIdx.x + s];
            sb[threadIdx.x] = sb[threadIdx.x] + sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = (float)((sqrt(sa[0]) + sqrt(sb[0])) / (double)n);
}

dim3 grid_rows(int W, int H, int bx = 256) { return dim3((W + bx - 1) / bx, H); }

int grid1d(size_t n) {
    size_t b = (n + 255) / 256;
    return (int)(b > 4096 ? 4096 : (b < 1 ? 1 : b));
}

// ---------------- host-side tables ----------------

double pil_filter(int filter, double x) {
    if (x < 0.0) x = -x;
    if (filter == 3) {
        const double a = -0.5;
        if (x < 1.0) return ((a + 2.0) * x - (a + 3.0)) * x * x + 1;
        if (x < 2.0) return (((x - 5) * x + 8) * x - 4) * a;
        return 0.0;
    }
    return x < 1.0 ? 1.0 - x : 0.0;
}

// Pillow's coefficient table for one axis (inSize -> outSize).
void pil_table(int inSize, int outSize, int filter, std::vector<double>& k, std::vector<int>& b,
               int& ks) {
    double scale = (double)inSize / outSize;
    double fs = scale < 1.0 ? 1.0 : scale;
    double support = (filter == 3 ? 2.0 : 1.0) * fs;
    ks = (int)std::ceil(support) * 2 + 1;
    k.assign((size_t)outSize * ks, 0.0);
    b.assign((size_t)outSize * 2, 0);
    for (int xx = 0; xx < outSize; ++xx) {
        double center = (xx + 0.5) * scale;
        double ss = 1.0 / fs;
        int xmin = (int)(center - support + 0.5);
        if (xmin < 0) xmin = 0;
        int xmax = (int)(center + support + 0.5);
        if (xmax > inSize) xmax = inSize;
        xmax -= xmin;
        double* kk = k.data() + (size_t)xx * ks;
        double ww = 0.0;
        for (int x = 0; x < xmax; ++x) {
            double w = pil_filter(filter, (x + xmin - center + 0.5) * ss);
            kk[x] = w;
            ww += w;
        }
        for (int x = 0; x < xmax; ++x)
            if (ww != 0.0) kk[x] /= ww;
        b[2 * xx] = xmin;
        b[2 * xx + 1] = xmax;
    }
}

int get_resize_table(ofri_ctx* ctx, int in, int out, int filter, ResizeTable** t) {
    auto key = std::make_tuple(in, out, filter);
    auto it = ctx->resize_tables.find(key);
    if (it != ctx->resize_tables.end()) {
        *t = &it->second;
        return OFRI_OK;
    }
    ResizeTable r{};
    r.in = in; r.out = out; r.filter = filter;
    std::vector<double> k;
    std::vector<int> b;
    pil_table(in, out, filter, k, b, r.ksize);
    r.first = b[0];
    r.last = b[2 * (out - 1)] + b[2 * (out - 1) + 1];
    OFRI_HIP(hipMalloc(&r.d_k, k.size() * sizeof(double)));
    OFRI_HIP(hipMalloc(&r.d_bounds, b.size() * sizeof(int)));
    OFRI_TRY(upload_table(ctx, r.d_k, k.data(), k.size() * sizeof(double), "resize weight table"));
    OFRI_TRY(upload_table(ctx, r.d_bounds, b.data(), b.size() * sizeof(int), "resize bounds table"));
    auto res = ctx->resize_tables.emplace(key, r);
    *t = &res.first->second;
    return OFRI_OK;
}

// Band of the not-a-knot interpolation weights, BAND coarse samples either
// side of the interval; the cardinal functions decay by (2 - sqrt 3) per
// sample, so the dropped tail is below 1e-18 of the kept weights.
constexpr int kSplineBand = 32;

void spline_weights(int n, int N, std::vector<double>& w, std::vector<int>& st, int& width) {
    // slopes S (n x n) of the not-a-knot spline through unit-spaced samples
    std::vector<double> S((size_t)n * n), cp(n), d(n);
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < n; ++i) d[i] = 0.0;
        if (j <= 2) d[0] = (j == 0 ? -5.0 : (j == 1 ? 4.0 : 1.0)) / 2.0;
        for (int i = 1; i < n - 1; ++i) d[i] = 3.0 * ((j == i + 1) - (j == i - 1));
        if (j >= n - 3) d[n - 1] = (j == n - 1 ? 5.0 : (j == n - 2 ? -4.0 : -1.0)) / 2.0;
        cp[0] = 2.0;
        for (int i = 1; i < n; ++i) {
            double a = (i == n - 1) ? 2.0 : 1.0;
            double bb = (i == n - 1) ? 1.0 : 4.0;
            double c = (i == n - 1) ? 0.0 : 1.0;
            double m = bb - a * cp[i - 1];
            cp[i] = c / m;
            d[i] = (d[i] - a * d[i - 1]) / m;
        }
        for (int i = n - 2; i >= 0; --i) d[i] -= cp[i] * d[i + 1];
        for (int i = 0; i < n; ++i) S[(size_t)i * n + j] = d[i];
    }
    width = n < 2 * kSplineBand + 2 ? n : 2 * kSplineBand + 2;
    w.assign((size_t)N * width, 0.0);
    st.assign(N, 0);
    std::vector<double> row(n);
    for (int k = 0; k < N; ++k) {
        bool over = (long long)k * n > (long long)(n - 1) * N;
        double u = over ? (double)(n - 1) : ((double)k * (double)n) / (double)N;
        int i = (int)std::floor(u);
        if (i > n - 2) i = n - 2;
        double t = u - i, t2 = t * t, t3 = t2 * t;
        double h00 = 2 * t3 - 3 * t2 + 1, h10 = t3 - 2 * t2 + t, h01 = -2 * t3 + 3 * t2, h11 = t3 - t2;
        for (int j = 0; j < n; ++j) row[j] = h10 * S[(size_t)i * n + j] + h11 * S[(size_t)(i + 1) * n + j];
        row[i] += h00;
        row[i + 1] += h01;
        int s0 = i - kSplineBand;
        if (s0 < 0) s0 = 0;
        if (s0 > n - width) s0 = n - width;
        st[k] = s0;
        for (int t2i = 0; t2i < width; ++t2i) w[(size_t)k * width + t2i] = row[s0 + t2i];
    }
}

int get_spline_table(ofri_ctx* ctx, int n, int N, SplineTable** t) {
    auto key = std::make_pair(n, N);
    auto it = ctx->spline_tables.find(key);
    if (it != ctx->spline_tables.end()) {
        *t = &it->second;
        return OFRI_OK;
    }
    SplineTable s{};
    s.n = n; s.N = N;
    std::vector<double> w;
    std::vector<int> st;
    spline_weights(n, N, w, st, s.width);
    std::vector<double> wt(w.size());
    for (int k = 0; k < N; ++k)
        for (int t = 0; t < s.width; ++t) wt[(size_t)t * N + k] = w[(size_t)k * s.width + t];
    OFRI_HIP(hipMalloc(&s.d_w, w.size() * sizeof(double)));
    OFRI_HIP(hipMalloc(&s.d_wT, wt.size() * sizeof(double)));
    OFRI_HIP(hipMalloc(&s.d_start, st.size() * sizeof(int)));
    OFRI_TRY(upload_table(ctx, s.d_w, w.data(), w.size() * sizeof(double), "spline weight table"));
    OFRI_TRY(upload_table(ctx, s.d_wT, wt.data(), wt.size() * sizeof(double), "spline weight table"));
    OFRI_TRY(upload_table(ctx, s.d_start, st.data(), st.size() * sizeof(int), "spline start table"));
    auto res = ctx->spline_tables.emplace(key, s);
    *t = &res.first->second;
    return OFRI_OK;
}

}  // namespace

int gauss_sep2(ofri_ctx* ctx, const float* src1, float* dst1, const float* src2, float* dst2, int H, int W,
               const float* w, int n) {
    if (n < 1 || n > kMaxTaps || (n & 1) == 0)
        return fail(OFRI_E_ARG, "gauss_sep: tap count %d must be odd and <= %d", n, kMaxTaps);
    if (H < n || W < n) return fail(OFRI_E_ARG, "gauss_sep: image %dx%d smaller than %d taps", H, W, n);
    Taps t{};
    for (int i = 0; i < n; ++i) t.w[i] = w[i];
    t.n = n;
    const int planes = src2 ? 2 : 1;
    // one launch unless an output is also an input (the staged tiles of other
    // workgroups would read what it writes): then the two passes via tmp
    const bool aliased = dst1 == src1 || (src2 && (dst1 == src2 || dst2 == src1 || dst2 == src2));
    if (n <= 2 * kGMaxH + 1 && !aliased) {
        dim3 g((W + kGTW - 1) / kGTW, (H + kGTH - 1) / kGTH, planes);
        gauss_fused<<<g, 256, 0, ctx->stream>>>(Planes2{{src1, src2}, {dst1, dst2}}, H, W, t);
        OFRI_HIP(hipGetLastError());
        return OFRI_OK;
    }
    void* tmp;
    OFRI_TRY(scratch(ctx, "gauss_tmp", (size_t)planes * H * W * sizeof(float), &tmp));
    float* t0 = (float*)tmp;
    float* t1 = t0 + (size_t)H * W;
    dim3 g((W + 255) / 256, H, planes);
    gauss_rows<<<g, 256, 0, ctx->stream>>>(Planes2{{src1, src2}, {t0, t1}}, H, W, t);
    gauss_cols<<<g, 256, 0, ctx->stream>>>(Planes2{{t0, t1}, {dst1, dst2}}, H, W, t);
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

int gauss_sep(ofri_ctx* ctx, const float* src, float* dst, int H, int W, const float* w, int n) {
    return gauss_sep2(ctx, src, dst, nullptr, nullptr, H, W, w, n);
}

int resize2(ofri_ctx* ctx, const float* src1, float* dst1, const float* src2, float* dst2, int H, int W, int H2,
            int W2, int filter) {
    if (filter != 2 && filter != 3) return fail(OFRI_E_ARG, "resize: filter %d not supported", filter);
    if (H < 1 || W < 1 || H2 < 1 || W2 < 1) return fail(OFRI_E_ARG, "resize: empty image");
    ResizeTable *th, *tv;
    OFRI_TRY(get_resize_table(ctx, W, W2, filter, &th));
    OFRI_TRY(get_resize_table(ctx, H, H2, filter, &tv));
    int Ht = tv->last - tv->first;
    const int planes = src2 ? 2 : 1;
    void* tmp;
    OFRI_TRY(scratch(ctx, "resize_tmp", (size_t)planes * Ht * W2 * sizeof(float), &tmp));
    float* t0 = (float*)tmp;
    float* t1 = t0 + (size_t)Ht * W2;
    resize_h<<<dim3((W2 + 255) / 256, Ht, planes), 256, 0, ctx->stream>>>(Planes2{{src1, src2}, {t0, t1}}, W, W2,
                                                                          Ht, tv->first, th->d_k, th->d_bounds,
                                                                          th->ksize);
    resize_v<<<dim3((W2 + 255) / 256, H2, planes), 256, 0, ctx->stream>>>(Planes2{{t0, t1}, {dst1, dst2}}, W2, H2,
                                                                          tv->first, tv->d_k, tv->d_bounds,
                                                                          tv->ksize);
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

int resize(ofri_ctx* ctx, const float* src, int H, int W, float* dst, int H2, int W2, int filter) {
    return resize2(ctx, src, dst, nullptr, nullptr, H, W, H2, W2, filter);
}

int spline_upsample2(ofri_ctx* ctx, const float* src_u, const float* src_v, int Hc, int Wc, float* dst_u,
                     float* dst_v, int H, int W, float scale_u, float scale_v) {
    if (Hc < 4 || Wc < 4) return fail(OFRI_E_ARG, "spline_upsample: coarse grid %dx%d below 4x4", Hc, Wc);
    if (H < 1 || W < 1) return fail(OFRI_E_ARG, "spline_upsample: empty output");
    // the x pass stages whole coarse rows in LDS (at most 64 KiB)
    const int nrows = std::min(kSplineRowsX, 16384 / Wc);
    if (nrows < 1) return fail(OFRI_E_ARG, "spline_upsample: coarse width %d above 16384", Wc);
    SplineTable *tx, *ty;
    OFRI_TRY(get_spline_table(ctx, Wc, W, &tx));
    OFRI_TRY(get_spline_table(ctx, Hc, H, &ty));
    if (tx->width > kSplineMaxWidth || ty->width > kSplineMaxWidth)
        return fail(OFRI_E_STATE, "spline_upsample: weight band wider than %d", kSplineMaxWidth);
    const int planes = src_v ? 2 : 1;
    void* tmp;
    OFRI_TRY(scratch(ctx, "spline_tmp", (size_t)planes * Hc * W * sizeof(double), &tmp));
    SplinePlanes p{};
    p.src[0] = src_u; p.src[1] = src_v;
    p.tmp[0] = (double*)tmp; p.tmp[1] = (double*)tmp + (size_t)Hc * W;
    p.dst[0] = dst_u; p.dst[1] = dst_v;
    p.scale[0] = scale_u; p.scale[1] = scale_v;
    if (W == 2 * Wc) {
        dim3 gx((Wc + 255) / 256, (Hc + nrows - 1) / nrows, planes);
        spline_x2<<<gx, 256, (size_t)nrows * Wc * sizeof(float), ctx->stream>>>(p, Wc, W, Hc, tx->d_wT, tx->d_start,
                                                                                tx->width, nrows);
    } else {
        dim3 gx((W + 255) / 256, (Hc + nrows - 1) / nrows, planes);
        spline_x<<<gx, 256, (size_t)nrows * Wc * sizeof(float), ctx->stream>>>(p, Wc, W, Hc, tx->d_wT, tx->d_start,
                                                                               tx->width, nrows);
    }
    if (H == 2 * Hc) {
        dim3 gy((W + 255) / 256, (Hc + kSplineRowsY - 1) / kSplineRowsY, planes);
        spline_y2<<<gy, 256, 0, ctx->stream>>>(p, W, H, Hc, ty->d_w, ty->d_start, ty->width);
    } else {
        dim3 gy((W + 255) / 256, (H + kSplineRowsY - 1) / kSplineRowsY, planes);
        spline_y<<<gy, 256, 0, ctx->stream>>>(p, W, H, ty->d_w, ty->d_start, ty->width);
    }
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

int spline_upsample(ofri_ctx* ctx, const float* src, int Hc, int Wc, float* dst, int H, int W,
                    float scale) {
    return spline_upsample2(ctx, src, nullptr, Hc, Wc, dst, nullptr, H, W, scale, 1.0f);
}

int warp_symmetric(ofri_ctx* ctx, const float* im1, const float* im2, const float* U,
                   const float* V, int H, int W, float* o1, float* o2) {
    if (H < 1 || W < 1) return fail(OFRI_E_ARG, "warp: empty image");
    warp_sym<<<grid_rows(W, H), 256, 0, ctx->stream>>>(im1, im2, U, V, H, W, o1, o2);
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

// ---- Liu-Shen warping mode (GenericPyramidalOpticalFlow.py:188-220) ----

// Scatter targets and fractional flow.  Source pixel i = (y, x) goes to
// (y + floor(v + 0.5), x + floor(u + 0.5)) as numpy's fancy assignment
// takes it: a negative target wraps once, a target still outside is an error
// (*err set), and of several sources the last in row-major order wins (the
// 1-based source index is max-reduced into owner, which starts at 0).
__global__ void lsw_target_k(const float* __restrict__ u, const float* __restrict__ v, int H, int W,
                             int* __restrict__ owner, int* __restrict__ err, float* __restrict__ du,
                             float* __restrict__ dv) {
    const size_t n = (size_t)H * W;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / W), x = (int)(i - (size_t)y * W);
        const float fu = floorf(u[i] + 0.5f), fv = floorf(v[i] + 0.5f);
        du[i] = u[i] - fu;
        dv[i] = v[i] - fv;
        if (!(fabsf(fu) < 1e9f) || !(fabsf(fv) < 1e9f)) {
            atomicOr(err, 1);
            continue;
        }
        long long tx = x + (long long)fu, ty = y + (long long)fv;
        if (tx < 0) tx += W;
        if (ty < 0) ty += H;
        if (tx < 0 || tx >= W || ty < 0 || ty >= H) {
            atomicOr(err, 1);
            continue;
        }
        atomicMax(&owner[(size_t)ty * W + tx], (int)i + 1);
    }
}

// The scattered image: each target takes its winning source, other pixels keep theirs.
__global__ void lsw_apply_k(const float* __restrict__ im, const int* __restrict__ owner, size_t n,
                            float* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int o = owner[i];
        out[i] = o ? im[o - 1] : im[i];
    }
}

// scipy's 'reflect' (half-sample symmetric) index of any position.
__device__ __forceinline__ int reflect_idx(int p, int L) {
    p %= 2 * L;
    if (p < 0) p += 2 * L;
    return p >= L ? 2 * L - 1 - p : p;
}

constexpr int kLswMaxR = 48;
struct LswTaps {
    double w[kLswMaxR + 1];   // w[0] centre, w[j] at offsets +-j
    int r;
};

// One axis of scipy's gaussian_filter on float32 data: float64 sum of the
// centre tap, then the symmetric pairs from the outermost inwards
// (NI_Correlate1D's order for a symmetric filter), rounded to float32.
__global__ void lsw_gauss_k(Planes2 p, int H, int W, int axis, LswTaps t) {
    const size_t n = (size_t)H * W;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float* __restrict__ src = p.s[blockIdx.z];
        const int y = (int)(i / W), x = (int)(i - (size_t)y * W);
        double acc = (double)src[i] * t.w[0];
        for (int j = t.r; j > 0; --j) {
            double a, b;
            if (axis == 0) {
                a = src[(size_t)reflect_idx(y - j, H) * W + x];
                b = src[(size_t)reflect_idx(y + j, H) * W + x];
            } else {
                a = src[(size_t)y * W + reflect_idx(x - j, W)];
                b = src[(size_t)y * W + reflect_idx(x + j, W)];
            }
            acc = acc + (a + b) * t.w[j];
        }
        p.d[blockIdx.z][i] = (float)acc;
    }
}

// Optical-flow-equation intensity correction of all but the last row and
// column, float32 without contraction: im - ((im_r du_r - im du) + (im_d dv_d - im dv)).
__global__ void lsw_correct_k(const float* __restrict__ im, const float* __restrict__ du,
                              const float* __restrict__ dv, int H, int W, float* __restrict__ out) {
    const size_t n = (size_t)H * W;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / W), x = (int)(i - (size_t)y * W);
        if (y < H - 1 && x < W - 1) {
            float a = im[i + 1] * du[i + 1];
            float b = im[i] * du[i];
            const float tdx = a - b;
            float c = im[i + W] * dv[i + W];
            float d = im[i] * dv[i];
            const float tdy = c - d;
            out[i] = im[i] - (tdx + tdy);
        } else {
            out[i] = im[i];
        }
    }
}

int warp_liushen(ofri_ctx* ctx, const float* im1, int H, int W, const float* u, const float* v,
                 const double* w, int r, float* out, int* d_err) {
    if (H < 1 || W < 1) return fail(OFRI_E_ARG, "warp_liushen: empty image");
    if (r < 0 || r > kLswMaxR) return fail(OFRI_E_ARG, "warp_liushen: Gaussian radius %d outside [0, %d]", r, kLswMaxR);
    const size_t n = (size_t)H * W;
    if (n >= (size_t)0x7fffffff) return fail(OFRI_E_ARG, "warp_liushen: %dx%d image too large", H, W);
    void* ws;
    OFRI_TRY(scratch(ctx, "lsw", 6 * n * sizeof(float), &ws));
    int* owner = (int*)ws;
    float* du = (float*)ws + n;
    float* dv = du + n;
    float* su = dv + n;
    float* sv = su + n;
    float* sc = sv + n;
    LswTaps t{};
    for (int j = 0; j <= r; ++j) t.w[j] = w[j];
    t.r = r;
    OFRI_TRY(fill_zero(ctx, (float*)owner, n));
    lsw_target_k<<<grid1d(n), 256, 0, ctx->stream>>>(u, v, H, W, owner, d_err, du, dv);
    lsw_apply_k<<<grid1d(n), 256, 0, ctx->stream>>>(im1, owner, n, sc);
    // gaussian_filter: axis 0 into (su, sv), then axis 1 back into (du, dv)
    lsw_gauss_k<<<dim3(grid1d(n), 1, 2), 256, 0, ctx->stream>>>(Planes2{{du, dv}, {su, sv}}, H, W, 0, t);
    lsw_gauss_k<<<dim3(grid1d(n), 1, 2), 256, 0, ctx->stream>>>(Planes2{{su, sv}, {du, dv}}, H, W, 1, t);
    lsw_correct_k<<<grid1d(n), 256, 0, ctx->stream>>>(sc, du, dv, H, W, out);
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

int warp_coords(ofri_ctx* ctx, const float* img, const float* cy, const float* cx, int H, int W,
                float* out) {
    if (H < 1 || W < 1) return fail(OFRI_E_ARG, "warp: empty image");
    warp_coords_k<<<grid_rows(W, H), 256, 0, ctx->stream>>>(img, cy, cx, H, W, out);
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

int accumulate2(ofri_ctx* ctx, float* acc1, const float* x1, float* acc2, const float* x2, size_t n) {
    if (n == 0) return OFRI_OK;
    accumulate_k<<<dim3(grid1d(n), 1, acc2 ? 2 : 1), 256, 0, ctx->stream>>>(Planes2{{x1, x2}, {acc1, acc2}}, n);
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

int accumulate(ofri_ctx* ctx, float* acc, const float* x, size_t n) {
    return accumulate2(ctx, acc, x, nullptr, nullptr, n);
}

int subtract(ofri_ctx* ctx, float* d, const float* a, const float* b, size_t n) {
    if (n == 0) return OFRI_OK;
    sub_k<<<grid1d(n), 256, 0, ctx->stream>>>(d, a, b, n);
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

int fill_zero2(ofri_ctx* ctx, float* p1, float* p2, size_t n) {
    if (n == 0) return OFRI_OK;
    fill_zero_k<<<dim3(grid1d(n), 1, p2 ? 2 : 1), 256, 0, ctx->stream>>>(Planes2{{nullptr, nullptr}, {p1, p2}}, n);
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

int fill_zero(ofri_ctx* ctx, float* p, size_t n) { return fill_zero2(ctx, p, nullptr, n); }

int select_copy2(ofri_ctx* ctx, const SolveCtrl* ctrl, const float* a0, const float* a1, float* ad,
                 const float* b0, const float* b1, float* bd, size_t n, bool add) {
    select_copy_k<<<dim3(grid1d(n), 1, bd ? 2 : 1), 256, 0, ctx->stream>>>(ctrl, Select2{{a0, b0}, {a1, b1}, {ad, bd}},
                                                                            n, add ? 1 : 0);
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

int select_copy(ofri_ctx* ctx, const SolveCtrl* ctrl, const float* s0, const float* s1, float* dst,
                size_t n) {
    return select_copy2(ctx, ctrl, s0, s1, dst, nullptr, nullptr, nullptr, n);
}

int reduce_max2(ofri_ctx* ctx, const float* x1, const float* x2, size_t n, float* d_out, float* ext_part) {
    static_assert(2 * kRedBlocks <= kMaxRedPart, "partial maxima");
    const int planes = x2 ? 2 : 1;
    void* part = ext_part;
    if (!part) OFRI_TRY(scratch(ctx, "max_part", 2 * kRedBlocks * sizeof(float), &part));
    int nb = (int)((n + kRedThreads - 1) / kRedThreads);
    if (nb > kRedBlocks) nb = kRedBlocks;
    if (nb < 1) nb = 1;
    max_stage1<<<dim3(nb, 1, planes), kRedThreads, 0, ctx->stream>>>(Planes2{{x1, x2}, {nullptr, nullptr}}, n,
                                                                       (float*)part);
    max_stage2<<<planes, kRedThreads, 0, ctx->stream>>>((const float*)part, nb, d_out);
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

int reduce_max(ofri_ctx* ctx, const float* x, size_t n, float* d_out) {
    return reduce_max2(ctx, x, nullptr, n, d_out);
}

int flow_epe(ofri_ctx* ctx, const float* U1, const float* V1, const float* U2, const float* V2, size_t n,
             float* d_out) {
    void* part;
    OFRI_TRY(scratch(ctx, "epe_part", kRedBlocks * sizeof(double), &part));
    int nb = (int)((n + kRedThreads - 1) / kRedThreads);
    if (nb > kRedBlocks) nb = kRedBlocks;
    if (nb < 1) nb = 1;
    epe_stage1<<<nb, kRedThreads, 0, ctx->stream>>>(U1, V1, U2, V2, n, (double*)part);
    epe_stage2<<<1, kRedThreads, 0, ctx->stream>>>((const double*)part, nb, n, d_out);
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

int diff_norm_err(ofri_ctx* ctx, const float* A, const float* A0, const float* B, const float* B0,
                  size_t n, float* d_out) {
    void* part;
    OFRI_TRY(scratch(ctx, "dnorm_part", 2 * kRedBlocks * sizeof(double), &part));
    int nb = (int)((n + kRedThreads - 1) / kRedThreads);
    if (nb > kRedBlocks) nb = kRedBlocks;
    if (nb < 1) nb = 1;
    dnorm_stage1<<<nb, kRedThreads, 0, ctx->stream>>>(A, A0, B, B0, n, (double*)part);
    dnorm_stage2<<<1, kRedThreads, 0, ctx->stream>>>((const double*)part, nb, n, d_out);
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

}  // namespace ofri
