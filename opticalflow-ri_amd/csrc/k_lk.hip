// Dense pyramidal Lucas-Kanade, one level (reference:
// denseLucasKanade_PyCL.py:113-169 with pyrlkDenseLargeW.cl:304-669).
//
// The reference launches one 8x8 workgroup per output pixel and reduces the
// window sums across its 64 work-items.  Here one thread owns one output
// pixel and a 16x16 workgroup owns a 16x16 block of outputs: the block's
// whole window region of I (block + 2*hw, plus the one-pixel Scharr ring,
// clamp-to-edge) is staged once in LDS together with its Scharr derivatives,
// and every pixel's window is read from there.  The structure tensor A and
// the per-iteration right-hand side b are summed in double; J is sampled
// bilinearly (the OpenCL LINEAR sampler at pixel centres, clamp-to-edge)
// with the pixel's fractional weights formed once per iteration, the four
// weighted samples accumulated with fused multiply-adds (the sampler's own
// precision is the hardware's; the oracle forms the same sums).
//
// The window is the reference's 32x32 grid of 4 tiles of 8 per axis with
// 0/1 weights (pyrlkDenseLargeW.cl:319-371): it is given here as one bit
// mask per axis (bit c: window column c, at offset c - hw from the pixel),
// so the asymmetric (vorticity) windows and the reference's rounding of the
// window to its tiles are the same code.
#include <climits>

#include "ofri_internal.h"

#pragma clang fp contract(off)

namespace ofri {
namespace {

constexpr int kTile = 16;
constexpr int kMaxHalf = 16;
constexpr int kReg = kTile + 32;                 // window region edge (window columns 0..31)
constexpr int kRegI = kReg + 2;                  // with the Scharr ring
constexpr int kStage = 72 * 72;                  // floats of J staged per Gauss-Newton step

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__global__ __launch_bounds__(kTile * kTile) void lk_dense_k(const float* __restrict__ I,
                                                            const float* __restrict__ J, int H, int W,
                                                            int hw, unsigned mx, unsigned my, int niter,
                                                            float* __restrict__ u, float* __restrict__ v,
                                                            unsigned char* __restrict__ status) {
    __shared__ float sI[kRegI * kRegI];
    __shared__ float sDx[kReg * kReg];
    __shared__ float sDy[kReg * kReg];
    const int wx = 32 - __builtin_clz(mx), wy = 32 - __builtin_clz(my);   // extent: last set bit + 1
    const unsigned full_x = wx >= 32 ? 0xFFFFFFFFu : (1u << wx) - 1u;        // every column counted
    const int regx = kTile + wx - 1, regy = kTile + wy - 1;                // region actually used
    const int regIx = regx + 2, regIy = regy + 2;
    const int bx0 = blockIdx.x * kTile, by0 = blockIdx.y * kTile;
    const int tid = threadIdx.y * kTile + threadIdx.x;
    // stage I over the region with its ring: image rows by0-hw-1 .. , clamped
    for (int k = tid; k < regIx * regIy; k += kTile * kTile) {
        int ry = k / regIx, rx = k % regIx;
        int y = clampi(by0 - hw - 1 + ry, 0, H - 1), x = clampi(bx0 - hw - 1 + rx, 0, W - 1);
        sI[ry * kRegI + rx] = I[(size_t)y * W + x];
    }
    __syncthreads();
    // Scharr derivatives at every region position from the staged (clamped) I
    for (int k = tid; k < regx * regy; k += kTile * kTile) {
        int ry = k / regx + 1, rx = k % regx + 1;
        const float* c = sI + ry * kRegI + rx;
        float dx = 3.0f * (c[-kRegI + 1] + c[kRegI + 1] - c[-kRegI - 1] - c[kRegI - 1]) +
                   10.0f * (c[1] - c[-1]);
        float dy = 3.0f * (c[kRegI - 1] + c[kRegI + 1] - c[-kRegI - 1] - c[-kRegI + 1]) +
                   10.0f * (c[kRegI] - c[-kRegI]);
        sDx[(ry - 1) * kReg + (rx - 1)] = dx;
        sDy[(ry - 1) * kReg + (rx - 1)] = dy;
    }
    __syncthreads();

    const int j = bx0 + threadIdx.x, i = by0 + threadIdx.y;
    const bool valid = j < W && i < H;
    const size_t g = valid ? (size_t)i * W + j : 0;
    // window of pixel (i, j) starts at region (threadIdx.y, threadIdx.x)
    const int o = threadIdx.y * kReg + threadIdx.x;
    const int oI = (threadIdx.y + 1) * kRegI + threadIdx.x + 1;
    // Window sums in double as the reference forms them, sum = sum + x * y
    // with x, y float32 values: their product is exact in double (24 + 24
    // significand bits of 53, exponents in range), so one rounded FMA gives
    // the same bits as the rounded product-then-add.
    double a11 = 0.0, a12 = 0.0, a22 = 0.0;
    for (int r = 0; r < wy; ++r) {
        if (!((my >> r) & 1u)) continue;
        const float* dxr = sDx + o + r * kReg;
        const float* dyr = sDy + o + r * kReg;
        if (mx == full_x) {   // every column counted: no mask test, loads issued ahead
#pragma unroll 4
            for (int c = 0; c < wx; ++c) {
                double dx = (double)dxr[c], dy = (double)dyr[c];
                a11 = fma(dx, dx, a11);
                a12 = fma(dx, dy, a12);
                a22 = fma(dy, dy, a22);
            }
        } else {
            for (int c = 0; c < wx; ++c) {
                if (!((mx >> c) & 1u)) continue;
                double dx = (double)dxr[c], dy = (double)dyr[c];
                a11 = fma(dx, dx, a11);
                a12 = fma(dx, dy, a12);
                a22 = fma(dy, dy, a22);
            }
        }
    }
    float A11 = (float)a11, A12 = (float)a12, A22 = (float)a22;
    float D = A11 * A22 - A12 * A12;
    unsigned char st = 1;
    // a singular tensor stops the pixel at once: flow left as it came in
    bool active = valid && !(D < 1.192092896e-07f);
    if (valid && !active) st = 0;
    A11 /= D;
    A12 /= D;
    A22 /= D;
    float px = 0.0f, py = 0.0f;
    if (active) {
        px = (float)j + u[g] - (float)hw;
        py = (float)i + v[g] - (float)hw;
    }
    // Every Gauss-Newton step samples J over the block's displaced windows:
    // the bounding box of the active pixels' windows is staged in LDS from
    // clamp-to-edge reads (indexed by unclamped positions, so every sample
    // reads what the clamped image read would), unless it exceeds the stage;
    // then the step reads J from global memory.  The whole block takes the
    // same number of steps (inactive pixels only keep the barriers).
    __shared__ float sJ[kStage];
    __shared__ int bb[4];   // min x0, min y0, max x0, max y0 of the active windows
    for (int k = 0; k < niter; ++k) {
        if (tid == 0) {
            bb[0] = INT_MAX;
            bb[1] = INT_MAX;
            bb[2] = INT_MIN;
            bb[3] = INT_MIN;
        }
        __syncthreads();   // also: the previous step's reads of sJ are done
        if (active && (px < (float)-hw || px >= (float)W || py < (float)-hw || py >= (float)H)) {
            st = 0;
            active = false;
        }
        float fx = 0.0f, fy = 0.0f;
        int x0 = 0, y0 = 0;
        if (active) {
            fx = floorf(px);
            fy = floorf(py);
            x0 = (int)fx;
            y0 = (int)fy;
            atomicMin(&bb[0], x0);
            atomicMin(&bb[1], y0);
            atomicMax(&bb[2], x0);
            atomicMax(&bb[3], y0);
        }
        __syncthreads();
        const int rx0 = bb[0], ry0 = bb[1];
        if (rx0 == INT_MAX) break;   // no active pixel left in the block
        // sample columns x0 .. x0 + wx (bilinear right neighbour included), rows likewise
        const int rw = bb[2] - rx0 + wx + 1, rh = bb[3] - ry0 + wy + 1;
        const bool staged = rw <= kStage && rh <= kStage && rw * rh <= kStage;   // no int overflow
        if (staged) {
            for (int q = tid; q < rw * rh; q += kTile * kTile) {
                const int ry = q / rw, rx = q - (q / rw) * rw;
                const int y = clampi(ry0 + ry, 0, H - 1), x = clampi(rx0 + rx, 0, W - 1);
                sJ[q] = J[(size_t)y * W + x];
            }
        }
        __syncthreads();
        if (!active) continue;
        const float a = px - fx, b = py - fy;
        const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b);
        const float w01 = (1.0f - a) * b, w11 = a * b;
        double b1 = 0.0, b2 = 0.0;   // (exact products: one FMA per term, as the window sums)
        if (staged) {
            for (int r = 0; r < wy; ++r) {
                if (!((my >> r) & 1u)) continue;
                const float* pr = sI + oI + r * kRegI;
                const float* dxr = sDx + o + r * kReg;
                const float* dyr = sDy + o + r * kReg;
                const float* Ja = sJ + (y0 + r - ry0) * rw + (x0 - rx0);
                const float* Jb = Ja + rw;
                // the bilinear right neighbours are carried to the next
                // column: two J reads per column instead of four.  A window
                // row with every column counted (symmetric windows) runs
                // without the per-column mask test, so the loads of the
                // unrolled columns can be issued ahead of their use.
                float ja0 = Ja[0], jb0 = Jb[0];
                if (mx == full_x) {
#pragma unroll 4
                    for (int c = 0; c < wx; ++c) {
                        const float ja1 = Ja[c + 1], jb1 = Jb[c + 1];
                        float t = w00 * ja0;
                        t = fmaf(w10, ja1, t);
                        t = fmaf(w01, jb0, t);
                        t = fmaf(w11, jb1, t);
                        float diff = t - pr[c];
                        b1 = fma((double)diff, (double)dxr[c], b1);
                        b2 = fma((double)diff, (double)dyr[c], b2);
                        ja0 = ja1;
                        jb0 = jb1;
                    }
                } else {
                    for (int c = 0; c < wx; ++c) {
                        const float ja1 = Ja[c + 1], jb1 = Jb[c + 1];
                        if ((mx >> c) & 1u) {
                            float t = w00 * ja0;
                            t = fmaf(w10, ja1, t);
                            t = fmaf(w01, jb0, t);
                            t = fmaf(w11, jb1, t);
                            float diff = t - pr[c];
                            b1 = fma((double)diff, (double)dxr[c], b1);
                            b2 = fma((double)diff, (double)dyr[c], b2);
                        }
                        ja0 = ja1;
                        jb0 = jb1;
                    }
                }
            }
        } else {
            const bool inside = x0 >= 0 && y0 >= 0 && x0 + wx < W && y0 + wy < H;
            for (int r = 0; r < wy; ++r) {
                if (!((my >> r) & 1u)) continue;
                const float* pr = sI + oI + r * kRegI;
                const float* dxr = sDx + o + r * kReg;
                const float* dyr = sDy + o + r * kReg;
                int ya = y0 + r, yb = ya + 1;
                if (!inside) {
                    ya = clampi(ya, 0, H - 1);
                    yb = clampi(yb, 0, H - 1);
                }
                const float* Ja = J + (size_t)ya * W;
                const float* Jb = J + (size_t)yb * W;
                for (int c = 0; c < wx; ++c) {
                    if (!((mx >> c) & 1u)) continue;
                    int xa = x0 + c, xb = xa + 1;
                    if (!inside) {
                        xa = clampi(xa, 0, W - 1);
                        xb = clampi(xb, 0, W - 1);
                    }
                    float t = w00 * Ja[xa];
                    t = fmaf(w10, Ja[xb], t);
                    t = fmaf(w01, Jb[xa], t);
                    t = fmaf(w11, Jb[xb], t);
                    float diff = t - pr[c];
                    b1 = fma((double)diff, (double)dxr[c], b1);
                    b2 = fma((double)diff, (double)dyr[c], b2);
                }
            }
        }
        float B1 = (float)b1, B2 = (float)b2;
        float ddx = (A12 * B2 - A22 * B1) * 32.0f;
        float ddy = (A12 * B1 - A11 * B2) * 32.0f;
        px = px + ddx;
        py = py + ddy;
        if (fabsf(ddx) < 0.01f && fabsf(ddy) < 0.01f) active = false;
    }
    if (!valid) return;
    if (!(D < 1.192092896e-07f)) {
        u[g] = px + (float)hw - (float)j;
        v[g] = py + (float)hw - (float)i;
    }
    if (status) status[g] = st;
}

}  // namespace

// Window mask of one axis (pyrlkDenseLargeW.cl:338-371): win = 2 hw + 1
// columns laid on 4 tiles of 8; tile 0 always counts, the first column of
// tile 1 counts unless `lo` (the asymmetric left/top flag), the rest of
// tiles 1-3 count below win - `hi` (right/bottom flag).  A window under 16
// columns uses tiles 0-1 only, and there the first column of tile 1 follows
// the `lo` flag alone.
unsigned lk_window_mask(int hw, int lo, int hi) {
    const int win = 2 * hw + 1;
    unsigned m = 0xFFu;
    for (int c = 8; c < 32; ++c) {
        bool on;
        if (win >= 16)
            on = c < 16 ? (c == 8 ? lo == 0 : true) : c < win - hi;
        else
            on = c < 16 ? (c == 8 ? lo == 0 : c < win - hi) : false;
        if (on) m |= 1u << c;
    }
    return m;
}

int lk_dense(ofri_ctx* ctx, const float* I, const float* J, int H, int W, int hw, int niter, float* u,
             float* v, unsigned char* status, const int* asym) {
    if (hw < 1 || hw > kMaxHalf) return fail(OFRI_E_ARG, "lk: half window %d outside 1..%d", hw, kMaxHalf);
    int L = 0, R = 0, T = 0, B = 0;
    if (asym) {
        L = asym[0], R = asym[1], T = asym[2], B = asym[3];
        if ((L | R | T | B) & ~1) return fail(OFRI_E_ARG, "lk: asymmetric window flags must be 0 or 1");
    }
    const unsigned mx = lk_window_mask(hw, L, R), my = lk_window_mask(hw, T, B);
    if (H < 1 || W < 1) return fail(OFRI_E_ARG, "lk: empty image");
    if (niter < 0) return fail(OFRI_E_ARG, "lk: negative iteration count");
    dim3 grid((W + kTile - 1) / kTile, (H + kTile - 1) / kTile);
    lk_dense_k<<<grid, dim3(kTile, kTile), 0, ctx->stream>>>(I, J, H, W, hw, mx, my, niter, u, v, status);
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

}  // namespace ofri
