// Horn-Schunck (reference: HornSchunck.py:51-126).
//
// Derivative setup is one per-pixel kernel.  The 600 Jacobi sweeps run as a
// temporally blocked kernel: a workgroup of WAVES wavefronts owns an extended
// tile 64 columns wide (one column per lane) and WAVES*R rows tall (R rows per
// lane, held in registers with fx, fy, ft and the denominator).  It fuses up
// to K sweeps per launch; the tile overlaps its neighbours by K pixels on
// every side, so after K sweeps the inner (64-2K) x (WAVES*R-2K) pixels are
// exact and are the only ones written.  Horizontal neighbours come from the
// adjacent lanes through DPP wave shifts, vertical ones from the lane's own
// registers, and only the two edge rows of each wavefront cross through LDS
// (one barrier per sweep).
//
// Numerics follow the reference exactly: the 3x3 average is formed in double
// from the float32 field and rounded to float32 (scipy.ndimage.convolve), the
// update is float32 without FMA contraction and with a correctly rounded
// quotient (HS_helper2; formed from a correctly rounded reciprocal and one
// FMA correction, which rounds identically to IEEE division).  The double sum
// is grouped symmetrically (vertical pairs, corners, edges), so a tile loaded
// as the mirror extension of the image stays an exact mirror extension under
// the sweeps and the border needs no special case.
//
// The Chebyshev-accelerated phase of the learned form (hs_cheb_tile) is not
// held to the plain method's iterates: it forms its sweep in float32 with
// fused multiply-adds, and plain sweeps finish every solve it starts.
#include <type_traits>

#include "ofri_internal.h"
#include "solve_check.h"
#include <cstring>

#pragma clang fp contract(off)

namespace ofri {
namespace {

// Value of the lane to the left (lane-1) / right (lane+1) of this lane; the
// vacated edge lane reads 0 (bound_ctrl), so the shift needs no copy of the
// old value and the compiler can fold it into a consuming add as a DPP
// operand.  Only a tile's outermost lanes see the 0, and they lie in the
// halo that is never written.
__device__ __forceinline__ float lane_left0(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float lane_right0(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xF, 0xF, true));
}

// Derivatives: HS(im2=I0, im1=I1) -> computeDerivatives(I1, I0)
// (HornSchunck.py:36,72,83,106-126), 2x2 stencils with mirror padding.
// What the first workgroup of the derivative launch also sets up for a
// solve with a stop test: its control word cleared (ctrl non-null).
struct SolveStart {
    SolveCtrl* ctrl;
};

__device__ __forceinline__ void clear_ctrl(const SolveStart& st) {
    SolveCtrl* c = st.ctrl;
    c->done = 0;
    c->sweeps = 0;
    c->final_buf = 0;
    c->redo = 0;
    c->redo_in = 0;
    c->last_err = 0.0f;
    c->limit = 0;
    c->base_buf = 0;
    c->last_nrm = 0.0f;
    c->slot[0] = SolveState{};
    c->slot[1] = SolveState{};
}

__global__ void hs_derivs_k(const float* __restrict__ I0, const float* __restrict__ I1, int H,
                            int W, float* __restrict__ fx, float* __restrict__ fy,
                            float* __restrict__ ft, SolveStart st) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    if (blockIdx.x == 0 && blockIdx.y == 0) {
        if (st.ctrl && threadIdx.x == 0) clear_ctrl(st);
    }
    if (j >= W) return;
    int i1 = (i + 1 < H) ? i + 1 : H - 2;
    int j1 = (j + 1 < W) ? j + 1 : W - 2;
    double a1 = I1[(size_t)i * W + j], b1 = I1[(size_t)i * W + j1];
    double c1 = I1[(size_t)i1 * W + j], d1 = I1[(size_t)i1 * W + j1];
    double a0 = I0[(size_t)i * W + j], b0 = I0[(size_t)i * W + j1];
    double c0 = I0[(size_t)i1 * W + j], d0 = I0[(size_t)i1 * W + j1];
    float cx1 = (float)(0.25 * a1 - 0.25 * b1 + 0.25 * c1 - 0.25 * d1);
    float cx0 = (float)(0.25 * a0 - 0.25 * b0 + 0.25 * c0 - 0.25 * d0);
    float cy1 = (float)(0.25 * a1 + 0.25 * b1 - 0.25 * c1 - 0.25 * d1);
    float cy0 = (float)(0.25 * a0 + 0.25 * b0 - 0.25 * c0 - 0.25 * d0);
    float ct1 = (float)(0.25 * a1 + 0.25 * b1 + 0.25 * c1 + 0.25 * d1);
    float ct0 = (float)(0.25 * a0 + 0.25 * b0 + 0.25 * c0 + 0.25 * d0);
    size_t q = (size_t)i * W + j;
    fx[q] = cx1 + cx0;
    fy[q] = cy1 + cy0;
    ft[q] = ct0 + (-ct1);
}

// The pre-filter and the derivatives in one pass over a 64 x 16 output
// tile.  Both frames are filtered into LDS over the tile plus one row and
// column on each side (the derivatives' 2 x 2 stencil and its mirror at the
// last row and column) exactly as gauss_fused filters them (row pass, then
// column pass, float32 in tap order, the reference's padding), then
// hs_derivs_k's arithmetic runs on the filtered values.  Taps: at most
// 2 * kDMaxH + 1.  512 threads: each half filters one frame.
constexpr int kDTW = 64, kDTH = 16, kDMaxH = 4;

__global__ __launch_bounds__(512) void hs_prefilter_derivs_k(const float* __restrict__ I0,
                                                             const float* __restrict__ I1, int H, int W, Taps t,
                                                             float* __restrict__ fx, float* __restrict__ fy,
                                                             float* __restrict__ ft, SolveStart st) {
    constexpr int FR = kDTH + 2, FC = kDTW + 2;   // filtered rows y0-1 .. y0+kDTH, columns x0-1 .. x0+kDTW
    __shared__ float sin_[2][FR + 2 * kDMaxH][FC + 2 * kDMaxH];
    __shared__ float srow[2][FR + 2 * kDMaxH][FC];
    __shared__ float filt[2][FR][FC];
    const int tid = threadIdx.x;
    const int img = tid >> 8, th = tid & 255;
    const int h = t.n / 2;
    const int x0 = blockIdx.x * kDTW, y0 = blockIdx.y * kDTH;
    if (blockIdx.x == 0 && blockIdx.y == 0) {
        if (st.ctrl && tid == 0) clear_ctrl(st);
    }
    const int R = FR + 2 * h, Cw = FC + 2 * h;
    {
        const float* __restrict__ src = img ? I1 : I0;
        // source positions [y0-1-h, y0-1+FR+h) x [x0-1-h, x0-1+FC+h); positions
        // outside the padding domain [-h, L+h) only feed filtered rows and
        // columns outside the image, which no derivative reads
        for (int i = th; i < R * Cw; i += 256) {
            const int rr = i / Cw, cc = i - rr * Cw;
            const int pr = y0 - 1 - h + rr, pc = x0 - 1 - h + cc;
            float v = 0.0f;
            if (pr >= -h && pr < H + h && pc >= -h && pc < W + h)
                v = src[(size_t)pf_index(pr, H, h) * W + pf_index(pc, W, h)];
            sin_[img][rr][cc] = v;
        }
        __syncthreads();
        for (int i = th; i < R * FC; i += 256) {
            const int rr = i / FC, cc = i - rr * FC;
            float acc = 0.0f;
            for (int j = 0; j < t.n; ++j) {
                float prod = t.w[j] * sin_[img][rr][cc + 2 * h - j];
                acc = acc + prod;
            }
            srow[img][rr][cc] = acc;
        }
        __syncthreads();
        for (int i = th; i < FR * FC; i += 256) {
            const int ry = i / FC, cc = i - ry * FC;
            float acc = 0.0f;
            for (int j = 0; j < t.n; ++j) {
                float prod = t.w[j] * srow[img][ry + 2 * h - j][cc];
                acc = acc + prod;
            }
            filt[img][ry][cc] = acc;
        }
        __syncthreads();
    }
    for (int i = tid; i < kDTH * kDTW; i += 512) {
        const int ry = i / kDTW, cx = i - ry * kDTW;
        const int y = y0 + ry, x = x0 + cx;
        if (y >= H || x >= W) continue;
        const int yi = y - y0 + 1, xi = x - x0 + 1;
        const int yi1 = ((y + 1 < H) ? y + 1 : H - 2) - y0 + 1;
        const int xi1 = ((x + 1 < W) ? x + 1 : W - 2) - x0 + 1;
        double a1 = filt[1][yi][xi], b1 = filt[1][yi][xi1];
        double c1 = filt[1][yi1][xi], d1 = filt[1][yi1][xi1];
        double a0 = filt[0][yi][xi], b0 = filt[0][yi][xi1];
        double c0 = filt[0][yi1][xi], d0 = filt[0][yi1][xi1];
        float cx1 = (float)(0.25 * a1 - 0.25 * b1 + 0.25 * c1 - 0.25 * d1);
        float cx0 = (float)(0.25 * a0 - 0.25 * b0 + 0.25 * c0 - 0.25 * d0);
        float cy1 = (float)(0.25 * a1 + 0.25 * b1 - 0.25 * c1 - 0.25 * d1);
        float cy0 = (float)(0.25 * a0 + 0.25 * b0 - 0.25 * c0 - 0.25 * d0);
        float ct1 = (float)(0.25 * a1 + 0.25 * b1 + 0.25 * c1 + 0.25 * d1);
        float ct0 = (float)(0.25 * a0 + 0.25 * b0 + 0.25 * c0 + 0.25 * d0);
        size_t q = (size_t)y * W + x;
        fx[q] = cx1 + cx0;
        fy[q] = cy1 + cy0;
        ft[q] = ct0 + (-ct1);
    }
}

struct HsPass {
    const float* in[2][2];   // [ping-pong][U, V]
    float* out[2][2];
    int src;                 // ping-pong index read (normal launch)
    // residual mode (hs_pass_k<..., true>): per-sweep norms and stop test
    double* part;            // two slots of [K][nblocks][2] squared-update sums
    size_t part_stride;      // doubles per slot
    int slot;                // launch parity within the phase (partials / state slot written)
    SolveCtrl* ctrl;
    int final_;              // 1: the phase's final launch (test, re-run, publish)
    int next_budget;         // final launch: > 0 arms a following phase of this many sweeps
    int rel_src;             // 1: buffer indices relative to ctrl->base_buf
    int zero_start;          // 1: the solve starts from U = V = 0 (the first sweep uses no input)
    double tol;
    int maxnum;
    size_t N;
    float* hist;             // per-sweep residuals (nullptr: not recorded)
    int check_last;          // 1: only a launch's last sweep is reduced and tested
    int rel;                 // 1: relative tolerance (result norms in row K of the partials)
    int copy_back;           // final launch with no re-run: the result, if in buffer 1, copied into buffer 0
    // Chebyshev mode (hs_pass_k<..., true, true>): previous iterate buffers,
    // the weight of every global sweep and the damping of the plain step
    const float* pin[2][2];
    float* pout[2][2];
    const float* omega;
    float gamma;
    int sweep0;              // global index of this launch's first sweep
    const float* fx;
    const float* fy;
    const float* ft;
    int H, W;
    int tiles_x, tiles_y;  // tile grid of this image (batched launches)
    int K;        // halo depth of the tiling (>= nsweeps)
    int nsweeps;  // sweeps in this launch
    float a2;     // alpha*alpha in float32
};

// Reflect-101 index of any integer position (period 2n-2): the extended
// tile is the mirror extension of the image, as scipy's 'mirror' mode sees it.
__device__ __forceinline__ int mirror_idx(int p, int n) {
    const int period = 2 * n - 2;
    p %= period;
    if (p < 0) p += period;
    return p < n ? p : period - p;
}

// A row's value at this lane and its left / right lane neighbours, in double.
struct Trip {
    double l, c, r;
};

__device__ __forceinline__ Trip trip(float x) {
    return Trip{(double)lane_left0(x), (double)x, (double)lane_right0(x)};
}

// Average of the 8 neighbours with weights wc (corners) and we (edges),
// summed in double and rounded to float.  Every addition pairs mirror-image
// operands, so a ghost cell's average equals its mirror pixel's bit for bit.
__device__ __forceinline__ float hs_avg(const Trip& up, const Trip& mid, const Trip& dn) {
    const double wc = (double)(float)(1.0 / 12.0);
    const double we = (double)(float)(1.0 / 6.0);
    double vl = up.l + dn.l;
    double vc = up.c + dn.c;
    double vr = up.r + dn.r;
    double corners = vl + vr;
    double edges = vc + (mid.l + mid.r);
    double s = wc * corners + we * edges;
    return (float)s;
}

// num / den correctly rounded, from rcp = RN(1/den): q = RN(num*rcp) is within
// an ulp of the quotient and one FMA residual correction rounds it exactly
// (Markstein's theorem for round-to-nearest binary arithmetic).  Tiny
// numerators, where the residual could underflow, take the IEEE division;
// the test for them is one branch for the whole wavefront, taken when any
// of its lanes has one.
__device__ __forceinline__ float quotient(float num, float den, float rcp) {
    float q = num * rcp;
    float res = fmaf(-den, q, num);
    q = fmaf(res, rcp, q);
    const bool tiny = fabsf(num) < 0x1p-100f;
    if (__builtin_expect(__any(tiny), 0)) {
        if (tiny) q = num / den;
    }
    return q;
}

// The extended tile is loaded as the mirror extension of the image (every
// field, including fx, fy, ft), so the image border needs no special case:
// the mirror symmetry of the extension is preserved exactly by every sweep.
// One workgroup's tile (bx, by) of an image whose launch has nblk tiles, this
// one being number blk; nsw sweeps from ping-pong buffer src.
template <int R, int WAVES, bool NORM>
__device__ __forceinline__ void hs_tile(const HsPass& a, int nsw, int src, int base, int bx, int by,
                                        int blk, int nblk) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int K = a.K, H = a.H, W = a.W;
    const int outW = 64 - 2 * K, outH = WAVES * R - 2 * K;
    const int gx = bx * outW - K + lane;
    const int gy0 = by * outH - K + w * R;
    const int cx = mirror_idx(gx, W);
    const bool own_x = lane >= K && lane < 64 - K && gx < W;

    float U[R], V[R], FX[R], FY[R], FT[R], DEN[R], RCP[R];
    src = __builtin_amdgcn_readfirstlane(src);
    base = __builtin_amdgcn_readfirstlane(base);
    const float* uin = a.in[src][0];
    const float* vin = a.in[src][1];
    // a zero start: the launch holding the first sweep (or its re-run) uses
    // no flow (the input buffers are read anyway, and dropped)
    const bool zs = a.zero_start && (NORM ? base : a.sweep0) == 0;
    // every load of the tile is issued before any value is used, so the
    // rows' memory latencies overlap instead of adding up
    size_t qs[R];
#pragma unroll
    for (int r = 0; r < R; ++r) qs[r] = (size_t)mirror_idx(gy0 + r, H) * W + cx;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        U[r] = uin[qs[r]];
        V[r] = vin[qs[r]];
        FX[r] = a.fx[qs[r]];
        FY[r] = a.fy[qs[r]];
        FT[r] = a.ft[qs[r]];
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (zs) {
            U[r] = 0.0f;
            V[r] = 0.0f;
        }
        float d = a.a2 + FX[r] * FX[r];
        DEN[r] = d + FY[r] * FY[r];
        RCP[r] = 1.0f / DEN[r];   // correctly rounded reciprocal
        // held, not re-derived: under register pressure the compiler would
        // otherwise redo this division (a dozen instructions) every sweep
        asm volatile("" : "+v"(RCP[r]));
    }

    __shared__ float edge[2][WAVES][2][2][64];  // [buf][wave][top,bottom][U,V][lane]
    // per-wave squared-update sums of a sweep, double-buffered: the next
    // sweep's barrier publishes them, so the norm costs no barrier of its own
    __shared__ double red[2][WAVES][2];
    // sweep sw reduces its update norm (every sweep, or a launch's last)
    auto normed = [&](int sw) { return NORM && !a.final_ && (!a.check_last || sw == nsw - 1); };
    double* part = NORM ? a.part + (size_t)a.slot * a.part_stride : nullptr;
    auto flush = [&](int sw) {
        double tu = 0.0, tv = 0.0;
        for (int q = 0; q < WAVES; ++q) {
            tu = tu + red[sw & 1][q][0];
            tv = tv + red[sw & 1][q][1];
        }
        part[((size_t)sw * nblk + blk) * 2] = tu;
        part[((size_t)sw * nblk + blk) * 2 + 1] = tv;
    };
    // one sweep; NS: reduce its update norm (compiled only into the sweeps
    // that test)
    auto sweep = [&](int s, auto ns_c) {
        constexpr bool norm = decltype(ns_c)::value;
        const int b = s & 1;
        edge[b][w][0][0][lane] = U[0];
        edge[b][w][0][1][lane] = V[0];
        edge[b][w][1][0][lane] = U[R - 1];
        edge[b][w][1][1][lane] = V[R - 1];
        __syncthreads();
        if (s > 0 && normed(s - 1) && threadIdx.x == 0) flush(s - 1);
        float uup = w > 0 ? edge[b][w - 1][1][0][lane] : U[0];
        float vup = w > 0 ? edge[b][w - 1][1][1][lane] : V[0];
        float udn = w < WAVES - 1 ? edge[b][w + 1][0][0][lane] : U[R - 1];
        float vdn = w < WAVES - 1 ? edge[b][w + 1][0][1][lane] : V[R - 1];

        Trip up_u = trip(uup), up_v = trip(vup);
        Trip cu = trip(U[0]), cv = trip(V[0]);
        double su = 0.0, sv = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            Trip dn_u = trip(r + 1 < R ? U[r + 1] : udn);
            Trip dn_v = trip(r + 1 < R ? V[r + 1] : vdn);
            float ub = hs_avg(up_u, cu, dn_u);
            float vb = hs_avg(up_v, cv, dn_v);
            float x = FX[r], y = FY[r];
            float num = x * ub;
            num = num + y * vb;
            num = num + FT[r];
            float den;
            if (norm) {
                // a sweep that reduces its update norm carries more live
                // state: the denominator is formed again (same operations,
                // same bits) instead of read, which keeps the tile in registers
                asm volatile("" : "+v"(x), "+v"(y));
                den = a.a2 + x * x;
                den = den + y * y;
            } else {
                den = DEN[r];
            }
            float der = quotient(num, den, RCP[r]);
            float un = ub - x * der;
            float vn = vb - y * der;
            if constexpr (norm) {
                const int ry = w * R + r, gy = gy0 + r;
                if (own_x && ry >= K && ry < WAVES * R - K && gy < H) {
                    float du = un - U[r], dv = vn - V[r];
                    su = su + (double)du * (double)du;
                    sv = sv + (double)dv * (double)dv;
                }
            }
            U[r] = un;
            V[r] = vn;
            up_u = cu; up_v = cv;
            cu = dn_u; cv = dn_v;
        }
        if constexpr (norm) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                su = su + __shfl_xor(su, off);
                sv = sv + __shfl_xor(sv, off);
            }
            if (lane == 0) {
                red[b][w][0] = su;
                red[b][w][1] = sv;
            }
        }
    };
    for (int s = 0; s < nsw; ++s) {
        if (NORM && normed(s)) sweep(s, std::true_type{});
        else sweep(s, std::false_type{});
    }
    if (nsw > 0 && normed(nsw - 1)) {
        __syncthreads();
        if (threadIdx.x == 0) flush(nsw - 1);
        if (a.rel) store_norm<R, WAVES>(U, V, own_x, w, gy0, K, H, part + ((size_t)K * nblk + blk) * 2);
    }

    float* uo = a.out[src ^ 1][0];
    float* vo = a.out[src ^ 1][1];
    if (own_x) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            int ry = w * R + r;
            int gy = gy0 + r;
            if (ry >= K && ry < WAVES * R - K && gy < H) {
                size_t q = (size_t)gy * W + gx;
                uo[q] = U[r];
                vo[q] = V[r];
            }
        }
    }
}


// The accelerated phase's tile: Chebyshev semi-iteration over the sweep T.
// With the damped step T_g(x) = x + gamma (T(x) - x), whose spectrum lies in
// [-sigma, sigma],
//     x_{k+1} = x_{k-1} + omega_{k+1} (T_g(x_k) - x_{k-1}),
// both multiply-adds fused; the previous iterate x_{k-1} rides along in
// registers.  These iterates are not the plain method's, and the phase only
// has to bring the residual |T(x_k) - x_k| to a multiple of the target before
// plain sweeps finish the solve, so T is formed in float32 throughout, scaled
// by 1/w (w = float32(1/12), the corner weight; the edge weight is 2w): the
// average's kernel is the separable [1,2,1] x [1,2,1] less its centre, so with
// the vertical pair sum p = x_up + x_down the column sum c = fma(2, x, p) and
// S = fma(2, p, c_l + c_r) = 12 x average (the centre never enters: no
// cancellation); with ft' = RN(ft / w), der' = fma(fx, S_u, fma(fy, S_v, ft'))
// times the held float32 reciprocal of den, t = fma(-fx, der', S_u) = T / w,
// and the update is T(x) - x = fma(w, t, -x).  Every addition pairs
// mirror-image operands, so the mirror-extended tile stays exact under the
// sweeps.  The residual tested is that of this T.  The re-run of a stopping
// launch makes its last sweep a plain one, x = RN(w t): the result is T(x_k)
// for the first iterate x_k that met the tolerance.
template <int R, int WAVES>
__device__ __forceinline__ void hs_cheb_tile(const HsPass& a, int nsw, int src, int base, int bx, int by,
                                             int blk, int nblk) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int K = a.K, H = a.H, W = a.W;
    const int outW = 64 - 2 * K, outH = WAVES * R - 2 * K;
    const int gx = bx * outW - K + lane;
    const int gy0 = by * outH - K + w * R;
    const int cx = mirror_idx(gx, W);
    const bool own_x = lane >= K && lane < 64 - K && gx < W;

    constexpr float wc = (float)(1.0 / 12.0);
    float U[R], V[R], UP[R], VP[R], FX[R], FY[R], FT[R], RCP[R];
    src = __builtin_amdgcn_readfirstlane(src);
    base = __builtin_amdgcn_readfirstlane(base);
    const float* uin = a.in[src][0];
    const float* vin = a.in[src][1];
    const float* upin = a.pin[src][0];
    const float* vpin = a.pin[src][1];
    // a zero start: the launch holding the first sweep (or its re-run) uses
    // no flow; before the first sweep x_{-1} = x_0 (the buffers are read
    // anyway, and dropped)
    const bool zs = a.zero_start && base == 0;
    // every load of the tile is issued before any value is used, so the
    // rows' memory latencies overlap instead of adding up
    size_t qs[R];
#pragma unroll
    for (int r = 0; r < R; ++r) qs[r] = (size_t)mirror_idx(gy0 + r, H) * W + cx;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        U[r] = uin[qs[r]];
        V[r] = vin[qs[r]];
        UP[r] = upin[qs[r]];
        VP[r] = vpin[qs[r]];
        FX[r] = a.fx[qs[r]];
        FY[r] = a.fy[qs[r]];
        FT[r] = a.ft[qs[r]];
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (zs) {
            U[r] = 0.0f;
            V[r] = 0.0f;
        }
        if (base == 0) {
            UP[r] = U[r];
            VP[r] = V[r];
        }
        FT[r] = FT[r] / wc;   // correctly rounded quotient
        RCP[r] = 1.0f / fmaf(FY[r], FY[r], fmaf(FX[r], FX[r], a.a2));
    }
    // the launch's Chebyshev weights, one per lane, broadcast per sweep
    float om_lane = 1.0f;
    if (lane < nsw) om_lane = a.omega[omega_index(base + lane)];

    __shared__ float edge[2][WAVES][2][2][64];  // [buf][wave][top,bottom][U,V][lane]
    __shared__ double red[2][WAVES][2];
    auto normed = [&](int sw) { return !a.final_ && (!a.check_last || sw == nsw - 1); };
    double* part = a.part + (size_t)a.slot * a.part_stride;
    auto flush = [&](int sw) {
        double tu = 0.0, tv = 0.0;
        for (int q = 0; q < WAVES; ++q) {
            tu = tu + red[sw & 1][q][0];
            tv = tv + red[sw & 1][q][1];
        }
        part[((size_t)sw * nblk + blk) * 2] = tu;
        part[((size_t)sw * nblk + blk) * 2 + 1] = tv;
    };
    const float g = a.gamma;
    // one sweep; NS: reduce its residual, PL: a plain sweep (x_{k+1} = T(x_k))
    auto sweep = [&](int s, auto ns_c, auto pl_c) {
        constexpr bool NS = decltype(ns_c)::value;
        constexpr bool PL = decltype(pl_c)::value;
        const int b = s & 1;
        edge[b][w][0][0][lane] = U[0];
        edge[b][w][0][1][lane] = V[0];
        edge[b][w][1][0][lane] = U[R - 1];
        edge[b][w][1][1][lane] = V[R - 1];
        __syncthreads();
        if (s > 0 && normed(s - 1) && threadIdx.x == 0) flush(s - 1);
        float up_u = w > 0 ? edge[b][w - 1][1][0][lane] : U[0];
        float up_v = w > 0 ? edge[b][w - 1][1][1][lane] : V[0];
        const float udn = w < WAVES - 1 ? edge[b][w + 1][0][0][lane] : U[R - 1];
        const float vdn = w < WAVES - 1 ? edge[b][w + 1][0][1][lane] : V[R - 1];
        const float om = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(om_lane), s));
        double qu = 0.0, qv = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const float pu = up_u + (r + 1 < R ? U[r + 1] : udn);
            const float pv = up_v + (r + 1 < R ? V[r + 1] : vdn);
            const float cu = fmaf(2.0f, U[r], pu);
            const float cv = fmaf(2.0f, V[r], pv);
            const float su = fmaf(2.0f, pu, lane_left0(cu) + lane_right0(cu));
            const float sv = fmaf(2.0f, pv, lane_left0(cv) + lane_right0(cv));
            const float x = FX[r], y = FY[r];
            const float der = fmaf(x, su, fmaf(y, sv, FT[r])) * RCP[r];
            const float tu = fmaf(-x, der, su);
            const float tv = fmaf(-y, der, sv);
            const float du = fmaf(wc, tu, -U[r]);
            const float dv = fmaf(wc, tv, -V[r]);
            if constexpr (NS) {
                const int ry = w * R + r, gy = gy0 + r;
                if (own_x && ry >= K && ry < WAVES * R - K && gy < H) {
                    qu = qu + (double)du * (double)du;
                    qv = qv + (double)dv * (double)dv;
                }
            }
            up_u = U[r];
            up_v = V[r];
            float nu, nv;
            if constexpr (PL) {
                nu = wc * tu;
                nv = wc * tv;
            } else {
                nu = fmaf(om, fmaf(g, du, U[r]) - UP[r], UP[r]);
                nv = fmaf(om, fmaf(g, dv, V[r]) - VP[r], VP[r]);
            }
            UP[r] = U[r];
            VP[r] = V[r];
            U[r] = nu;
            V[r] = nv;
        }
        if constexpr (NS) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                qu = qu + __shfl_xor(qu, off);
                qv = qv + __shfl_xor(qv, off);
            }
            if (lane == 0) {
                red[b][w][0] = qu;
                red[b][w][1] = qv;
            }
        }
    };
    using T_ = std::true_type;
    using F_ = std::false_type;
    for (int s = 0; s < nsw; ++s) {
        if (normed(s)) sweep(s, T_{}, F_{});
        else if (a.final_ && s == nsw - 1) sweep(s, F_{}, T_{});
        else sweep(s, F_{}, F_{});
    }
    if (nsw > 0 && normed(nsw - 1)) {
        __syncthreads();
        if (threadIdx.x == 0) flush(nsw - 1);
        if (a.rel) store_norm<R, WAVES>(U, V, own_x, w, gy0, K, H, part + ((size_t)K * nblk + blk) * 2);
    }

    float* uo = a.out[src ^ 1][0];
    float* vo = a.out[src ^ 1][1];
    if (own_x) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            int ry = w * R + r;
            int gy = gy0 + r;
            if (ry >= K && ry < WAVES * R - K && gy < H) {
                size_t q = (size_t)gy * W + gx;
                uo[q] = U[r];
                vo[q] = V[r];
                a.pout[src ^ 1][0][q] = UP[r];
                a.pout[src ^ 1][1][q] = VP[r];
            }
        }
    }
}

// The owned pixels of tile (bx, by) copied from buffer 1 into buffer 0.
template <int R, int WAVES>
__device__ __forceinline__ void hs_copy_owned(const HsPass& a, int bx, int by) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int K = a.K;
    const int outW = 64 - 2 * K, outH = WAVES * R - 2 * K;
    const int gx = bx * outW - K + lane;
    const int gy0 = by * outH - K + w * R;
    if (!(lane >= K && lane < 64 - K && gx < a.W)) return;
    // every value is read before any is written (rows outside the owned band
    // read row 0 and are dropped), so the reads' latencies overlap
    size_t q[R];
    bool own[R];
    float u[R], v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int ry = w * R + r, gy = gy0 + r;
        own[r] = ry >= K && ry < WAVES * R - K && gy < a.H;
        q[r] = own[r] ? (size_t)gy * a.W + gx : (size_t)gx;
        u[r] = a.in[1][0][q[r]];
        v[r] = a.in[1][1][q[r]];
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (own[r]) {
            a.out[0][0][q[r]] = u[r];
            a.out[0][1][q[r]] = v[r];
        }
}

template <int R, int WAVES, bool NORM, bool CHEB>
__global__ __launch_bounds__(64 * WAVES) void hs_pass_k(HsPass a) {
    int nsw = a.nsweeps, src = a.src, base = 0, fin = 0;
    const int blk = blockIdx.y * gridDim.x + blockIdx.x, nblk = gridDim.x * gridDim.y;
    int bx, by;
    xcd_tile(blk, nblk, gridDim.x, bx, by);
    if (NORM && !launch_args(a, blk, nblk, CHEB && !a.check_last, nsw, src, base, &fin)) {
        // a final launch with nothing to re-run: it delivers the result
        if (a.final_ && a.copy_back && fin == 1) hs_copy_owned<R, WAVES>(a, bx, by);
        return;
    }
    // partial sums are kept per tile (row-major), so their reduction order
    // does not depend on which workgroup ran which tile
    const int tile = by * gridDim.x + bx;
    if constexpr (CHEB) hs_cheb_tile<R, WAVES>(a, nsw, src, base, bx, by, tile, nblk);
    else hs_tile<R, WAVES, NORM>(a, nsw, src, base, bx, by, tile, nblk);
}


// Many images in one launch: workgroup b belongs to the image p with
// blk0[p] <= b < blk0[p+1] (blk0 ascending, blk0[0] = 0) and is tile
// b - blk0[p] of it, row-major over its tiles_x columns of tiles.  Every image
// shares the tile shape, depth and sweep count; pointers and sizes come from
// its own HsPass in the device table.
template <int R, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void hs_batch_k(const HsPass* __restrict__ tab,
                                                         const int* __restrict__ blk0, int n,
                                                         int nsweeps, int src) {
    // the tile order that keeps neighbouring tiles on one XCD
    const int b = xcd_run(blockIdx.x, gridDim.x);
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (blk0[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const HsPass& a = tab[lo];
    const int local = b - blk0[lo];
    hs_tile<R, WAVES, false>(a, nsweeps, src, 0, local % a.tiles_x, local / a.tiles_x, local,
                             a.tiles_x * a.tiles_y);
}

// Tile shapes (rows per lane, wavefronts per workgroup) compiled in; the
// tuning picks one by index.
struct HsVariant {
    int R, waves;
    void (*kernel)(HsPass);   // fixed sweep count
    void (*normed)(HsPass);   // with per-sweep residuals and the stop test
    void (*cheb)(HsPass);     // Chebyshev-accelerated, with the stop test
    void (*batched)(const HsPass*, const int*, int, int, int);
};
#define HS_SHAPE(R, WV)                                                                     \
    {R, WV, hs_pass_k<R, WV, false, false>, hs_pass_k<R, WV, true, false>, hs_pass_k<R, WV, true, true>, \
     hs_batch_k<R, WV>}
const HsVariant kHsVariants[] = {
    HS_SHAPE(16, 4), HS_SHAPE(8, 4), HS_SHAPE(8, 8), HS_SHAPE(12, 4), HS_SHAPE(16, 2),
    HS_SHAPE(4, 8), HS_SHAPE(24, 2), HS_SHAPE(12, 8),
    // same tile areas spread over 16 wavefronts: fewer rows per lane, more waves in flight
    HS_SHAPE(6, 16), HS_SHAPE(4, 16), HS_SHAPE(2, 16), HS_SHAPE(8, 16),
};
#undef HS_SHAPE
constexpr int kNumHsVariants = sizeof(kHsVariants) / sizeof(kHsVariants[0]);
// Defaults measured on gfx950 (600-sweep solves over every compiled shape
// and depth 4..12, after the XCD tile order and the batched tile loads):
// about 1M pixels and up favour 6-row lanes in 16-wavefront tiles fused 5
// sweeps deep (1.85 ms at 1024^2; 8 x 16 fused 10 deep, the earlier
// default, 1.93 ms: the narrower halo now outweighs the extra launches);
// smaller images 4-row lanes in 8-wavefront tiles, 5 deep (0.91 ms at 512^2).
// Chebyshev defaults (measured on gfx950 over every compiled shape and depth
// 6..14): 8 rows x 16 wavefronts fused 10 deep from about 1M pixels (2.93 us
// per sweep at 1024^2), 4 rows x 16 wavefronts fused 12 deep below (1.66 us
// per sweep at 512^2; 14 deep runs 1.55 us per sweep but tests the stop rule
// only every 14 sweeps); the finishing plain sweeps of a solve keep its shape.
void hs_cheb_default(size_t npx, int* variant, int* depth) {
    *variant = npx >= (size_t)700000 ? 11 : 9;
    *depth = npx >= (size_t)700000 ? 10 : 12;
}

void hs_default(size_t npx, int* variant, int* depth) {
    if (npx >= (size_t)700000) {
        *variant = 8;
        *depth = 5;
    } else {
        *variant = 5;
        *depth = 5;
    }
}

}  // namespace

int omega_table(ofri_ctx* ctx, double sigma, const float** out) {
    uint64_t key;
    static_assert(sizeof key == sizeof sigma, "sigma key");
    memcpy(&key, &sigma, sizeof key);
    DevBuf& b = ctx->omega_tabs[key];
    if (!b.p) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        OFRI_HIP(hipStreamIsCapturing(ctx->stream, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return fail(OFRI_E_STATE, "hs: Chebyshev weights for a new bound requested while capturing a graph");
        float t[kOmegaTab];
        const double s2 = sigma * sigma;
        double om = 1.0;
        for (int k = 0; k < kOmegaTab; ++k) {
            if (k == 1) om = 2.0 / (2.0 - s2);
            else if (k > 1) om = 1.0 / (1.0 - 0.25 * s2 * om);
            t[k] = (float)om;
        }
        void* p = nullptr;
        if (hipMalloc(&p, sizeof t) != hipSuccess) return fail(OFRI_E_NOMEM, "hs: Chebyshev weight table");
        if (upload_table(ctx, p, t, sizeof t, "hs: Chebyshev weight table") != OFRI_OK) {
            (void)hipFree(p);
            return OFRI_E_HIP;
        }
        b.p = p;
        b.bytes = sizeof t;
    }
    *out = (const float*)b.p;
    return OFRI_OK;
}

int hs_solve(ofri_ctx* ctx, const float* im1, const float* im2, int H, int W, float alpha,
             int niter, float* U, float* V, float* d_err, double tol, SolveCtrl* d_ctrl,
             float* d_hist, float cheb_b, float cheb_switch, bool check_last, int fin_budget,
             bool zero_start, const float* pre_w, int pre_n, bool rel, int fin_first) {
    if (pre_w && pre_n > 0 && (pre_n > kMaxTaps || (pre_n & 1) == 0 || H < pre_n || W < pre_n))
        return fail(OFRI_E_ARG, "hs: pre-filter of %d taps on a %dx%d image", pre_n, H, W);
    if (H < 2 || W < 2) return fail(OFRI_E_ARG, "hs: image %dx%d too small", H, W);
    if (niter < 0) return fail(OFRI_E_ARG, "hs: negative iteration count");
    if (zero_start && d_err) return fail(OFRI_E_ARG, "hs: a zero start keeps no initial flow for the error");
    const bool cheb = cheb_b > 0.0f;
    if (cheb && (!d_ctrl || !(cheb_b <= 0.999f)))
        return fail(OFRI_E_ARG, "hs: Chebyshev mode needs a stop test and a bound in (0, 0.999], got %g", cheb_b);
    if (rel && (!d_ctrl || !check_last || d_hist))
        return fail(OFRI_E_ARG, "hs: a relative tolerance needs the per-launch stop test (and no history)");
    if (fin_first < 0 || (fin_first > 0 && (!cheb || !(cheb_switch > 1.0f))))
        return fail(OFRI_E_ARG, "hs: a first finishing launch of %d sweeps needs Chebyshev sweeps with a "
                                "finishing phase", fin_first);
    size_t N = (size_t)H * W;
    int vi, depth;
    if (cheb) {
        hs_cheb_default((size_t)H * W, &vi, &depth);
        if (ctx->tuning.hs_depth > 0) depth = ctx->tuning.hs_depth;
        if (ctx->tuning.hs_variant > 0) vi = ctx->tuning.hs_variant - 1;
    } else {
        hs_default((size_t)H * W, &vi, &depth);
        shape_override(ctx, (size_t)H * W, true, &vi, &depth);
    }
    if (vi >= kNumHsVariants) return fail(OFRI_E_ARG, "hs: tile variant %d not compiled", vi + 1);
    const HsVariant& var = kHsVariants[vi];
    if (2 * depth >= var.R * var.waves || 2 * depth >= 64)
        return fail(OFRI_E_ARG, "hs: depth %d too large for tile %dx%d", depth, 64, var.R * var.waves);
    const int outW = 64 - 2 * depth, outH = var.waves * var.R - 2 * depth;
    dim3 grid((W + outW - 1) / outW, (H + outH - 1) / outH);
    const int nblk = grid.x * grid.y;

    void* buf;
    OFRI_TRY(scratch(ctx, "hs_fields", 7 * N * sizeof(float), &buf));
    float* fx = (float*)buf;
    float* fy = fx + N;
    float* ft = fy + N;
    float* Ub = ft + N;
    float* Vb = Ub + N;
    float* U0 = Vb + N;
    float* V0 = U0 + N;
    void* pbuf = nullptr;
    // per-sweep partial sums, two launch-parity slots (row `depth`: the
    // result norms of a relative test)
    const size_t pstride = (size_t)(depth + 1) * nblk * 2;
    if (d_ctrl) OFRI_TRY(scratch(ctx, "hs_part", 2 * pstride * sizeof(double), &pbuf));
    float* P0 = nullptr;
    const float* omega = nullptr;
    if (cheb) {
        void* cb;
        OFRI_TRY(scratch(ctx, "hs_cheb", 4 * N * sizeof(float), &cb));
        P0 = (float*)cb;
    }
    // spectrum of one sweep's linear part lies in [-1/3, b]: the 8-neighbour
    // average has eigenvalues in [-1/3, 1] and the per-pixel data projection
    // only shrinks them; b < 1 is the caller's bound on the slowest mode.
    // (The first accelerated launch takes the previous iterate from U, V.)
    const double lo = -1.0 / 3.0, hi = cheb_b;
    SolveStart start{};
    start.ctrl = d_ctrl;
    if (cheb) OFRI_TRY(omega_table(ctx, (hi - lo) / (2.0 - lo - hi), &omega));
    // computeDerivatives(I1=im2, I0=im1) of the pre-filtered frames (pre_w:
    // the filter, fused into the same pass when it has at most 9 taps); the
    // first workgroup also clears the control word
    const bool pre = pre_w && pre_n > 0;
    if (pre && pre_n <= 2 * kDMaxH + 1) {
        Taps t{};
        for (int i = 0; i < pre_n; ++i) t.w[i] = pre_w[i];
        t.n = pre_n;
        dim3 gd((W + kDTW - 1) / kDTW, (H + kDTH - 1) / kDTH);
        hs_prefilter_derivs_k<<<gd, 512, 0, ctx->stream>>>(im1, im2, H, W, t, fx, fy, ft, start);
    } else {
        if (pre) {
            void* fb;
            OFRI_TRY(scratch(ctx, "hs_prefiltered", 2 * N * sizeof(float), &fb));
            OFRI_TRY(gauss_sep2(ctx, im1, (float*)fb, im2, (float*)fb + N, H, W, pre_w, pre_n));
            im1 = (const float*)fb;
            im2 = (const float*)fb + N;
        }
        dim3 g((W + 255) / 256, H);
        hs_derivs_k<<<g, 256, 0, ctx->stream>>>(im1, im2, H, W, fx, fy, ft, start);
    }
    OFRI_HIP(hipGetLastError());
    if (d_err) {
        OFRI_HIP(hipMemcpyAsync(U0, U, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
        OFRI_HIP(hipMemcpyAsync(V0, V, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    }
    HsPass p{};
    p.in[0][0] = U; p.in[0][1] = V; p.in[1][0] = Ub; p.in[1][1] = Vb;
    p.out[0][0] = U; p.out[0][1] = V; p.out[1][0] = Ub; p.out[1][1] = Vb;
    p.fx = fx; p.fy = fy; p.ft = ft;
    p.H = H; p.W = W;
    p.a2 = alpha * alpha;
    p.K = depth;
    p.part = (double*)pbuf;
    p.part_stride = pstride;
    p.ctrl = d_ctrl;
    p.tol = tol;
    p.maxnum = niter;
    p.N = N;
    p.hist = d_hist;
    p.check_last = check_last ? 1 : 0;
    p.rel = rel ? 1 : 0;
    if (cheb) {
        p.pin[0][0] = P0; p.pin[0][1] = P0 + N; p.pin[1][0] = P0 + 2 * N; p.pin[1][1] = P0 + 3 * N;
        p.pout[0][0] = P0; p.pout[0][1] = P0 + N; p.pout[1][0] = P0 + 2 * N; p.pout[1][1] = P0 + 3 * N;
        p.omega = omega;
        p.gamma = (float)(2.0 / (2.0 - lo - hi));
    }
    // one phase: up to `sweeps` sweeps (fewer when the control word's shared
    // count reaches the budget), then the final launch (test of the last
    // launch, re-run of a launch stopped inside, result published).  Launch
    // parities start at slot0; next_budget > 0 makes the final launch arm a
    // following phase instead of copying the result back into U, V.  first
    // > 0: the first launch runs at most that many sweeps.  Returns the
    // number of regular launches in *nl.
    auto phase = [&](void (*kern)(HsPass), double t, int sweeps, int slot0, int next_budget, int first,
                     int* nl) -> int {
        p.tol = t;
        int done = 0, src = 0, j = 0;
        while (done < sweeps) {
            int ns = sweeps - done < depth ? sweeps - done : depth;
            if (j == 0 && first > 0 && ns > first) ns = first;
            p.nsweeps = ns;
            p.src = src;
            p.sweep0 = done;
            p.slot = (slot0 + j) & 1;
            p.final_ = 0;
            p.next_budget = 0;
            kern<<<grid, 64 * var.waves, 0, ctx->stream>>>(p);
            OFRI_HIP(hipGetLastError());
            src ^= 1;
            done += ns;
            ++j;
        }
        if (nl) *nl = j;
        if (d_ctrl) {
            if (niter > 0) {
                p.slot = (slot0 + j) & 1;
                p.final_ = 1;
                p.next_budget = next_budget;
                // with the test at launch ends only, no launch is ever re-run:
                // the final launch itself copies the result into U, V
                p.copy_back = next_budget <= 0 && check_last ? 1 : 0;
                kern<<<grid, 64 * var.waves, 0, ctx->stream>>>(p);
                OFRI_HIP(hipGetLastError());
                p.copy_back = 0;
                if (next_budget <= 0 && !check_last) OFRI_TRY(select_copy2(ctx, d_ctrl, U, Ub, U, V, Vb, V, N));
            }
        } else if (src != 0) {
            OFRI_HIP(hipMemcpyAsync(U, Ub, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
            OFRI_HIP(hipMemcpyAsync(V, Vb, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
        }
        return OFRI_OK;
    };
    p.rel_src = 0;
    p.zero_start = zero_start ? 1 : 0;
    if (!cheb) {
        OFRI_TRY(phase(d_ctrl ? var.normed : var.kernel, tol, niter, 0, 0, 0, nullptr));
    } else if (!(cheb_switch > 1.0f)) {
        OFRI_TRY(phase(var.cheb, tol, niter, 0, 0, 0, nullptr));
    } else {
        // Chebyshev sweeps stall at float32 rounding level, where plain sweeps
        // still settle: accelerate to cheb_switch * tol in at most niter - fin
        // sweeps, then finish with at most fin (fin_budget, default niter / 2)
        // plain sweeps to tol, on one sweep count and history (niter sweeps
        // in all).  The finishing phase starts from the accelerated phase's
        // result buffer (its launches take buffers relative to it) and the
        // opposite parity.  fin_first > 0: the finishing phase's first
        // launch runs at most that many sweeps (the plain sweeps after the
        // hand-over often need only a few).
        int fin = fin_budget > 0 ? fin_budget : (niter / 2 > 0 ? niter / 2 : 1);
        if (fin > niter) fin = niter;
        int nl = 0;
        OFRI_TRY(phase(var.cheb, tol * (double)cheb_switch, niter - fin, 0, niter > 0 ? fin : 0, 0, &nl));
        p.rel_src = 1;
        p.zero_start = 0;
        OFRI_TRY(phase(var.normed, tol, fin, nl + 1, 0, fin_first, nullptr));
        p.rel_src = 0;
    }
    if (d_err) OFRI_TRY(diff_norm_err(ctx, U, U0, V, V0, N, d_err));
    return OFRI_OK;
}

}  // namespace ofri

namespace ofri {

size_t hs_batch_table_bytes(int n) {
    return align256((size_t)n * sizeof(HsPass)) + align256((size_t)(n + 1) * sizeof(int));
}

// Batched Horn-Schunck: derivatives per image, then every sweep launch covers
// the tiles of all n images (one tile shape and depth for the batch, chosen
// for the batch's total pixel count unless tuned).
int hs_batch(ofri_ctx* ctx, const BatchImage* im, int n, float alpha, int niter, void* d_tab,
             bool upload) {
    if (n < 1) return fail(OFRI_E_ARG, "hs_batch: no images");
    if (niter < 0) return fail(OFRI_E_ARG, "hs_batch: negative iteration count");
    size_t total = 0;
    for (int i = 0; i < n; ++i) {
        if (im[i].H < 2 || im[i].W < 2)
            return fail(OFRI_E_ARG, "hs_batch: image %d is %dx%d", i, im[i].H, im[i].W);
        total += (size_t)im[i].H * im[i].W;
    }
    int vi, depth;
    hs_default(total, &vi, &depth);
    shape_override(ctx, total, true, &vi, &depth);
    if (vi >= kNumHsVariants) return fail(OFRI_E_ARG, "hs: tile variant %d not compiled", vi + 1);
    const HsVariant& var = kHsVariants[vi];
    if (2 * depth >= var.R * var.waves || 2 * depth >= 64)
        return fail(OFRI_E_ARG, "hs: depth %d too large for tile %dx%d", depth, 64, var.R * var.waves);
    const int outW = 64 - 2 * depth, outH = var.waves * var.R - 2 * depth;

    void* buf;
    OFRI_TRY(scratch(ctx, "hsb_fields", 7 * total * sizeof(float), &buf));
    std::vector<HsPass> tab(n);
    std::vector<int> blk0(n + 1);
    float* f = (float*)buf;
    blk0[0] = 0;
    for (int i = 0; i < n; ++i) {
        const int H = im[i].H, W = im[i].W;
        const size_t N = (size_t)H * W;
        float *fx = f, *fy = f + N, *ft = f + 2 * N, *Ub = f + 3 * N, *Vb = f + 4 * N;
        float *U0 = f + 5 * N, *V0 = f + 6 * N;
        f += 7 * N;
        dim3 g((W + 255) / 256, H);
        hs_derivs_k<<<g, 256, 0, ctx->stream>>>(im[i].im1, im[i].im2, H, W, fx, fy, ft, SolveStart{});
        OFRI_HIP(hipGetLastError());
        if (im[i].d_err) {
            OFRI_HIP(hipMemcpyAsync(U0, im[i].U, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
            OFRI_HIP(hipMemcpyAsync(V0, im[i].V, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
        }
        HsPass& p = tab[i];
        p = HsPass{};
        p.in[0][0] = im[i].U; p.in[0][1] = im[i].V; p.in[1][0] = Ub; p.in[1][1] = Vb;
        p.out[0][0] = im[i].U; p.out[0][1] = im[i].V; p.out[1][0] = Ub; p.out[1][1] = Vb;
        p.fx = fx; p.fy = fy; p.ft = ft;
        p.H = H; p.W = W; p.N = N;
        p.a2 = alpha * alpha;
        p.K = depth;
        p.tiles_x = (W + outW - 1) / outW;
        p.tiles_y = (H + outH - 1) / outH;
        blk0[i + 1] = blk0[i] + p.tiles_x * p.tiles_y;
    }
    HsPass* d_pass = (HsPass*)d_tab;
    int* d_blk0 = (int*)((char*)d_tab + align256((size_t)n * sizeof(HsPass)));
    if (upload) {
        OFRI_HIP(hipMemcpyAsync(d_pass, tab.data(), n * sizeof(HsPass), hipMemcpyHostToDevice, ctx->stream));
        OFRI_HIP(hipMemcpyAsync(d_blk0, blk0.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        OFRI_HIP(hipStreamSynchronize(ctx->stream));
    }
    int done = 0, src = 0;
    while (done < niter) {
        const int ns = niter - done < depth ? niter - done : depth;
        var.batched<<<blk0[n], 64 * var.waves, 0, ctx->stream>>>(d_pass, d_blk0, n, ns, src);
        OFRI_HIP(hipGetLastError());
        src ^= 1;
        done += ns;
    }
    f = (float*)buf;
    for (int i = 0; i < n; ++i) {
        const size_t N = (size_t)im[i].H * im[i].W;
        if (src != 0) {
            OFRI_HIP(hipMemcpyAsync(im[i].U, f + 3 * N, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
            OFRI_HIP(hipMemcpyAsync(im[i].V, f + 4 * N, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
        }
        if (im[i].d_err) OFRI_TRY(diff_norm_err(ctx, im[i].U, f + 5 * N, im[i].V, f + 6 * N, N, im[i].d_err));
        f += 7 * N;
    }
    return OFRI_OK;
}

// Effective Horn-Schunck launch shape for an image of npx pixels (1-based variant).
void hs_effective(ofri_ctx* ctx, size_t npx, int* variant1, int* depth) {
    int vi, d;
    hs_default(npx, &vi, &d);
    shape_override(ctx, npx, true, &vi, &d);
    *variant1 = vi + 1;
    *depth = d;
}
}  // namespace ofri
