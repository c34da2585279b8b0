// Liu-Shen physics-based refinement (reference:
// PhysicsBasedOpticalFlowLiuShen.py:47-158).
//
// Setup: one per-pixel kernel forms the normalised first frame I and the
// temporal difference It (the images divided by their own maxima).  Sweeps:
// temporally blocked like Horn-Schunck (64-column extended tiles, R rows per
// lane in registers with the eight per-pixel coefficients IIx, IIy, II, Ixt,
// Iyt and generate_invmatrix's 2x2 inverse B11, B12, B22 -- formed from I and
// It in each launch's load phase, as the setup of the reference forms them
// --, DPP for horizontal neighbours, LDS for the edge rows of each
// wavefront), with 'nearest' (clamp) and 'constant' (zero) boundaries in
// the workgroups whose tile touches the image border (interior tiles skip
// them): the column tests once per row entering a lane's register window,
// the image's first and last rows in a branch of their own.  Each launch also reduces, per fused sweep, the
// squared update of the pixels it owns; the last workgroup to finish turns
// the per-sweep sums into the reference's norm (||du||_2 + ||dv||_2)/(r*c),
// applies `while err > tol and k < maxnum` on the device, and marks a re-run
// of the stopping launch with fewer sweeps when the test held inside it.
// The launch sequence is fixed, so a whole solve can be captured in a HIP
// graph.
#include <cmath>
#include <map>
#include <mutex>

#include <type_traits>

#include "ofri_internal.h"
#include "solve_check.h"

// The allocation-free entry point's device memory (ofri_ls_noalloc): one
// array per GPU the library runs on, held until the process ends.  Host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device (ofri::noalloc_scratch): the array's name is a host-side
// stand-in, not a device address.
constexpr size_t kNoAllocBytes = (size_t)42 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

#pragma clang fp contract(off)

namespace ofri {
namespace {

// Value of the lane to the left (lane-1) / right (lane+1); the vacated edge
// lane reads 0 (bound_ctrl: no copy of the old value).  Only a tile's
// outermost lanes see the 0, and they lie in the halo that is never written
// (the halo is at least one column wide).
__device__ __forceinline__ float from_left(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float from_right(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xF, 0xF, true));
}
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Setup (PhysicsBasedOpticalFlowLiuShen.py:96-97,124-128): planes I and It
// of F (kLsPlanes planes of N).
__global__ void ls_setup_k(const float* __restrict__ im1, const float* __restrict__ im2, int H,
                           int W, const float* __restrict__ mx, float* __restrict__ F, size_t N,
                           SolveCtrl* ctrl) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    if (ctrl && i == 0 && j == 0) {
        // the solve's control word starts cleared
        ctrl->done = 0;
        ctrl->sweeps = 0;
        ctrl->final_buf = 0;
        ctrl->redo = 0;
        ctrl->redo_in = 0;
        ctrl->last_err = 0.0f;
        ctrl->limit = 0;
        ctrl->base_buf = 0;
        ctrl->last_nrm = 0.0f;
        ctrl->slot[0] = SolveState{};
        ctrl->slot[1] = SolveState{};
    }
    if (j >= W) return;
    const float m1 = mx[0], m2 = mx[1];
    auto I = [&](int a, int b) -> float { return im1[(size_t)a * W + b] / m1; };
    auto It = [&](int a, int b) -> float {
        float i2 = im2[(size_t)a * W + b] / m2;
        return i2 - I(a, b);
    };
    const size_t q = (size_t)i * W + j;
    F[q] = I(i, j);
    F[N + q] = It(i, j);
}

struct LsPass {
    const float* in[2][2];  // [ping-pong][u, v]
    float* out[2][2];
    const float* F;         // kLsPlanes planes of N: I, It
    size_t N;
    int H, W;
    int K;
    int nsweeps;
    int src;                // ping-pong index read (normal launch)
    float h;
    double* part;           // two slots of [K][nblocks][2] squared-update sums
    size_t part_stride;     // doubles per slot
    int slot;               // launch parity within the solve (partials / state slot written)
    double tol;             // stopping tolerance on the update norm
    int maxnum;             // sweep budget
    float* hist;            // per-sweep norms (nullptr: not recorded)
    SolveCtrl* ctrl;
    int final_;             // 1: the final launch (test, re-run, publish)
    int next_budget;        // (a following phase: unused here)
    int rel_src;            // (relative buffers: unused here)
    int tiles_x, tiles_y;   // tile grid of this image (batched launches)
    int check_last;         // 1: only a launch's last sweep is reduced and tested
    int rel;                // 1: relative tolerance (result norms in row K of the partials)
    // final launch with no re-run: the result (u, v of the buffer holding it)
    // written into dst (dst_add: added); dst[0] null: left where it is
    float* dst[2];
    int dst_add;
    // Chebyshev mode (ls_pass_k<..., true>): previous iterate buffers, the
    // weight of every global sweep and the damping of the plain step
    const float* pin[2][2];
    float* pout[2][2];
    const float* omega;
    float gamma;
    int sweep0;             // global index of this launch's first sweep
};

struct Raw {
    float l, c, r;
};

__device__ __forceinline__ Raw raw(float x) { return Raw{from_left(x), x, from_right(x)}; }

// One row of u or v around a pixel as the update reads it: with the
// 'nearest' boundary (clamped columns: the derivative and corner terms) and
// with the 'constant' one (zero outside the image: the 8-neighbour sum).  In a
// tile away from the image border both are the lane neighbours as they are;
// in a border tile the lane's column tests are applied once, when the row
// enters the register window.  Rows above and below the image are replaced
// where a pixel row reads them (border_rows).
struct Row {
    Raw n, z;
    // the values the double-precision sums read, converted once per row
    double nl, nr, zl, zc, zr;
};

template <bool EDGE>
__device__ __forceinline__ Row make_row(float x, bool el, bool er) {
    const Raw q = raw(x);
    Row w;
    if constexpr (!EDGE) {
        w.n = q;
        w.z = q;
    } else {
        w.n = Raw{el ? x : q.l, x, er ? x : q.r};
        w.z = Raw{el ? 0.0f : q.l, x, er ? 0.0f : q.r};
    }
    w.nl = (double)w.n.l;
    w.nr = (double)w.n.r;
    w.zl = (double)w.z.l;
    w.zc = (double)x;
    w.zr = (double)w.z.r;
    return w;
}

// The rows above and below pixel row gy as its update reads them: at the
// image's first (last) row the row above (below) is the row itself with the
// 'nearest' boundary and zero with the 'constant' one.
__device__ __forceinline__ Row border_row(const Row& c) {
    Row w;
    w.n = c.n;
    w.z = Raw{0.0f, 0.0f, 0.0f};
    w.nl = c.nl;
    w.nr = c.nr;
    w.zl = 0.0;
    w.zc = 0.0;
    w.zr = 0.0;
    return w;
}

// One LS update at a pixel from the rows above (p), at (c) and below (n) of
// u and v.
//   Dr, Dc, Fr, Fc: two-term sums scaled by a power of two are formed in
//   float32, which rounds exactly as the double sum rounded to float does;
//   M and the 8-neighbour sum H0 are summed in double as the reference does.
__device__ __forceinline__ void ls_update(const Row& pu, const Row& cu, const Row& nu, const Row& pv,
                                          const Row& cv, const Row& nv, float h, float iix, float iiy, float ii2,
                                          float ixt, float iyt, float b11, float b12, float b22, float& un,
                                          float& vn) {
    const Raw &uu = pu.n, &ud = nu.n, &vu = pv.n, &vd = nv.n;
    float dru = 0.5f * (ud.c - uu.c);
    float dcu = 0.5f * (cu.n.r - cu.n.l);
    float drv = 0.5f * (vd.c - vu.c);
    float dcv = 0.5f * (cv.n.r - cv.n.l);
    float fru = uu.c + ud.c;
    float fcv = cv.n.l + cv.n.r;
    // the reference's double sum of quarter-weighted corners, with the exact
    // power-of-two scale taken out of the sum (it commutes with rounding)
    float mu = (float)(((pu.nl - pu.nr) - nu.nl) + nu.nr) * 0.25f;
    float mv = (float)(((pv.nl - pv.nr) - nv.nl) + nv.nr) * 0.25f;
    // 8-neighbour sums with zero outside the image
    double hu = 0.0;
    hu = hu + pu.zl; hu = hu + pu.zc; hu = hu + pu.zr;
    hu = hu + cu.zl; hu = hu + cu.zr;
    hu = hu + nu.zl; hu = hu + nu.zc; hu = hu + nu.zr;
    double hv = 0.0;
    hv = hv + pv.zl; hv = hv + pv.zc; hv = hv + pv.zr;
    hv = hv + cv.zl; hv = hv + cv.zr;
    hv = hv + nv.zl; hv = hv + nv.zc; hv = hv + nv.zr;
    float h0u = (float)hu, h0v = (float)hv;
    float bu = (2.0f * iix) * dru;
    bu = bu + iix * dcv;
    bu = bu + iiy * drv;
    bu = bu + ii2 * fru;
    bu = bu + ii2 * mv;
    bu = bu + h * h0u;
    bu = bu + ixt;
    float bv = iiy * dru;
    bv = bv + iix * dcu;
    bv = bv + (2.0f * iiy) * dcv;
    bv = bv + ii2 * mu;
    bv = bv + ii2 * fcv;
    bv = bv + h * h0v;
    bv = bv + iyt;
    float t1 = b11 * bu, t2 = b12 * bv;
    un = -(t1 + t2);
    float t3 = b12 * bu, t4 = b22 * bv;
    vn = -(t3 + t4);
}

// The same update in float32 with fused multiply-adds, for the sweeps the
// learned form runs between its tests: full differences with the halves
// folded into the coefficients (0.5 iix, 0.5 iiy), the corner term as a
// difference of row differences, the 8-neighbour sum in pairs.
__device__ __forceinline__ void ls_update_fast(const Row& pu, const Row& cu, const Row& nu, const Row& pv,
                                               const Row& cv, const Row& nv, float h, float iix, float iiy,
                                               float ii2, float ixt, float iyt, float b11, float b12, float b22,
                                               float& un, float& vn) {
    const Raw &uu = pu.n, &ud = nu.n, &vu = pv.n, &vd = nv.n;
    const float dru = ud.c - uu.c, dcu = cu.n.r - cu.n.l;   // twice Dr, Dc
    const float drv = vd.c - vu.c, dcv = cv.n.r - cv.n.l;
    const float fru = uu.c + ud.c, fcv = cv.n.l + cv.n.r;
    const float mu = 0.25f * ((uu.l - uu.r) - (ud.l - ud.r));
    const float mv = 0.25f * ((vu.l - vu.r) - (vd.l - vd.r));
    const float h0u = ((pu.z.l + pu.z.r) + (nu.z.l + nu.z.r)) + ((pu.z.c + nu.z.c) + (cu.z.l + cu.z.r));
    const float h0v = ((pv.z.l + pv.z.r) + (nv.z.l + nv.z.r)) + ((pv.z.c + nv.z.c) + (cv.z.l + cv.z.r));
    const float hix = 0.5f * iix, hiy = 0.5f * iiy;
    float bu = fmaf(iix, dru, ixt);
    bu = fmaf(hix, dcv, bu);
    bu = fmaf(hiy, drv, bu);
    bu = fmaf(ii2, fru + mv, bu);
    bu = fmaf(h, h0u, bu);
    float bv = fmaf(hiy, dru, iyt);
    bv = fmaf(hix, dcu, bv);
    bv = fmaf(iiy, dcv, bv);
    bv = fmaf(ii2, mu + fcv, bv);
    bv = fmaf(h, h0v, bv);
    un = -fmaf(b11, bu, b12 * bv);
    vn = -fmaf(b12, bu, b22 * bv);
}

// CHEB: Chebyshev semi-iteration over the same sweep T (as in the
// Horn-Schunck kernel): x_{k+1} = x_{k-1} + omega_{k+1} ((x_k + gamma
// (T(x_k) - x_k)) - x_{k-1}), the previous iterate in registers; the
// reduced norm is the plain one, |T(x_k) - x_k|.  FAST: every sweep but a
// launch's last runs ls_update_fast; the last, the one a launch tests, is
// the reference update.
template <int R, int WAVES, bool EDGE, bool CHEB, bool FAST>
__device__ __forceinline__ void ls_sweeps(const LsPass& a, int nsw, int base, int blk, int nblk, int lane, int w,
                                          int gx, int gy0, float (&u)[R], float (&v)[R], float (&up)[R],
                                          float (&vp)[R], const float (&IIX)[R],
                                          const float (&IIY)[R], const float (&II)[R], const float (&IXT)[R],
                                          const float (&IYT)[R], const float (&B11)[R], const float (&B12)[R],
                                          const float (&B22)[R]) {
    __shared__ float edge[2][WAVES][2][2][64];
    // per-wave squared-update sums, double-buffered: the next sweep's barrier
    // publishes them (one barrier per sweep in all)
    __shared__ double red[2][WAVES][2];
    const int K = a.K, H = a.H, W = a.W;
    const bool el = gx == 0, er = gx == W - 1;
    const bool own_x = lane >= K && lane < 64 - K && gx < W;
    const float h = a.h;
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
    // one sweep; F: the float32 update (no norm reduced)
    auto sweep = [&](int s, auto f_c) {
        constexpr bool F = decltype(f_c)::value;
        const int b = s & 1;
        edge[b][w][0][0][lane] = u[0];
        edge[b][w][0][1][lane] = v[0];
        edge[b][w][1][0][lane] = u[R - 1];
        edge[b][w][1][1][lane] = v[R - 1];
        __syncthreads();
        if (s > 0 && normed(s - 1) && threadIdx.x == 0) flush(s - 1);
        const bool norm = !F && normed(s);
        float uup = w > 0 ? edge[b][w - 1][1][0][lane] : u[0];
        float vup = w > 0 ? edge[b][w - 1][1][1][lane] : v[0];
        float udn = w < WAVES - 1 ? edge[b][w + 1][0][0][lane] : u[R - 1];
        float vdn = w < WAVES - 1 ? edge[b][w + 1][0][1][lane] : v[R - 1];
        Row pu = make_row<EDGE>(uup, el, er), pv = make_row<EDGE>(vup, el, er);
        Row cu = make_row<EDGE>(u[0], el, er), cv = make_row<EDGE>(v[0], el, er);
        double su = 0.0, sv = 0.0;
        const float om = CHEB ? a.omega[omega_index(base + s)] : 1.0f;
        // the wavefront's first row, opaque to the compiler: the border tests
        // on it are formed per sweep in scalar registers rather than hoisted
        // out of the sweep loop and held in vector lanes
        int gys = __builtin_amdgcn_readfirstlane(gy0);
        if (EDGE) asm volatile("" : "+s"(gys));
#pragma unroll
        for (int r = 0; r < R; ++r) {
            Row nu = make_row<EDGE>(r + 1 < R ? u[r + 1] : udn, el, er);
            Row nv = make_row<EDGE>(r + 1 < R ? v[r + 1] : vdn, el, er);
            const int gy = gys + r;
            float un, vn;
            auto update = [&](const Row& pu_, const Row& nu_, const Row& pv_, const Row& nv_) {
                if constexpr (F)
                    ls_update_fast(pu_, cu, nu_, pv_, cv, nv_, h, IIX[r], IIY[r], II[r], IXT[r], IYT[r], B11[r],
                                   B12[r], B22[r], un, vn);
                else
                    ls_update(pu_, cu, nu_, pv_, cv, nv_, h, IIX[r], IIY[r], II[r], IXT[r], IYT[r], B11[r], B12[r],
                              B22[r], un, vn);
            };
            // gy is the same for the whole wavefront: the image's first and
            // last rows take their own branch
            if (EDGE && (gy == 0 || gy == H - 1)) {
                const bool top = gy == 0, bot = gy == H - 1;
                update(top ? border_row(cu) : pu, bot ? border_row(cu) : nu, top ? border_row(cv) : pv,
                       bot ? border_row(cv) : nv);
            } else {
                update(pu, nu, pv, nv);
            }
            if (norm) {
                const int ry = w * R + r;
                if (own_x && ry >= K && ry < WAVES * R - K && gy < H) {
                    float du = un - u[r], dv = vn - v[r];
                    su = su + (double)du * (double)du;
                    sv = sv + (double)dv * (double)dv;
                }
            }
            if (CHEB) {
                const float du = un - u[r], dv = vn - v[r];
                const float tu = u[r] + a.gamma * du, tv = v[r] + a.gamma * dv;
                un = up[r] + om * (tu - up[r]);
                vn = vp[r] + om * (tv - vp[r]);
                up[r] = u[r];
                vp[r] = v[r];
            }
            u[r] = un;
            v[r] = vn;
            pu = cu; pv = cv;
            cu = nu; cv = nv;
        }
        if (norm) {
            // fixed-order wave reduction, then waves in order
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
        if (FAST && s < nsw - 1) sweep(s, std::true_type{});
        else sweep(s, std::false_type{});
    }
    if (nsw > 0 && normed(nsw - 1)) {
        __syncthreads();
        if (threadIdx.x == 0) flush(nsw - 1);
    }
}

// One workgroup's tile (bx, by) of an image whose launch has nblk tiles, this
// one being number blk.
template <int R, int WAVES, bool CHEB = false, bool FAST = false>
__device__ __forceinline__ void ls_tile(const LsPass& a, int nsw, int src, int base, int bx, int by, int blk,
                                        int nblk) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int K = a.K, H = a.H, W = a.W;
    const int outW = 64 - 2 * K, outH = WAVES * R - 2 * K;
    const int x0 = bx * outW - K, y0 = by * outH - K;
    const int gx = x0 + lane;
    const int gy0 = y0 + w * R;
    const int cx = clampi(gx, 0, W - 1);
    const size_t N = a.N;
    // does any pixel of the extended tile lie on the image border?
    const bool edge_tile = x0 <= 0 || y0 <= 0 || x0 + 63 >= W - 1 || y0 + WAVES * R - 1 >= H - 1;

    float u[R], v[R], IIX[R], IIY[R], II[R], IXT[R], IYT[R], B11[R], B12[R], B22[R];
    float up[R], vp[R];
    const float* uin = a.in[src][0];
    const float* vin = a.in[src][1];
    // I and It on the lane's rows and one row either side (clamped rows, as
    // the setup's neighbour indices are)
    float Ir[R + 2], Tr[R + 2];
#pragma unroll
    for (int m = 0; m < R + 2; ++m) {
        const size_t q = (size_t)clampi(gy0 - 1 + m, 0, H - 1) * W + cx;
        Ir[m] = a.F[q];
        Tr[m] = a.F[N + q];
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int cy = clampi(gy0 + r, 0, H - 1);
        size_t q = (size_t)cy * W + cx;
        u[r] = uin[q];
        v[r] = vin[q];
        if (CHEB) {
            up[r] = a.pin[src][0][q];
            vp[r] = a.pin[src][1][q];
        }
        // the coefficients (PhysicsBasedOpticalFlowLiuShen.py:124-128,138 and
        // generate_invmatrix, :47-73) with the setup's operations: the row
        // neighbours are the loaded rows, the column neighbours the adjacent
        // lanes' values (a lane's column is clamped into the image, so at the
        // border its neighbour lane holds the clamped column).  Exact
        // wherever the sweeps use them (the tile's outermost lanes, whose
        // neighbours are missing, lie in the halo that is never written).
        const float x = Ir[r + 1], t = Tr[r + 1];
        const float xl = from_left(x), xr = from_right(x);
        const float dr = (float)(-0.5 * (double)Ir[r] + 0.5 * (double)Ir[r + 2]);
        const float dc = (float)(-0.5 * (double)xl + 0.5 * (double)xr);
        const float dtr = (float)(-0.5 * (double)Tr[r] + 0.5 * (double)Tr[r + 2]);
        const float dtc = (float)(-0.5 * (double)from_left(t) + 0.5 * (double)from_right(t));
        IIX[r] = x * dr;
        IIY[r] = x * dc;
        II[r] = x * x;
        IXT[r] = x * dtr;
        IYT[r] = x * dtc;
        // generate_invmatrix: the 2x2 system of the pixel and its inverse
        const int cnt = (cy > 0) + (cy < H - 1), cnt_c = (cx > 0) + (cx < W - 1);
        const float cm = (float)((cnt + 1) * (cnt_c + 1) - 1);
        const float d2r = (float)(1.0 * (double)Ir[r] - 2.0 * (double)x + 1.0 * (double)Ir[r + 2]);
        const float d2c = (float)(1.0 * (double)xl - 2.0 * (double)x + 1.0 * (double)xr);
        const float mm = (float)(0.25 * (double)from_left(Ir[r]) - 0.25 * (double)from_right(Ir[r]) -
                                 0.25 * (double)from_left(Ir[r + 2]) + 0.25 * (double)from_right(Ir[r + 2]));
        const float two_x = 2.0f * x;
        const float hc = a.h * cm;
        const float A11 = x * (d2r - two_x) - hc;
        const float A22 = x * (d2c - two_x) - hc;
        const float A12 = x * mm;
        const float det = A11 * A22 - A12 * A12;
        B11[r] = A22 / det;
        B12[r] = (-A12) / det;
        B22[r] = A11 / det;
    }
    if (edge_tile)
        ls_sweeps<R, WAVES, true, CHEB, FAST>(a, nsw, base, blk, nblk, lane, w, gx, gy0, u, v, up, vp, IIX, IIY,
                                              II, IXT, IYT, B11, B12, B22);
    else
        ls_sweeps<R, WAVES, false, CHEB, FAST>(a, nsw, base, blk, nblk, lane, w, gx, gy0, u, v, up, vp, IIX, IIY,
                                               II, IXT, IYT, B11, B12, B22);
    if (a.rel && nsw > 0 && !a.final_)
        store_norm<R, WAVES>(u, v, lane >= K && lane < 64 - K && gx < W, w, gy0, K, H,
                             a.part + (size_t)a.slot * a.part_stride + ((size_t)K * nblk + blk) * 2);

    float* uo = a.out[src ^ 1][0];
    float* vo = a.out[src ^ 1][1];
    if (lane >= K && lane < 64 - K && gx < W) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            int ry = w * R + r, gy = gy0 + r;
            if (ry >= K && ry < WAVES * R - K && gy < H) {
                size_t q = (size_t)gy * W + gx;
                uo[q] = u[r];
                vo[q] = v[r];
                if (CHEB) {
                    a.pout[src ^ 1][0][q] = up[r];
                    a.pout[src ^ 1][1][q] = vp[r];
                }
            }
        }
    }
}

// The owned pixels of tile (bx, by) of buffer `from` written (or added)
// into dst.
template <int R, int WAVES>
__device__ __forceinline__ void ls_deliver(const LsPass& a, int from, int bx, int by) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int K = a.K;
    const int outW = 64 - 2 * K, outH = WAVES * R - 2 * K;
    const int gx = bx * outW - K + lane;
    const int gy0 = by * outH - K + w * R;
    if (!(lane >= K && lane < 64 - K && gx < a.W)) return;
    const float* su = a.in[from][0];
    const float* sv = a.in[from][1];
    // every value is read before any is written (rows outside the owned
    // band read row 0 and are dropped), so the reads' latencies overlap
    size_t q[R];
    bool own[R];
    float u[R], v[R], du[R], dv[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int ry = w * R + r, gy = gy0 + r;
        own[r] = ry >= K && ry < WAVES * R - K && gy < a.H;
        q[r] = own[r] ? (size_t)gy * a.W + gx : (size_t)gx;
    }
    if (a.dst_add) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            u[r] = su[q[r]];
            v[r] = sv[q[r]];
            du[r] = a.dst[0][q[r]];
            dv[r] = a.dst[1][q[r]];
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            u[r] = du[r] + u[r];
            v[r] = dv[r] + v[r];
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            u[r] = su[q[r]];
            v[r] = sv[q[r]];
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (own[r]) {
            a.dst[0][q[r]] = u[r];
            a.dst[1][q[r]] = v[r];
        }
}

template <int R, int WAVES, bool CHEB, bool FAST = false>
__global__ __launch_bounds__(64 * WAVES) void ls_pass_k(LsPass a) {
    int nsw, src, base, fin = 0;
    const int blk = blockIdx.y * gridDim.x + blockIdx.x, nblk = gridDim.x * gridDim.y;
    int bx, by;
    xcd_tile(blk, nblk, gridDim.x, bx, by);
    if (!launch_args(a, blk, nblk, false, nsw, src, base, &fin)) {
        // a final launch with nothing to re-run: it delivers the result
        if (a.final_ && a.dst[0] && (a.dst[0] != a.in[fin][0] || a.dst_add))
            ls_deliver<R, WAVES>(a, fin, bx, by);
        return;
    }
    // partial sums are kept per tile (row-major), so their reduction order
    // does not depend on which workgroup ran which tile
    ls_tile<R, WAVES, CHEB, FAST>(a, nsw, src, base, bx, by, by * gridDim.x + bx, nblk);
}

// Many images in one launch (workgroup b is tile b - blk0[p] of image p,
// blk0 ascending from 0): each image has its own coefficient fields, buffers,
// partial sums and control word in its LsPass, so each stops on its own; the
// launch-wide sweep count, input buffer and re-run flag come as arguments.
template <int R, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void ls_batch_k(const LsPass* __restrict__ tab,
                                                         const int* __restrict__ blk0, int n,
                                                         int nsweeps, int src0, int slot, int final_) {
    // the tile order that keeps neighbouring tiles on one XCD
    const int b = xcd_run(blockIdx.x, gridDim.x);
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (blk0[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    LsPass a = tab[lo];
    a.nsweeps = nsweeps;
    a.src = src0;
    a.slot = slot;
    a.final_ = final_;
    int nsw, src, base;
    const int local = b - blk0[lo], nloc = a.tiles_x * a.tiles_y;
    if (!launch_args(a, local, nloc, false, nsw, src, base)) return;
    ls_tile<R, WAVES>(a, nsw, src, base, local % a.tiles_x, local / a.tiles_x, local, nloc);
}


struct LsVariant {
    int R, waves;
    void (*kernel)(LsPass);
    void (*batched)(const LsPass*, const int*, int, int, int, int, int);
    void (*cheb)(LsPass);
    void (*fast)(LsPass);   // float32 sweeps between the tests
};
#define LS_SHAPE(R, WV) \
    {R, WV, ls_pass_k<R, WV, false>, ls_batch_k<R, WV>, ls_pass_k<R, WV, true>, ls_pass_k<R, WV, false, true>}
const LsVariant kLsVariants[] = {
    LS_SHAPE(8, 8), LS_SHAPE(8, 4), LS_SHAPE(4, 8), LS_SHAPE(12, 4),
    LS_SHAPE(2, 16), LS_SHAPE(4, 16), LS_SHAPE(6, 8),
};
#undef LS_SHAPE
constexpr int kNumLsVariants = sizeof(kLsVariants) / sizeof(kLsVariants[0]);
// Defaults measured on gfx950 (60-sweep solves over every compiled shape and
// depth 3..8): 12 rows x 4 wavefronts 4 deep from about 0.7 Mpx (0.53 ms at
// 1024^2), 4 rows x 8 wavefronts 5 deep below (0.27 ms at 512^2, 12 % less
// than 8 rows x 4 wavefronts).
void ls_default(size_t npx, int* variant, int* depth) {
    if (npx >= (size_t)700000) {
        *variant = 3;
        *depth = 4;
    } else {
        *variant = 2;
        *depth = 5;
    }
}

// The batched launch from about 0.7 Mpx in all: 8 rows x 8 wavefronts 5 deep
// (measured on gfx950 over 64 uneven LS(HS) pairs of 256..2048-pixel sides:
// 5 % less per batch than the single-image shape, whose batched kernel holds
// one workgroup per CU; depths 4..6 within 0.5 %).
void ls_batch_default(size_t npx, int* variant, int* depth) {
    ls_default(npx, variant, depth);
    if (npx >= (size_t)700000) {
        *variant = 0;
        *depth = 5;
    }
}

}  // namespace

// Liu-Shen's setup planes for (im1, im2) into F (kLsPlanes planes of H * W),
// the images' maxima into mx (2 floats); the control word starts cleared.
static int ls_prepare(ofri_ctx* ctx, const float* im1, const float* im2, int H, int W, float* F,
                      float* mx, SolveCtrl* d_ctrl, float* max_part = nullptr) {
    const size_t N = (size_t)H * W;
    OFRI_TRY(reduce_max2(ctx, im1, im2, N, mx, max_part));
    dim3 g((W + 255) / 256, H);
    ls_setup_k<<<g, 256, 0, ctx->stream>>>(im1, im2, H, W, mx, F, N, d_ctrl);
    OFRI_HIP(hipGetLastError());
    return OFRI_OK;
}

int ls_solve(ofri_ctx* ctx, const float* im1, const float* im2, int H, int W, float h,
             int maxnum, double tol, float* u, float* v, SolveCtrl* d_ctrl, bool check_last, float cheb_lo,
             float cheb_hi, float* sink_u, float* sink_v, bool sink_add, bool fast, bool rel,
             const LsWorkspace* ws) {
    if ((sink_u == nullptr) != (sink_v == nullptr)) return fail(OFRI_E_ARG, "ls: one result sink given without the other");
    if (H < 2 || W < 2) return fail(OFRI_E_ARG, "ls: image %dx%d too small", H, W);
    if (maxnum < 0) return fail(OFRI_E_ARG, "ls: negative sweep budget");
    if (rel && (!d_ctrl || !check_last)) return fail(OFRI_E_ARG, "ls: a relative tolerance needs the per-launch test");
    const bool cheb = cheb_hi > 0.0f;
    if (ws && (cheb || fast)) return fail(OFRI_E_ARG, "ls: a caller workspace serves plain sweeps only");
    if (cheb && !(cheb_lo > -1.0f && cheb_lo < 0.0f && cheb_hi <= 0.999f))
        return fail(OFRI_E_ARG, "ls: Chebyshev interval [%g, %g] outside (-1, 0) x (0, 0.999]", cheb_lo, cheb_hi);
    size_t N = (size_t)H * W;
    int vi, depth;
    ls_default((size_t)H * W, &vi, &depth);
    // the float32 sweeps' registers leave room for 8 wavefronts of 4 rows
    // below about 0.7 Mpx (measured: 8 % less per sweep than 8 rows x 4
    // wavefronts at 512^2; the depth, and so the tested sweeps, unchanged)
    if (fast && (size_t)H * W < (size_t)700000) vi = 2;
    if (cheb) {
        vi = 2;      // 4 rows x 8 wavefronts: room for the previous iterate in registers
        depth = 6;
    }
    if (fast || cheb) {
        if (ctx->tuning.ls_depth > 0) depth = ctx->tuning.ls_depth;
        if (ctx->tuning.ls_variant > 0) vi = ctx->tuning.ls_variant - 1;
    } else {
        shape_override(ctx, (size_t)H * W, false, &vi, &depth);
    }
    if (vi >= kNumLsVariants) return fail(OFRI_E_ARG, "ls: tile variant %d not compiled", vi + 1);
    const LsVariant& var = kLsVariants[vi];
    if (2 * depth >= var.R * var.waves || 2 * depth >= 64)
        return fail(OFRI_E_ARG, "ls: depth %d too large for tile %dx%d", depth, 64, var.R * var.waves);
    const int outW = 64 - 2 * depth, outH = var.waves * var.R - 2 * depth;
    dim3 grid((W + outW - 1) / outW, (H + outH - 1) / outH);
    int nblk = grid.x * grid.y;
    void *fbuf, *pbuf, *mbuf;
    // per-sweep partial sums, two launch-parity slots (row `depth`: the result
    // norms of a relative test)
    const size_t pstride = (size_t)(depth + 1) * nblk * 2;
    if (ws) {
        if (ws->part_doubles < 2 * pstride)
            return fail(OFRI_E_ARG, "ls: workspace holds %zu partial sums, %zu needed", ws->part_doubles, 2 * pstride);
        fbuf = ws->fields;
        pbuf = ws->part;
        mbuf = ws->mx;
    } else {
        OFRI_TRY(scratch(ctx, "ls_fields", (kLsFieldPlanes + (cheb ? 4 : 0)) * N * sizeof(float), &fbuf));
        OFRI_TRY(scratch(ctx, "ls_part", 2 * pstride * sizeof(double), &pbuf));
        OFRI_TRY(scratch(ctx, "ls_max", 2 * sizeof(float), &mbuf));
    }
    float* F = (float*)fbuf;
    float* ub = F + kLsPlanes * N;
    float* vb = F + (kLsPlanes + 1) * N;
    OFRI_TRY(ls_prepare(ctx, im1, im2, H, W, F, (float*)mbuf, d_ctrl, ws ? ws->max_part : nullptr));

    LsPass p{};
    p.in[0][0] = u; p.in[0][1] = v; p.in[1][0] = ub; p.in[1][1] = vb;
    p.out[0][0] = u; p.out[0][1] = v; p.out[1][0] = ub; p.out[1][1] = vb;
    p.F = F; p.N = N; p.H = H; p.W = W; p.K = depth; p.h = h;
    p.part = (double*)pbuf;
    p.part_stride = pstride;
    p.ctrl = d_ctrl;
    p.tol = tol;
    p.maxnum = maxnum;
    p.check_last = check_last ? 1 : 0;
    p.rel = rel ? 1 : 0;
    if (fast && (cheb || !check_last)) return fail(OFRI_E_ARG, "ls: float32 sweeps need the per-launch test alone");
    auto kern = cheb ? var.cheb : (fast ? var.fast : var.kernel);
    if (cheb) {
        // previous-iterate buffers (the first step's previous iterate is the
        // start) and the weights for [cheb_lo, cheb_hi]
        float* P0 = F + kLsFieldPlanes * N;
        OFRI_HIP(hipMemcpyAsync(P0, u, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
        OFRI_HIP(hipMemcpyAsync(P0 + N, v, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
        const double lo = cheb_lo, hi = cheb_hi;
        const float* omega = nullptr;
        OFRI_TRY(omega_table(ctx, (hi - lo) / (2.0 - lo - hi), &omega));
        p.pin[0][0] = P0; p.pin[0][1] = P0 + N; p.pin[1][0] = P0 + 2 * N; p.pin[1][1] = P0 + 3 * N;
        p.pout[0][0] = P0; p.pout[0][1] = P0 + N; p.pout[1][0] = P0 + 2 * N; p.pout[1][1] = P0 + 3 * N;
        p.omega = omega;
        p.gamma = (float)(2.0 / (2.0 - lo - hi));
    }
    int done = 0, src = 0, j = 0;
    while (done < maxnum) {
        int ns = maxnum - done < depth ? maxnum - done : depth;
        p.nsweeps = ns;
        p.src = src;
        p.sweep0 = done;
        p.slot = j & 1;
        p.final_ = 0;
        kern<<<grid, 64 * var.waves, 0, ctx->stream>>>(p);
        OFRI_HIP(hipGetLastError());
        src ^= 1;
        done += ns;
        ++j;
    }
    if (maxnum > 0) {
        // tests the last launch, re-runs a launch the test stopped inside,
        // publishes the result; with the test at launch ends only no launch
        // is ever re-run, and the final launch itself delivers the result
        p.slot = j & 1;
        p.final_ = 1;
        if (check_last) {
            p.dst[0] = sink_u ? sink_u : u;
            p.dst[1] = sink_u ? sink_v : v;
            p.dst_add = sink_u && sink_add ? 1 : 0;
        }
        kern<<<grid, 64 * var.waves, 0, ctx->stream>>>(p);
        OFRI_HIP(hipGetLastError());
        if (check_last) return OFRI_OK;
        if (!sink_u) OFRI_TRY(select_copy2(ctx, d_ctrl, u, ub, u, v, vb, v, N));
    }
    // the result (the start when no sweep ran: the control word says buffer 0)
    if (sink_u) OFRI_TRY(select_copy2(ctx, d_ctrl, u, ub, sink_u, v, vb, sink_v, N, sink_add));
    return OFRI_OK;
}

}  // namespace ofri

namespace ofri {

size_t ls_batch_table_bytes(int n) {
    return align256((size_t)n * sizeof(LsPass)) + align256((size_t)(n + 1) * sizeof(int));
}

// Batched Liu-Shen: setup per image, then every sweep launch covers the tiles
// of all n images; each image runs its own stop test on its own control word
// (a finished image's workgroups return at once), and one re-run launch
// finishes every image that stopped inside a fused launch.
int ls_batch(ofri_ctx* ctx, const BatchImage* im, int n, float h, int maxnum, double tol,
             SolveCtrl* d_ctrls, void* d_tab, bool upload) {
    if (n < 1) return fail(OFRI_E_ARG, "ls_batch: no images");
    if (maxnum < 0) return fail(OFRI_E_ARG, "ls_batch: negative sweep budget");
    size_t total = 0;
    for (int i = 0; i < n; ++i) {
        if (im[i].H < 2 || im[i].W < 2)
            return fail(OFRI_E_ARG, "ls_batch: image %d is %dx%d", i, im[i].H, im[i].W);
        total += (size_t)im[i].H * im[i].W;
    }
    int vi, depth;
    ls_batch_default(total, &vi, &depth);
    shape_override(ctx, total, false, &vi, &depth);
    if (vi >= kNumLsVariants) return fail(OFRI_E_ARG, "ls: tile variant %d not compiled", vi + 1);
    const LsVariant& var = kLsVariants[vi];
    if (2 * depth >= var.R * var.waves || 2 * depth >= 64)
        return fail(OFRI_E_ARG, "ls: depth %d too large for tile %dx%d", depth, 64, var.R * var.waves);
    const int outW = 64 - 2 * depth, outH = var.waves * var.R - 2 * depth;
    size_t nparts = 0;
    for (int i = 0; i < n; ++i) {
        const size_t tiles = (size_t)((im[i].W + outW - 1) / outW) * ((im[i].H + outH - 1) / outH);
        nparts += 2 * (size_t)depth * tiles * 2;
    }
    void *fbuf, *pbuf, *mbuf;
    OFRI_TRY(scratch(ctx, "lsb_fields", kLsFieldPlanes * total * sizeof(float), &fbuf));
    OFRI_TRY(scratch(ctx, "lsb_part", nparts * sizeof(double), &pbuf));
    OFRI_TRY(scratch(ctx, "lsb_max", 2 * (size_t)n * sizeof(float), &mbuf));
    std::vector<LsPass> tab(n);
    std::vector<int> blk0(n + 1);
    float* f = (float*)fbuf;
    double* part = (double*)pbuf;
    blk0[0] = 0;
    for (int i = 0; i < n; ++i) {
        const int H = im[i].H, W = im[i].W;
        const size_t N = (size_t)H * W;
        float* F = f;
        float *ub = F + kLsPlanes * N, *vb = F + (kLsPlanes + 1) * N;
        f += kLsFieldPlanes * N;
        float* mx = (float*)mbuf + 2 * i;
        OFRI_TRY(reduce_max2(ctx, im[i].im1, im[i].im2, N, mx));
        dim3 g((W + 255) / 256, H);
        ls_setup_k<<<g, 256, 0, ctx->stream>>>(im[i].im1, im[i].im2, H, W, mx, F, N, d_ctrls + i);
        OFRI_HIP(hipGetLastError());
        LsPass& p = tab[i];
        p = LsPass{};
        p.in[0][0] = im[i].U; p.in[0][1] = im[i].V; p.in[1][0] = ub; p.in[1][1] = vb;
        p.out[0][0] = im[i].U; p.out[0][1] = im[i].V; p.out[1][0] = ub; p.out[1][1] = vb;
        p.F = F; p.N = N; p.H = H; p.W = W; p.K = depth; p.h = h;
        p.tol = tol;
        p.maxnum = maxnum;
        p.ctrl = d_ctrls + i;
        p.tiles_x = (W + outW - 1) / outW;
        p.tiles_y = (H + outH - 1) / outH;
        p.part = part;
        p.part_stride = (size_t)depth * p.tiles_x * p.tiles_y * 2;
        part += 2 * p.part_stride;
        blk0[i + 1] = blk0[i] + p.tiles_x * p.tiles_y;
    }
    LsPass* d_pass = (LsPass*)d_tab;
    int* d_blk0 = (int*)((char*)d_tab + align256((size_t)n * sizeof(LsPass)));
    if (upload) {
        OFRI_HIP(hipMemcpyAsync(d_pass, tab.data(), n * sizeof(LsPass), hipMemcpyHostToDevice, ctx->stream));
        OFRI_HIP(hipMemcpyAsync(d_blk0, blk0.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        OFRI_HIP(hipStreamSynchronize(ctx->stream));
    }
    int done = 0, src = 0, j = 0;
    while (done < maxnum) {
        const int ns = maxnum - done < depth ? maxnum - done : depth;
        var.batched<<<blk0[n], 64 * var.waves, 0, ctx->stream>>>(d_pass, d_blk0, n, ns, src, j & 1, 0);
        OFRI_HIP(hipGetLastError());
        src ^= 1;
        done += ns;
        ++j;
    }
    if (maxnum > 0) {
        var.batched<<<blk0[n], 64 * var.waves, 0, ctx->stream>>>(d_pass, d_blk0, n, 0, 0, j & 1, 1);
        OFRI_HIP(hipGetLastError());
        f = (float*)fbuf;
        for (int i = 0; i < n; ++i) {
            const size_t N = (size_t)im[i].H * im[i].W;
            OFRI_TRY(select_copy2(ctx, d_ctrls + i, im[i].U, f + kLsPlanes * N, im[i].U, im[i].V,
                                  f + (kLsPlanes + 1) * N, im[i].V, N));
            f += kLsFieldPlanes * N;
        }
    }
    return OFRI_OK;
}

size_t ls_part_doubles(ofri_ctx* ctx, int H, int W) {
    int vi, depth;
    ls_default((size_t)H * W, &vi, &depth);
    shape_override(ctx, (size_t)H * W, false, &vi, &depth);
    if (vi >= kNumLsVariants) return 0;
    const LsVariant& var = kLsVariants[vi];
    const int outW = 64 - 2 * depth, outH = var.waves * var.R - 2 * depth;
    if (outW <= 0 || outH <= 0) return 0;
    const size_t nblk = (size_t)((W + outW - 1) / outW) * ((H + outH - 1) / outH);
    return 2 * (size_t)(depth + 1) * nblk * 2;
}

namespace {
std::mutex g_noalloc_mu;
std::map<int, void*> g_noalloc_ptr;   // device -> address of that device's copy of ::scratch
}  // namespace

int noalloc_scratch(int device, void** ptr, size_t* bytes) {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    auto it = g_noalloc_ptr.find(device);
    if (it == g_noalloc_ptr.end()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
        if (e != hipSuccess || !p)
            return fail(OFRI_E_HIP, "no-alloc scratch of device %d: hipGetSymbolAddress failed: %s", device,
                        hipGetErrorString(e));
        it = g_noalloc_ptr.emplace(device, p).first;
    }
    *ptr = it->second;
    *bytes = kNoAllocBytes;
    return OFRI_OK;
}

// Effective Liu-Shen launch shape for an image of npx pixels (1-based variant).
void ls_effective(ofri_ctx* ctx, size_t npx, int* variant1, int* depth) {
    int vi, d;
    ls_default(npx, &vi, &d);
    shape_override(ctx, npx, false, &vi, &d);
    *variant1 = vi + 1;
    *depth = d;
}
}  // namespace ofri
