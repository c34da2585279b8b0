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
// update is float32 without FMA contraction and with IEEE division
// (HS_helper2).  The double sum is grouped symmetrically (vertical pairs,
// corners, edges), which also makes the mirror boundary an exact select.
#include "ofri_internal.h"

#pragma clang fp contract(off)

namespace ofri {
namespace {

// value of the lane to the left (lane-1) / right (lane+1) of this lane
__device__ __forceinline__ float from_left(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(x), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float from_right(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(x), 0x130, 0xF, 0xF, false));
}

// Derivatives: HS(im2=I0, im1=I1) -> computeDerivatives(I1, I0)
// (HornSchunck.py:36,72,83,106-126), 2x2 stencils with mirror padding.
__global__ void hs_derivs_k(const float* __restrict__ I0, const float* __restrict__ I1, int H,
                            int W, float* __restrict__ fx, float* __restrict__ fy,
                            float* __restrict__ ft) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
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

struct HsPass {
    const float* Uin;
    const float* Vin;
    float* Uout;
    float* Vout;
    const float* fx;
    const float* fy;
    const float* ft;
    int H, W;
    int K;        // halo depth of the tiling (>= nsweeps)
    int nsweeps;  // sweeps in this launch
    float a2;     // alpha*alpha in float32
};

struct Trip {
    float l, c, r;  // left / centre / right neighbour of one row, mirrored at the image edge
};

__device__ __forceinline__ Trip trip(float x, bool edge_l, bool edge_r) {
    Trip t;
    float l = from_left(x), r = from_right(x);
    t.c = x;
    t.l = edge_l ? r : l;   // column -1 mirrors to column 1
    t.r = edge_r ? l : r;   // column W mirrors to column W-2
    return t;
}

// Average of the 8 neighbours with weights wc (corners) and we (edges),
// summed in double and rounded to float.
__device__ __forceinline__ float hs_avg(const Trip& up, const Trip& mid, const Trip& dn) {
    const double wc = (double)(float)(1.0 / 12.0);
    const double we = (double)(float)(1.0 / 6.0);
    double vl = (double)up.l + (double)dn.l;
    double vc = (double)up.c + (double)dn.c;
    double vr = (double)up.r + (double)dn.r;
    double corners = vl + vr;
    double edges = vc + ((double)mid.l + (double)mid.r);
    double s = wc * corners + we * edges;
    return (float)s;
}

template <int R, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void hs_pass_k(HsPass a) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int K = a.K, H = a.H, W = a.W;
    const int outW = 64 - 2 * K, outH = WAVES * R - 2 * K;
    const int gx = (int)blockIdx.x * outW - K + lane;
    const int gy0 = (int)blockIdx.y * outH - K + w * R;
    const int cx = gx < 0 ? 0 : (gx >= W ? W - 1 : gx);
    const bool edge_l = gx == 0, edge_r = gx == W - 1;

    float U[R], V[R], FX[R], FY[R], FT[R], DEN[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int gy = gy0 + r;
        int cy = gy < 0 ? 0 : (gy >= H ? H - 1 : gy);
        size_t q = (size_t)cy * W + cx;
        U[r] = a.Uin[q];
        V[r] = a.Vin[q];
        FX[r] = a.fx[q];
        FY[r] = a.fy[q];
        FT[r] = a.ft[q];
        float d = a.a2 + FX[r] * FX[r];
        DEN[r] = d + FY[r] * FY[r];
    }

    __shared__ float edge[2][WAVES][2][2][64];  // [buf][wave][top,bottom][U,V][lane]
    for (int s = 0; s < a.nsweeps; ++s) {
        const int b = s & 1;
        edge[b][w][0][0][lane] = U[0];
        edge[b][w][0][1][lane] = V[0];
        edge[b][w][1][0][lane] = U[R - 1];
        edge[b][w][1][1][lane] = V[R - 1];
        __syncthreads();
        float uup = w > 0 ? edge[b][w - 1][1][0][lane] : U[0];
        float vup = w > 0 ? edge[b][w - 1][1][1][lane] : V[0];
        float udn = w < WAVES - 1 ? edge[b][w + 1][0][0][lane] : U[R - 1];
        float vdn = w < WAVES - 1 ? edge[b][w + 1][0][1][lane] : V[R - 1];

        Trip up_u = trip(uup, edge_l, edge_r), up_v = trip(vup, edge_l, edge_r);
        Trip cu = trip(U[0], edge_l, edge_r), cv = trip(V[0], edge_l, edge_r);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float nxu = r + 1 < R ? U[r + 1] : udn;
            float nxv = r + 1 < R ? V[r + 1] : vdn;
            Trip dn_u = trip(nxu, edge_l, edge_r), dn_v = trip(nxv, edge_l, edge_r);
            const int gy = gy0 + r;
            // mirror at the top / bottom image rows (wave-uniform test)
            const Trip& tu = gy == 0 ? dn_u : up_u;
            const Trip& tv = gy == 0 ? dn_v : up_v;
            const Trip& bu = gy == H - 1 ? up_u : dn_u;
            const Trip& bv = gy == H - 1 ? up_v : dn_v;
            float ub = hs_avg(tu, cu, bu);
            float vb = hs_avg(tv, cv, bv);
            float x = FX[r], y = FY[r];
            float num = x * ub;
            num = num + y * vb;
            num = num + FT[r];
            float der = num / DEN[r];
            U[r] = ub - x * der;
            V[r] = vb - y * der;
            up_u = cu; up_v = cv;
            cu = dn_u; cv = dn_v;
        }
    }

    if (lane >= K && lane < 64 - K && gx < W) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            int ry = w * R + r;
            int gy = gy0 + r;
            if (ry >= K && ry < WAVES * R - K && gy < H) {
                size_t q = (size_t)gy * W + gx;
                a.Uout[q] = U[r];
                a.Vout[q] = V[r];
            }
        }
    }
}

constexpr int kR = 16;
constexpr int kWaves = 4;
constexpr int kDefaultDepth = 6;

}  // namespace

int hs_solve(ofri_ctx* ctx, const float* im1, const float* im2, int H, int W, float alpha,
             int niter, float* U, float* V, float* d_err) {
    if (H < 2 || W < 2) return fail(OFRI_E_ARG, "hs: image %dx%d too small", H, W);
    if (niter < 0) return fail(OFRI_E_ARG, "hs: negative iteration count");
    size_t N = (size_t)H * W;
    void* buf;
    OFRI_TRY(scratch(ctx, "hs_fields", 7 * N * sizeof(float), &buf));
    float* fx = (float*)buf;
    float* fy = fx + N;
    float* ft = fy + N;
    float* Ub = ft + N;
    float* Vb = Ub + N;
    float* U0 = Vb + N;
    float* V0 = U0 + N;
    dim3 g((W + 255) / 256, H);
    // computeDerivatives(I1=im2, I0=im1)
    hs_derivs_k<<<g, 256, 0, ctx->stream>>>(im1, im2, H, W, fx, fy, ft);
    OFRI_HIP(hipGetLastError());
    if (d_err) {
        OFRI_HIP(hipMemcpyAsync(U0, U, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
        OFRI_HIP(hipMemcpyAsync(V0, V, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    }
    int depth = ctx->tuning.hs_depth > 0 ? ctx->tuning.hs_depth : kDefaultDepth;
    if (depth > 24) depth = 24;
    HsPass p{};
    p.fx = fx; p.fy = fy; p.ft = ft;
    p.H = H; p.W = W;
    p.a2 = alpha * alpha;
    p.K = depth;
    const int outW = 64 - 2 * depth, outH = kWaves * kR - 2 * depth;
    dim3 grid((W + outW - 1) / outW, (H + outH - 1) / outH);
    float *cu = U, *cv = V, *nu = Ub, *nv = Vb;
    int done = 0;
    while (done < niter) {
        int ns = niter - done < depth ? niter - done : depth;
        p.Uin = cu; p.Vin = cv; p.Uout = nu; p.Vout = nv;
        p.nsweeps = ns;
        hs_pass_k<kR, kWaves><<<grid, 64 * kWaves, 0, ctx->stream>>>(p);
        OFRI_HIP(hipGetLastError());
        std::swap(cu, nu);
        std::swap(cv, nv);
        done += ns;
    }
    if (cu != U) {
        OFRI_HIP(hipMemcpyAsync(U, cu, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
        OFRI_HIP(hipMemcpyAsync(V, cv, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    }
    if (d_err) OFRI_TRY(diff_norm_err(ctx, U, U0, V, V0, N, d_err));
    return OFRI_OK;
}

}  // namespace ofri
