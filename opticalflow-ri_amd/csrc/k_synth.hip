// Synthetic PIV frame pairs rendered on the device (training data for the
// warm-start network; the particle model of the host generator in
// ofri/synth.py, with counter-based random numbers).
//
// Particles live in 8x8-pixel cells with 16 slots each; a slot holds a
// particle with probability density*64/16, at a uniform position in its
// cell, diameter d ~ N(2.5, 0.5) clipped to [1, 4] and peak I0 ~ U[150, 255].
// Every random number is a hash of (seed, slot, draw), so a pair is a pure
// function of its seed.  Frame 2 moves each particle by the flow at its
// position (one explicit step).  A pixel sums I0 exp(-8 r^2 / d^2) over the
// particles whose 10x10 window (-4..+5 around the particle's integer
// position) covers it, clips to [0, 255] and adds N(0, noise) noise.
//
// Flow families (u along x, v along y, pixels):
//   0 Poiseuille   u = 4 umax y (H - y) / H^2, v = 0            par: umax
//   1 Lamb-Oseen   vortex at (cx, cy), core rc, peak speed vmax   par: cx cy rc vmax
//   2 turbulence   shear (y/H - 0.5) + sum_k a_k cos(k.x + phi_k) (ky, -kx),
//                  scaled so max |d| over a 4-px grid is dmax     par: shear dmax
#include "ofri_internal.h"

namespace ofri {
namespace {

constexpr int kCell = 8, kSlots = 16;

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
// uniform in [0, 1) from (seed, stream, draw)
__device__ __forceinline__ float unif(unsigned long long seed, unsigned long long s, unsigned d) {
    unsigned long long h = mix64(seed ^ mix64(s * 0x100000001b3ull + d));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}
__device__ __forceinline__ float normal(unsigned long long seed, unsigned long long s, unsigned d) {
    float u1 = fmaxf(unif(seed, s, d), 1e-7f), u2 = unif(seed, s, d + 1);
    return sqrtf(-2.0f * logf(u1)) * cosf(6.28318530718f * u2);
}

struct FlowPar {
    int family, H, W, nmodes;
    float p[4];
    const float* modes;   // kx[n], ky[n], phi[n], amp[n]
    const float* scale;   // device scalar (turbulence)
};

__device__ void flow_raw(const FlowPar& f, float x, float y, float& u, float& v) {
    if (f.family == 0) {
        u = 4.0f * f.p[0] * y * ((float)f.H - y) / ((float)f.H * (float)f.H);
        v = 0.0f;
    } else if (f.family == 1) {
        const float dx = x - f.p[0], dy = y - f.p[1], rc = f.p[2];
        const float r2 = fmaxf(dx * dx + dy * dy, 1e-12f), r = sqrtf(r2);
        const float rp = 1.1209f * rc;
        const float peak = (1.0f - expf(-(rp * rp) / (rc * rc))) / rp;
        const float vt = (1.0f - expf(-r2 / (rc * rc))) / r * (f.p[3] / peak);
        u = -vt * dy / r;
        v = vt * dx / r;
    } else {
        u = f.p[0] * (y / (float)f.H - 0.5f);
        v = 0.0f;
        const int n = f.nmodes;
        for (int k = 0; k < n; ++k) {
            const float kx = f.modes[k], ky = f.modes[n + k];
            const float c = cosf(kx * x + ky * y + f.modes[2 * n + k]) * f.modes[3 * n + k];
            u += c * ky;
            v -= c * kx;
        }
    }
}

__device__ __forceinline__ void flow_at(const FlowPar& f, float x, float y, float& u, float& v) {
    flow_raw(f, x, y, u, v);
    if (f.family == 2) {
        const float s = *f.scale;
        u *= s;
        v *= s;
    }
}

// max |d| of the unscaled turbulence field on a 4-pixel grid (float bits
// of non-negative values order like integers)
__global__ void flow_max_k(FlowPar f, unsigned* out) {
    const int gw = (f.W + 3) / 4, gh = (f.H + 3) / 4;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float m = 0.0f;
    if (i < gw * gh) {
        float u, v;
        flow_raw(f, (float)(4 * (i % gw)), (float)(4 * (i / gw)), u, v);
        m = sqrtf(u * u + v * v);
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

__global__ void flow_scale_k(const unsigned* mx, float dmax, float* scale) {
    const float m = __uint_as_float(*mx);
    *scale = m > 0.0f ? dmax / m : 1.0f;
}

struct Particle {
    float x, y, x2, y2, inv, i0;   // inv = 8 / d^2; i0 = 0: empty slot
};

__global__ void particles_k(FlowPar f, unsigned long long seed, float density, int cw, int ch, Particle* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cw * ch * kSlots) return;
    const int cell = i / kSlots;
    const int cx = cell % cw, cy = cell / cw;
    const unsigned long long s = (unsigned long long)i;
    Particle p{0, 0, 0, 0, 1, 0};
    const float prob = density * (float)(kCell * kCell) / (float)kSlots;
    const float x = (float)(cx * kCell) + unif(seed, s, 1) * kCell;
    const float y = (float)(cy * kCell) + unif(seed, s, 2) * kCell;
    if (unif(seed, s, 0) < prob && x < (float)f.W && y < (float)f.H) {
        const float d = fminf(fmaxf(2.5f + 0.5f * normal(seed, s, 3), 1.0f), 4.0f);
        p.x = x;
        p.y = y;
        p.inv = 8.0f / (d * d);
        p.i0 = 150.0f + 105.0f * unif(seed, s, 5);
        float u, v;
        flow_at(f, x, y, u, v);
        p.x2 = x + u;
        p.y2 = y + v;
    }
    out[i] = p;
}

// Frame `which` (0: positions, 1: displaced positions); particles can sit up
// to 2 cells away from the pixel's cell in frame 2 (|d| <= 8 px).
__global__ void render_k(const Particle* P, int cw, int ch, int H, int W, int which, unsigned long long seed,
                         float noise, float* img) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    const int cx = x / kCell, cy = y / kCell;
    float acc = 0.0f;
    for (int j = cy - 2; j <= cy + 2; ++j) {
        if (j < 0 || j >= ch) continue;
        for (int i = cx - 2; i <= cx + 2; ++i) {
            if (i < 0 || i >= cw) continue;
            const Particle* q = P + (size_t)(j * cw + i) * kSlots;
            for (int k = 0; k < kSlots; ++k) {
                const Particle p = q[k];
                if (p.i0 == 0.0f) continue;
                const float px = which ? p.x2 : p.x, py = which ? p.y2 : p.y;
                const float fx = floorf(px), fy = floorf(py);
                const float ox = (float)x - fx, oy = (float)y - fy;
                if (ox < -4.0f || ox > 5.0f || oy < -4.0f || oy > 5.0f) continue;
                const float dx = (float)x - px, dy = (float)y - py;
                acc += p.i0 * expf(-(dx * dx + dy * dy) * p.inv);
            }
        }
    }
    acc = fminf(fmaxf(acc, 0.0f), 255.0f);
    if (noise > 0.0f)
        acc += noise * normal(seed ^ 0x5bd1e995ull, ((unsigned long long)which << 40) + (unsigned long long)y * W + x, 0);
    img[(size_t)y * W + x] = acc;
}

__global__ void flow_field_k(FlowPar f, float* u, float* v) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= f.W) return;
    float a, b;
    flow_at(f, (float)x, (float)y, a, b);
    u[(size_t)y * f.W + x] = a;
    v[(size_t)y * f.W + x] = b;
}

}  // namespace

int render_pair(ofri_ctx* ctx, int H, int W, int family, float density, unsigned long long seed, const float* par,
                int npar, const float* h_modes, int nmodes, float noise, float* im1, float* im2, float* u, float* v) {
    if (H < 8 || W < 8) return fail(OFRI_E_ARG, "render: %dx%d too small", H, W);
    if (family < 0 || family > 2) return fail(OFRI_E_ARG, "render: flow family %d", family);
    if (!(density > 0.0f && density <= 0.25f)) return fail(OFRI_E_ARG, "render: density %g outside (0, 0.25]", density);
    const int need = family == 0 ? 1 : (family == 1 ? 4 : 2);
    if (npar < need || !par) return fail(OFRI_E_ARG, "render: family %d needs %d parameters", family, need);
    if (family == 2 && (nmodes < 1 || !h_modes)) return fail(OFRI_E_ARG, "render: turbulence needs modes");
    // the gather covers particles up to 8 px from their frame-1 position
    const float dmax = family == 0 ? fabsf(par[0]) : (family == 1 ? fabsf(par[3]) : fabsf(par[1]));
    if (!(dmax <= 8.0f)) return fail(OFRI_E_ARG, "render: displacement scale %g above 8 px", dmax);
    if (family == 1 && !(par[2] > 0.0f)) return fail(OFRI_E_ARG, "render: core radius %g", par[2]);
    FlowPar f{};
    f.family = family;
    f.H = H;
    f.W = W;
    for (int i = 0; i < need; ++i) f.p[i] = par[i];
    const int cw = (W + kCell - 1) / kCell, ch = (H + kCell - 1) / kCell;
    const size_t nslots = (size_t)cw * ch * kSlots;
    void* buf;
    OFRI_TRY(scratch(ctx, "synth", nslots * sizeof(Particle) + (size_t)(4 * nmodes + 64) * sizeof(float), &buf));
    Particle* P = (Particle*)buf;
    float* aux = (float*)(P + nslots);   // [0] max bits, [1] scale, [16..] modes
    f.scale = aux + 1;
    if (family == 2) {
        f.nmodes = nmodes;
        f.modes = aux + 16;
        OFRI_HIP(hipMemcpyAsync(aux + 16, h_modes, 4 * (size_t)nmodes * sizeof(float), hipMemcpyHostToDevice,
                                ctx->stream));
        OFRI_HIP(hipMemsetAsync(aux, 0, sizeof(float), ctx->stream));
        const int gpts = ((W + 3) / 4) * ((H + 3) / 4);
        flow_max_k<<<(gpts + 255) / 256, 256, 0, ctx->stream>>>(f, (unsigned*)aux);
        flow_scale_k<<<1, 1, 0, ctx->stream>>>((const unsigned*)aux, par[1], aux + 1);
        OFRI_HIP(hipGetLastError());
    }
    particles_k<<<(unsigned)((nslots + 255) / 256), 256, 0, ctx->stream>>>(f, seed, density, cw, ch, P);
    dim3 g((W + 127) / 128, H);
    render_k<<<g, 128, 0, ctx->stream>>>(P, cw, ch, H, W, 0, seed, noise, im1);
    render_k<<<g, 128, 0, ctx->stream>>>(P, cw, ch, H, W, 1, seed, noise, im2);
    if (u && v) flow_field_k<<<g, 128, 0, ctx->stream>>>(f, u, v);
    OFRI_HIP(hipGetLastError());
    // the host memcpy above reads h_modes asynchronously
    OFRI_HIP(hipStreamSynchronize(ctx->stream));
    return OFRI_OK;
}

}  // namespace ofri
