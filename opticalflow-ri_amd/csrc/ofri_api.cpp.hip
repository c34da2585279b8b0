// C ABI of libofri: context, errors, device memory, the per-operation entry
// points and the whole-pyramid LS(HS) driver.
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>

#include <algorithm>

#include "ofri_internal.h"

namespace ofri {

static thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

void clear_error() { g_err.clear(); }

int scratch(ofri_ctx* ctx, const std::string& name, size_t bytes, void** out) {
    DevBuf& b = ctx->ws[name];
    if (b.bytes < bytes) {
        if (b.p) {
            OFRI_HIP(hipStreamSynchronize(ctx->stream));
            OFRI_HIP(hipFree(b.p));
            b.p = nullptr;
            b.bytes = 0;
        }
        size_t sz = (bytes + 255) & ~(size_t)255;
        if (hipMalloc(&b.p, sz) != hipSuccess) {
            (void)hipGetLastError();
            return fail(OFRI_E_NOMEM, "scratch '%s': cannot allocate %zu bytes", name.c_str(), sz);
        }
        b.bytes = sz;
        ++ctx->ws_gen;
    }
    *out = b.p;
    return OFRI_OK;
}

int upload_table(ofri_ctx* ctx, void* dst, const void* src, size_t bytes, const char* what) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    OFRI_HIP(hipStreamIsCapturing(ctx->stream, &cs));
    if (cs != hipStreamCaptureStatusNone) return fail(OFRI_E_STATE, "%s requested while capturing a graph", what);
    OFRI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    OFRI_HIP(hipStreamSynchronize(ctx->stream));
    return OFRI_OK;
}

// Level sizes of the pyramid (GenericPyramidalOpticalFlow.py:291,335-347,414):
// level l < L is resized to round(dim * 2^-(L-l)) (half-to-even), level L keeps
// the input size.
static void level_sizes(int H, int W, int L, std::vector<int>& hs, std::vector<int>& ws) {
    double scale = 1.0 / std::pow(2.0, L - 1);
    hs.clear();
    ws.clear();
    for (int level = 1; level <= L; ++level) {
        if (scale < 1.0 && level != L) {
            hs.push_back((int)std::nearbyint(H * scale));
            ws.push_back((int)std::nearbyint(W * scale));
        } else {
            hs.push_back(H);
            ws.push_back(W);
        }
        scale *= 2;
    }
}

}  // namespace ofri

using namespace ofri;

#define CHECK_CTX(ctx) \
    if (!(ctx)) return fail(OFRI_E_STATE, "null context")

extern "C" {

const char* ofri_last_error(void) { return g_err.c_str(); }

int ofri_version(void) { return 100; }

}  // extern "C"

namespace {

// A context on `device` (made current): stream, timer and ordering events,
// the allocation-free entry point's scratch address.
int new_ctx(int device, ofri_ctx** out) {
    OFRI_HIP(hipSetDevice(device));
    ofri_ctx* c = new ofri_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&c->order_ev, hipEventDisableTiming) != hipSuccess) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        if (c->ev0) (void)hipEventDestroy(c->ev0);
        if (c->ev1) (void)hipEventDestroy(c->ev1);
        delete c;
        (void)hipGetLastError();
        return fail(OFRI_E_HIP, "ofri_create: stream/event creation failed");
    }
    // the allocation-free entry point's scratch on this device (device
    // current); a failure is kept for that entry point to report
    if (noalloc_scratch(device, &c->noalloc, &c->noalloc_bytes) != OFRI_OK) {
        c->noalloc = nullptr;
        c->noalloc_bytes = 0;
        c->noalloc_err = ofri_last_error();
        clear_error();
    }
    *out = c;
    return OFRI_OK;
}

void free_ctx(ofri_ctx* ctx) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->ws)
        if (kv.second.p) (void)hipFree(kv.second.p);
    for (auto& kv : ctx->resize_tables) {
        (void)hipFree(kv.second.d_k);
        (void)hipFree(kv.second.d_bounds);
    }
    for (auto& kv : ctx->spline_tables) {
        (void)hipFree(kv.second.d_w);
        (void)hipFree(kv.second.d_wT);
        (void)hipFree(kv.second.d_start);
    }
    for (auto& kv : ctx->graphs) {
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        if (kv.second.tab.p) (void)hipFree(kv.second.tab.p);
    }
    for (auto& kv : ctx->omega_tabs)
        if (kv.second.p) (void)hipFree(kv.second.p);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipEventDestroy(ctx->order_ev);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

}  // namespace

extern "C" {

int ofri_create(int device, ofri_ctx** out) {
    if (!out) return fail(OFRI_E_ARG, "ofri_create: null output");
    int n = 0;
    OFRI_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(OFRI_E_ARG, "ofri_create: device %d of %d", device, n);
    return new_ctx(device, out);
}

int ofri_destroy(ofri_ctx* ctx) {
    if (!ctx) return OFRI_OK;
    if (ctx->parent) return fail(OFRI_E_STATE, "ofri_destroy: a lane is destroyed with its parent context");
    for (ofri_ctx* l : ctx->lanes) free_ctx(l);
    ctx->lanes.clear();
    free_ctx(ctx);
    return OFRI_OK;
}

int ofri_lane(ofri_ctx* ctx, int k, ofri_ctx** out) {
    CHECK_CTX(ctx);
    if (!out) return fail(OFRI_E_ARG, "ofri_lane: null output");
    if (ctx->parent) return fail(OFRI_E_STATE, "ofri_lane: a lane has no lanes of its own");
    if (k < 0 || k >= OFRI_MAX_LANES) return fail(OFRI_E_ARG, "ofri_lane: lane %d outside 0..%d", k, OFRI_MAX_LANES - 1);
    while ((int)ctx->lanes.size() <= k) {
        ofri_ctx* l = nullptr;
        OFRI_TRY(new_ctx(ctx->device, &l));
        l->parent = ctx;
        ctx->lanes.push_back(l);
    }
    ofri_ctx* l = ctx->lanes[k];
    l->tuning = ctx->tuning;
    l->graphs_on = ctx->graphs_on;
    l->classes = ctx->classes;
    l->classes_gen = ctx->classes_gen;
    *out = l;
    return OFRI_OK;
}

int ofri_fork(ofri_ctx* ctx, int n) {
    CHECK_CTX(ctx);
    if (ctx->parent) return fail(OFRI_E_STATE, "ofri_fork: called on a lane");
    if (n < 0 || n > (int)ctx->lanes.size())
        return fail(OFRI_E_ARG, "ofri_fork: %d lanes asked, %zu made", n, ctx->lanes.size());
    OFRI_HIP(hipSetDevice(ctx->device));
    OFRI_HIP(hipEventRecord(ctx->order_ev, ctx->stream));
    for (int k = 0; k < n; ++k) {
        ofri_ctx* l = ctx->lanes[k];
        l->tuning = ctx->tuning;
        l->graphs_on = ctx->graphs_on;
        l->classes = ctx->classes;
        l->classes_gen = ctx->classes_gen;
        OFRI_HIP(hipStreamWaitEvent(l->stream, ctx->order_ev, 0));
    }
    return OFRI_OK;
}

int ofri_join(ofri_ctx* ctx, int n) {
    CHECK_CTX(ctx);
    if (ctx->parent) return fail(OFRI_E_STATE, "ofri_join: called on a lane");
    if (n < 0 || n > (int)ctx->lanes.size())
        return fail(OFRI_E_ARG, "ofri_join: %d lanes asked, %zu made", n, ctx->lanes.size());
    OFRI_HIP(hipSetDevice(ctx->device));
    for (int k = 0; k < n; ++k) {
        ofri_ctx* l = ctx->lanes[k];
        OFRI_HIP(hipEventRecord(l->order_ev, l->stream));
        OFRI_HIP(hipStreamWaitEvent(ctx->stream, l->order_ev, 0));
    }
    return OFRI_OK;
}

int ofri_sync(ofri_ctx* ctx) {
    CHECK_CTX(ctx);
    OFRI_HIP(hipStreamSynchronize(ctx->stream));
    return OFRI_OK;
}

int ofri_malloc(ofri_ctx* ctx, size_t bytes, void** d_out) {
    CHECK_CTX(ctx);
    if (!d_out) return fail(OFRI_E_ARG, "ofri_malloc: null output");
    OFRI_HIP(hipSetDevice(ctx->device));
    if (hipMalloc(d_out, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return fail(OFRI_E_NOMEM, "ofri_malloc: %zu bytes", bytes);
    }
    return OFRI_OK;
}

int ofri_free(ofri_ctx* ctx, void* d_ptr) {
    CHECK_CTX(ctx);
    if (d_ptr) {
        OFRI_HIP(hipStreamSynchronize(ctx->stream));
        OFRI_HIP(hipFree(d_ptr));
    }
    return OFRI_OK;
}

int ofri_upload(ofri_ctx* ctx, void* d_dst, const void* h_src, size_t bytes) {
    CHECK_CTX(ctx);
    OFRI_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    OFRI_HIP(hipStreamSynchronize(ctx->stream));
    return OFRI_OK;
}

int ofri_download(ofri_ctx* ctx, void* h_dst, const void* d_src, size_t bytes) {
    CHECK_CTX(ctx);
    OFRI_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    OFRI_HIP(hipStreamSynchronize(ctx->stream));
    return OFRI_OK;
}

int ofri_copy(ofri_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
    CHECK_CTX(ctx);
    OFRI_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return OFRI_OK;
}

int ofri_memset(ofri_ctx* ctx, void* d_dst, int value, size_t bytes) {
    CHECK_CTX(ctx);
    OFRI_HIP(hipMemsetAsync(d_dst, value, bytes, ctx->stream));
    return OFRI_OK;
}

int ofri_timer_start(ofri_ctx* ctx) {
    CHECK_CTX(ctx);
    OFRI_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    return OFRI_OK;
}

int ofri_timer_stop(ofri_ctx* ctx, float* ms_out) {
    CHECK_CTX(ctx);
    OFRI_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    OFRI_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    OFRI_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (ms_out) *ms_out = ms;
    return OFRI_OK;
}

int ofri_set_tuning(ofri_ctx* ctx, const ofri_tuning* t) {
    CHECK_CTX(ctx);
    if (!t) return fail(OFRI_E_ARG, "ofri_set_tuning: null");
    if (t->hs_depth < 0 || t->hs_depth > 24 || t->ls_depth < 0 || t->ls_depth > 24)
        return fail(OFRI_E_ARG, "ofri_set_tuning: depth outside 0..24");
    if (t->hs_variant < 0 || t->hs_variant >= 64 || t->ls_variant < 0 || t->ls_variant >= 64)
        return fail(OFRI_E_ARG, "ofri_set_tuning: variant index outside 0..63");
    ctx->tuning = *t;
    return OFRI_OK;
}

int ofri_get_tuning(ofri_ctx* ctx, ofri_tuning* t) {
    CHECK_CTX(ctx);
    if (!t) return fail(OFRI_E_ARG, "ofri_get_tuning: null");
    *t = ctx->tuning;
    return OFRI_OK;
}

int ofri_set_tuning_classes(ofri_ctx* ctx, const ofri_tuning_class* classes, int n) {
    CHECK_CTX(ctx);
    if (n < 0 || n > 16 || (n > 0 && !classes))
        return fail(OFRI_E_ARG, "ofri_set_tuning_classes: %d classes (at most 16)", n);
    std::vector<ofri_tuning_class> v(classes, classes + n);
    for (const auto& c : v) {
        const ofri_tuning& t = c.t;
        if (c.min_px < 0 || t.hs_depth < 0 || t.hs_depth > 24 || t.ls_depth < 0 || t.ls_depth > 24 ||
            t.hs_variant < 0 || t.hs_variant >= 64 || t.ls_variant < 0 || t.ls_variant >= 64)
            return fail(OFRI_E_ARG, "ofri_set_tuning_classes: class at %lld pixels out of range", c.min_px);
    }
    std::sort(v.begin(), v.end(),
              [](const ofri_tuning_class& a, const ofri_tuning_class& b) { return a.min_px < b.min_px; });
    ctx->classes = v;
    ctx->classes_gen += 1;
    return OFRI_OK;
}

int ofri_arch(ofri_ctx* ctx, char* out, int len) {
    CHECK_CTX(ctx);
    if (!out || len < 1) return fail(OFRI_E_ARG, "ofri_arch: no room");
    hipDeviceProp_t prop;
    OFRI_HIP(hipGetDeviceProperties(&prop, ctx->device));
    int i = 0;
    for (; i < len - 1 && prop.gcnArchName[i] && prop.gcnArchName[i] != ':'; ++i) out[i] = prop.gcnArchName[i];
    out[i] = 0;
    return OFRI_OK;
}

int ofri_query_tuning(ofri_ctx* ctx, int H, int W, ofri_tuning* out) {
    CHECK_CTX(ctx);
    if (!out || H < 1 || W < 1) return fail(OFRI_E_ARG, "ofri_query_tuning: bad argument");
    size_t n = (size_t)H * W;
    hs_effective(ctx, n, &out->hs_variant, &out->hs_depth);
    ls_effective(ctx, n, &out->ls_variant, &out->ls_depth);
    return OFRI_OK;
}

int ofri_gauss_sep(ofri_ctx* ctx, const float* d_src, float* d_dst, int H, int W, const float* h_w,
                   int n) {
    CHECK_CTX(ctx);
    if (!d_src || !d_dst || !h_w) return fail(OFRI_E_ARG, "ofri_gauss_sep: null pointer");
    return gauss_sep(ctx, d_src, d_dst, H, W, h_w, n);
}

int ofri_resize(ofri_ctx* ctx, const float* d_src, int H, int W, float* d_dst, int H2, int W2,
                int filter) {
    CHECK_CTX(ctx);
    if (!d_src || !d_dst) return fail(OFRI_E_ARG, "ofri_resize: null pointer");
    return resize(ctx, d_src, H, W, d_dst, H2, W2, filter);
}

int ofri_spline_upsample(ofri_ctx* ctx, const float* d_src, int Hc, int Wc, float* d_dst, int H,
                         int W, float scale) {
    CHECK_CTX(ctx);
    if (!d_src || !d_dst) return fail(OFRI_E_ARG, "ofri_spline_upsample: null pointer");
    return spline_upsample(ctx, d_src, Hc, Wc, d_dst, H, W, scale);
}

int ofri_warp_symmetric(ofri_ctx* ctx, const float* d_im1, const float* d_im2, const float* d_U,
                        const float* d_V, int H, int W, float* d_out1, float* d_out2) {
    CHECK_CTX(ctx);
    if (!d_im1 || !d_im2 || !d_U || !d_V || !d_out1 || !d_out2)
        return fail(OFRI_E_ARG, "ofri_warp_symmetric: null pointer");
    return warp_symmetric(ctx, d_im1, d_im2, d_U, d_V, H, W, d_out1, d_out2);
}

int ofri_warp_coords(ofri_ctx* ctx, const float* d_img, const float* d_cy, const float* d_cx, int H,
                     int W, float* d_out) {
    CHECK_CTX(ctx);
    if (!d_img || !d_cy || !d_cx || !d_out) return fail(OFRI_E_ARG, "ofri_warp_coords: null pointer");
    return warp_coords(ctx, d_img, d_cy, d_cx, H, W, d_out);
}

int ofri_accumulate(ofri_ctx* ctx, float* d_acc, const float* d_x, size_t n) {
    CHECK_CTX(ctx);
    if (!d_acc || !d_x) return fail(OFRI_E_ARG, "ofri_accumulate: null pointer");
    return accumulate(ctx, d_acc, d_x, n);
}

int ofri_flow_epe(ofri_ctx* ctx, const float* d_U1, const float* d_V1, const float* d_U2, const float* d_V2,
                  size_t n, float* d_out) {
    CHECK_CTX(ctx);
    if (!d_U1 || !d_V1 || !d_U2 || !d_V2 || !d_out) return fail(OFRI_E_ARG, "ofri_flow_epe: null pointer");
    if (n == 0) return fail(OFRI_E_ARG, "ofri_flow_epe: empty flow");
    return flow_epe(ctx, d_U1, d_V1, d_U2, d_V2, n, d_out);
}

int ofri_hs(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W, float alpha,
            int niter, float* d_U, float* d_V, float* h_err) {
    CHECK_CTX(ctx);
    if (!d_im1 || !d_im2 || !d_U || !d_V) return fail(OFRI_E_ARG, "ofri_hs: null pointer");
    void* e = nullptr;
    if (h_err) OFRI_TRY(scratch(ctx, "api_err", sizeof(float), &e));
    OFRI_TRY(hs_solve(ctx, d_im1, d_im2, H, W, alpha, niter, d_U, d_V, (float*)e));
    if (h_err) {
        OFRI_HIP(hipMemcpyAsync(h_err, e, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        OFRI_HIP(hipStreamSynchronize(ctx->stream));
    }
    return OFRI_OK;
}

int ofri_hs_tol(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W, float alpha,
                int maxnum, double tol, float* d_U, float* d_V, float* d_hist, float* h_err,
                int* h_iters) {
    CHECK_CTX(ctx);
    if (!d_im1 || !d_im2 || !d_U || !d_V) return fail(OFRI_E_ARG, "ofri_hs_tol: null pointer");
    void* c;
    OFRI_TRY(scratch(ctx, "api_ctrl", sizeof(SolveCtrl), &c));
    OFRI_TRY(hs_solve(ctx, d_im1, d_im2, H, W, alpha, maxnum, d_U, d_V, nullptr, tol,
                      (SolveCtrl*)c, d_hist));
    if (h_err || h_iters) {
        SolveCtrl hc;
        OFRI_HIP(hipMemcpyAsync(&hc, c, sizeof hc, hipMemcpyDeviceToHost, ctx->stream));
        OFRI_HIP(hipStreamSynchronize(ctx->stream));
        if (h_err) *h_err = hc.last_err;
        if (h_iters) *h_iters = hc.sweeps;
    }
    return OFRI_OK;
}

int ofri_hs_accel(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W, float alpha,
                  int maxnum, double tol, float bound, float finish, float* d_U, float* d_V,
                  float* d_hist, float* h_err, int* h_iters, int launch_test) {
    CHECK_CTX(ctx);
    if (!d_im1 || !d_im2 || !d_U || !d_V) return fail(OFRI_E_ARG, "ofri_hs_accel: null pointer");
    if (!(bound > 0.0f && bound <= 0.999f))
        return fail(OFRI_E_ARG, "ofri_hs_accel: bound %g not in (0, 0.999]", bound);
    if (launch_test & ~3) return fail(OFRI_E_ARG, "ofri_hs_accel: unknown flags 0x%x", launch_test);
    if ((launch_test & 1) && d_hist)
        return fail(OFRI_E_ARG, "ofri_hs_accel: the per-launch stop test records no per-sweep history");
    if ((launch_test & 2) && !(launch_test & 1))
        return fail(OFRI_E_ARG, "ofri_hs_accel: a relative tolerance needs the per-launch stop test");
    void* c;
    OFRI_TRY(scratch(ctx, "api_ctrl", sizeof(SolveCtrl), &c));
    OFRI_TRY(hs_solve(ctx, d_im1, d_im2, H, W, alpha, maxnum, d_U, d_V, nullptr, tol, (SolveCtrl*)c, d_hist,
                      bound, finish, (launch_test & 1) != 0, 0, false, nullptr, 0, (launch_test & 2) != 0));
    if (h_err || h_iters) {
        SolveCtrl hc;
        OFRI_HIP(hipMemcpyAsync(&hc, c, sizeof hc, hipMemcpyDeviceToHost, ctx->stream));
        OFRI_HIP(hipStreamSynchronize(ctx->stream));
        if (h_err) *h_err = hc.last_err;
        if (h_iters) *h_iters = hc.sweeps;
    }
    return OFRI_OK;
}

int ofri_ls(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W, float h,
            int maxnum, double tol, float* d_u, float* d_v, float* h_err, int* h_iters) {
    CHECK_CTX(ctx);
    if (!d_im1 || !d_im2 || !d_u || !d_v) return fail(OFRI_E_ARG, "ofri_ls: null pointer");
    void* c;
    OFRI_TRY(scratch(ctx, "api_ctrl", sizeof(SolveCtrl), &c));
    OFRI_TRY(ls_solve(ctx, d_im1, d_im2, H, W, h, maxnum, tol, d_u, d_v, (SolveCtrl*)c));
    if (h_err || h_iters) {
        SolveCtrl hc;
        OFRI_HIP(hipMemcpyAsync(&hc, c, sizeof hc, hipMemcpyDeviceToHost, ctx->stream));
        OFRI_HIP(hipStreamSynchronize(ctx->stream));
        if (h_err) *h_err = hc.last_err;
        if (h_iters) *h_iters = hc.sweeps;
    }
    return OFRI_OK;
}

int ofri_ls_noalloc(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W, float h,
                    int maxnum, double tol, float* d_u, float* d_v, float* h_err, int* h_iters) {
    CHECK_CTX(ctx);
    if (!d_im1 || !d_im2 || !d_u || !d_v) return fail(OFRI_E_ARG, "ofri_ls_noalloc: null pointer");
    if (H < 2 || W < 2) return fail(OFRI_E_ARG, "ofri_ls_noalloc: image %dx%d too small", H, W);
    if (!ctx->noalloc)
        return fail(OFRI_E_STATE, "ofri_ls_noalloc: no device scratch on device %d (%s)", ctx->device,
                    ctx->noalloc_err.c_str());
    // layout of the device scratch: control word, maxima, partial maxima,
    // stop-test partial sums, then the setup's planes and the ping-pong pair
    // (kLsFieldPlanes planes); everything is written before it is read (the setup
    // kernel clears the control word), so nothing needs clearing
    const size_t N = (size_t)H * W;
    const size_t nparts = ls_part_doubles(ctx, H, W);
    const size_t o_mx = align256(sizeof(SolveCtrl));
    const size_t o_red = o_mx + 256;
    const size_t o_part = o_red + align256(kMaxRedPart * sizeof(float));
    const size_t o_f = o_part + align256(nparts * sizeof(double));
    const size_t need = o_f + kLsFieldPlanes * N * sizeof(float);
    if (nparts == 0 || need > ctx->noalloc_bytes)
        return fail(OFRI_E_ARG,
                    "ofri_ls_noalloc: a %dx%d image needs %zu bytes of device scratch, %zu available "
                    "(about %zu pixels at most)",
                    H, W, need, ctx->noalloc_bytes, (ctx->noalloc_bytes - o_f) / (kLsFieldPlanes * sizeof(float)));
    char* base = (char*)ctx->noalloc;
    SolveCtrl* ctrl = (SolveCtrl*)base;
    LsWorkspace ws{(float*)(base + o_f), (double*)(base + o_part), nparts, (float*)(base + o_mx),
                   (float*)(base + o_red)};
    OFRI_TRY(ls_solve(ctx, d_im1, d_im2, H, W, h, maxnum, tol, d_u, d_v, ctrl, false, 0.0f, 0.0f, nullptr, nullptr,
                      false, false, false, &ws));
    if (h_err || h_iters) {
        SolveCtrl hc;
        OFRI_HIP(hipMemcpyAsync(&hc, ctrl, sizeof hc, hipMemcpyDeviceToHost, ctx->stream));
        OFRI_HIP(hipStreamSynchronize(ctx->stream));
        if (h_err) *h_err = hc.last_err;
        if (h_iters) *h_iters = hc.sweeps;
    }
    return OFRI_OK;
}

int ofri_warp_liushen(ofri_ctx* ctx, const float* d_im1, int H, int W, const float* d_u, const float* d_v,
                      const double* h_w, int radius, float* d_out, int* h_status) {
    CHECK_CTX(ctx);
    if (!d_im1 || !d_u || !d_v || !h_w || !d_out) return fail(OFRI_E_ARG, "ofri_warp_liushen: null pointer");
    void* e;
    OFRI_TRY(scratch(ctx, "lsw_err", sizeof(int), &e));
    OFRI_TRY(fill_zero(ctx, (float*)e, 1));
    OFRI_TRY(warp_liushen(ctx, d_im1, H, W, d_u, d_v, h_w, radius, d_out, (int*)e));
    int st = 0;
    OFRI_HIP(hipMemcpyAsync(&st, e, sizeof st, hipMemcpyDeviceToHost, ctx->stream));
    OFRI_HIP(hipStreamSynchronize(ctx->stream));
    if (h_status) *h_status = st ? 1 : 0;
    return OFRI_OK;
}

int ofri_ls_learned(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W, float h,
                    int maxnum, double tol, float* d_u, float* d_v, float* h_err, int* h_iters) {
    CHECK_CTX(ctx);
    if (!d_im1 || !d_im2 || !d_u || !d_v) return fail(OFRI_E_ARG, "ofri_ls_learned: null pointer");
    void* c;
    OFRI_TRY(scratch(ctx, "api_ctrl", sizeof(SolveCtrl), &c));
    OFRI_TRY(ls_solve(ctx, d_im1, d_im2, H, W, h, maxnum, tol, d_u, d_v, (SolveCtrl*)c, true, 0.0f, 0.0f, nullptr,
                      nullptr, false, true));
    if (h_err || h_iters) {
        SolveCtrl hc;
        OFRI_HIP(hipMemcpyAsync(&hc, c, sizeof hc, hipMemcpyDeviceToHost, ctx->stream));
        OFRI_HIP(hipStreamSynchronize(ctx->stream));
        if (h_err) *h_err = hc.last_err;
        if (h_iters) *h_iters = hc.sweeps;
    }
    return OFRI_OK;
}

int ofri_lk_dense(ofri_ctx* ctx, const float* d_I, const float* d_J, int H, int W, float* d_u,
                  float* d_v, int half_win, int niter, unsigned char* d_status, const int* h_asym) {
    CHECK_CTX(ctx);
    if (!d_I || !d_J || !d_u || !d_v) return fail(OFRI_E_ARG, "ofri_lk_dense: null pointer");
    return lk_dense(ctx, d_I, d_J, H, W, half_win, niter, d_u, d_v, d_status, h_asym);
}

int ofri_farneback(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W,
                   float* d_flowx, float* d_flowy, const ofri_fb_params* p) {
    CHECK_CTX(ctx);
    if (!d_im1 || !d_im2 || !d_flowx || !d_flowy || !p || !p->g || !p->xg || !p->xxg || !p->ig ||
        !p->pre_kernel || (!p->win_kernel && !p->box_filter))
        return fail(OFRI_E_ARG, "ofri_farneback: null pointer");
    if (p->iters < 0) return fail(OFRI_E_ARG, "ofri_farneback: negative iteration count");
    return farneback(ctx, d_im1, d_im2, H, W, d_flowx, d_flowy, p);
}

int ofri_fb_blur(ofri_ctx* ctx, const float* d_src, int H, int W, float* d_dst, const float* half_kernel,
                 int kh) {
    CHECK_CTX(ctx);
    if (!d_src || !d_dst || !half_kernel) return fail(OFRI_E_ARG, "ofri_fb_blur: null pointer");
    return fb_blur(ctx, d_src, d_dst, H, W, half_kernel, kh);
}

}  // extern "C"

namespace ofri {

// Enqueues the whole LS(HS) pyramid on the context stream.  Per-level
// statistics land in the "pyr_stats" scratch (L Liu-Shen control words, L
// HS control words, L HS errors).
// d_taps (may be null): per level, planar [HS u, HS v, LS u - HS u, LS v -
// HS v] at the level's size, levels one after another.  record_res: the HS
// sweeps also reduce their update norms (the last one lands in the stats).
// lrn (may be null): the learned form.  Every level's HS runs accelerated
// to the level's residual target; then the correction network maps the
// level's Liu-Shen input pair and HS's flow (4 planes) to the correction
// Liu-Shen would add (fine head fields 0, 1), which is added to HS's flow to
// seed Liu-Shen; Liu-Shen runs to its own target.
struct LearnedRun {
    const ofri_learned_params* lp;
    Net* net;               // null: Liu-Shen seeded with HS's flow alone
};

static int pyr_lshs_enqueue(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W,
                            const ofri_lshs_params* p, float* d_Uacc, float* d_Vacc, float* d_taps = nullptr,
                            bool record_res = false, const LearnedRun* lrn = nullptr,
                            const int* ls_tap_back = nullptr) {
    const int L = p->levels;
    std::vector<int> lh, lw;
    level_sizes(H, W, L, lh, lw);
    for (int l = 0; l < L; ++l)
        if (lh[l] < 4 || lw[l] < 4)
            return fail(OFRI_E_ARG, "ofri_pyr_lshs: level %d size %dx%d below 4x4", l + 1, lh[l], lw[l]);
    const size_t NF = (size_t)H * W;
    // workspace: n1 n2 (resized) w1 w2 (warped) (2 planes unused: the HS
    // pre-filter is fused into its derivative pass) o1 o2 (opt-filtered)
    //            U V (solver) A B (accumulated, ping-pong with upsampled)
    void* ws;
    OFRI_TRY(scratch(ctx, "pyr_ws", 14 * NF * sizeof(float), &ws));
    float* base = (float*)ws;
    float *n1 = base, *n2 = base + NF, *w1 = base + 2 * NF, *w2 = base + 3 * NF;
    float *o1 = base + 6 * NF, *o2 = base + 7 * NF;
    float *U = base + 8 * NF, *V = base + 9 * NF;
    float *Ua = base + 10 * NF, *Va = base + 11 * NF, *Ub = base + 12 * NF, *Vb = base + 13 * NF;
    void* st;
    OFRI_TRY(scratch(ctx, "pyr_stats", (size_t)L * (2 * sizeof(SolveCtrl) + 16), &st));
    SolveCtrl* ctrls = (SolveCtrl*)st;
    SolveCtrl* hs_ctrls = ctrls + L;
    float* hs_errs = (float*)(hs_ctrls + L);

    // Liu-Shen writes (adds) its result straight into the accumulators
    // unless the taps need it in U, V
    const bool ls_sink = p->ls_alpha > 0.0f && !d_taps;
    float *acc_u = Ua, *acc_v = Va;
    float* tap = d_taps;
    int ph = 0, pw = 0;
    for (int l = 0; l < L; ++l) {
        const int h = lh[l], w = lw[l];
        const size_t N = (size_t)h * w;
        const bool last = l == L - 1;
        const float *c1 = d_im1, *c2 = d_im2;
        if (h != H || w != W) {
            OFRI_TRY(resize2(ctx, d_im1, n1, d_im2, n2, H, W, h, w, 3));
            c1 = n1;
            c2 = n2;
        }
        const float *s1 = c1, *s2 = c2;  // images the main solver sees
        float* out_u = last ? d_Uacc : (acc_u == Ua ? Ub : Ua);
        float* out_v = last ? d_Vacc : (acc_v == Va ? Vb : Va);
        if (l > 0) {
            // updateNextPyramidalLevel with warping, biLinear and scaling
            if (h != ph || w != pw) {
                float sx = (float)w / (float)pw, sy = (float)h / (float)ph;
                OFRI_TRY(spline_upsample2(ctx, acc_u, acc_v, ph, pw, out_u, out_v, h, w, sx, sy));
            } else {
                OFRI_HIP(hipMemcpyAsync(out_u, acc_u, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
                OFRI_HIP(hipMemcpyAsync(out_v, acc_v, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
            }
            OFRI_TRY(warp_symmetric(ctx, c1, c2, out_u, out_v, h, w, w1, w2));
            s1 = w1;
            s2 = w2;
        } else if (!ls_sink) {
            OFRI_TRY(fill_zero2(ctx, out_u, out_v, N));
        }
        // (the pre-filter runs inside HS's derivative pass; the learned
        // form's HS starts from zero without reading U, V)
        const float* hpw = p->filt_n > 0 ? p->filt_w : nullptr;
        const int hpn = hpw ? p->filt_n : 0;
        if (!lrn) OFRI_TRY(fill_zero2(ctx, U, V, N));
        if (lrn) {
            const ofri_learned_params* lp = lrn->lp;
            // (HS's own error value, the change from the initial flow, is not formed here)
            // (the first finishing launch's length applies where Chebyshev
            // sweeps hand over to a finishing phase)
            const bool phases = lp->cheb_bound[l] > 0.0f && lp->cheb_finish[l] > 1.0f;
            OFRI_TRY(hs_solve(ctx, s1, s2, h, w, p->alphas[l], lp->hs_budget[l], U, V, nullptr, lp->hs_tol[l],
                              hs_ctrls + l, nullptr, lp->cheb_bound[l], lp->cheb_finish[l], true,
                              lp->hs_finish_budget[l], true, hpw, hpn, lp->relative != 0,
                              phases ? lp->hs_finish_first[l] : 0));
        } else if (record_res) {
            OFRI_TRY(hs_solve(ctx, s1, s2, h, w, p->alphas[l], p->niter, U, V, hs_errs + l, -1.0, hs_ctrls + l,
                              nullptr, 0.0f, 0.0f, false, 0, false, hpw, hpn));
        } else {
            OFRI_TRY(hs_solve(ctx, s1, s2, h, w, p->alphas[l], p->niter, U, V, hs_errs + l, -1.0, nullptr, nullptr,
                              0.0f, 0.0f, false, 0, false, hpw, hpn));
        }
        if (tap) {
            OFRI_HIP(hipMemcpyAsync(tap, U, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
            OFRI_HIP(hipMemcpyAsync(tap + N, V, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
        }
        if (p->ls_alpha > 0.0f) {
            const float *q1 = c1, *q2 = c2;
            if (p->opt_w && p->opt_n > 0) {
                OFRI_TRY(gauss_sep2(ctx, c1, o1, c2, o2, h, w, p->opt_w, p->opt_n));
                q1 = o1;
                q2 = o2;
            }
            if (lrn && lrn->net && h % 16 == 0 && w % 16 == 0) {
                // the correction network: (Liu-Shen pair, HS flow) -> Liu-Shen's correction
                const float* planes[4] = {q1, q2, U, V};
                OFRI_TRY(net_correct(ctx, lrn->net, planes, h, w, U, V));
            }
            // adapter swap (PhysicsBasedOpticalFlowLiuShen.py:38-39): solver u = V, v = U.
            // Without taps the result lands in the accumulators directly
            // (copied at the first level, whose accumulators start empty).
            float* su = ls_sink ? out_v : nullptr;
            float* sv = ls_sink ? out_u : nullptr;
            const int back = ls_tap_back && tap ? ls_tap_back[l] : 0;
            if (lrn) {
                // fixed sweep count: the test never stops it (tolerance 0),
                // the check after the solve accepts or rejects the level
                const ofri_learned_params* lp = lrn->lp;
                const bool fixed = lp->ls_sweeps[l] > 0;
                OFRI_TRY(ls_solve(ctx, q1, q2, h, w, p->ls_alpha, fixed ? lp->ls_sweeps[l] : lp->ls_budget,
                                  fixed ? 0.0 : (double)lp->ls_tol[l], V, U, ctrls + l, true, lp->ls_cheb_lo,
                                  lp->ls_cheb_hi, su, sv, l > 0, lp->ls_cheb_hi <= 0.0f, lp->relative != 0));
            } else if (back > 0 && back < p->ls_maxnum) {
                // the tap k sweeps before the stop: ls_maxnum - k sweeps, the
                // Liu-Shen planes exported, then the remaining k sweeps from there
                OFRI_TRY(ls_solve(ctx, q1, q2, h, w, p->ls_alpha, p->ls_maxnum - back, p->ls_tol, V, U, ctrls + l,
                                  false, 0.0f, 0.0f, nullptr, nullptr, false));
                OFRI_TRY(subtract(ctx, tap + 2 * N, U, tap, N));
                OFRI_TRY(subtract(ctx, tap + 3 * N, V, tap + N, N));
                OFRI_TRY(ls_solve(ctx, q1, q2, h, w, p->ls_alpha, back, p->ls_tol, V, U, ctrls + l, false, 0.0f,
                                  0.0f, nullptr, nullptr, false));
            } else {
                OFRI_TRY(ls_solve(ctx, q1, q2, h, w, p->ls_alpha, p->ls_maxnum, p->ls_tol, V, U, ctrls + l, false,
                                  0.0f, 0.0f, su, sv, l > 0));
            }
        }
        if (tap) {
            const bool tapped = p->ls_alpha > 0.0f && ls_tap_back && ls_tap_back[l] > 0 &&
                                ls_tap_back[l] < p->ls_maxnum;
            if (!tapped) {
                OFRI_TRY(subtract(ctx, tap + 2 * N, U, tap, N));
                OFRI_TRY(subtract(ctx, tap + 3 * N, V, tap + N, N));
            }
            tap += 4 * N;
        }
        if (!ls_sink) OFRI_TRY(accumulate2(ctx, out_u, U, out_v, V, N));
        acc_u = out_u;
        acc_v = out_v;
        ph = h;
        pw = w;
    }
    return OFRI_OK;
}

// Key of a pyramid call: everything that is baked into its launch sequence.
static std::string lshs_key(ofri_ctx* ctx, const float* a, const float* b, int H, int W,
                            const ofri_lshs_params* p, const float* u, const float* v) {
    std::string k;
    auto put = [&](const void* x, size_t n) { k.append((const char*)x, n); };
    const void* ptrs[4] = {a, b, u, v};
    put(ptrs, sizeof ptrs);
    put(&H, sizeof H);
    put(&W, sizeof W);
    put(&p->levels, sizeof p->levels);
    put(p->alphas, sizeof(float) * p->levels);
    put(&p->niter, sizeof p->niter);
    put(&p->filt_n, sizeof p->filt_n);
    if (p->filt_w && p->filt_n > 0) put(p->filt_w, sizeof(float) * p->filt_n);
    put(&p->opt_n, sizeof p->opt_n);
    if (p->opt_w && p->opt_n > 0) put(p->opt_w, sizeof(float) * p->opt_n);
    put(&p->ls_alpha, sizeof p->ls_alpha);
    put(&p->ls_maxnum, sizeof p->ls_maxnum);
    put(&p->ls_tol, sizeof p->ls_tol);
    put(&ctx->tuning, sizeof ctx->tuning);
    put(&ctx->classes_gen, sizeof ctx->classes_gen);
    return k;
}

// Enqueues a batch of LS(HS) pyramids.  Every pair is taken through the same
// steps as pyr_lshs_enqueue; the light per-pixel steps run per pair, the
// Horn-Schunck and Liu-Shen sweeps of a level in launches shared by all pairs.
// Statistics land in "batch_stats": [L][n] SolveCtrl, then [L][n] HS errors.
static int batch_lshs_enqueue(ofri_ctx* ctx, const ofri_pair* pr, int n, const ofri_lshs_params* p,
                              void* d_tab, bool upload) {
    const int L = p->levels;
    std::vector<std::vector<int>> lh(n), lw(n);
    size_t totalF = 0;
    for (int i = 0; i < n; ++i) {
        level_sizes(pr[i].H, pr[i].W, L, lh[i], lw[i]);
        totalF += (size_t)pr[i].H * pr[i].W;
    }
    void* ws;
    OFRI_TRY(scratch(ctx, "batch_ws", 14 * totalF * sizeof(float), &ws));
    void* st;
    OFRI_TRY(scratch(ctx, "batch_stats", (size_t)L * n * (sizeof(SolveCtrl) + sizeof(float)), &st));
    SolveCtrl* ctrls = (SolveCtrl*)st;
    float* hs_errs = (float*)(ctrls + (size_t)L * n);
    const size_t tab_level = align256(hs_batch_table_bytes(n)) + align256(ls_batch_table_bytes(n));

    struct PairWs {
        float *n1, *n2, *w1, *w2, *f1, *f2, *o1, *o2, *U, *V, *Ua, *Va, *Ub, *Vb;
        float *acc_u, *acc_v;
        const float *c1, *c2;
        int ph, pw;
    };
    std::vector<PairWs> P(n);
    float* base = (float*)ws;
    for (int i = 0; i < n; ++i) {
        const size_t NF = (size_t)pr[i].H * pr[i].W;
        float** planes[14] = {&P[i].n1, &P[i].n2, &P[i].w1, &P[i].w2, &P[i].f1, &P[i].f2, &P[i].o1,
                              &P[i].o2, &P[i].U,  &P[i].V,  &P[i].Ua, &P[i].Va, &P[i].Ub, &P[i].Vb};
        for (int k = 0; k < 14; ++k) *planes[k] = base + k * NF;
        base += 14 * NF;
        P[i].acc_u = P[i].Ua;
        P[i].acc_v = P[i].Va;
        P[i].ph = P[i].pw = 0;
    }
    std::vector<BatchImage> hsim(n), lsim(n);
    for (int l = 0; l < L; ++l) {
        const bool last = l == L - 1;
        for (int i = 0; i < n; ++i) {
            PairWs& q = P[i];
            const int H = pr[i].H, W = pr[i].W, h = lh[i][l], w = lw[i][l];
            const size_t N = (size_t)h * w;
            q.c1 = pr[i].im1;
            q.c2 = pr[i].im2;
            if (h != H || w != W) {
                OFRI_TRY(resize2(ctx, pr[i].im1, q.n1, pr[i].im2, q.n2, H, W, h, w, 3));
                q.c1 = q.n1;
                q.c2 = q.n2;
            }
            const float *s1 = q.c1, *s2 = q.c2;
            float* out_u = last ? pr[i].U : (q.acc_u == q.Ua ? q.Ub : q.Ua);
            float* out_v = last ? pr[i].V : (q.acc_v == q.Va ? q.Vb : q.Va);
            if (l > 0) {
                if (h != q.ph || w != q.pw) {
                    float sx = (float)w / (float)q.pw, sy = (float)h / (float)q.ph;
                    OFRI_TRY(spline_upsample2(ctx, q.acc_u, q.acc_v, q.ph, q.pw, out_u, out_v, h, w, sx, sy));
                } else {
                    OFRI_HIP(hipMemcpyAsync(out_u, q.acc_u, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
                    OFRI_HIP(hipMemcpyAsync(out_v, q.acc_v, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
                }
                OFRI_TRY(warp_symmetric(ctx, q.c1, q.c2, out_u, out_v, h, w, q.w1, q.w2));
                s1 = q.w1;
                s2 = q.w2;
            } else {
                OFRI_TRY(fill_zero2(ctx, out_u, out_v, N));
            }
            if (p->filt_w && p->filt_n > 0) {
                OFRI_TRY(gauss_sep2(ctx, s1, q.f1, s2, q.f2, h, w, p->filt_w, p->filt_n));
                s1 = q.f1;
                s2 = q.f2;
            }
            OFRI_TRY(fill_zero2(ctx, q.U, q.V, N));
            hsim[i] = BatchImage{s1, s2, h, w, q.U, q.V, hs_errs + (size_t)l * n + i};
            q.acc_u = out_u;  // the accumulation target of this level
            q.acc_v = out_v;
        }
        char* tab = (char*)d_tab + l * tab_level;
        OFRI_TRY(hs_batch(ctx, hsim.data(), n, p->alphas[l], p->niter, tab, upload));
        if (p->ls_alpha > 0.0f) {
            for (int i = 0; i < n; ++i) {
                PairWs& q = P[i];
                const int h = lh[i][l], w = lw[i][l];
                const float *q1 = q.c1, *q2 = q.c2;
                if (p->opt_w && p->opt_n > 0) {
                    OFRI_TRY(gauss_sep2(ctx, q.c1, q.o1, q.c2, q.o2, h, w, p->opt_w, p->opt_n));
                    q1 = q.o1;
                    q2 = q.o2;
                }
                // adapter swap (PhysicsBasedOpticalFlowLiuShen.py:38-39): solver u = V, v = U
                lsim[i] = BatchImage{q1, q2, h, w, q.V, q.U, nullptr};
            }
            OFRI_TRY(ls_batch(ctx, lsim.data(), n, p->ls_alpha, p->ls_maxnum, p->ls_tol, ctrls + (size_t)l * n,
                              tab + align256(hs_batch_table_bytes(n)), upload));
        }
        for (int i = 0; i < n; ++i) {
            PairWs& q = P[i];
            const size_t N = (size_t)lh[i][l] * lw[i][l];
            OFRI_TRY(accumulate2(ctx, q.acc_u, q.U, q.acc_v, q.V, N));
            q.ph = lh[i][l];
            q.pw = lw[i][l];
        }
    }
    return OFRI_OK;
}

// The learned form's convergence flag on the device: every level's HS and
// Liu-Shen residuals (the published control words of the "pyr_stats"
// scratch: 2 Liu-Shen, then 2 HS) against their targets, as the host test.
// Relative targets scale with the norm the test published (the same double
// comparison as the stop test).
__device__ __host__ inline bool meets(float err, float tol, float nrm, int rel) {
    const double t = rel ? (double)tol * (double)nrm : (double)tol;
    return !((double)err > t);
}

__global__ void learned_check_k(const SolveCtrl* c, float hs0, float hs1, float ls0, float ls1, int with_ls,
                                int rel, int* out) {
    int ok = 1;
    if (!meets(c[2].last_err, hs0, c[2].last_nrm, rel) || !meets(c[3].last_err, hs1, c[3].last_nrm, rel)) ok = 0;
    if (with_ls && (!meets(c[0].last_err, ls0, c[0].last_nrm, rel) || !meets(c[1].last_err, ls1, c[1].last_nrm, rel)))
        ok = 0;
    *out = ok;
}

// Per-level statistics of the last pyramid run (the "pyr_stats" scratch).
static int pyr_stats_download(ofri_ctx* ctx, int L, const std::vector<int>& lh, const std::vector<int>& lw,
                              const ofri_lshs_params* p, bool with_res, ofri_level_stats* out,
                              bool with_hs_err = true) {
    void* st;
    OFRI_TRY(scratch(ctx, "pyr_stats", (size_t)L * (2 * sizeof(SolveCtrl) + 16), &st));
    SolveCtrl* ctrls = (SolveCtrl*)st;
    std::vector<SolveCtrl> hc(2 * (size_t)L);
    std::vector<float> he(L);
    OFRI_HIP(hipMemcpyAsync(hc.data(), ctrls, 2 * L * sizeof(SolveCtrl), hipMemcpyDeviceToHost, ctx->stream));
    OFRI_HIP(hipMemcpyAsync(he.data(), ctrls + 2 * L, L * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    OFRI_HIP(hipStreamSynchronize(ctx->stream));
    for (int l = 0; l < L; ++l) {
        out[l].H = lh[l];
        out[l].W = lw[l];
        out[l].hs_err = with_hs_err ? he[l] : 0.0f;
        out[l].ls_err = p->ls_alpha > 0.0f ? hc[l].last_err : 0.0f;
        out[l].ls_iters = p->ls_alpha > 0.0f ? hc[l].sweeps : 0;
        out[l].hs_res = with_res ? hc[L + l].last_err : 0.0f;
        out[l].hs_iters = with_res ? hc[L + l].sweeps : p->niter;
        out[l].hs_nrm = with_res ? hc[L + l].last_nrm : 0.0f;
        out[l].ls_nrm = p->ls_alpha > 0.0f ? hc[l].last_nrm : 0.0f;
    }
    return OFRI_OK;
}

// Runs a launch sequence through the context's graph cache: the first call
// with an entry (and any call after a scratch buffer moved) runs eagerly,
// which also sizes every scratch buffer; the next captures the sequence into
// an executable graph; later calls replay it.  enqueue(capturing) enqueues
// the sequence.  stable_only: an eager run that itself moved a scratch
// buffer does not count towards capture (work it wrote before the move,
// such as job tables, is stale).  An executable graph is replaced only
// after the stream has drained: its last launch may still be in flight, and
// destroying it would free the argument and node storage that launch (and
// any tracing tool following it) still uses.
template <class F>
static int run_cached(ofri_ctx* ctx, GraphEntry& g, F&& enqueue, bool stable_only = false) {
    if (g.exec && g.ws_gen == ctx->ws_gen) {
        OFRI_HIP(hipGraphLaunch(g.exec, ctx->stream));
        return OFRI_OK;
    }
    if (g.exec) {
        OFRI_HIP(hipStreamSynchronize(ctx->stream));
        (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
        g.eager_runs = 0;
    }
    if (g.eager_runs == 0 || g.ws_gen != ctx->ws_gen) {
        const uint64_t gen0 = ctx->ws_gen;
        OFRI_TRY(enqueue(false));
        g.eager_runs = stable_only && ctx->ws_gen != gen0 ? 0 : g.eager_runs + 1;
        g.ws_gen = ctx->ws_gen;
        return OFRI_OK;
    }
    hipGraph_t graph = nullptr;
    OFRI_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int st = enqueue(true);
    const hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
    if (st != OFRI_OK || ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (st != OFRI_OK) return st;
        return fail(OFRI_E_HIP, "graph capture failed: %s", hipGetErrorString(ce));
    }
    const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        g.exec = nullptr;
        return fail(OFRI_E_HIP, "graph instantiate failed: %s", hipGetErrorString(ie));
    }
    g.ws_gen = ctx->ws_gen;
    OFRI_HIP(hipGraphLaunch(g.exec, ctx->stream));
    return OFRI_OK;
}

// The generic level loop (GenericPyramidalOpticalFlow.py:328-414) with a
// Farneback or dense-LK main solver and an optional Liu-Shen stage, all on
// device buffers.  Liu-Shen's control words land in "gpyr_stats" (L words).
static int pyr_generic_enqueue(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W,
                               const ofri_pyr_params* p, const ofri_pyr_main* m, float* d_Uacc, float* d_Vacc) {
    const int L = p->levels;
    std::vector<int> lh, lw;
    level_sizes(H, W, L, lh, lw);
    const size_t NF = (size_t)H * W;
    // workspace: n1 n2 (resized) w1 w2 (warped) f1 f2 (FILTER) o1 o2 (FILTER_OPT)
    //            U V (solver) A B / C D (accumulators, ping-pong with the upsampled flow)
    void* ws;
    OFRI_TRY(scratch(ctx, "gpyr_ws", 14 * NF * sizeof(float), &ws));
    float* base = (float*)ws;
    float *n1 = base, *n2 = base + NF, *w1 = base + 2 * NF, *w2 = base + 3 * NF;
    float *f1 = base + 4 * NF, *f2 = base + 5 * NF, *o1 = base + 6 * NF, *o2 = base + 7 * NF;
    float *U = base + 8 * NF, *V = base + 9 * NF;
    float *Ua = base + 10 * NF, *Va = base + 11 * NF, *Ub = base + 12 * NF, *Vb = base + 13 * NF;
    void* st;
    OFRI_TRY(scratch(ctx, "gpyr_stats", (size_t)L * sizeof(SolveCtrl), &st));
    SolveCtrl* ctrls = (SolveCtrl*)st;
    float *acc_u = Ua, *acc_v = Va;
    int ph = 0, pw = 0;
    for (int l = 0; l < L; ++l) {
        const int h = lh[l], w = lw[l];
        const size_t N = (size_t)h * w;
        const bool last = l == L - 1;
        const bool scale_here = last ? p->final_scaling != 0 : p->inter_scaling != 0;
        const float *c1 = d_im1, *c2 = d_im2;
        if (h != H || w != W) {
            OFRI_TRY(resize2(ctx, d_im1, n1, d_im2, n2, H, W, h, w, 3));
            c1 = n1;
            c2 = n2;
        }
        float* out_u = last ? d_Uacc : (acc_u == Ua ? Ub : Ua);
        float* out_v = last ? d_Vacc : (acc_v == Va ? Vb : Va);
        const float *s1 = c1, *s2 = c2;
        if (l > 0) {
            // updateNextPyramidalLevel (:117-234): the accumulated flow upsampled
            // (and scaled) into U, V (no warping: the initial flow) or into the
            // level's accumulator (warping: the pair is warped by it and the
            // solver starts from zero)
            float* up_u = p->warping ? out_u : U;
            float* up_v = p->warping ? out_v : V;
            if (h != ph || w != pw) {
                const float sx = scale_here ? (float)w / (float)pw : 1.0f;
                const float sy = scale_here ? (float)h / (float)ph : 1.0f;
                OFRI_TRY(spline_upsample2(ctx, acc_u, acc_v, ph, pw, up_u, up_v, h, w, sx, sy));
            } else {
                // same size: no resampling, and the size ratio is exactly 1
                OFRI_HIP(hipMemcpyAsync(up_u, acc_u, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
                OFRI_HIP(hipMemcpyAsync(up_v, acc_v, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
            }
            if (p->warping) {
                OFRI_TRY(warp_symmetric(ctx, c1, c2, out_u, out_v, h, w, w1, w2));
                s1 = w1;
                s2 = w2;
                OFRI_TRY(fill_zero2(ctx, U, V, N));
            } else {
                OFRI_TRY(fill_zero2(ctx, out_u, out_v, N));
            }
        } else {
            OFRI_TRY(fill_zero2(ctx, U, V, N));
            OFRI_TRY(fill_zero2(ctx, out_u, out_v, N));
        }
        if (p->filt_w && p->filt_n > 0) {
            OFRI_TRY(gauss_sep2(ctx, s1, f1, s2, f2, h, w, p->filt_w, p->filt_n));
            s1 = f1;
            s2 = f2;
        }
        if (m->kind == 0) {
            OFRI_TRY(farneback(ctx, s1, s2, h, w, U, V, m->fb));
        } else {
            OFRI_TRY(lk_dense(ctx, s1, s2, h, w, m->lk_half_win, m->lk_niter, U, V, last ? m->d_lk_status : nullptr,
                              nullptr));
        }
        if (p->ls_alpha > 0.0f) {
            const float *q1 = c1, *q2 = c2;
            if (p->opt_w && p->opt_n > 0) {
                OFRI_TRY(gauss_sep2(ctx, c1, o1, c2, o2, h, w, p->opt_w, p->opt_n));
                q1 = o1;
                q2 = o2;
            }
            // adapter swap (PhysicsBasedOpticalFlowLiuShen.py:38-39): solver u = V, v = U
            OFRI_TRY(ls_solve(ctx, q1, q2, h, w, p->ls_alpha, p->ls_maxnum, p->ls_tol, V, U, ctrls + l));
        }
        OFRI_TRY(accumulate2(ctx, out_u, U, out_v, V, N));
        acc_u = out_u;
        acc_v = out_v;
        ph = h;
        pw = w;
    }
    return OFRI_OK;
}

static std::string generic_key(ofri_ctx* ctx, const float* a, const float* b, int H, int W, const ofri_pyr_params* p,
                               const ofri_pyr_main* m, const float* u, const float* v) {
    std::string k = "generic";
    auto put = [&](const void* x, size_t n) { k.append((const char*)x, n); };
    const void* ptrs[5] = {a, b, u, v, m->d_lk_status};
    put(ptrs, sizeof ptrs);
    put(&H, sizeof H);
    put(&W, sizeof W);
    put(&p->levels, sizeof p->levels);
    put(&p->filt_n, sizeof p->filt_n);
    if (p->filt_w && p->filt_n > 0) put(p->filt_w, sizeof(float) * p->filt_n);
    put(&p->opt_n, sizeof p->opt_n);
    if (p->opt_w && p->opt_n > 0) put(p->opt_w, sizeof(float) * p->opt_n);
    put(&p->warping, sizeof(int) * 3);
    put(&p->ls_alpha, sizeof p->ls_alpha);
    put(&p->ls_maxnum, sizeof p->ls_maxnum);
    put(&p->ls_tol, sizeof p->ls_tol);
    put(&m->kind, sizeof m->kind);
    if (m->kind == 0) {
        const ofri_fb_params* f = m->fb;
        put(&f->poly_n, sizeof f->poly_n);
        put(f->g, sizeof(float) * (f->poly_n + 1));
        put(f->xg, sizeof(float) * (f->poly_n + 1));
        put(f->xxg, sizeof(float) * (f->poly_n + 1));
        put(f->ig, sizeof(float) * 4);
        put(&f->pre_kh, sizeof f->pre_kh);
        put(f->pre_kernel, sizeof(float) * (f->pre_kh + 1));
        put(&f->win_kh, sizeof f->win_kh);
        if (!f->box_filter) put(f->win_kernel, sizeof(float) * (f->win_kh + 1));
        put(&f->iters, sizeof f->iters);
        put(&f->box_filter, sizeof f->box_filter);
    } else {
        put(&m->lk_half_win, sizeof m->lk_half_win);
        put(&m->lk_niter, sizeof m->lk_niter);
    }
    put(&ctx->tuning, sizeof ctx->tuning);
    put(&ctx->classes_gen, sizeof ctx->classes_gen);
    return k;
}

static std::string batch_key(ofri_ctx* ctx, const ofri_pair* pr, int n, const ofri_lshs_params* p) {
    std::string k = "batch";
    k.append((const char*)pr, sizeof(ofri_pair) * n);
    k += lshs_key(ctx, nullptr, nullptr, 0, 0, p, nullptr, nullptr);
    return k;
}

}  // namespace ofri

extern "C" {

int ofri_set_graphs(ofri_ctx* ctx, int on) {
    CHECK_CTX(ctx);
    ctx->graphs_on = on != 0;
    return OFRI_OK;
}

int ofri_pyr_lshs(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W,
                  const ofri_lshs_params* p, float* d_Uacc, float* d_Vacc,
                  ofri_level_stats* h_stats) {
    CHECK_CTX(ctx);
    if (!p || !d_im1 || !d_im2 || !d_Uacc || !d_Vacc) return fail(OFRI_E_ARG, "ofri_pyr_lshs: null pointer");
    if (p->levels < 1 || p->levels > 12) return fail(OFRI_E_ARG, "ofri_pyr_lshs: levels %d", p->levels);
    if (!p->alphas) return fail(OFRI_E_ARG, "ofri_pyr_lshs: alphas missing");
    const int L = p->levels;
    std::vector<int> lh, lw;
    level_sizes(H, W, L, lh, lw);
    for (int l = 0; l < L; ++l)
        if (lh[l] < 4 || lw[l] < 4)
            return fail(OFRI_E_ARG, "ofri_pyr_lshs: level %d size %dx%d below 4x4", l + 1, lh[l], lw[l]);
    if (!ctx->graphs_on) {
        OFRI_TRY(pyr_lshs_enqueue(ctx, d_im1, d_im2, H, W, p, d_Uacc, d_Vacc));
    } else {
        GraphEntry& g = ctx->graphs[lshs_key(ctx, d_im1, d_im2, H, W, p, d_Uacc, d_Vacc)];
        OFRI_TRY(run_cached(ctx, g, [&](bool) { return pyr_lshs_enqueue(ctx, d_im1, d_im2, H, W, p, d_Uacc, d_Vacc); }));
    }
    if (h_stats) OFRI_TRY(pyr_stats_download(ctx, L, lh, lw, p, false, h_stats));
    return OFRI_OK;
}

int ofri_pyr_generic(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W,
                     const ofri_pyr_params* p, const ofri_pyr_main* m, float* d_Uacc, float* d_Vacc,
                     ofri_level_stats* h_stats) {
    CHECK_CTX(ctx);
    if (!p || !m || !d_im1 || !d_im2 || !d_Uacc || !d_Vacc) return fail(OFRI_E_ARG, "ofri_pyr_generic: null pointer");
    if (p->levels < 1 || p->levels > 12) return fail(OFRI_E_ARG, "ofri_pyr_generic: levels %d", p->levels);
    if (m->kind == 0) {
        if (!m->fb || !m->fb->g || !m->fb->xg || !m->fb->xxg || !m->fb->ig || !m->fb->pre_kernel ||
            (!m->fb->box_filter && !m->fb->win_kernel))
            return fail(OFRI_E_ARG, "ofri_pyr_generic: Farneback parameters missing");
        if (m->fb->flow_scale != 1.0f) return fail(OFRI_E_ARG, "ofri_pyr_generic: Farneback flow_scale must be 1");
    } else if (m->kind != 1) {
        return fail(OFRI_E_ARG, "ofri_pyr_generic: main solver kind %d", m->kind);
    }
    if ((p->filt_w == nullptr) != (p->filt_n <= 0) || (p->opt_w == nullptr) != (p->opt_n <= 0))
        return fail(OFRI_E_ARG, "ofri_pyr_generic: filter weights and counts disagree");
    const int L = p->levels;
    std::vector<int> lh, lw;
    level_sizes(H, W, L, lh, lw);
    for (int l = 0; l < L; ++l)
        if (lh[l] < 4 || lw[l] < 4)
            return fail(OFRI_E_ARG, "ofri_pyr_generic: level %d size %dx%d below 4x4", l + 1, lh[l], lw[l]);
    auto enqueue = [&](bool) { return pyr_generic_enqueue(ctx, d_im1, d_im2, H, W, p, m, d_Uacc, d_Vacc); };
    if (!ctx->graphs_on) OFRI_TRY(enqueue(false));
    else OFRI_TRY(run_cached(ctx, ctx->graphs[generic_key(ctx, d_im1, d_im2, H, W, p, m, d_Uacc, d_Vacc)], enqueue));
    if (h_stats) {
        void* st;
        OFRI_TRY(scratch(ctx, "gpyr_stats", (size_t)L * sizeof(SolveCtrl), &st));
        std::vector<SolveCtrl> hc(L);
        OFRI_HIP(hipMemcpyAsync(hc.data(), st, L * sizeof(SolveCtrl), hipMemcpyDeviceToHost, ctx->stream));
        OFRI_HIP(hipStreamSynchronize(ctx->stream));
        for (int l = 0; l < L; ++l) {
            ofri_level_stats& o = h_stats[l];
            o = ofri_level_stats{};
            o.H = lh[l];
            o.W = lw[l];
            o.ls_err = p->ls_alpha > 0.0f ? hc[l].last_err : 0.0f;
            o.ls_iters = p->ls_alpha > 0.0f ? hc[l].sweeps : 0;
        }
    }
    return OFRI_OK;
}

int ofri_pyr_lshs_taps(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W,
                       const ofri_lshs_params* p, float* d_Uacc, float* d_Vacc, float* d_taps,
                       const int* h_ls_tap_back, ofri_level_stats* h_stats) {
    CHECK_CTX(ctx);
    if (!p || !d_im1 || !d_im2 || !d_Uacc || !d_Vacc) return fail(OFRI_E_ARG, "ofri_pyr_lshs_taps: null pointer");
    if (p->levels < 1 || p->levels > 12) return fail(OFRI_E_ARG, "ofri_pyr_lshs_taps: levels %d", p->levels);
    if (!p->alphas) return fail(OFRI_E_ARG, "ofri_pyr_lshs_taps: alphas missing");
    const int L = p->levels;
    std::vector<int> lh, lw;
    level_sizes(H, W, L, lh, lw);
    if (h_ls_tap_back)
        for (int l = 0; l < L; ++l)
            if (h_ls_tap_back[l] < 0) return fail(OFRI_E_ARG, "ofri_pyr_lshs_taps: negative tap offset at level %d", l);
    OFRI_TRY(pyr_lshs_enqueue(ctx, d_im1, d_im2, H, W, p, d_Uacc, d_Vacc, d_taps, true, nullptr, h_ls_tap_back));
    if (h_stats) OFRI_TRY(pyr_stats_download(ctx, L, lh, lw, p, true, h_stats));
    else OFRI_HIP(hipStreamSynchronize(ctx->stream));
    return OFRI_OK;
}

int ofri_batch_lshs(ofri_ctx* ctx, const ofri_pair* pairs, int n, const ofri_lshs_params* p,
                    ofri_level_stats* h_stats) {
    CHECK_CTX(ctx);
    if (!p || !pairs) return fail(OFRI_E_ARG, "ofri_batch_lshs: null pointer");
    if (n < 1) return fail(OFRI_E_ARG, "ofri_batch_lshs: %d pairs", n);
    if (p->levels < 1 || p->levels > 12) return fail(OFRI_E_ARG, "ofri_batch_lshs: levels %d", p->levels);
    if (!p->alphas) return fail(OFRI_E_ARG, "ofri_batch_lshs: alphas missing");
    const int L = p->levels;
    for (int i = 0; i < n; ++i) {
        const ofri_pair& q = pairs[i];
        if (!q.im1 || !q.im2 || !q.U || !q.V) return fail(OFRI_E_ARG, "ofri_batch_lshs: pair %d: null pointer", i);
        std::vector<int> lh, lw;
        level_sizes(q.H, q.W, L, lh, lw);
        for (int l = 0; l < L; ++l)
            if (lh[l] < 4 || lw[l] < 4)
                return fail(OFRI_E_ARG, "ofri_batch_lshs: pair %d level %d size %dx%d below 4x4", i, l + 1, lh[l],
                            lw[l]);
    }
    const std::string key = batch_key(ctx, pairs, n, p);
    GraphEntry& g = ctx->graphs[key];
    const size_t tab_bytes = (size_t)L * (align256(hs_batch_table_bytes(n)) + align256(ls_batch_table_bytes(n)));
    if (g.tab.bytes < tab_bytes) {
        OFRI_HIP(hipStreamSynchronize(ctx->stream));
        if (g.tab.p) OFRI_HIP(hipFree(g.tab.p));
        g.tab.p = nullptr;
        g.tab.bytes = 0;
        OFRI_HIP(hipMalloc(&g.tab.p, tab_bytes));
        g.tab.bytes = tab_bytes;
    }
    if (!ctx->graphs_on) {
        OFRI_TRY(batch_lshs_enqueue(ctx, pairs, n, p, g.tab.p, true));
    } else {
        // the job tables hold scratch addresses: they are written by eager
        // runs only (never while capturing), and an eager run that moved a
        // scratch buffer is run again before any capture
        OFRI_TRY(run_cached(ctx, g, [&](bool capturing) {
            return batch_lshs_enqueue(ctx, pairs, n, p, g.tab.p, !capturing);
        }, true));
    }
    if (h_stats) {
        void* st;
        OFRI_TRY(scratch(ctx, "batch_stats", (size_t)L * n * (sizeof(SolveCtrl) + sizeof(float)), &st));
        std::vector<SolveCtrl> hc((size_t)L * n);
        std::vector<float> he((size_t)L * n);
        OFRI_HIP(hipMemcpyAsync(hc.data(), st, hc.size() * sizeof(SolveCtrl), hipMemcpyDeviceToHost, ctx->stream));
        OFRI_HIP(hipMemcpyAsync(he.data(), (SolveCtrl*)st + hc.size(), he.size() * sizeof(float),
                                hipMemcpyDeviceToHost, ctx->stream));
        OFRI_HIP(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < n; ++i) {
            std::vector<int> lh, lw;
            level_sizes(pairs[i].H, pairs[i].W, L, lh, lw);
            for (int l = 0; l < L; ++l) {
                ofri_level_stats& o = h_stats[(size_t)i * L + l];
                const size_t k = (size_t)l * n + i;
                o.H = lh[l];
                o.W = lw[l];
                o.hs_err = he[k];
                o.ls_err = p->ls_alpha > 0.0f ? hc[k].last_err : 0.0f;
                o.ls_iters = p->ls_alpha > 0.0f ? hc[k].sweeps : 0;
                o.hs_res = 0.0f;
                o.hs_iters = p->niter;
            }
        }
    }
    return OFRI_OK;
}

// ---- warm-start network ----
#define NET(p) reinterpret_cast<ofri::Net*>(p)

int ofri_net_create(ofri_ctx* ctx, int planes, ofri_net** out) {
    CHECK_CTX(ctx);
    if (!out) return fail(OFRI_E_ARG, "ofri_net_create: null pointer");
    Net* n = nullptr;
    OFRI_TRY(net_create(ctx, planes, &n));
    *out = reinterpret_cast<ofri_net*>(n);
    return OFRI_OK;
}

int ofri_net_destroy(ofri_net* net) {
    net_destroy(NET(net));
    return OFRI_OK;
}

int ofri_net_num_params(ofri_net* net, size_t* n) {
    if (!net || !n) return fail(OFRI_E_ARG, "ofri_net_num_params: null pointer");
    *n = net_num_params(NET(net));
    return OFRI_OK;
}

int ofri_net_flops(ofri_net* net, int H, int W, double* flops) {
    if (!net || !flops) return fail(OFRI_E_ARG, "ofri_net_flops: null pointer");
    if (H <= 0 || W <= 0 || H % 16 || W % 16) return fail(OFRI_E_ARG, "ofri_net_flops: H, W must be positive multiples of 16");
    *flops = net_forward_flops(NET(net), H, W);
    return OFRI_OK;
}

int ofri_net_set_params(ofri_ctx* ctx, ofri_net* net, const float* h_params, size_t n) {
    CHECK_CTX(ctx);
    if (!net || !h_params) return fail(OFRI_E_ARG, "ofri_net_set_params: null pointer");
    return net_set_params(ctx, NET(net), h_params, n);
}

int ofri_net_init(ofri_ctx* ctx, ofri_net* net, unsigned long long seed) {
    CHECK_CTX(ctx);
    if (!net) return fail(OFRI_E_ARG, "ofri_net_init: null pointer");
    return net_init_params(ctx, NET(net), seed);
}

int ofri_net_get_params(ofri_ctx* ctx, ofri_net* net, float* h_params, size_t n) {
    CHECK_CTX(ctx);
    if (!net || !h_params) return fail(OFRI_E_ARG, "ofri_net_get_params: null pointer");
    return net_get_params(ctx, NET(net), h_params, n);
}

int ofri_net_get_grads(ofri_ctx* ctx, ofri_net* net, float* h_grads, size_t n) {
    CHECK_CTX(ctx);
    if (!net || !h_grads) return fail(OFRI_E_ARG, "ofri_net_get_grads: null pointer");
    return net_get_grads(ctx, NET(net), h_grads, n);
}

int ofri_net_correct(ofri_ctx* ctx, ofri_net* net, const float* const* d_planes, int H, int W, float* d_u,
                     float* d_v) {
    CHECK_CTX(ctx);
    if (!net || !d_planes || !d_u || !d_v) return fail(OFRI_E_ARG, "ofri_net_correct: null pointer");
    if (H <= 0 || W <= 0 || H % 16 || W % 16)
        return fail(OFRI_E_ARG, "ofri_net_correct: H, W must be positive multiples of 16 (got %dx%d)", H, W);
    const int P = net_planes(NET(net));
    for (int i = 0; i < P; ++i)
        if (!d_planes[i]) return fail(OFRI_E_ARG, "ofri_net_correct: input plane %d is null", i);
    return net_correct(ctx, NET(net), d_planes, H, W, d_u, d_v);
}

int ofri_net_predict(ofri_ctx* ctx, ofri_net* net, const float* d_in, int B, int H, int W, float* d_coarse,
                     float* d_fine) {
    CHECK_CTX(ctx);
    if (!net || !d_in || !d_coarse || !d_fine) return fail(OFRI_E_ARG, "ofri_net_predict: null pointer");
    return net_predict(ctx, NET(net), d_in, B, H, W, d_coarse, d_fine);
}

int ofri_net_train(ofri_ctx* ctx, ofri_net* net, const float* d_in, int B, int H, int W,
                   const float* d_lab_coarse, const float* d_lab_fine, const float* h_field_w, float lr,
                   float* h_loss) {
    CHECK_CTX(ctx);
    if (!net || !d_in || !d_lab_coarse || !d_lab_fine || !h_field_w)
        return fail(OFRI_E_ARG, "ofri_net_train: null pointer");
    void* l;
    OFRI_TRY(scratch(ctx, "net_loss", sizeof(float), &l));
    OFRI_TRY(net_train(ctx, NET(net), d_in, B, H, W, d_lab_coarse, d_lab_fine, h_field_w, lr, (float*)l));
    if (h_loss) {
        OFRI_HIP(hipMemcpyAsync(h_loss, l, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        OFRI_HIP(hipStreamSynchronize(ctx->stream));
    }
    return OFRI_OK;
}

int ofri_render_pair(ofri_ctx* ctx, int H, int W, int family, float density, unsigned long long seed,
                     const float* h_par, int npar, const float* h_modes, int nmodes, float noise, float* d_im1,
                     float* d_im2, float* d_u, float* d_v) {
    CHECK_CTX(ctx);
    if (!d_im1 || !d_im2) return fail(OFRI_E_ARG, "ofri_render_pair: null pointer");
    if ((d_u == nullptr) != (d_v == nullptr)) return fail(OFRI_E_ARG, "ofri_render_pair: give both d_u and d_v");
    return render_pair(ctx, H, W, family, density, seed, h_par, npar, h_modes, nmodes, noise, d_im1, d_im2, d_u,
                       d_v);
}

int ofri_pyr_lshs_learned(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W,
                          const ofri_lshs_params* p, const ofri_learned_params* lp, ofri_net* net,
                          float* d_Uacc, float* d_Vacc, ofri_level_stats* h_stats, int* h_converged,
                          int* d_converged) {
    CHECK_CTX(ctx);
    if (!p || !lp || !d_im1 || !d_im2 || !d_Uacc || !d_Vacc)
        return fail(OFRI_E_ARG, "ofri_pyr_lshs_learned: null pointer");
    if (p->levels != 2) return fail(OFRI_E_ARG, "ofri_pyr_lshs_learned: 2 levels only (got %d)", p->levels);
    if (!p->alphas || !lp->hs_tol || !lp->ls_tol) return fail(OFRI_E_ARG, "ofri_pyr_lshs_learned: missing arrays");
    if (H < 8 || W < 8) return fail(OFRI_E_ARG, "ofri_pyr_lshs_learned: %dx%d too small", H, W);
    for (int l = 0; l < 2; ++l)
        if (lp->hs_budget[l] < 1 || lp->hs_finish_budget[l] < 0 || !(lp->cheb_finish[l] >= 0.0f) ||
            lp->hs_finish_first[l] < 0)
            return fail(OFRI_E_ARG, "ofri_pyr_lshs_learned: HS budget / finish at level %d", l);
    if (lp->ls_budget < 0) return fail(OFRI_E_ARG, "ofri_pyr_lshs_learned: Liu-Shen budget");
    for (int l = 0; l < 2; ++l)
        if (lp->ls_sweeps[l] < 0) return fail(OFRI_E_ARG, "ofri_pyr_lshs_learned: Liu-Shen sweeps at level %d", l);
    for (int l = 0; l < 2; ++l)
        if (lp->cheb_bound[l] != 0.0f && !(lp->cheb_bound[l] > 0.0f && lp->cheb_bound[l] <= 0.999f))
            return fail(OFRI_E_ARG, "ofri_pyr_lshs_learned: Chebyshev bound %g at level %d", lp->cheb_bound[l], l);
    if (lp->ls_cheb_hi != 0.0f &&
        !(lp->ls_cheb_lo > -1.0f && lp->ls_cheb_lo < 0.0f && lp->ls_cheb_hi > 0.0f && lp->ls_cheb_hi <= 0.999f))
        return fail(OFRI_E_ARG, "ofri_pyr_lshs_learned: Liu-Shen interval [%g, %g]", lp->ls_cheb_lo, lp->ls_cheb_hi);
    const int L = 2;
    std::vector<int> lh, lw;
    level_sizes(H, W, L, lh, lw);
    if (net && net_planes(NET(net)) != 4)
        return fail(OFRI_E_ARG, "ofri_pyr_lshs_learned: the correction network takes 4 input planes");
    LearnedRun run{lp, NET(net)};
    std::string key = "learned" + lshs_key(ctx, d_im1, d_im2, H, W, p, d_Uacc, d_Vacc);
    {
        const void* np = net;
        key.append((const char*)&np, sizeof np);
        key.append((const char*)lp->hs_tol, 2 * sizeof(float));
        key.append((const char*)lp->ls_tol, 2 * sizeof(float));
        key.append((const char*)&lp->hs_budget, sizeof(ofri_learned_params) - offsetof(ofri_learned_params, hs_budget));
    }
    auto enqueue = [&]() { return pyr_lshs_enqueue(ctx, d_im1, d_im2, H, W, p, d_Uacc, d_Vacc, nullptr, false, &run); };
    if (!ctx->graphs_on) {
        OFRI_TRY(enqueue());
    } else {
        OFRI_TRY(run_cached(ctx, ctx->graphs[key], [&](bool) { return enqueue(); }));
    }
    if (d_converged) {
        void* stp;
        OFRI_TRY(scratch(ctx, "pyr_stats", (size_t)L * (2 * sizeof(SolveCtrl) + 16), &stp));
        learned_check_k<<<1, 1, 0, ctx->stream>>>((const SolveCtrl*)stp, lp->hs_tol[0], lp->hs_tol[1], lp->ls_tol[0],
                                                  lp->ls_tol[1], p->ls_alpha > 0.0f ? 1 : 0, lp->relative ? 1 : 0,
                                                  d_converged);
        OFRI_HIP(hipGetLastError());
    }
    if (h_stats || h_converged) {
        ofri_level_stats st[2];
        OFRI_TRY(pyr_stats_download(ctx, L, lh, lw, p, true, st, false));
        int ok = 1;
        const int rel = lp->relative ? 1 : 0;
        for (int l = 0; l < L; ++l) {
            if (!meets(st[l].hs_res, lp->hs_tol[l], st[l].hs_nrm, rel)) ok = 0;
            if (p->ls_alpha > 0.0f && !meets(st[l].ls_err, lp->ls_tol[l], st[l].ls_nrm, rel)) ok = 0;
            if (h_stats) h_stats[l] = st[l];
        }
        if (h_converged) *h_converged = ok;
    }
    return OFRI_OK;
}

}  // extern "C"
