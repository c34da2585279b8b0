// Internal declarations shared by the libofri translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/ofri.h"

namespace ofri {

int fail(int code, const char* fmt, ...);
void clear_error();

#define OFRI_HIP(call)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ::ofri::fail(OFRI_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

#define OFRI_TRY(expr)              \
    do {                            \
        int s_ = (expr);            \
        if (s_ != OFRI_OK) return s_; \
    } while (0)

// Tile of workgroup blk among nblk, for a tile grid tiles_x wide.  The
// dispatcher deals workgroups to the chip's 8 XCDs round-robin by linear
// index (XCD blk % 8), and each XCD has its own L2.  Neighbouring tiles read
// overlapping halos (and share the 128-byte lines their unaligned rows
// straddle), so XCD j takes a contiguous row-major run of tiles: run j holds
// the nblk / 8 (+1 for the first nblk % 8 runs) tiles after those of runs
// 0..j-1, and its workgroups take them in dispatch order.  A bijection of
// [0, nblk) whatever the XCD count; results do not depend on it.
__device__ __forceinline__ int xcd_run(int blk, int nblk) {
    const int q = nblk >> 3, r = nblk & 7, j = blk & 7;
    return j * q + (j < r ? j : r) + (blk >> 3);
}

__device__ __forceinline__ void xcd_tile(int blk, int nblk, int tiles_x, int& bx, int& by) {
    const int t = xcd_run(blk, nblk);
    by = t / tiles_x;
    bx = t - by * tiles_x;
}

// Device-side control word of a solver run with a convergence test.
// Stop-test state of a Jacobi solve as one launch hands it to the next: the
// decision of every test evaluated so far, plus the launch whose per-sweep
// partial sums are still untested (pend_nsw > 0).
struct SolveState {
    int done;        // 1 once the stop condition held or the sweep budget ran out
    int sweeps;      // sweeps taken so far
    int final_buf;   // ping-pong index holding the result
    int redo;        // >0: the stopping launch must be re-run with this many sweeps
    int redo_in;     // ping-pong index the re-run reads
    float last_err;  // last update norm (float32, like the reference)
    int limit;       // >0: the solve also stops at this many sweeps (a later phase's budget)
    int pend_nsw;    // sweeps of the untested launch (0: none)
    int pend_dst;    // its output buffer
    int pend_s0;     // its first tested sweep (earlier ones are counted, not tested)
    float last_nrm;  // relative mode: (||u||_2 + ||v||_2) / (H*W) of the tested launch's result
};

// Control word of one solve.  The leading fields are the published result
// (written by a phase's final launch); launch j of a phase reads slot[(j+1)&1]
// and its workgroup 0 writes slot[j&1], so no launch reads what it writes.
struct SolveCtrl {
    int done;
    int sweeps;
    int final_buf;
    int redo;
    int redo_in;
    float last_err;
    int limit;
    int base_buf;    // ping-pong index a following phase starts from (its launches' buffers are relative to it)
    float last_nrm;  // relative mode: the solution norm last_err was compared against (0 otherwise)
    SolveState slot[2];
};

// A device allocation owned by the context, grown on demand.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct ResizeTable {
    int in, out, filter, ksize, first, last;
    double* d_k = nullptr;
    int* d_bounds = nullptr;
};

struct SplineTable {
    int n, N, width;
    double* d_w = nullptr;  // N x width
    double* d_wT = nullptr; // width x N (the same weights transposed)
    int* d_start = nullptr;  // N
};

}  // namespace ofri

namespace ofri {
// A captured launch sequence, valid while no scratch buffer is reallocated.
struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    uint64_t ws_gen = 0;
    int eager_runs = 0;  // calls seen with this key before capture
    DevBuf tab;          // device job tables of a batched call (owned)
};
}  // namespace ofri

struct ofri_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ofri_tuning tuning{0, 0, 0, 0};
    bool graphs_on = true;
    uint64_t ws_gen = 1;  // bumped whenever a scratch buffer is (re)allocated
    std::map<std::string, ofri::GraphEntry> graphs;
    std::map<std::string, ofri::DevBuf> ws;
    std::map<std::tuple<int, int, int>, ofri::ResizeTable> resize_tables;
    std::map<std::pair<int, int>, ofri::SplineTable> spline_tables;
    std::map<uint64_t, ofri::DevBuf> omega_tabs;  // Chebyshev weight tables by sigma (bits)
    // the allocation-free entry point's device scratch on this device
    // (the address hipGetSymbolAddress gave for it; null if that failed)
    void* noalloc = nullptr;
    size_t noalloc_bytes = 0;
    std::string noalloc_err;
    // lanes: child contexts on the same device with streams of their own,
    // owned by this context (a lane has a parent and no lanes)
    ofri_ctx* parent = nullptr;
    std::vector<ofri_ctx*> lanes;
    // launch-shape size classes, ascending min_px (ofri_set_tuning_classes);
    // the generation enters every graph key
    std::vector<ofri_tuning_class> classes;
    uint64_t classes_gen = 0;
    hipEvent_t order_ev = nullptr;   // fork / join ordering point of this stream
};

namespace ofri {

// Taps of the separable pre-filter (odd count, float32 weights).
constexpr int kMaxTaps = 127;
struct Taps {
    float w[kMaxTaps];
    int n;
};

// Padding of a line of length L for a filter of half width h, as the
// reference pads (gaussian_filter.py:63-66): p < 0 -> -1-p, p >= L -> p-h.
__device__ __forceinline__ int pf_index(int p, int L, int h) {
    return p < 0 ? -1 - p : (p >= L ? p - h : p);
}

// Launch shape of a plain sweep over npx pixels (an image, or a batch's
// total): the size class's fields, then the context's tuning, over the
// built-in default in *vi (0-based variant) and *depth.
inline void shape_override(const ofri_ctx* ctx, size_t npx, bool hs, int* vi, int* depth) {
    const ofri_tuning* c = nullptr;
    for (const auto& k : ctx->classes)
        if ((long long)npx >= k.min_px) c = &k.t;
    const ofri_tuning* ts[2] = {c, &ctx->tuning};
    for (const ofri_tuning* t : ts) {
        if (!t) continue;
        const int d = hs ? t->hs_depth : t->ls_depth, v = hs ? t->hs_variant : t->ls_variant;
        if (d > 0) *depth = d;
        if (v > 0) *vi = v - 1;
    }
}

// Named scratch buffer of at least `bytes` bytes (grows, never shrinks).
int scratch(ofri_ctx* ctx, const std::string& name, size_t bytes, void** out);
// Host table -> device, on the context's stream and finished on return (no
// copy goes through the device's null stream, which would take a hardware
// queue of its own); refused while the stream captures a graph.  `what`
// names the table in the error.
int upload_table(ofri_ctx* ctx, void* dst, const void* src, size_t bytes, const char* what);

// ---- image kernels (k_image.hip) ----
// "2" variants: the same operation on two planes in one launch (second plane
// null: one plane).
int gauss_sep2(ofri_ctx* ctx, const float* src1, float* dst1, const float* src2, float* dst2, int H, int W,
               const float* w, int n);
int gauss_sep(ofri_ctx* ctx, const float* src, float* dst, int H, int W, const float* w, int n);
int resize2(ofri_ctx* ctx, const float* src1, float* dst1, const float* src2, float* dst2, int H, int W, int H2,
            int W2, int filter);
int resize(ofri_ctx* ctx, const float* src, int H, int W, float* dst, int H2, int W2, int filter);
// U and V in one pair of launches (src_v == nullptr: one plane)
int spline_upsample2(ofri_ctx* ctx, const float* src_u, const float* src_v, int Hc, int Wc, float* dst_u,
                     float* dst_v, int H, int W, float scale_u, float scale_v);
int spline_upsample(ofri_ctx* ctx, const float* src, int Hc, int Wc, float* dst, int H, int W,
                    float scale);
int warp_symmetric(ofri_ctx* ctx, const float* im1, const float* im2, const float* U,
                   const float* V, int H, int W, float* o1, float* o2);
int warp_coords(ofri_ctx* ctx, const float* img, const float* cy, const float* cx, int H, int W,
                float* out);
// The Liu-Shen warping mode (integer scatter, smoothed fractional flow,
// intensity correction) of im1 by (u, v) into out; w[0..r] the symmetric
// Gaussian's half taps (w[0] the centre); *d_err (device int, cleared by the
// caller) set when a target lies outside the image.
int warp_liushen(ofri_ctx* ctx, const float* im1, int H, int W, const float* u, const float* v,
                 const double* w, int r, float* out, int* d_err);
int accumulate2(ofri_ctx* ctx, float* acc1, const float* x1, float* acc2, const float* x2, size_t n);
int accumulate(ofri_ctx* ctx, float* acc, const float* x, size_t n);
int fill_zero2(ofri_ctx* ctx, float* p1, float* p2, size_t n);
int fill_zero(ofri_ctx* ctx, float* p, size_t n);
// d = a - b (float32), n elements; d may alias neither input
int subtract(ofri_ctx* ctx, float* d, const float* a, const float* b, size_t n);
// max over n floats -> d_out[0]; part (optional): kMaxRedPart floats of
// caller memory for the partial maxima (else a scratch buffer)
constexpr int kMaxRedPart = 1024;
int reduce_max2(ofri_ctx* ctx, const float* x1, const float* x2, size_t n, float* d_out, float* part = nullptr);
int reduce_max(ofri_ctx* ctx, const float* x, size_t n, float* d_out);
// d_out[0] = (||A-A0|| + ||B-B0||) / n as float32 (differences in float32)
int diff_norm_err(ofri_ctx* ctx, const float* A, const float* A0, const float* B, const float* B0,
                  size_t n, float* d_out);
// d_out[0] = mean over n of sqrt((U1-U2)^2 + (V1-V2)^2) (float32 terms, double sum)
int flow_epe(ofri_ctx* ctx, const float* U1, const float* V1, const float* U2, const float* V2, size_t n,
             float* d_out);
// dst[i] = src[ctrl->final_buf][i] for a ping-pong pair
// add: ad += selected, bd += selected instead of copies
int select_copy2(ofri_ctx* ctx, const SolveCtrl* ctrl, const float* a0, const float* a1, float* ad,
                 const float* b0, const float* b1, float* bd, size_t n, bool add = false);
int select_copy(ofri_ctx* ctx, const SolveCtrl* ctrl, const float* s0, const float* s1, float* dst,
                size_t n);

// ---- Horn-Schunck (k_hs.hip) ----
// With d_ctrl: per-sweep update norms, stop once the norm <= tol (tol < 0:
// never), the sweep history in d_hist (niter floats, may be null).
// cheb_b in (0, 1) (needs d_ctrl): Chebyshev-accelerated sweeps for a
// spectrum bound b on the slowest mode; 0 = plain sweeps.  cheb_switch > 1:
// accelerate to cheb_switch * tol, then at most niter / 2 plain sweeps to
// tol (niter sweeps in all).  check_last: the stop test runs once per fused
// launch, on its last sweep (no re-run of the stopping launch).
// rel (needs check_last): tol is relative -- a launch's tested norm is
// compared with tol * (||U||_2 + ||V||_2) / (H*W) of that launch's result.
// With Chebyshev sweeps and a finishing phase: fin_first > 0 caps the
// finishing phase's first launch at that many sweeps.
int hs_solve(ofri_ctx* ctx, const float* im1, const float* im2, int H, int W, float alpha,
             int niter, float* U, float* V, float* d_err, double tol = -1.0,
             SolveCtrl* d_ctrl = nullptr, float* d_hist = nullptr, float cheb_b = 0.0f,
             float cheb_switch = 0.0f, bool check_last = false, int fin_budget = 0, bool zero_start = false,
             const float* pre_w = nullptr, int pre_n = 0, bool rel = false, int fin_first = 0);

void hs_effective(ofri_ctx* ctx, size_t npx, int* variant1, int* depth);
void ls_effective(ofri_ctx* ctx, size_t npx, int* variant1, int* depth);

// ---- Farneback (k_fb.hip) ----
int farneback(ofri_ctx* ctx, const float* im1, const float* im2, int H, int W, float* flx, float* fly,
              const ofri_fb_params* p);
int fb_blur(ofri_ctx* ctx, const float* src, float* dst, int H, int W, const float* half, int kh);

// ---- dense Lucas-Kanade (k_lk.hip) ----
int lk_dense(ofri_ctx* ctx, const float* I, const float* J, int H, int W, int hw, int niter, float* u,
             float* v, unsigned char* status, const int* asym = nullptr);
unsigned lk_window_mask(int hw, int lo, int hi);

// ---- Liu-Shen (k_ls.hip) ----
// check_last: the stop test runs once per fused launch, on its last sweep
// (the solve may run up to depth - 1 sweeps past the first one that met tol)
// cheb_hi > 0: Chebyshev-accelerated sweeps for a spectrum in
// [cheb_lo, cheb_hi] (cheb_lo in (-1, 0)).
// sink_u / sink_v (optional): the result goes there instead of into u, v
// (added when sink_add, else copied); u, v then hold the start only.
// rel (needs check_last): tol is relative, as for hs_solve; the control word
// also publishes the solution norm of the tested launch (last_nrm).
// ws (optional): caller-provided device memory instead of the context's
// scratch buffers (plain sweeps only: no Chebyshev, no float32 sweeps).
// Planes of a Liu-Shen solve's field buffer: the setup's I and It, then the
// u / v ping-pong pair.
constexpr int kLsPlanes = 2;
constexpr int kLsFieldPlanes = kLsPlanes + 2;
struct LsWorkspace {
    float* fields;        // kLsFieldPlanes * H * W floats: the setup's planes, u / v ping-pong
    double* part;         // the stop test's partial sums, part_doubles of them
    size_t part_doubles;
    float* mx;            // the two frames' maxima
    float* max_part;      // kMaxRedPart floats
};
int ls_solve(ofri_ctx* ctx, const float* im1, const float* im2, int H, int W, float h,
             int maxnum, double tol, float* u, float* v, SolveCtrl* d_ctrl, bool check_last = false,
             float cheb_lo = 0.0f, float cheb_hi = 0.0f, float* sink_u = nullptr, float* sink_v = nullptr,
             bool sink_add = false, bool fast = false, bool rel = false, const LsWorkspace* ws = nullptr);
// Doubles of stop-test partial sums a plain ls_solve of an H x W image uses.
size_t ls_part_doubles(ofri_ctx* ctx, int H, int W);
// The library's allocation-free device scratch of this device (the device
// must be current): its address and size.
int noalloc_scratch(int device, void** ptr, size_t* bytes);

// ---- batched solves: every image's tiles in the same launches ----
// One image of a batched solve.  HS: U, V in/out, d_err (may be null) gets
// HS's returned error.  LS: U, V are the solver's u (row) and v (column).
struct BatchImage {
    const float* im1;
    const float* im2;
    int H, W;
    float* U;
    float* V;
    float* d_err;
};
inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
// Bytes of the device job table of n images (HS / LS).  The table is written
// when `upload` is true (never while a graph is being captured; a captured
// launch sequence reads the table its eager run wrote).
size_t hs_batch_table_bytes(int n);
size_t ls_batch_table_bytes(int n);
int hs_batch(ofri_ctx* ctx, const BatchImage* im, int n, float alpha, int niter, void* d_tab,
             bool upload);
// d_ctrls: n control words (results: sweeps, last norm)
int ls_batch(ofri_ctx* ctx, const BatchImage* im, int n, float h, int maxnum, double tol,
             SolveCtrl* d_ctrls, void* d_tab, bool upload);

// ---- synthetic pairs (k_synth.hip) ----
int render_pair(ofri_ctx* ctx, int H, int W, int family, float density, unsigned long long seed, const float* par,
                int npar, const float* h_modes, int nmodes, float noise, float* im1, float* im2, float* u, float* v);

// ---- warm-start network (k_cnn.hip) ----
struct Net;
int net_create(ofri_ctx* ctx, int planes, Net** out);
int net_planes(const Net* n);
void net_destroy(Net* n);
size_t net_num_params(const Net* n);
double net_forward_flops(const Net* n, int H, int W);
int net_set_params(ofri_ctx* ctx, Net* n, const float* h, size_t count);
int net_init_params(ofri_ctx* ctx, Net* n, unsigned long long seed);
int net_get_params(ofri_ctx* ctx, Net* n, float* h, size_t count);
int net_get_grads(ofri_ctx* ctx, Net* n, float* h, size_t count);
// input: [B][planes][H][W] planar float32; planar outputs: coarse
// [B][4][H/2][W/2], fine [B][4][H][W]
int net_predict(ofri_ctx* ctx, Net* n, const float* in, int B, int H, int W, float* coarse, float* fine);
// one input as separate planes; fine head fields 0, 1 added into add_u, add_v
int net_correct(ofri_ctx* ctx, Net* n, const float* const* planes, int H, int W, float* add_u, float* add_v);
int net_train(ofri_ctx* ctx, Net* n, const float* in, int B, int H, int W, const float* lab_c, const float* lab_f,
              const float* h_field_w, float lr, float* d_loss);

}  // namespace ofri
