/*
 * libofri: MI355X (gfx950) dense optical-flow kernels for PIV, C ABI.
 *
 * Every entry point returns an int status: OFRI_OK (0) or a negative error
 * class; ofri_last_error() gives the message of the calling thread's last
 * failure.  Images and flows are H x W row-major float32 arrays.  Pointers
 * named d_* are device pointers (from ofri_malloc or any HIP allocation on
 * the context's device); pointers named h_* are host pointers.  All work is
 * enqueued on the context's stream; entry points returning values to the host
 * (errors, iteration counts) synchronise that stream first.
 *
 * Each operation cites the reference code path it replaces (paths relative
 * to the OpticalFlow-RI checkout).
 */
#ifndef OFRI_H
#define OFRI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OFRI_OK 0
#define OFRI_E_ARG (-1)      /* invalid argument (shape, size, null pointer) */
#define OFRI_E_HIP (-2)      /* HIP runtime failure */
#define OFRI_E_NOMEM (-3)    /* device allocation failed */
#define OFRI_E_STATE (-4)    /* invalid context / plan state */
#define OFRI_E_NUMERIC (-5)  /* NaN / zero-division condition the reference raises on */

typedef struct ofri_ctx ofri_ctx;

/* ---- context, errors, device memory -------------------------------------- */
const char* ofri_last_error(void);
int ofri_version(void);                                   /* 10000*major + 100*minor + patch */
int ofri_create(int device, ofri_ctx** out);              /* one HIP stream per context */
int ofri_destroy(ofri_ctx* ctx);
int ofri_sync(ofri_ctx* ctx);

/* Lanes (no counterpart in the reference, which runs one pair at a time on
 * the CPU or one OpenCL queue): child contexts of a context on the same
 * device, each with a HIP stream (default priority) and scratch, graph and
 * table caches of its own, so independent pairs queued on different lanes
 * run concurrently and fill the GPU where one pair's launches do not.  Any
 * entry point takes a lane in place of a context.  A lane is created on
 * first request, takes its parent's tuning and graph setting at every
 * ofri_lane / ofri_fork call, is destroyed with its parent (ofri_destroy on
 * a lane fails) and has no lanes of its own.  Device buffers are shared by
 * the whole device: ordering between streams is the caller's, through
 * ofri_fork (the parent's work queued so far precedes lanes 0..n-1's later
 * work) and ofri_join (lanes 0..n-1's work so far precedes the parent's
 * later work). */
#define OFRI_MAX_LANES 8
int ofri_lane(ofri_ctx* ctx, int k, ofri_ctx** out);       /* 0 <= k < OFRI_MAX_LANES */
int ofri_fork(ofri_ctx* ctx, int n);
int ofri_join(ofri_ctx* ctx, int n);
int ofri_malloc(ofri_ctx* ctx, size_t bytes, void** d_out);
int ofri_free(ofri_ctx* ctx, void* d_ptr);
int ofri_upload(ofri_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int ofri_download(ofri_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);
int ofri_copy(ofri_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);
int ofri_memset(ofri_ctx* ctx, void* d_dst, int value, size_t bytes);
/* Device-side timing of the context stream: start/stop an event pair, then read ms. */
int ofri_timer_start(ofri_ctx* ctx);
int ofri_timer_stop(ofri_ctx* ctx, float* ms_out);

/* ---- pre-filter ---------------------------------------------------------- */
/* Separable Gaussian with the reference's padding and float32 tap-order
 * accumulation; weights h_w[n] (n odd, n <= 127) are the float32 kernel.
 * d_src may equal d_dst.
 * Replaces gaussian_filter.py:54-94 (convolveSeparableFilter / gaussian_filterPx). */
int ofri_gauss_sep(ofri_ctx* ctx, const float* d_src, float* d_dst, int H, int W,
                   const float* h_w, int n);

/* ---- pyramid helpers ----------------------------------------------------- */
/* Pillow-compatible antialiased resize of a float image, filter 3 = BICUBIC,
 * 2 = BILINEAR.  Replaces GenericPyramidalOpticalFlow.py:66-67 (imresize) and
 * the Farneback_PyCL.py:62 resize. */
int ofri_resize(ofri_ctx* ctx, const float* d_src, int H, int W, float* d_dst, int H2, int W2,
                int filter);
/* Not-a-knot bicubic spline upsample of a flow component on arange(n)/n
 * grids with clamped extrapolation, then multiplication by `scale` (pass 1.0f
 * for none).  Requires Hc, Wc >= 4.
 * Replaces GenericPyramidalOpticalFlow.py:151-171 (RectBivariateSpline + scaling). */
int ofri_spline_upsample(ofri_ctx* ctx, const float* d_src, int Hc, int Wc, float* d_dst, int H,
                         int W, float scale);
/* Symmetric bilinear warp: out1 = im1 at (y - V/2, x - U/2), out2 = im2 at
 * (y + V/2, x + U/2), round-half-even integer part, clamped neighbours.
 * Replaces GenericPyramidalOpticalFlow.py:69-115 as called at :197-200. */
int ofri_warp_symmetric(ofri_ctx* ctx, const float* d_im1, const float* d_im2, const float* d_U,
                        const float* d_V, int H, int W, float* d_out1, float* d_out2);
/* Bilinear sample of d_img at float32 coordinates (d_cy, d_cx) with the same
 * rounding and clamping.  Replaces GenericPyramidalOpticalFlow.py:69-115
 * (doBiLinearWarping) for arbitrary coordinates. */
int ofri_warp_coords(ofri_ctx* ctx, const float* d_img, const float* d_cy, const float* d_cx, int H,
                     int W, float* d_out);
/* d_acc += d_x (float32), n elements.  Replaces GenericPyramidalOpticalFlow.py:412-413. */
int ofri_accumulate(ofri_ctx* ctx, float* d_acc, const float* d_x, size_t n);
/* Mean endpoint error of two flows, mean over n of sqrt((U1-U2)^2 +
 * (V1-V2)^2), written to device memory (*d_out, float) without host
 * synchronisation: the learned form's EPE-to-plain judge on its check
 * sample.  No counterpart in the reference. */
int ofri_flow_epe(ofri_ctx* ctx, const float* d_U1, const float* d_V1, const float* d_U2, const float* d_V2,
                  size_t n, float* d_out);

/* ---- Horn-Schunck -------------------------------------------------------- */
/* niter Jacobi sweeps from the flow in d_U, d_V (in/out).  d_im1 is the frame
 * at t, d_im2 at t+1.  *h_err (may be NULL) receives (||U-U0||_F +
 * ||V-V0||_F)/(H*W), the reference's returned error.
 * Replaces HornSchunck.py:72-126 (HS, computeDerivatives, HS_helper,
 * HS_helper2) behind HSOpticalFlowAlgoAdapter.compute (HornSchunck.py:34-36). */
int ofri_hs(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W, float alpha,
            int niter, float* d_U, float* d_V, float* h_err);

/* Horn-Schunck with a stopping rule: sweeps until the per-sweep update norm
 * (||dU||_2 + ||dV||_2)/(H*W) <= tol, or maxnum sweeps (tol < 0: exactly
 * maxnum).  d_hist (device, maxnum floats, may be NULL) receives each sweep's
 * norm; *h_err / *h_iters (may be NULL) the last norm and the sweep count.
 * The sweeps are those of ofri_hs bit for bit: after k sweeps the flow equals
 * ofri_hs with niter = k.  The reference's HS always runs Niter sweeps
 * (HornSchunck.py:86-104); the rule is the Liu-Shen one
 * (PhysicsBasedOpticalFlowLiuShen.py:79,141) applied to the same update, for
 * callers that solve to a residual instead of a sweep count (warm starts). */
int ofri_hs_tol(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W, float alpha,
                int maxnum, double tol, float* d_U, float* d_V, float* d_hist, float* h_err,
                int* h_iters);

/* ofri_hs_tol with Chebyshev semi-iteration: each step combines the damped
 * sweep with the previous iterate, for a spectrum of the sweep's linear part
 * in [-1/3, bound] (0 < bound <= 0.999; the 8-neighbour average bounds it
 * below by -1/3, and a bound below the true slowest mode only slows that
 * mode down).  The accelerated sweep T is the HS update formed in float32
 * with fused multiply-adds (the same operator as ofri_hs's to float32
 * rounding).  Sweep k records |T(x_k) - x_k|, and the accelerated phase
 * returns T(x_k) for the first iterate x_k with that norm <= its tolerance.
 * finish > 1: the accelerated phase runs to finish * tol and the reference
 * sweeps of ofri_hs_tol continue to tol (float32 Chebyshev iterates stall
 * near the rounding level, plain sweeps settle); finish <= 1: accelerated to
 * tol.  maxnum bounds the sweeps of both phases together: the plain phase
 * gets at most maxnum / 2 of them, the accelerated phase the rest; d_hist,
 * *h_err and *h_iters cover both.
 * launch_test != 0: the stop test runs only at the last sweep of each launch
 * (cheaper; d_hist must be NULL): the accelerated phase then ends with the
 * iterate that launch produced, the plain phase with T(x_k) as before, and
 * the sweep count is a whole number of launches per phase.  launch_test is
 * a flag word: bit 0 the per-launch test, bit 1 (with bit 0) a relative
 * tolerance -- a launch's norm is compared with tol * (||U||_2 + ||V||_2) /
 * (H*W) of that launch's result.  No counterpart in the reference (an
 * accelerated form for solving to a residual). */
int ofri_hs_accel(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W, float alpha,
                  int maxnum, double tol, float bound, float finish, float* d_U, float* d_V,
                  float* d_hist, float* h_err, int* h_iters, int launch_test);

/* ---- Liu-Shen ------------------------------------------------------------ */
/* Liu-Shen physics-based refinement in the solver's own orientation: d_u is
 * the row (y) component, d_v the column (x) component, both in/out.  Sweeps
 * until the update norm (||du||_2 + ||dv||_2)/(H*W) <= tol or maxnum sweeps.
 * *h_err / *h_iters (may be NULL) receive the last norm and the sweep count.
 * As in the reference, an all-zero frame (maximum 0) yields NaN flow.
 * Replaces PhysicsBasedOpticalFlowLiuShen.py:47-158 (generate_invmatrix,
 * helper, physicsBasedOpticalFlowLiuShen); the adapter at :37-39 swaps U, V. */
int ofri_ls(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W, float h,
            int maxnum, double tol, float* d_u, float* d_v, float* h_err, int* h_iters);
/* ofri_ls without any device allocation: the coefficient fields, the
 * ping-pong pair, the stop test's partial sums, the maxima and the control
 * word live in a fixed 42 MiB device array of the library (one per GPU,
 * held until the process ends), so the call makes no hipMalloc / hipFree
 * and allocates nothing else either; images above about 2,700,000 pixels
 * are rejected (OFRI_E_ARG, the message says how much they need).  Same
 * arguments, results and sweeps as ofri_ls, bit for bit.  Calls that
 * overlap on one GPU (several contexts, or threads, of one process) share
 * that array and race.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid allocation). */
int ofri_ls_noalloc(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W, float h,
                    int maxnum, double tol, float* d_u, float* d_v, float* h_err, int* h_iters);
/* Liu-Shen as the learned form runs it: the update norm is tested once per
 * fused launch, at its last sweep, which is the reference update; the
 * sweeps before it in the launch are the same update in float32 with fused
 * multiply-adds.  Stops after the first launch whose tested norm is <= tol
 * (or maxnum sweeps); *h_err = that norm, *h_iters = sweeps run.  No
 * counterpart in the reference (the learned form's Liu-Shen step). */
int ofri_ls_learned(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W, float h,
                    int maxnum, double tol, float* d_u, float* d_v, float* h_err, int* h_iters);

/* ---- Farneback ----------------------------------------------------------- */
typedef struct ofri_fb_params {
    int poly_n;               /* polyN (5 or 7) */
    const float* g;           /* poly_n+1 taps of the expansion Gaussian, g[0] at the centre */
    const float* xg;          /* x * g */
    const float* xxg;         /* x^2 * g */
    const float* ig;          /* 4 entries of the inverse Gram matrix: ig11, ig03, ig33, ig55 */
    const float* pre_kernel;  /* pre_kh+1 half taps of the pre-blur (centre first) */
    int pre_kh;
    const float* win_kernel;  /* win_kh+1 half taps of the window blur */
    int win_kh;
    int iters;                /* flow iterations */
    float flow_scale;         /* the initial flow is multiplied by this first (float32); 1 = as given */
    int box_filter;           /* 1: box window of half size win_kh (useGaussian=False, replicate
                                 borders; win_kernel unused); 0: Gaussian window win_kernel */
} ofri_fb_params;

/* One Farneback level (the reference's compute() at one internal pyramid
 * level): flow *= flow_scale, reflect-101 pre-blur, polynomial expansion of
 * both frames, then `iters` rounds of window blur (Gaussian, or box for
 * useGaussian=False: updateFlowBoxFilter, Farneback_PyCL.py:424-439 with
 * the .cl boxFilter5), flow update and matrices update.  d_flowx / d_flowy: initial flow in, result out.  With pre_kh = 0
 * and pre_kernel = {1} the frames are taken as already pre-filtered (the
 * coarser levels, whose blur runs at full resolution before the resize).
 * Replaces the per-level body of Farneback_PyCL.compute
 * (Farneback_PyCL.py:515-596) and the kernels of optical_flow_farneback.cl;
 * the level loop (Farneback_PyCL.py:475-482,507-520) stays on the host. */
int ofri_farneback(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W,
                   float* d_flowx, float* d_flowy, const ofri_fb_params* p);

/* Reflect-101 separable Gaussian blur of one float plane with a half kernel
 * (kh+1 taps, centre first; kh <= 192), vertical pass then horizontal.
 * Replaces Farneback_PyCL.gaussianBlur (Farneback_PyCL.py:210-240) with the
 * .cl gaussianBlur kernel; the level loop blurs each full-resolution frame
 * with the level's kernel before resizing it (Farneback_PyCL.py:560-580). */
int ofri_fb_blur(ofri_ctx* ctx, const float* d_src, int H, int W, float* d_dst, const float* half_kernel,
                 int kh);

/* ---- dense Lucas-Kanade --------------------------------------------------- */
/* The Liu-Shen warping mode of updateNextPyramidalLevel
 * (GenericPyramidalOpticalFlow.py:188-220, biLinear=False): d_out = d_im1 with
 * every pixel moved by its flow (d_u column, d_v row component) rounded half
 * up, as numpy's fancy assignment moves it (a negative target wraps once, of
 * several sources the last in row-major order wins), then the fractional
 * flow smoothed by the symmetric Gaussian h_w[0..radius] (h_w[0] the centre:
 * scipy.ndimage.gaussian_filter's float64 weights, radius <= 48, 'reflect'
 * borders, float64 sums) and the optical-flow-equation intensity correction
 * applied to all but the last row and column.  *h_status = 1 when a target
 * lies past the image (the reference raises IndexError; d_out is then
 * undefined), else 0.  The second frame is not warped in this mode. */
int ofri_warp_liushen(ofri_ctx* ctx, const float* d_im1, int H, int W, const float* d_u, const float* d_v,
                      const double* h_w, int radius, float* d_out, int* h_status);


/* One level of dense pyramidal LK over the reference's window (half_win
 * 1..16; the 2*half_win+1 columns laid on a 32-wide grid of four 8-column
 * tiles, with the reference's tile rules: tile 0 always counts, so windows
 * under 16 columns cover 9, and columns past 32 do not count): Scharr
 * structure tensor of I over the window, up to niter Gauss-Newton steps
 * with J sampled bilinearly at the displaced window, step 32*A^-1*b, early
 * exit when both step components are below 0.01 px.  d_u / d_v: initial flow
 * in, result out (left unchanged where the tensor is singular).  d_status
 * (may be NULL): 1 = tracked, 0 = singular tensor or the window left the
 * image.  h_asym (may be NULL = {0,0,0,0}): the asymmetric window flags
 * {left, right, top, bottom}, each 0 or 1, of the vorticity enhancement
 * (denseLucasKanade_PyCL.py:75-92): left/top drop window column/row 8 (the
 * first of tile 1), right/bottom drop the last column/row.  Replaces
 * denseLucasKanade_PyCl.compute (denseLucasKanade_PyCL.py:113-169) and
 * pyrlkDenseLargeW.cl lkDense (:304-669). */
int ofri_lk_dense(ofri_ctx* ctx, const float* d_I, const float* d_J, int H, int W, float* d_u,
                  float* d_v, int half_win, int niter, unsigned char* d_status, const int* h_asym);

/* ---- whole LS(HS) pyramid ------------------------------------------------ */
typedef struct ofri_lshs_params {
    int levels;            /* pyramidalLevels */
    const float* alphas;   /* HS alpha per level, in the order popped (coarsest first) */
    int niter;             /* HS sweeps per level */
    const float* filt_w;   /* FILTER kernel (float32), NULL = no pre-filter */
    int filt_n;
    const float* opt_w;    /* FILTER_OPT kernel, NULL = unfiltered */
    int opt_n;
    float ls_alpha;        /* Liu-Shen h; <= 0 = no Liu-Shen stage */
    int ls_maxnum;         /* 60 in the reference */
    double ls_tol;         /* 1e-8 in the reference */
} ofri_lshs_params;

/* Per-level statistics written by ofri_pyr_lshs (may be NULL). */
typedef struct ofri_level_stats {
    int H, W;
    float hs_err;    /* HS's returned error (change from its initial flow) */
    float ls_err;    /* Liu-Shen's last update norm */
    int ls_iters;    /* Liu-Shen sweeps */
    float hs_res;    /* HS's last update norm (0 unless recorded) */
    int hs_iters;    /* HS sweeps */
    float hs_nrm;    /* relative tests: (||U||_2 + ||V||_2)/(H*W) of HS's tested result (0 otherwise) */
    float ls_nrm;    /* relative tests: the same norm of Liu-Shen's tested result (0 otherwise) */
} ofri_level_stats;

/* Coarse-to-fine LS(HS): bicubic downsample, spline upsample + x2 scaling,
 * symmetric bilinear warp, FILTER on the warped pair, HS, FILTER_OPT on the
 * un-warped pair, Liu-Shen seeded with HS's flow, accumulation; kLevels = 1.
 * Device in/out.  Replaces genericPyramidalOpticalFlow
 * (GenericPyramidalOpticalFlow.py:237-415) driven with
 * HSOpticalFlowAlgoAdapter as main and LiuShenOpticalFlowAlgoAdapter as the
 * optional adapter (examples/LiuSE_PyHSchunck_Fs3_4_PyrLvls2.py:133-139). */
int ofri_pyr_lshs(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W,
                  const ofri_lshs_params* p, float* d_Uacc, float* d_Vacc,
                  ofri_level_stats* h_stats);

/* ofri_pyr_lshs run eagerly with every level's solver results exported: d_taps
 * (device) receives per level, levels in order, the planes HS u, HS v,
 * LS u - HS u, LS v - HS v at the level's size (4 * sum of level pixels
 * floats), and the HS sweeps record their update norms (hs_res in h_stats).
 * h_ls_tap_back (may be NULL): per level, k > 0 takes the Liu-Shen planes
 * k sweeps before Liu-Shen's stop (at sweep ls_maxnum - k; Liu-Shen then
 * runs its remaining k sweeps from there, so the pyramid's result is
 * unchanged unless the tolerance stops it earlier; the level's ls_iters then
 * counts those k sweeps).  The warm-start network's training labels.  No
 * counterpart in the reference. */
int ofri_pyr_lshs_taps(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W,
                       const ofri_lshs_params* p, float* d_Uacc, float* d_Vacc, float* d_taps,
                       const int* h_ls_tap_back, ofri_level_stats* h_stats);

/* One image pair of a batch: device frames at t and t+1 (H x W) and the
 * device outputs U, V (H x W) the pair's accumulated flow is written to. */
typedef struct ofri_pair {
    const float* im1;
    const float* im2;
    int H, W;
    float* U;
    float* V;
} ofri_pair;

/* Batched LS(HS): ofri_pyr_lshs for n pairs of any sizes in one call, with
 * the same parameters for every pair.  Each pair's result equals what
 * ofri_pyr_lshs returns for it alone, bit for bit.  The Horn-Schunck and
 * Liu-Shen sweeps of all pairs at a level run in shared launches (each pair
 * keeps its own Liu-Shen stop test), and repeated calls on the same buffers
 * replay one HIP graph of the whole batch.  h_stats (may be NULL) receives
 * n * levels entries, pair-major.  The reference has no batch entry point: a
 * caller loops genericPyramidalOpticalFlow (GenericPyramidalOpticalFlow.py:237)
 * over the pairs. */
int ofri_batch_lshs(ofri_ctx* ctx, const ofri_pair* pairs, int n, const ofri_lshs_params* p,
                    ofri_level_stats* h_stats);

/* ---- generic pyramid with a Farneback or dense-LK main solver ------------ */
typedef struct ofri_pyr_params {
    int levels;            /* pyramidalLevels */
    const float* filt_w;   /* FILTER kernel (3 taps, float32) on the (warped) pair; NULL = none */
    int filt_n;
    const float* opt_w;    /* FILTER_OPT kernel (5 taps) on the un-warped pair; NULL = unfiltered */
    int opt_n;
    int warping;           /* 1: symmetric bilinear warp by the upsampled flow; 0: no warping */
    int inter_scaling;     /* pyramidalIntermediateScaling (levels before the last) */
    int final_scaling;     /* pyramidalScaling (the last level) */
    float ls_alpha;        /* optional Liu-Shen refinement: h; <= 0 = none */
    int ls_maxnum;         /* 60 in the reference */
    double ls_tol;         /* 1e-8 in the reference */
} ofri_pyr_params;

/* The main solver of ofri_pyr_generic. */
typedef struct ofri_pyr_main {
    int kind;                       /* 0: Farneback (one internal level, `fb`); 1: dense LK */
    const ofri_fb_params* fb;       /* kind 0: flow_scale must be 1 */
    int lk_half_win, lk_niter;      /* kind 1: halfWindow, Niter (no vorticity enhancement) */
    unsigned char* d_lk_status;     /* kind 1 (may be NULL): the last level's status map, H x W */
} ofri_pyr_main;

/* The generic coarse-to-fine driver on the device with a Farneback
 * (Farneback_PyCL with pyramidalLevels = 1) or dense-LK
 * (denseLucasKanade_PyCl without vorticity enhancement) main adapter and
 * an optional Liu-Shen adapter: bicubic downsample, spline upsample with
 * the level's scaling flag, symmetric bilinear warp or no warping, FILTER
 * on the (warped) pair, main solver, FILTER_OPT on the un-warped pair,
 * Liu-Shen seeded with the main solver's flow, accumulation; kLevels = 1.
 * The caller resolves the adapter defaults (warping, scaling) first.
 * h_stats (may be NULL): per level, Liu-Shen's last norm and sweeps.
 * Repeated calls replay a HIP graph.  Replaces genericPyramidalOpticalFlow
 * (GenericPyramidalOpticalFlow.py:237-415) as driven by
 * examples/LiuSE_Farneback_Fs0_0_PyrLvls2.py:51-76 and
 * examples/denseLK_Fs2_0_PyrLvls2.py:49-73, with every adapter call
 * (Farneback_PyCL.py:461-603, denseLucasKanade_PyCL.py:113-169,
 * PhysicsBasedOpticalFlowLiuShen.py:37-39) on device buffers. */
int ofri_pyr_generic(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W,
                     const ofri_pyr_params* p, const ofri_pyr_main* m, float* d_Uacc, float* d_Vacc,
                     ofri_level_stats* h_stats);

/* ---- Synthetic PIV pairs (training data) --------------------------------- */
/* Renders a particle-image pair on the device: particles at `density` per
 * pixel (8x8-pixel cells of 16 slots), d ~ N(2.5, 0.5) in [1, 4], I0 ~
 * U[150, 255], I0 exp(-8 r^2 / d^2) summed and clipped to [0, 255], frame 2
 * moved by the flow at each particle, N(0, noise) added.  family 0
 * Poiseuille (par: umax), 1 Lamb-Oseen (par: cx, cy, rc, vmax), 2 shear +
 * turbulence (par: shear, dmax; h_modes: kx[n], ky[n], phase[n], amp[n]),
 * displacements at most 8 px.  A pure function of the seed.  d_u, d_v (may
 * be NULL) receive the flow field.  The synthetic workloads of the
 * reference's examples are host-generated; no counterpart in the reference. */
int ofri_render_pair(ofri_ctx* ctx, int H, int W, int family, float density, unsigned long long seed,
                     const float* h_par, int npar, const float* h_modes, int nmodes, float noise, float* d_im1,
                     float* d_im2, float* d_u, float* d_v);

/* ---- Correction network (the learned form) ------------------------------ */
/* A fixed 4-level convolutional encoder-decoder (3x3 convolutions, BF16
 * weights and activations, FP32 accumulation on MFMA, about 0.57M
 * parameters).  Its input is a batch of B planar float32 H x W stacks of
 * `planes` planes (H, W multiples of 16): planes 0, 1 are frames (any grey
 * range, scaled inside by (I - 128) / 64), planes 2, 3 (planes = 4) flow
 * components in pixels.  In the learned LS(HS) form (planes = 4) it maps a
 * level's Liu-Shen input pair and the HS flow to the correction Liu-Shen's
 * sweeps would add (fine head, fields 0, 1).  No counterpart in the
 * reference.  Parameters are exchanged packed: layer by layer, weights
 * [cout][tap][cin] then bias [cout], float32. */
typedef struct ofri_net ofri_net;
int ofri_net_create(ofri_ctx* ctx, int planes, ofri_net** out);
int ofri_net_destroy(ofri_net* net);
int ofri_net_num_params(ofri_net* net, size_t* n);
/* Floating-point operations (2 per multiply-add, real channels) of one
 * forward pass over one H x W input; H, W multiples of 16. */
int ofri_net_flops(ofri_net* net, int H, int W, double* flops);
int ofri_net_set_params(ofri_ctx* ctx, ofri_net* net, const float* h_params, size_t n);
int ofri_net_get_params(ofri_ctx* ctx, ofri_net* net, float* h_params, size_t n);
/* Random initial parameters (He-normal for the leaky-ReLU layers, heads
 * scaled down, zero biases) from a seed; resets the optimiser state. */
int ofri_net_init(ofri_ctx* ctx, ofri_net* net, unsigned long long seed);
/* Planar outputs of the two heads: d_coarse [B][4][H/2][W/2] and d_fine
 * [B][4][H][W] (4 fields each). */
int ofri_net_predict(ofri_ctx* ctx, ofri_net* net, const float* d_in, int B, int H, int W, float* d_coarse,
                     float* d_fine);
/* One training step: forward, weighted Charbonnier loss against planar
 * labels laid out like ofri_net_predict's outputs (h_field_w: 8 weights,
 * coarse fields then fine fields; weight 0 leaves a field untrained),
 * backward, and with lr > 0 an Adam step.  *h_loss (may be NULL) receives
 * the loss (synchronises). */
/* The correction step on one H x W input given as separate device planes
 * (d_planes: `planes` device pointers, the layout ofri_net_predict reads as
 * [planes][H][W]): the fine head's fields 0, 1 are added into d_u, d_v
 * (H x W each, may be the flow planes among the inputs); the coarse head is
 * not computed.  H, W multiples of 16.  Same values as ofri_net_predict's
 * fine fields 0, 1 added to d_u, d_v. */
int ofri_net_correct(ofri_ctx* ctx, ofri_net* net, const float* const* d_planes, int H, int W, float* d_u,
                     float* d_v);
int ofri_net_train(ofri_ctx* ctx, ofri_net* net, const float* d_in, int B, int H, int W,
                   const float* d_lab_coarse, const float* d_lab_fine, const float* h_field_w, float lr,
                   float* h_loss);
/* Gradients of the last ofri_net_train call, packed like the parameters. */
int ofri_net_get_grads(ofri_ctx* ctx, ofri_net* net, float* h_grads, size_t n);

/* The learned form of the LS(HS) 2-level pyramid: at each level HS runs
 * (Chebyshev-accelerated when cheb_bound > 0) from zero until its update
 * norm is <= hs_tol[level] (relative != 0: <= hs_tol[level] * (||U||_2 +
 * ||V||_2)/(H*W) of the tested iterate) or hs_budget[level] sweeps; the
 * correction network (4 input planes) maps the level's Liu-Shen pair and
 * HS's flow to a correction of HS's flow, and Liu-Shen runs from HS's flow
 * plus that correction.  ls_sweeps[level] > 0: exactly that many Liu-Shen
 * sweeps, after which the level is accepted when the last update norm is
 * <= ls_tol[level] (relative as above); ls_sweeps[level] = 0: Liu-Shen runs
 * until its norm is <= ls_tol[level] or ls_budget sweeps.
 * net NULL: Liu-Shen from HS's flow alone (the method without the network).  Both solvers test
 * their targets once per fused launch (at its last sweep).  *h_converged (may be
 * NULL) is 1 when every level met its targets; otherwise the caller falls
 * back to ofri_pyr_lshs.  d_converged (may be NULL): the same flag written to
 * device memory (int) with no host synchronisation, so that many pairs can
 * be queued and their flags read once.  h_stats (may be NULL): 2 entries (hs_err, the
 * plain HS's change-from-start value, is not formed: 0).  Levels whose
 * size is not a multiple of 16 skip the network.  Repeated calls replay a
 * HIP graph.  No counterpart in the reference. */
typedef struct ofri_learned_params {
    const float* hs_tol;      /* per level (coarse, fine) */
    const float* ls_tol;
    int hs_budget[2];         /* HS sweeps per level, both phases */
    int ls_budget;
    float cheb_bound[2];      /* HS per level (coarse, fine): e.g. 0.99, 0.985; 0 = plain sweeps */
    float cheb_finish[2];     /* HS per level: e.g. 4, 2; see ofri_hs_accel */
    float ls_cheb_lo;         /* Liu-Shen Chebyshev interval [lo, hi] (lo in (-1, 0)); */
    float ls_cheb_hi;         /* hi = 0: plain Liu-Shen sweeps */
    int hs_finish_budget[2];  /* HS plain finishing sweeps per level at most (0: hs_budget / 2), within hs_budget */
    int relative;             /* 1: hs_tol / ls_tol are relative to the tested iterate's norm */
    int ls_sweeps[2];         /* per level: > 0 = exactly this many Liu-Shen sweeps, ls_tol accepts */
    int hs_finish_first[2];   /* per level: > 0 = the first plain finishing launch runs at most this many
                               * sweeps (0: the fused depth); used where cheb_bound > 0 and cheb_finish > 1 */
} ofri_learned_params;
int ofri_pyr_lshs_learned(ofri_ctx* ctx, const float* d_im1, const float* d_im2, int H, int W,
                          const ofri_lshs_params* p, const ofri_learned_params* lp, ofri_net* net,
                          float* d_Uacc, float* d_Vacc, ofri_level_stats* h_stats, int* h_converged,
                          int* d_converged);

/* Tuning knobs of the temporally blocked sweeps.  0 = the library's
 * per-image-size default (measured on gfx950); a variant k >= 1 selects the
 * k-th compiled tile shape. */
typedef struct ofri_tuning {
    int hs_depth;    /* sweeps fused per launch, Horn-Schunck */
    int ls_depth;    /* sweeps fused per launch, Liu-Shen */
    int hs_variant;  /* Horn-Schunck tile shape, 1-based (0 = default) */
    int ls_variant;  /* Liu-Shen tile shape, 1-based (0 = default) */
} ofri_tuning;
int ofri_set_tuning(ofri_ctx* ctx, const ofri_tuning* t);
int ofri_get_tuning(ofri_ctx* ctx, ofri_tuning* t);
/* The depths and 1-based tile shapes the sweeps actually use for an H x W image. */
int ofri_query_tuning(ofri_ctx* ctx, int H, int W, ofri_tuning* out);
/* Size classes of launch shapes (no counterpart in the reference: the
 * autotuner's per-size results, which ofri.autotune caches per GPU
 * architecture).  For a plain Horn-Schunck or Liu-Shen solve of an image, or
 * a batch by its total pixel count, the class with the largest min_px at or
 * below the pixel count supplies the depth and tile shape fields it sets
 * (> 0); ofri_set_tuning's fields (> 0) still override.  The learned form's
 * accelerated sweeps keep their own shapes.  n = 0 clears; at most 16. */
typedef struct ofri_tuning_class {
    long long min_px;
    ofri_tuning t;
} ofri_tuning_class;
int ofri_set_tuning_classes(ofri_ctx* ctx, const ofri_tuning_class* classes, int n);
/* The device's architecture name (e.g. "gfx950"), NUL-terminated in out[len]. */
int ofri_arch(ofri_ctx* ctx, char* out, int len);

/* HIP-graph replay of ofri_pyr_lshs (default on): the first call with a given
 * set of pointers, sizes, parameters and tuning runs eagerly, the second
 * captures the launch sequence, later calls replay it.  A graph is rebuilt
 * after any internal scratch buffer is reallocated. */
int ofri_set_graphs(ofri_ctx* ctx, int on);

#ifdef __cplusplus
}
#endif
#endif /* OFRI_H */
