/*
 * OpticalFlow-RI CPU oracle.  TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's dense-flow numerics, used by the
 * tests, by the smoke check and by the benchmark's CPU-baseline leg as the
 * checker the HIP library is compared against.  The shipped library never
 * links or loads it and has no CPU fallback to it.
 *
 * Two instances of every solver are built from oracle_impl.h:
 *   *_f32  the reference dataflow (float32 storage, float64 stencil sums);
 *   *_f64  the same method in float64 throughout.
 *
 * Reference files followed (paths relative to the reference checkout):
 *   src/gaussian_filter.py:24-94            orc_gauss_sep_*
 *   src/HornSchunck.py:51-126               orc_hs_derivs_*, orc_hs_sweep_*, orc_hs_*
 *   src/PhysicsBasedOpticalFlowLiuShen.py:47-158
 *                                           orc_ls_setup_*, orc_ls_sweep_*, orc_ls_*
 *   src/GenericPyramidalOpticalFlow.py:66-67   orc_resize_bicubic_* (+ orc_pil_coeffs)
 *   src/Farneback_PyCL.py:61-62             orc_resize_* (BILINEAR, filter 2)
 *   src/GenericPyramidalOpticalFlow.py:151-171 orc_spline_upsample_* (+ orc_spline_matrix)
 *   src/GenericPyramidalOpticalFlow.py:69-115,197-200  orc_warp_symmetric_*
 *   src/GenericPyramidalOpticalFlow.py:237-415 orc_pyr_lshs_*
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp, no -march, so
 * float arithmetic is plain SSE with no FMA contraction).
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { ORC_FILTER_BILINEAR = 2, ORC_FILTER_BICUBIC = 3 };

void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Pillow's filters (a = -0.5 bicubic, triangle bilinear), support 2 and 1. */
static double orc_bicubic(double x) {
    const double a = -0.5;
    if (x < 0.0) x = -x;
    if (x < 1.0) return ((a + 2.0) * x - (a + 3.0)) * x * x + 1;
    if (x < 2.0) return (((x - 5) * x + 8) * x - 4) * a;
    return 0.0;
}
static double orc_bilinear(double x) {
    if (x < 0.0) x = -x;
    if (x < 1.0) return 1.0 - x;
    return 0.0;
}

/* Pillow's per-output-sample coefficient table for a resize of one axis     */
/* (inSize -> outSize): window [int(c - s + .5), int(c + s + .5)) clipped to */
/* the input, weights f((x + xmin - c + .5) / fs) normalised by their sum,   */
/* with c = (xx + .5) * scale, fs = max(scale, 1), s = support * fs.         */
/* Returns a malloc'ed outSize x ksize table; bounds[2*xx] = xmin,           */
/* bounds[2*xx+1] = number of taps.  Pinned by Pillow impulse fixtures.      */
double* orc_pil_coeffs(int inSize, int outSize, int filter, int* ksize_out, int* bounds) {
    double support0 = filter == ORC_FILTER_BICUBIC ? 2.0 : 1.0;
    double scale = (double)inSize / outSize;
    double fs = scale < 1.0 ? 1.0 : scale;
    double support = support0 * fs;
    int ksize = (int)ceil(support) * 2 + 1;
    double* kk = (double*)calloc((size_t)outSize * ksize, sizeof(double));
    for (int xx = 0; xx < outSize; ++xx) {
        double center = (xx + 0.5) * scale;
        double ss = 1.0 / fs;
        int xmin = (int)(center - support + 0.5);
        if (xmin < 0) xmin = 0;
        int xmax = (int)(center + support + 0.5);
        if (xmax > inSize) xmax = inSize;
        xmax -= xmin;
        double* k = kk + (size_t)xx * ksize;
        double ww = 0.0;
        for (int x = 0; x < xmax; ++x) {
            double arg = (x + xmin - center + 0.5) * ss;
            double w = filter == ORC_FILTER_BICUBIC ? orc_bicubic(arg) : orc_bilinear(arg);
            k[x] = w;
            ww += w;
        }
        for (int x = 0; x < xmax; ++x)
            if (ww != 0.0) k[x] /= ww;
        bounds[2 * xx] = xmin;
        bounds[2 * xx + 1] = xmax;
    }
    *ksize_out = ksize;
    return kk;
}

/* Not-a-knot cubic interpolation matrix on n unit-spaced samples, evaluated */
/* at u_k = k * n / N (k = 0..N-1), u clamped to [0, n-1].  out is N x n,     */
/* row-major: value(u_k) = sum_j out[k*n + j] * y_j.  Slopes s = S y solve    */
/*   s0 + 2 s1            = (-5 y0 + 4 y1 + y2) / 2                          */
/*   s_{i-1} + 4 s_i + s_{i+1} = 3 (y_{i+1} - y_{i-1})        (0 < i < n-1)   */
/*   2 s_{n-2} + s_{n-1}  = (-y_{n-3} - 4 y_{n-2} + 5 y_{n-1}) / 2           */
/* (third derivative continuous at the second and second-last sample), then */
/* each interval is the cubic Hermite piece.  Requires n >= 4.               */
void orc_spline_matrix(int n, int N, double* out) {
    double* S = (double*)calloc((size_t)n * n, sizeof(double)); /* S[i*n + j] = ds_i/dy_j */
    double* cp = (double*)malloc(sizeof(double) * n);
    double* d = (double*)malloc(sizeof(double) * n);
    for (int j = 0; j < n; ++j) {
        /* rhs for y = e_j */
        for (int i = 0; i < n; ++i) d[i] = 0.0;
        if (j <= 2) d[0] = (j == 0 ? -5.0 : (j == 1 ? 4.0 : 1.0)) / 2.0;
        for (int i = 1; i < n - 1; ++i) d[i] = 3.0 * ((j == i + 1) - (j == i - 1));
        if (j >= n - 3) d[n - 1] = (j == n - 1 ? 5.0 : (j == n - 2 ? -4.0 : -1.0)) / 2.0;
        /* Thomas: row0 (b=1,c=2), rows (a=1,b=4,c=1), last (a=2,b=1) */
        cp[0] = 2.0 / 1.0;
        d[0] = d[0] / 1.0;
        for (int i = 1; i < n; ++i) {
            double a = (i == n - 1) ? 2.0 : 1.0;
            double b = (i == n - 1) ? 1.0 : 4.0;
            double c = (i == n - 1) ? 0.0 : 1.0;
            double m = b - a * cp[i - 1];
            cp[i] = c / m;
            d[i] = (d[i] - a * d[i - 1]) / m;
        }
        for (int i = n - 2; i >= 0; --i) d[i] -= cp[i] * d[i + 1];
        for (int i = 0; i < n; ++i) S[(size_t)i * n + j] = d[i];
    }
    for (int k = 0; k < N; ++k) {
        double* row = out + (size_t)k * n;
        memset(row, 0, sizeof(double) * n);
        int over = (long long)k * n > (long long)(n - 1) * N;
        double u = over ? (double)(n - 1) : ((double)k * (double)n) / (double)N;
        int i = (int)floor(u);
        if (i > n - 2) i = n - 2;
        double t = u - i;
        double t2 = t * t, t3 = t2 * t;
        double h00 = 2 * t3 - 3 * t2 + 1, h10 = t3 - 2 * t2 + t;
        double h01 = -2 * t3 + 3 * t2, h11 = t3 - t2;
        for (int j = 0; j < n; ++j)
            row[j] = h10 * S[(size_t)i * n + j] + h11 * S[(size_t)(i + 1) * n + j];
        row[i] += h00;
        row[i + 1] += h01;
    }
    free(S); free(cp); free(d);
}

#define REAL float
#define NAME(x) x##_f32
#include "oracle_impl.h"
#undef REAL
#undef NAME

#define REAL double
#define NAME(x) x##_f64
#include "oracle_impl.h"
#undef REAL
#undef NAME

/* Cross-precision helpers used by tests: widen / narrow whole arrays. */
void orc_f32_to_f64(const float* a, double* b, size_t n) { for (size_t i = 0; i < n; ++i) b[i] = a[i]; }
void orc_f64_to_f32(const double* a, float* b, size_t n) { for (size_t i = 0; i < n; ++i) b[i] = (float)a[i]; }
