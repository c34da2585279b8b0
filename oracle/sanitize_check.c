/*
 * Memory and undefined-behaviour check of the CPU oracle.  TEST
 * INFRASTRUCTURE ONLY: built with -fsanitize=address,undefined (Makefile
 * target `sanitize`) and run by tests/test_oracle.py; it drives every oracle
 * entry point over small, odd and edge-case shapes (images barely larger
 * than the stencils, flows that push samples out of the image, windows
 * wider than the image) in both instances and exits 0.  Any out-of-bounds
 * access, leak or undefined operation makes the sanitizers abort it.
 */
#include "ofri_oracle.c"

#include <stdio.h>

static unsigned g_state = 12345u;
static double rnd(void) {
    g_state = g_state * 1664525u + 1013904223u;
    return (double)(g_state >> 8) / 16777216.0;
}

#define CHECK_INSTANCE(R, S)                                                                              \
    static void check_##S(void) {                                                                         \
        const int sizes[][2] = {{5, 7}, {9, 16}, {33, 40}, {64, 49}};                                     \
        for (int si = 0; si < 4; ++si) {                                                                  \
            const int H = sizes[si][0], W = sizes[si][1];                                                 \
            const size_t N = (size_t)H * W;                                                               \
            R *a = malloc(sizeof(R) * N), *b = malloc(sizeof(R) * N), *u = malloc(sizeof(R) * N);         \
            R *v = malloc(sizeof(R) * N), *o1 = malloc(sizeof(R) * N), *o2 = malloc(sizeof(R) * N);       \
            for (size_t i = 0; i < N; ++i) {                                                              \
                a[i] = (R)(255 * rnd());                                                                  \
                b[i] = (R)(255 * rnd());                                                                  \
                u[i] = (R)(6 * rnd() - 3);                                                                \
                v[i] = (R)(6 * rnd() - 3);                                                                \
            }                                                                                             \
            const R k3[3] = {0.25, 0.5, 0.25}, k5[5] = {0.1, 0.2, 0.4, 0.2, 0.1};                         \
            if (H >= 5 && W >= 5) {                                                                       \
                memcpy(o1, a, sizeof(R) * N);                                                             \
                orc_gauss_sep_##S(o1, H, W, k3, 3);                                                       \
                orc_gauss_sep_##S(o1, H, W, k5, 5);                                                       \
            }                                                                                             \
            memcpy(o1, u, sizeof(R) * N);                                                                 \
            memcpy(o2, v, sizeof(R) * N);                                                                 \
            orc_hs_##S(a, b, H, W, (R)21, 7, o1, o2);                                                     \
            int it = 0;                                                                                   \
            orc_ls_##S(a, b, H, W, (R)5, 6, 1e-8, o1, o2, &it);                                           \
            orc_warp_symmetric_##S(a, b, u, v, H, W, o1, o2);                                             \
            const int H2 = (H + 1) / 2, W2 = (W + 1) / 2;                                                 \
            R* r = malloc(sizeof(R) * (size_t)(2 * H + 3) * (2 * W + 3));                                 \
            orc_resize_##S(a, H, W, r, H2, W2, 3);                                                        \
            orc_resize_##S(a, H, W, r, H2, W2, 2);                                                        \
            orc_resize_##S(a, H, W, r, 2 * H + 3, 2 * W + 3, 3);                                          \
            orc_resize_##S(a, H, W, r, H, W, 2);                                                          \
            if (H >= 4 && W >= 4) {                                                                       \
                double* d64 = malloc(sizeof(double) * (size_t)(2 * H + 1) * (2 * W + 1));                 \
                orc_spline_upsample_##S(a, H, W, r, 2 * H + 1, 2 * W + 1, (R)2, d64);                     \
                orc_spline_upsample_##S(a, H, W, r, 2 * H, 2 * W, (R)1, NULL);                            \
                free(d64);                                                                                \
            }                                                                                             \
            if (H >= 8 && W >= 8) {                                                                       \
                R* U = malloc(sizeof(R) * N);                                                             \
                R* V = malloc(sizeof(R) * N);                                                             \
                double st[6];                                                                             \
                const R al[2] = {45, 21};                                                                 \
                orc_pyr_lshs_##S(a, b, H, W, 2, al, 5, k3, 3, k5, 5, (R)5, 4, 1e-8, U, V, st);            \
                orc_pyr_lshs_##S(a, b, H, W, 1, al, 3, NULL, 0, NULL, 0, (R)0, 0, 1e-8, U, V, NULL);      \
                free(U);                                                                                  \
                free(V);                                                                                  \
            }                                                                                             \
            /* Farneback: polyN 7, a window half size (16) wider than the small images */                 \
            {                                                                                             \
                R g[8], xg[8], xxg[8], ig[4] = {0.1, -0.05, 0.2, 0.3}, wk[17], pk[2] = {0.5, 0.25};       \
                for (int i = 0; i < 8; ++i) {                                                             \
                    g[i] = (R)exp(-i * i / 4.5);                                                          \
                    xg[i] = i * g[i];                                                                     \
                    xxg[i] = i * i * g[i];                                                                \
                }                                                                                         \
                for (int i = 0; i < 17; ++i) wk[i] = (R)exp(-i * i / 56.0) / 10;                          \
                R* big = malloc(sizeof(R) * 5 * N);                                                       \
                R* big2 = malloc(sizeof(R) * 5 * N);                                                      \
                orc_fb_blur_##S(a, o1, H, W, 1, wk, 16);                                                  \
                orc_fb_polyexp_##S(a, H, W, g, xg, xxg, 7, ig, big);                                      \
                orc_fb_box_##S(big, big2, H, W, 5, 16);                                                   \
                memcpy(o1, u, sizeof(R) * N);                                                             \
                memcpy(o2, v, sizeof(R) * N);                                                             \
                orc_farneback_##S(a, b, H, W, pk, 1, g, xg, xxg, 7, ig, wk, 16, 3, o1, o2);              \
                orc_farneback_##S(a, b, H, W, pk, 1, g, xg, xxg, 7, ig, NULL, 16, 2, o1, o2);            \
                free(big);                                                                                \
                free(big2);                                                                               \
            }                                                                                             \
            /* dense LK: window 27 wider than the small images, flows pushing samples out */              \
            {                                                                                             \
                unsigned char w32[32], st8[64 * 49];                                                      \
                for (int i = 0; i < 32; ++i) w32[i] = i < 27 ? 1 : 0;                                     \
                memcpy(o1, u, sizeof(R) * N);                                                             \
                memcpy(o2, v, sizeof(R) * N);                                                             \
                orc_lk_dense_##S(a, b, H, W, 13, 5, o1, o2, st8, w32, w32);                               \
            }                                                                                             \
            free(r);                                                                                      \
            free(a);                                                                                      \
            free(b);                                                                                      \
            free(u);                                                                                      \
            free(v);                                                                                      \
            free(o1);                                                                                     \
            free(o2);                                                                                     \
        }                                                                                                 \
    }

CHECK_INSTANCE(float, f32)
CHECK_INSTANCE(double, f64)

int main(void) {
    check_f32();
    check_f64();
    int ks = 0;
    int bounds[2 * 33];
    double* k = orc_pil_coeffs(64, 33, 3, &ks, bounds);
    free(k);
    double* sm = malloc(sizeof(double) * 9 * 20);
    orc_spline_matrix(9, 20, sm);
    free(sm);
    printf("oracle sanitizer check: ok\n");
    return 0;
}
