/*
 * Type-generic body of the CPU oracle.  TEST INFRASTRUCTURE ONLY: nothing in
 * the shipped library links, loads or calls this code; it is the checker the
 * GPU path is compared against.
 *
 * This header is included twice by ofri_oracle.c:
 *   REAL = float  , NAME(x) = x##_f32  -> the reference dataflow: every field is
 *       stored as float32, every scipy.ndimage.convolve stencil is summed in
 *       float64 and rounded to float32 (scipy computes in double and returns
 *       the input dtype), every element-wise numpy / numba expression is
 *       float32 arithmetic evaluated left to right without FMA contraction.
 *   REAL = double , NAME(x) = x##_f64  -> the same method run entirely in
 *       float64 ("FP64 run of the same method"), with the stencil weights taken
 *       in double (1/12, 1/6, ...) rather than rounded to float32.
 *
 * Conventions: images and flows are H x W row-major; index (i, j) = row, col.
 */

#define IDX(i, j) ((size_t)(i) * (size_t)W + (size_t)(j))

/* mirror (scipy 'mirror'): -1 -> 1, N -> N-2 */
static inline int NAME(mir)(int i, int n) {
    if (i < 0) return -i;
    if (i >= n) return 2 * n - 2 - i;
    return i;
}
/* nearest (scipy 'nearest'): clamp */
static inline int NAME(clp)(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

/* ------------------------------------------------------------------------ */
/* Pre-filter.  gaussian_filter.py:54-85 (convolveSeparableFilter) with the  */
/* inner product of gaussian_filter.py:24-45 (convolveDirect).               */
/* In-place, x (row) pass over every row first, then y (column) pass on the */
/* row-filtered image.  Padding of a line of length L with half width h:    */
/*   position p < 0  -> line[-1-p]       (gaussian_filter.py:64-65, :76-77)   */
/*   position p >= L -> line[p-h]        (gaussian_filter.py:66,   :78)       */
/* (for h = 1 this is half-sample symmetric; for h = 2 the far edge repeats */
/* "L-2, L-1" instead of mirroring, exactly as the reference writes it).    */
/* Output m = sum_{j=0}^{n-1} k[j] * line(m + h - j), accumulated in REAL   */
/* in that order starting from 0 (gaussian_filter.py:37-40).                */
/* ------------------------------------------------------------------------ */
static inline REAL NAME(pf_at)(const REAL* line, ptrdiff_t stride, int L, int h, int p) {
    int q = p < 0 ? -1 - p : (p >= L ? p - h : p);
    return line[(ptrdiff_t)q * stride];
}

void NAME(orc_gauss_sep)(REAL* img, int H, int W, const REAL* k, int n) {
    int h = n / 2;
#pragma omp parallel
    {
        REAL* buf = (REAL*)malloc(sizeof(REAL) * (size_t)(H > W ? H : W));
#pragma omp for schedule(static)
        for (int i = 0; i < H; ++i) {
            REAL* row = img + IDX(i, 0);
            for (int m = 0; m < W; ++m) {
                REAL acc = 0;
                for (int j = 0; j < n; ++j) {
                    REAL prod = k[j] * NAME(pf_at)(row, 1, W, h, m + h - j);
                    acc = acc + prod;
                }
                buf[m] = acc;
            }
            memcpy(row, buf, sizeof(REAL) * (size_t)W);
        }
#pragma omp barrier
#pragma omp for schedule(static)
        for (int j0 = 0; j0 < W; ++j0) {
            REAL* col = img + j0;
            for (int m = 0; m < H; ++m) {
                REAL acc = 0;
                for (int j = 0; j < n; ++j) {
                    REAL prod = k[j] * NAME(pf_at)(col, W, H, h, m + h - j);
                    acc = acc + prod;
                }
                buf[m] = acc;
            }
            for (int m = 0; m < H; ++m) col[(size_t)m * W] = buf[m];
        }
        free(buf);
    }
}

/* ------------------------------------------------------------------------ */
/* Horn-Schunck.  HornSchunck.py:72-126.                                     */
/* HS(im2, im1, ...) is called as HS(I0, I1, ...) (HornSchunck.py:36,72), so */
/* computeDerivatives sees (I1, I0).  2x2 kernels through scipy convolve    */
/* with mirror padding give, for X in {I1, I0}:                             */
/*   CX[X](i,j) = .25*( X[i,j] - X[i,j+1] + X[i+1,j] - X[i+1,j+1])           */
/*   CY[X](i,j) = .25*( X[i,j] + X[i,j+1] - X[i+1,j] - X[i+1,j+1])           */
/*   CT[X](i,j) = .25*( X[i,j] + X[i,j+1] + X[i+1,j] + X[i+1,j+1])           */
/* each summed in double and rounded to REAL; then (HornSchunck.py:118-124) */
/*   fx = CX[I1] + CX[I0],  fy = CY[I1] + CY[I0],  ft = CT[I0] - CT[I1].     */
/* ------------------------------------------------------------------------ */
void NAME(orc_hs_derivs)(const REAL* I0, const REAL* I1, int H, int W,
                         REAL* fx, REAL* fy, REAL* ft) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < H; ++i) {
        int i1 = NAME(mir)(i + 1, H);
        for (int j = 0; j < W; ++j) {
            int j1 = NAME(mir)(j + 1, W);
            double a1 = I1[IDX(i, j)], b1 = I1[IDX(i, j1)], c1 = I1[IDX(i1, j)], d1 = I1[IDX(i1, j1)];
            double a0 = I0[IDX(i, j)], b0 = I0[IDX(i, j1)], c0 = I0[IDX(i1, j)], d0 = I0[IDX(i1, j1)];
            REAL cx1 = (REAL)(0.25 * a1 - 0.25 * b1 + 0.25 * c1 - 0.25 * d1);
            REAL cx0 = (REAL)(0.25 * a0 - 0.25 * b0 + 0.25 * c0 - 0.25 * d0);
            REAL cy1 = (REAL)(0.25 * a1 + 0.25 * b1 - 0.25 * c1 - 0.25 * d1);
            REAL cy0 = (REAL)(0.25 * a0 + 0.25 * b0 - 0.25 * c0 - 0.25 * d0);
            REAL ct1 = (REAL)(0.25 * a1 + 0.25 * b1 + 0.25 * c1 + 0.25 * d1);
            REAL ct0 = (REAL)(0.25 * a0 + 0.25 * b0 + 0.25 * c0 + 0.25 * d0);
            fx[IDX(i, j)] = cx1 + cx0;
            fy[IDX(i, j)] = cy1 + cy0;
            ft[IDX(i, j)] = ct0 + (-ct1);
        }
    }
}

/* One Jacobi sweep, HornSchunck.py:61-70 with HS_helper2 :51-58.            */
/*   ubar = K * U (K = [[1/12,1/6,1/12],[1/6,0,1/6],[1/12,1/6,1/12]], mirror)*/
/*   der  = (fx*ubar + fy*vbar + ft) / (alpha^2 + fx^2 + fy^2)               */
/*   U'   = ubar - fx*der,  V' = vbar - fy*der                               */
void NAME(orc_hs_sweep)(const REAL* fx, const REAL* fy, const REAL* ft, REAL alpha,
                        int H, int W, const REAL* U, const REAL* V, REAL* Un, REAL* Vn) {
    const double wc = (sizeof(REAL) == 4) ? (double)(float)(1.0 / 12.0) : 1.0 / 12.0;
    const double we = (sizeof(REAL) == 4) ? (double)(float)(1.0 / 6.0) : 1.0 / 6.0;
    const REAL a2 = alpha * alpha;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < H; ++i) {
        int im = NAME(mir)(i - 1, H), ip = NAME(mir)(i + 1, H);
        for (int j = 0; j < W; ++j) {
            int jm = NAME(mir)(j - 1, W), jp = NAME(mir)(j + 1, W);
            /* The 8-term double sum is grouped as vertical pairs, then corners /
             * edges, then the two weights (K has only two distinct weights and
             * is symmetric); any grouping is within an f64 ulp of scipy's. */
            double su, sv;
            {
                double vl = (double)U[IDX(im, jm)] + (double)U[IDX(ip, jm)];
                double vc = (double)U[IDX(im, j)] + (double)U[IDX(ip, j)];
                double vr = (double)U[IDX(im, jp)] + (double)U[IDX(ip, jp)];
                double corners = vl + vr;
                double edges = vc + ((double)U[IDX(i, jm)] + (double)U[IDX(i, jp)]);
                su = wc * corners + we * edges;
            }
            {
                double vl = (double)V[IDX(im, jm)] + (double)V[IDX(ip, jm)];
                double vc = (double)V[IDX(im, j)] + (double)V[IDX(ip, j)];
                double vr = (double)V[IDX(im, jp)] + (double)V[IDX(ip, jp)];
                double corners = vl + vr;
                double edges = vc + ((double)V[IDX(i, jm)] + (double)V[IDX(i, jp)]);
                sv = wc * corners + we * edges;
            }
            REAL ub = (REAL)su, vb = (REAL)sv;
            REAL x = fx[IDX(i, j)], y = fy[IDX(i, j)], t = ft[IDX(i, j)];
            REAL num = x * ub;
            num = num + y * vb;
            num = num + t;
            REAL den = a2 + x * x;
            den = den + y * y;
            REAL der = num / den;
            Un[IDX(i, j)] = ub - x * der;
            Vn[IDX(i, j)] = vb - y * der;
        }
    }
}

/* (||A - B||_F) with the difference formed in REAL and the sum in double. */
static double NAME(diffnorm)(const REAL* A, const REAL* B, size_t n) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (size_t q = 0; q < n; ++q) {
        REAL d = A[q] - B[q];
        s += (double)d * (double)d;
    }
    return sqrt(s);
}

/* HS driver: HornSchunck.py:72-104.  U, V: initial flow in, result out.     */
/* Returns total_error = (||U-U0||_F + ||V-V0||_F) / (H*W) (HornSchunck.py:99)*/
double NAME(orc_hs)(const REAL* I0, const REAL* I1, int H, int W, REAL alpha, int niter,
                    REAL* U, REAL* V) {
    size_t N = (size_t)H * W;
    REAL* fx = (REAL*)malloc(sizeof(REAL) * N);
    REAL* fy = (REAL*)malloc(sizeof(REAL) * N);
    REAL* ft = (REAL*)malloc(sizeof(REAL) * N);
    REAL* U0 = (REAL*)malloc(sizeof(REAL) * N);
    REAL* V0 = (REAL*)malloc(sizeof(REAL) * N);
    REAL* Ua = (REAL*)malloc(sizeof(REAL) * N);
    REAL* Va = (REAL*)malloc(sizeof(REAL) * N);
    REAL* Ub = (REAL*)malloc(sizeof(REAL) * N);
    REAL* Vb = (REAL*)malloc(sizeof(REAL) * N);
    NAME(orc_hs_derivs)(I0, I1, H, W, fx, fy, ft);
    memcpy(U0, U, sizeof(REAL) * N);
    memcpy(V0, V, sizeof(REAL) * N);
    memcpy(Ua, U, sizeof(REAL) * N);
    memcpy(Va, V, sizeof(REAL) * N);
    for (int s = 0; s < niter; ++s) {
        NAME(orc_hs_sweep)(fx, fy, ft, alpha, H, W, Ua, Va, Ub, Vb);
        REAL* t = Ua; Ua = Ub; Ub = t;
        t = Va; Va = Vb; Vb = t;
    }
    memcpy(U, Ua, sizeof(REAL) * N);
    memcpy(V, Va, sizeof(REAL) * N);
    double e = (NAME(diffnorm)(U, U0, N) + NAME(diffnorm)(V, V0, N)) / (double)N;
    free(fx); free(fy); free(ft); free(U0); free(V0); free(Ua); free(Va); free(Ub); free(Vb);
    return (double)(REAL)e;
}

/* ------------------------------------------------------------------------ */
/* Liu-Shen.  PhysicsBasedOpticalFlowLiuShen.py:47-158.                      */
/* Inside the solver u is the row-axis (y) component and v the column-axis  */
/* (x) component; the adapter swaps them (:38-39).  Stencils, after the     */
/* kernel flips at :117-121 and :53-55, are correlations with the kernels as */
/* written; 'nearest' clamps, 'constant' pads with zero:                     */
/*   Dr[X]=(X[i+1,j]-X[i-1,j])/2   Dc[X]=(X[i,j+1]-X[i,j-1])/2               */
/*   M[X]=(X[i-1,j-1]-X[i-1,j+1]-X[i+1,j-1]+X[i+1,j+1])/4                    */
/*   Fr[X]=X[i-1,j]+X[i+1,j]       Fc[X]=X[i,j-1]+X[i,j+1]                   */
/*   D2r[X]=X[i-1,j]-2X[i,j]+X[i+1,j]   D2c = column analogue                */
/*   H0[X]=8-neighbour sum, zero outside                                     */
/* ------------------------------------------------------------------------ */
typedef struct {
    REAL *IIx, *IIy, *II, *Ixt, *Iyt, *B11, *B12, *B22;
} NAME(ls_fields);

/* Setup: normalisation (:96-97), coefficient fields (:124-128) and the     */
/* per-pixel inverse matrix (generate_invmatrix, :47-73).                    */
void NAME(orc_ls_setup)(const REAL* im1, const REAL* im2, int H, int W, REAL h,
                        REAL* IIx, REAL* IIy, REAL* II, REAL* Ixt, REAL* Iyt,
                        REAL* B11, REAL* B12, REAL* B22) {
    size_t N = (size_t)H * W;
    REAL m1 = im1[0], m2 = im2[0];
    for (size_t q = 1; q < N; ++q) {
        if (im1[q] > m1) m1 = im1[q];
        if (im2[q] > m2) m2 = im2[q];
    }
    REAL* I = (REAL*)malloc(sizeof(REAL) * N);
    REAL* It = (REAL*)malloc(sizeof(REAL) * N);
    for (size_t q = 0; q < N; ++q) {
        I[q] = im1[q] / m1;
        REAL i2 = im2[q] / m2;
        It[q] = i2 - I[q];
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < H; ++i) {
        int im = NAME(clp)(i - 1, H), ip = NAME(clp)(i + 1, H);
        for (int j = 0; j < W; ++j) {
            int jm = NAME(clp)(j - 1, W), jp = NAME(clp)(j + 1, W);
            size_t c = IDX(i, j);
            REAL x = I[c];
            REAL dr = (REAL)(-0.5 * I[IDX(im, j)] + 0.5 * I[IDX(ip, j)]);
            REAL dc = (REAL)(-0.5 * I[IDX(i, jm)] + 0.5 * I[IDX(i, jp)]);
            REAL dtr = (REAL)(-0.5 * It[IDX(im, j)] + 0.5 * It[IDX(ip, j)]);
            REAL dtc = (REAL)(-0.5 * It[IDX(i, jm)] + 0.5 * It[IDX(i, jp)]);
            IIx[c] = x * dr;
            IIy[c] = x * dc;
            II[c] = x * x;
            Ixt[c] = x * dtr;
            Iyt[c] = x * dtc;
            /* generate_invmatrix */
            int cnt = (i > 0) + (i < H - 1);              /* rows above/below  */
            int cnt_c = (j > 0) + (j < W - 1);
            REAL cm = (REAL)((cnt + 1) * (cnt_c + 1) - 1); /* in-image 8-nbrs */
            REAL d2r = (REAL)(1.0 * I[IDX(im, j)] - 2.0 * x + 1.0 * I[IDX(ip, j)]);
            REAL d2c = (REAL)(1.0 * I[IDX(i, jm)] - 2.0 * x + 1.0 * I[IDX(i, jp)]);
            REAL mm = (REAL)(0.25 * I[IDX(im, jm)] - 0.25 * I[IDX(im, jp)]
                           - 0.25 * I[IDX(ip, jm)] + 0.25 * I[IDX(ip, jp)]);
            REAL two_x = 2 * x;
            REAL hc = h * cm;
            REAL A11 = x * (d2r - two_x) - hc;
            REAL A22 = x * (d2c - two_x) - hc;
            REAL A12 = x * mm;
            REAL det = A11 * A22 - A12 * A12;
            B11[c] = A22 / det;
            B12[c] = (-A12) / det;
            B22[c] = A11 / det;
        }
    }
    free(I);
    free(It);
}

/* One sweep (:142-150), Jacobi in (u, v).  Returns the update norm         */
/* (||u'-u||_2 + ||v'-v||_2) / (r*c) (helper, :76-80).                       */
double NAME(orc_ls_sweep)(const REAL* IIx, const REAL* IIy, const REAL* II, const REAL* Ixt,
                          const REAL* Iyt, const REAL* B11, const REAL* B12, const REAL* B22,
                          REAL h, int H, int W, const REAL* u, const REAL* v, REAL* un, REAL* vn) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < H; ++i) {
        int im = NAME(clp)(i - 1, H), ip = NAME(clp)(i + 1, H);
        for (int j = 0; j < W; ++j) {
            int jm = NAME(clp)(j - 1, W), jp = NAME(clp)(j + 1, W);
            size_t c = IDX(i, j);
            REAL dru = (REAL)(-0.5 * u[IDX(im, j)] + 0.5 * u[IDX(ip, j)]);
            REAL dcu = (REAL)(-0.5 * u[IDX(i, jm)] + 0.5 * u[IDX(i, jp)]);
            REAL drv = (REAL)(-0.5 * v[IDX(im, j)] + 0.5 * v[IDX(ip, j)]);
            REAL dcv = (REAL)(-0.5 * v[IDX(i, jm)] + 0.5 * v[IDX(i, jp)]);
            REAL fru = (REAL)(1.0 * u[IDX(im, j)] + 1.0 * u[IDX(ip, j)]);
            REAL fcv = (REAL)(1.0 * v[IDX(i, jm)] + 1.0 * v[IDX(i, jp)]);
            REAL mu = (REAL)(0.25 * u[IDX(im, jm)] - 0.25 * u[IDX(im, jp)]
                           - 0.25 * u[IDX(ip, jm)] + 0.25 * u[IDX(ip, jp)]);
            REAL mv = (REAL)(0.25 * v[IDX(im, jm)] - 0.25 * v[IDX(im, jp)]
                           - 0.25 * v[IDX(ip, jm)] + 0.25 * v[IDX(ip, jp)]);
            /* H0 with zero padding: only in-image neighbours contribute */
            double hu = 0.0, hv = 0.0;
            for (int a = -1; a <= 1; ++a) {
                int ii = i + a;
                if (ii < 0 || ii >= H) continue;
                for (int b = -1; b <= 1; ++b) {
                    int jj = j + b;
                    if ((a == 0 && b == 0) || jj < 0 || jj >= W) continue;
                    hu += (double)u[IDX(ii, jj)];
                    hv += (double)v[IDX(ii, jj)];
                }
            }
            REAL h0u = (REAL)hu, h0v = (REAL)hv;
            REAL iix = IIx[c], iiy = IIy[c], ii2 = II[c];
            REAL bu = (2 * iix) * dru;
            bu = bu + iix * dcv;
            bu = bu + iiy * drv;
            bu = bu + ii2 * fru;
            bu = bu + ii2 * mv;
            bu = bu + h * h0u;
            bu = bu + Ixt[c];
            REAL bv = iiy * dru;
            bv = bv + iix * dcu;
            bv = bv + (2 * iiy) * dcv;
            bv = bv + ii2 * mu;
            bv = bv + ii2 * fcv;
            bv = bv + h * h0v;
            bv = bv + Iyt[c];
            REAL b11 = B11[c], b12 = B12[c], b22 = B22[c];
            REAL t1 = b11 * bu, t2 = b12 * bv;
            un[c] = -(t1 + t2);
            REAL t3 = b12 * bu, t4 = b22 * bv;
            vn[c] = -(t3 + t4);
        }
    }
    size_t N = (size_t)H * W;
    return (NAME(diffnorm)(un, u, N) + NAME(diffnorm)(vn, v, N)) / (double)N;
}

/* LS driver (:82-158): while err > tol and k < maxnum.  u (row comp.) and  */
/* v (column comp.) in/out.  Returns the last err; *iters = sweeps done.     */
double NAME(orc_ls)(const REAL* im1, const REAL* im2, int H, int W, REAL h, int maxnum,
                    double tol, REAL* u, REAL* v, int* iters) {
    size_t N = (size_t)H * W;
    REAL* f = (REAL*)malloc(sizeof(REAL) * N * 12);
    REAL *IIx = f, *IIy = f + N, *II = f + 2 * N, *Ixt = f + 3 * N, *Iyt = f + 4 * N;
    REAL *B11 = f + 5 * N, *B12 = f + 6 * N, *B22 = f + 7 * N;
    REAL *ua = f + 8 * N, *va = f + 9 * N, *ub = f + 10 * N, *vb = f + 11 * N;
    NAME(orc_ls_setup)(im1, im2, H, W, h, IIx, IIy, II, Ixt, Iyt, B11, B12, B22);
    memcpy(ua, u, sizeof(REAL) * N);
    memcpy(va, v, sizeof(REAL) * N);
    double err = 100000000.0, last = 0.0;
    int k = 0;
    while (err > tol && k < maxnum) {
        err = (double)(REAL)NAME(orc_ls_sweep)(IIx, IIy, II, Ixt, Iyt, B11, B12, B22, h, H, W,
                                               ua, va, ub, vb);
        REAL* t = ua; ua = ub; ub = t;
        t = va; va = vb; vb = t;
        last = err;
        ++k;
    }
    memcpy(u, ua, sizeof(REAL) * N);
    memcpy(v, va, sizeof(REAL) * N);
    free(f);
    if (iters) *iters = k;
    return last;
}

/* ------------------------------------------------------------------------ */
/* Pillow BICUBIC / BILINEAR resize of a float image (GenericPyramidal-    */
/* OpticalFlow.py:66-67 and Farneback_PyCL.py:61-62).  Pillow's separable   */
/* antialiased resampler: horizontal pass into a REAL intermediate covering */
/* the rows the vertical pass reads, then the vertical pass; coefficients   */
/* and accumulation in double.  An unchanged axis has unit coefficients, so */
/* skipping its pass (as Pillow does) gives the same values.                */
/* ------------------------------------------------------------------------ */
void NAME(orc_resize)(const REAL* src, int H, int W, REAL* dst, int H2, int W2, int filter) {
    int ksh, ksv;
    int* bh = (int*)malloc(sizeof(int) * 2 * (size_t)W2);
    int* bv = (int*)malloc(sizeof(int) * 2 * (size_t)H2);
    double* kh = orc_pil_coeffs(W, W2, filter, &ksh, bh);
    double* kv = orc_pil_coeffs(H, H2, filter, &ksv, bv);
    int yfirst = bv[0];
    int ylast = bv[2 * (H2 - 1)] + bv[2 * (H2 - 1) + 1];
    int Ht = ylast - yfirst;
    REAL* tmp = (REAL*)malloc(sizeof(REAL) * (size_t)Ht * W2);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < Ht; ++y) {
        for (int xx = 0; xx < W2; ++xx) {
            int xmin = bh[2 * xx], xn = bh[2 * xx + 1];
            const double* k = kh + (size_t)xx * ksh;
            double ss = 0.0;
            for (int x = 0; x < xn; ++x) ss += (double)src[(size_t)(y + yfirst) * W + xmin + x] * k[x];
            tmp[(size_t)y * W2 + xx] = (REAL)ss;
        }
    }
#pragma omp parallel for schedule(static)
    for (int yy = 0; yy < H2; ++yy) {
        int ymin = bv[2 * yy] - yfirst, yn = bv[2 * yy + 1];
        const double* k = kv + (size_t)yy * ksv;
        for (int xx = 0; xx < W2; ++xx) {
            double ss = 0.0;
            for (int y = 0; y < yn; ++y) ss += (double)tmp[(size_t)(y + ymin) * W2 + xx] * k[y];
            dst[(size_t)yy * W2 + xx] = (REAL)ss;
        }
    }
    free(tmp); free(kh); free(kv); free(bh); free(bv);
}

void NAME(orc_resize_bicubic)(const REAL* src, int H, int W, REAL* dst, int H2, int W2) {
    NAME(orc_resize)(src, H, W, dst, H2, W2, ORC_FILTER_BICUBIC);
}

/* ------------------------------------------------------------------------ */
/* Spline upsample, GenericPyramidalOpticalFlow.py:151-171: RectBivariate-  */
/* Spline(k=3, s=0) on grids arange(n)/n, evaluated on the finer grid with  */
/* positions past the last sample clamped, i.e. a tensor-product not-a-knot */
/* cubic interpolant applied separably (x, then y), in double.  Then the    */
/* result is rounded to REAL and multiplied by `scale` in REAL (:166-171).  */
/* ------------------------------------------------------------------------ */
void NAME(orc_spline_upsample)(const REAL* src, int Hc, int Wc, REAL* dst, int H, int W,
                               REAL scale, double* dst64) {
    double* Wx = (double*)malloc(sizeof(double) * (size_t)W * Wc);
    double* Wy = (double*)malloc(sizeof(double) * (size_t)H * Hc);
    orc_spline_matrix(Wc, W, Wx);
    orc_spline_matrix(Hc, H, Wy);
    double* t = (double*)malloc(sizeof(double) * (size_t)Hc * W);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < Hc; ++r)
        for (int x = 0; x < W; ++x) {
            double s = 0.0;
            for (int j = 0; j < Wc; ++j) s += Wx[(size_t)x * Wc + j] * (double)src[(size_t)r * Wc + j];
            t[(size_t)r * W + x] = s;
        }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double s = 0.0;
            for (int r = 0; r < Hc; ++r) s += Wy[(size_t)y * Hc + r] * t[(size_t)r * W + x];
            if (dst64) dst64[(size_t)y * W + x] = s;
            REAL o = (REAL)s;
            dst[(size_t)y * W + x] = o * scale;
        }
    free(t); free(Wx); free(Wy);
}

/* ------------------------------------------------------------------------ */
/* Symmetric bilinear warp, GenericPyramidalOpticalFlow.py:69-115 as called */
/* at :199-200: out1 = img1 at (y - V/2, x - U/2), out2 = img2 at           */
/* (y + V/2, x + U/2).  Coordinates are formed in double from the REAL half */
/* flow and rounded to REAL (np.float32(ysMesh - vsNew/2.0)); the integer   */
/* part is round-half-even, the neighbour is +1 or -1 by the sign of the   */
/* fraction, both indices are clamped, the blend is double.                 */
/* ------------------------------------------------------------------------ */
static inline REAL NAME(warp_px)(const REAL* img, int H, int W, REAL cy, REAL cx) {
    int iy = (int)rint((double)cy), ix = (int)rint((double)cx);
    double dy = (double)cy - (double)iy, dx = (double)cx - (double)ix;
    int iyn = dy < 0 ? iy - 1 : iy + 1;
    int ixn = dx < 0 ? ix - 1 : ix + 1;
    dy = fabs(dy);
    dx = fabs(dx);
    iy = NAME(clp)(iy, H); iyn = NAME(clp)(iyn, H);
    ix = NAME(clp)(ix, W); ixn = NAME(clp)(ixn, W);
    double r = (1.0 - dy) * (1.0 - dx) * (double)img[IDX(iy, ix)];
    r = r + (1.0 - dy) * dx * (double)img[IDX(iy, ixn)];
    r = r + dy * (1.0 - dx) * (double)img[IDX(iyn, ix)];
    r = r + dy * dx * (double)img[IDX(iyn, ixn)];
    return (REAL)r;
}

void NAME(orc_warp_symmetric)(const REAL* im1, const REAL* im2, const REAL* U, const REAL* V,
                              int H, int W, REAL* out1, REAL* out2) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            REAL hu = U[IDX(i, j)] / 2, hv = V[IDX(i, j)] / 2;
            REAL cy1 = (REAL)((double)i - (double)hv), cx1 = (REAL)((double)j - (double)hu);
            REAL cy2 = (REAL)((double)i + (double)hv), cx2 = (REAL)((double)j + (double)hu);
            out1[IDX(i, j)] = NAME(warp_px)(im1, H, W, cy1, cx1);
            out2[IDX(i, j)] = NAME(warp_px)(im2, H, W, cy2, cx2);
        }
}

/* ------------------------------------------------------------------------ */
/* LS(HS) pyramid, GenericPyramidalOpticalFlow.py:237-415 with the HS        */
/* adapter as main (defaults warping/biLinear/scaling = True, HornSchunck.py*/
/* :44-49) and the LS adapter as the optional one, kLevels = 1.             */
/*   alphas[0..L-1] : HS alpha per level in the order popped (coarse first) */
/*   ls_alpha <= 0  : no Liu-Shen stage                                     */
/*   filt_w/nf      : FILTER weights (nf = 0 -> no pre-filter)              */
/*   opt_w/no       : FILTER_OPT weights (no = 0 -> unfiltered copy)        */
/*   stats (may be NULL): per level {hs_err, ls_err, ls_iters}              */
/* ------------------------------------------------------------------------ */
void NAME(orc_pyr_lshs)(const REAL* im1, const REAL* im2, int H, int W, int levels,
                        const REAL* alphas, int niter, const REAL* filt_w, int nf,
                        const REAL* opt_w, int no, REAL ls_alpha, int ls_maxnum, double ls_tol,
                        REAL* Uacc_out, REAL* Vacc_out, double* stats) {
    double scale = 1.0 / pow(2.0, levels - 1);
    int ph = 0, pw = 0;
    REAL *Uacc = NULL, *Vacc = NULL;
    for (int level = 1; level <= levels; ++level) {
        int h = H, w = W;
        REAL *n1 = NULL, *n2 = NULL; /* im1IterNew, im2IterNew */
        if (scale < 1.0 && level != levels) {
            w = (int)rint(W * scale);
            h = (int)rint(H * scale);
            n1 = (REAL*)malloc(sizeof(REAL) * (size_t)h * w);
            n2 = (REAL*)malloc(sizeof(REAL) * (size_t)h * w);
            NAME(orc_resize_bicubic)(im1, H, W, n1, h, w);
            NAME(orc_resize_bicubic)(im2, H, W, n2, h, w);
        } else {
            n1 = (REAL*)malloc(sizeof(REAL) * (size_t)h * w);
            n2 = (REAL*)malloc(sizeof(REAL) * (size_t)h * w);
            memcpy(n1, im1, sizeof(REAL) * (size_t)h * w);
            memcpy(n2, im2, sizeof(REAL) * (size_t)h * w);
        }
        size_t N = (size_t)h * w;
        REAL* w1 = (REAL*)malloc(sizeof(REAL) * N);
        REAL* w2 = (REAL*)malloc(sizeof(REAL) * N);
        REAL* U = (REAL*)calloc(N, sizeof(REAL));
        REAL* V = (REAL*)calloc(N, sizeof(REAL));
        if (level > 1) {
            REAL* us = (REAL*)malloc(sizeof(REAL) * N);
            REAL* vs = (REAL*)malloc(sizeof(REAL) * N);
            REAL sx = 1, sy = 1;
            if (ph != h || pw != w) {
                /* scaling (True at every level for HS) */
                sx = (REAL)w / (REAL)pw;
                sy = (REAL)h / (REAL)ph;
                NAME(orc_spline_upsample)(Uacc, ph, pw, us, h, w, sx, NULL);
                NAME(orc_spline_upsample)(Vacc, ph, pw, vs, h, w, sy, NULL);
            } else {
                memcpy(us, Uacc, sizeof(REAL) * N);
                memcpy(vs, Vacc, sizeof(REAL) * N);
            }
            NAME(orc_warp_symmetric)(n1, n2, us, vs, h, w, w1, w2);
            free(Uacc); free(Vacc);
            Uacc = us; Vacc = vs;
        } else {
            memcpy(w1, n1, sizeof(REAL) * N);
            memcpy(w2, n2, sizeof(REAL) * N);
            Uacc = (REAL*)calloc(N, sizeof(REAL));
            Vacc = (REAL*)calloc(N, sizeof(REAL));
        }
        if (nf > 0) {
            NAME(orc_gauss_sep)(w1, h, w, filt_w, nf);
            NAME(orc_gauss_sep)(w2, h, w, filt_w, nf);
        }
        double hs_err = NAME(orc_hs)(w1, w2, h, w, alphas[level - 1], niter, U, V);
        double ls_err = 0.0;
        int ls_it = 0;
        if (ls_alpha > 0) {
            if (no > 0) {
                NAME(orc_gauss_sep)(n1, h, w, opt_w, no);
                NAME(orc_gauss_sep)(n2, h, w, opt_w, no);
            }
            /* adapter swap: solver u = caller V, solver v = caller U */
            ls_err = NAME(orc_ls)(n1, n2, h, w, ls_alpha, ls_maxnum, ls_tol, V, U, &ls_it);
        }
        for (size_t q = 0; q < N; ++q) {
            Uacc[q] = Uacc[q] + U[q];
            Vacc[q] = Vacc[q] + V[q];
        }
        if (stats) {
            stats[3 * (level - 1) + 0] = hs_err;
            stats[3 * (level - 1) + 1] = ls_err;
            stats[3 * (level - 1) + 2] = ls_it;
        }
        free(n1); free(n2); free(w1); free(w2); free(U); free(V);
        ph = h; pw = w;
        scale *= 2;
    }
    memcpy(Uacc_out, Uacc, sizeof(REAL) * (size_t)H * W);
    memcpy(Vacc_out, Vacc, sizeof(REAL) * (size_t)H * W);
    free(Uacc); free(Vacc);
}


/* ------------------------------------------------------------------------ */
/* Farneback (Farneback_PyCL.py:123-603, optical_flow_farneback.cl:72-429). */
/* Float arithmetic as the OpenCL source is written, with the `acc +=      */
/* a*b` accumulations fused (OpenCL contracts by default): FMA(a, b, c).    */
/* Planes of the 5-plane buffers are stored one after another (5H x W).    */
/* ------------------------------------------------------------------------ */
static inline REAL NAME(fma_)(REAL a, REAL b, REAL c) {
    return sizeof(REAL) == 4 ? (REAL)fmaf((float)a, (float)b, (float)c) : (REAL)fma(a, b, c);
}
/* reflect-101 of the OpenCL helpers idx_row_low / idx_row_high / idx_col (.cl:135-158) */
static inline int NAME(r101_low)(int y, int last) { return abs(y) % (last + 1); }
static inline int NAME(r101_high)(int y, int last) { return abs(last - abs(last - y)) % (last + 1); }
static inline int NAME(r101_col)(int x, int last) {
    return NAME(r101_low)(NAME(r101_high)(x, last), last);
}

/* gaussianBlur over `planes` stacked planes (.cl:160-253): vertical pass    */
/* v = s*g0, v = fma(s[y-j] + s[y+j], g[j], v); then horizontal pass on the */
/* vertical results with reflect-101 columns.  gk: kh+1 half-kernel taps.   */
void NAME(orc_fb_blur)(const REAL* src, REAL* dst, int H, int W, int planes, const REAL* gk, int kh) {
    REAL* t = (REAL*)malloc(sizeof(REAL) * (size_t)H * W);
    for (int k = 0; k < planes; ++k) {
        const REAL* s = src + (size_t)k * H * W;
        REAL* d = dst + (size_t)k * H * W;
#pragma omp parallel for schedule(static)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                REAL v = s[IDX(y, x)] * gk[0];
                for (int j = 1; j <= kh; ++j) {
                    REAL a = s[IDX(NAME(r101_low)(y - j, H - 1), x)] + s[IDX(NAME(r101_high)(y + j, H - 1), x)];
                    v = NAME(fma_)(a, gk[j], v);
                }
                t[IDX(y, x)] = v;
            }
#pragma omp parallel for schedule(static)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                REAL r = t[IDX(y, x)] * gk[0];
                for (int i = 1; i <= kh; ++i) {
                    REAL a = t[IDX(y, NAME(r101_col)(x - i, W - 1))] + t[IDX(y, NAME(r101_col)(x + i, W - 1))];
                    r = NAME(fma_)(a, gk[i], r);
                }
                d[IDX(y, x)] = r;
            }
    }
    free(t);
}

/* polynomialExpansion (.cl:72-133), clamped borders.  g/xg/xxg: n+1 taps,  */
/* ig[4] = ig11, ig03, ig33, ig55.  dst: 5 planes.                           */
void NAME(orc_fb_polyexp)(const REAL* src, int H, int W, const REAL* g, const REAL* xg,
                          const REAL* xxg, int n, const REAL* ig, REAL* dst) {
    REAL* r0 = (REAL*)malloc(sizeof(REAL) * (size_t)H * W * 3);
    REAL* r1 = r0 + (size_t)H * W;
    REAL* r2 = r1 + (size_t)H * W;
    size_t N = (size_t)H * W;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            REAL a = src[IDX(y, x)] * g[0], b = 0, c = 0;
            for (int k = 1; k <= n; ++k) {
                REAL t0 = src[IDX(y - k < 0 ? 0 : y - k, x)];
                REAL t1 = src[IDX(y + k > H - 1 ? H - 1 : y + k, x)];
                a = NAME(fma_)(g[k], t0 + t1, a);
                b = NAME(fma_)(xg[k], t1 - t0, b);
                c = NAME(fma_)(xxg[k], t0 + t1, c);
            }
            r0[IDX(y, x)] = a; r1[IDX(y, x)] = b; r2[IDX(y, x)] = c;
        }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            REAL b1 = g[0] * r0[IDX(y, x)], b3 = g[0] * r1[IDX(y, x)], b5 = g[0] * r2[IDX(y, x)];
            REAL b2 = 0, b4 = 0, b6 = 0;
            for (int k = 1; k <= n; ++k) {
                int xp = x + k > W - 1 ? W - 1 : x + k, xm = x - k < 0 ? 0 : x - k;
                REAL p0 = r0[IDX(y, xp)], m0 = r0[IDX(y, xm)];
                REAL p1 = r1[IDX(y, xp)], m1 = r1[IDX(y, xm)];
                REAL p2 = r2[IDX(y, xp)], m2 = r2[IDX(y, xm)];
                b1 = NAME(fma_)(p0 + m0, g[k], b1);
                b4 = NAME(fma_)(p0 + m0, xxg[k], b4);
                b2 = NAME(fma_)(p0 - m0, xg[k], b2);
                b3 = NAME(fma_)(p1 + m1, g[k], b3);
                b6 = NAME(fma_)(p1 - m1, xg[k], b6);
                b5 = NAME(fma_)(p2 + m2, g[k], b5);
            }
            size_t q = IDX(y, x);
            dst[q] = b3 * ig[0];
            dst[N + q] = b2 * ig[0];
            dst[2 * N + q] = NAME(fma_)(b1, ig[1], b5 * ig[2]);
            dst[3 * N + q] = NAME(fma_)(b1, ig[1], b4 * ig[2]);
            dst[4 * N + q] = b6 * ig[3];
        }
    free(r0);
}

/* updateMatrices (.cl:256-348). */
void NAME(orc_fb_update_matrices)(const REAL* fxw, const REAL* fyw, const REAL* R0, const REAL* R1,
                                  int H, int W, REAL* M) {
    static const float border[6] = {0.14f, 0.14f, 0.4472f, 0.4472f, 0.4472f, 1.f};
    size_t N = (size_t)H * W;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            size_t q = IDX(y, x);
            REAL dx = fxw[q], dy = fyw[q];
            REAL fx = (REAL)x + dx, fy = (REAL)y + dy;
            REAL flx = floor(fx), fly = floor(fy);
            int x1 = (int)flx, y1 = (int)fly;
            fx -= (REAL)x1;
            fy -= (REAL)y1;
            REAL r2, r3, r4, r5, r6;
            if (x1 >= 0 && y1 >= 0 && x1 < W - 1 && y1 < H - 1) {
                REAL a00 = (1 - fx) * (1 - fy), a01 = fx * (1 - fy), a10 = (1 - fx) * fy, a11 = fx * fy;
                REAL r[5];
                for (int k = 0; k < 5; ++k) {
                    const REAL* P = R1 + (size_t)k * N;
                    REAL s = a00 * P[IDX(y1, x1)];
                    s = NAME(fma_)(a01, P[IDX(y1, x1 + 1)], s);
                    s = NAME(fma_)(a10, P[IDX(y1 + 1, x1)], s);
                    s = NAME(fma_)(a11, P[IDX(y1 + 1, x1 + 1)], s);
                    r[k] = s;
                }
                r2 = r[0]; r3 = r[1];
                r4 = (R0[2 * N + q] + r[2]) * (REAL)0.5;
                r5 = (R0[3 * N + q] + r[3]) * (REAL)0.5;
                r6 = (R0[4 * N + q] + r[4]) * (REAL)0.25;
            } else {
                r2 = r3 = 0;
                r4 = R0[2 * N + q];
                r5 = R0[3 * N + q];
                r6 = R0[4 * N + q] * (REAL)0.5;
            }
            r2 = (R0[q] - r2) * (REAL)0.5;
            r3 = (R0[N + q] - r3) * (REAL)0.5;
            r2 = r2 + NAME(fma_)(r4, dy, r6 * dx);
            r3 = r3 + NAME(fma_)(r6, dy, r5 * dx);
            int bx0 = x < 5 ? x : 5, by0 = y < 5 ? y : 5;
            int bx1 = W - x - 1 < 5 ? W - x - 1 : 5, by1 = H - y - 1 < 5 ? H - y - 1 : 5;
            REAL sc = (REAL)border[bx0] * (REAL)border[by0];
            sc = sc * (REAL)border[bx1];
            sc = sc * (REAL)border[by1];
            r2 *= sc; r3 *= sc; r4 *= sc; r5 *= sc; r6 *= sc;
            M[q] = NAME(fma_)(r4, r4, r6 * r6);
            M[N + q] = (r4 + r5) * r6;
            M[2 * N + q] = NAME(fma_)(r5, r5, r6 * r6);
            M[3 * N + q] = NAME(fma_)(r4, r2, r6 * r3);
            M[4 * N + q] = NAME(fma_)(r6, r2, r5 * r3);
        }
}

/* updateFlow (.cl:408-429). */
void NAME(orc_fb_update_flow)(const REAL* M, int H, int W, REAL* fxw, REAL* fyw) {
    size_t N = (size_t)H * W;
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; ++q) {
        REAL g11 = M[q], g12 = M[N + q], g22 = M[2 * N + q], h1 = M[3 * N + q], h2 = M[4 * N + q];
        REAL det = NAME(fma_)(g11, g22, -(g12 * g12)) + (REAL)1e-3f;
        REAL inv = (REAL)1 / det;
        fxw[q] = NAME(fma_)(g11, h2, -(g12 * h1)) * inv;
        fyw[q] = NAME(fma_)(g22, h1, -(g12 * h2)) * inv;
    }
}

/* boxFilter5 over `planes` stacked planes (Farneback_PyCL.py:273-308 with  */
/* .cl:350-406): replicate borders; vertical v = s[y], v += (s[y-j] +       */
/* s[y+j]) for j = 1..kh; the same horizontally on the vertical results;    */
/* times 1 / (2 kh + 1)^2 formed in REAL.                                  */
void NAME(orc_fb_box)(const REAL* src, REAL* dst, int H, int W, int planes, int kh) {
    REAL* t = (REAL*)malloc(sizeof(REAL) * (size_t)H * W);
    const REAL inv = (REAL)1 / (REAL)((1 + 2 * kh) * (1 + 2 * kh));
    for (int k = 0; k < planes; ++k) {
        const REAL* s = src + (size_t)k * H * W;
        REAL* d = dst + (size_t)k * H * W;
#pragma omp parallel for schedule(static)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                REAL v = s[IDX(y, x)];
                for (int j = 1; j <= kh; ++j)
                    v = v + (s[IDX(NAME(clp)(y - j, H), x)] + s[IDX(NAME(clp)(y + j, H), x)]);
                t[IDX(y, x)] = v;
            }
#pragma omp parallel for schedule(static)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                REAL r = t[IDX(y, x)];
                for (int i = 1; i <= kh; ++i)
                    r = r + (t[IDX(y, NAME(clp)(x - i, W))] + t[IDX(y, NAME(clp)(x + i, W))]);
                d[IDX(y, x)] = r * inv;
            }
    }
    free(t);
}

/* One Farneback compute() with pyramidalLevels = 1 (Farneback_PyCL.py:461-603):
 * pre-blur both frames with pk (pkh+1 half taps), polynomial expansion,
 * updateMatrices from the initial flow, then `iters` x [window blur (wk,
 * wkh+1 half taps; wk NULL: box window of half size wkh, the useGaussian =
 * False path of Farneback_PyCL.py:424-439) -> updateFlow -> updateMatrices
 * (all but the last)].  fxw / fyw: initial flow in, result out. */
void NAME(orc_farneback)(const REAL* im1, const REAL* im2, int H, int W, const REAL* pk, int pkh,
                         const REAL* g, const REAL* xg, const REAL* xxg, int n, const REAL* ig,
                         const REAL* wk, int wkh, int iters, REAL* fxw, REAL* fyw) {
    size_t N = (size_t)H * W;
    REAL* buf = (REAL*)malloc(sizeof(REAL) * N * 21);
    REAL *bl = buf, *RA = buf + N, *RB = RA + 5 * N, *M = RB + 5 * N, *Mb = M + 5 * N;
    NAME(orc_fb_blur)(im1, bl, H, W, 1, pk, pkh);
    NAME(orc_fb_polyexp)(bl, H, W, g, xg, xxg, n, ig, RA);
    NAME(orc_fb_blur)(im2, bl, H, W, 1, pk, pkh);
    NAME(orc_fb_polyexp)(bl, H, W, g, xg, xxg, n, ig, RB);
    NAME(orc_fb_update_matrices)(fxw, fyw, RA, RB, H, W, M);
    for (int i = 0; i < iters; ++i) {
        if (wk)
            NAME(orc_fb_blur)(M, Mb, H, W, 5, wk, wkh);
        else
            NAME(orc_fb_box)(M, Mb, H, W, 5, wkh);
        REAL* t = M; M = Mb; Mb = t;
        NAME(orc_fb_update_flow)(M, H, W, fxw, fyw);
        if (i < iters - 1) NAME(orc_fb_update_matrices)(fxw, fyw, RA, RB, H, W, M);
    }
    free(buf);
}

/* ------------------------------------------------------------------------ */
/* Dense pyramidal-LK, one level (denseLucasKanade_PyCL.py:113-169 with     */
/* pyrlkDenseLargeW.cl:304-669).  The window is the kernel's 32 x 32 grid  */
/* (4 tiles of 8 work-items per axis) with 0/1 weights wym[r] * wxm[c]; a   */
/* zero weight removes the sample from A and b alike (.cl:247-248,260).     */
/* For output pixel (i, j): window pixels (i-hw+r, j-hw+c); the patch I and */
/* its Scharr derivatives come from clamp-to-edge reads (the LINEAR sampler */
/* at pixel centres), Dx = 3(I[-1,+1]+I[+1,+1]-I[-1,-1]-I[+1,-1]) +         */
/* 10(I[0,+1]-I[0,-1]), Dy likewise; A = sum of the derivative products;   */
/* D = A11 A22 - A12^2, stop (status 0, flow unchanged) if D < FLT_EPSILON; */
/* up to niter Gauss-Newton steps with J sampled bilinearly (OpenCL LINEAR, */
/* clamp-to-edge) at the displaced window, delta = 32 A^-1 b, early exit    */
/* when |delta| < 0.01 in both components.  Window sums in double.          */
/* ------------------------------------------------------------------------ */
void NAME(orc_lk_dense)(const REAL* I, const REAL* J, int H, int W, int hw, int niter, REAL* u, REAL* v,
                        unsigned char* status, const unsigned char* wxm, const unsigned char* wym) {
    const int win = 32;   /* the window grid; wxm / wym: 0/1 weight of each grid column / row */
#pragma omp parallel
    {
        REAL* P = (REAL*)malloc(sizeof(REAL) * (size_t)win * win * 3);
        REAL *Dx = P + (size_t)win * win, *Dy = Dx + (size_t)win * win;
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < H; ++i)
            for (int j = 0; j < W; ++j) {
                size_t g = IDX(i, j);
                double a11 = 0, a12 = 0, a22 = 0;
                for (int r = 0; r < win; ++r)
                    for (int c = 0; c < win; ++c) {
                        if (!wxm[c] || !wym[r]) continue;
                        int y = i - hw + r, x = j - hw + c;
                        int ym = NAME(clp)(y - 1, H), y0 = NAME(clp)(y, H), yp = NAME(clp)(y + 1, H);
                        int xm = NAME(clp)(x - 1, W), x0 = NAME(clp)(x, W), xp = NAME(clp)(x + 1, W);
                        REAL dx = (REAL)3 * (I[IDX(ym, xp)] + I[IDX(yp, xp)] - I[IDX(ym, xm)] - I[IDX(yp, xm)]) +
                                  (REAL)10 * (I[IDX(y0, xp)] - I[IDX(y0, xm)]);
                        REAL dy = (REAL)3 * (I[IDX(yp, xm)] + I[IDX(yp, xp)] - I[IDX(ym, xm)] - I[IDX(ym, xp)]) +
                                  (REAL)10 * (I[IDX(yp, x0)] - I[IDX(ym, x0)]);
                        size_t k = (size_t)r * win + c;
                        P[k] = I[IDX(y0, x0)];
                        Dx[k] = dx;
                        Dy[k] = dy;
                        a11 += (double)dx * dx;
                        a12 += (double)dx * dy;
                        a22 += (double)dy * dy;
                    }
                REAL A11 = (REAL)a11, A12 = (REAL)a12, A22 = (REAL)a22;
                REAL D = A11 * A22 - A12 * A12;
                if (status) status[g] = 1;
                if (D < (REAL)1.192092896e-07f) {
                    if (status) status[g] = 0;
                    continue;
                }
                A11 /= D; A12 /= D; A22 /= D;
                REAL px = (REAL)j + u[g] - (REAL)hw, py = (REAL)i + v[g] - (REAL)hw;
                for (int k = 0; k < niter; ++k) {
                    if (px < (REAL)-hw || px >= (REAL)W || py < (REAL)-hw || py >= (REAL)H) {
                        if (status) status[g] = 0;
                        break;
                    }
                    /* bilinear weights are those of the window origin: sample
                     * (r, c) sits at integer offsets from it */
                    REAL flx = floor(px), fly = floor(py);
                    REAL fa = px - flx, fb = py - fly;
                    int x0 = (int)flx, y0 = (int)fly;
                    REAL w00 = (1 - fa) * (1 - fb), w10 = fa * (1 - fb), w01 = (1 - fa) * fb, w11 = fa * fb;
                    double b1 = 0, b2 = 0;
                    for (int r = 0; r < win; ++r) {
                        if (!wym[r]) continue;
                        int ya = NAME(clp)(y0 + r, H), yb = NAME(clp)(y0 + r + 1, H);
                        for (int c = 0; c < win; ++c) {
                            if (!wxm[c]) continue;
                            size_t q = (size_t)r * win + c;
                            int xa = NAME(clp)(x0 + c, W), xb = NAME(clp)(x0 + c + 1, W);
                            /* the four weighted samples accumulated with fused
                             * multiply-adds, as the kernel does */
                            REAL t = w00 * J[IDX(ya, xa)];
                            t = NAME(fma_)(w10, J[IDX(ya, xb)], t);
                            t = NAME(fma_)(w01, J[IDX(yb, xa)], t);
                            t = NAME(fma_)(w11, J[IDX(yb, xb)], t);
                            REAL diff = t - P[q];
                            b1 += (double)diff * Dx[q];
                            b2 += (double)diff * Dy[q];
                        }
                    }
                    REAL B1 = (REAL)b1, B2 = (REAL)b2;
                    REAL ddx = (A12 * B2 - A22 * B1) * (REAL)32;
                    REAL ddy = (A12 * B1 - A11 * B2) * (REAL)32;
                    px += ddx;
                    py += ddy;
                    if (fabs((double)ddx) < 0.01 && fabs((double)ddy) < 0.01) break;
                }
                u[g] = px + (REAL)hw - (REAL)j;
                v[g] = py + (REAL)hw - (REAL)i;
            }
        free(P);
    }
}

#undef IDX
