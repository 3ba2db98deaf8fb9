/*
 * aldic_oracle.c -- FP64 CPU restatement of the AL-DIC correlation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: the tests, the smoke check
 * and the CPU-baseline leg of the benchmark load it as the reference the HIP library
 * is compared against.  Nothing in the product links or calls it.
 *
 * Each function restates one reference function (MATLAB, jyang526843/2D_ALDIC) and
 * cites the file:line it follows.  Conventions shared with the reference:
 *   - images are row-major [y][x] (x fastest) == the reference's column-major image
 *     after its load-time transpose (dim 1 = x, ReadImage.m:79);
 *   - pixel coordinates are 1-based: pixel (x, y) lives at img[(y-1)*nx + (x-1)];
 *   - nodal vectors are interleaved: U = [u1 v1 u2 v2 ...], F = [F11 F21 F12 F22 ...].
 * Per-subset loops run under OpenMP, the analogue of the reference's parfor.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   orc_normalize     closed form (mean 0 / std 1 over the ROI)
 *   orc_gradient      exact on polynomials of degree <= 6 (quadratic image test)
 *   orc_bicubic       Keys identities: integer nodes, linear/quadratic reproduction, clamp
 *   orc_ncc_*         brute-force definition on tiny inputs, NCC==1 at the embedded lag,
 *                     integer translations give integer peaks, value range [-1, 1]
 *   orc_findpeak      paraboloid recovery (closed form), edge rule, 1e-3 rounding
 *   orc_icgn6/2       zero-deformation convergence, analytic translation/affine recovery,
 *                     H symmetric PSD, compose == 3x3 matrix product
 *   orc_subpb2_*      patch test, dense re-assembly, scipy spsolve on small meshes
 *   orc_nodal_strain  linear field -> exact constant gradient in the swapped slot order
 *   PPE histogram     parity unpinned (MATLAB histcounts bin edges are not visible here)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define IDX(x, y, nx) ((size_t)((y) - 1) * (size_t)(nx) + (size_t)((x) - 1))

/* MATLAB round(): half away from zero. */
static double mround(double a) { return a < 0 ? -floor(-a + 0.5) : floor(a + 0.5); }

/* ------------------------------------------------------------------------------
 * a1  funNormalizeImg.m:22-40 -- (I - mean_ROI) / std_ROI, std with N-1, ROI
 *     clipped to the image.  roi = {x1, x2, y1, y2}, 1-based inclusive.
 * ---------------------------------------------------------------------------- */
void orc_normalize(const double *img, int nx, int ny, const int *roi, double *out,
                   double *mean_std)
{
    int x1 = roi[0] < 1 ? 1 : roi[0], x2 = roi[1] > nx ? nx : roi[1];
    int y1 = roi[2] < 1 ? 1 : roi[2], y2 = roi[3] > ny ? ny : roi[3];
    double s = 0.0;
    long n = (long)(x2 - x1 + 1) * (y2 - y1 + 1);
    for (int y = y1; y <= y2; ++y)
        for (int x = x1; x <= x2; ++x) s += img[IDX(x, y, nx)];
    double mean = s / (double)n, ss = 0.0;
    for (int y = y1; y <= y2; ++y)
        for (int x = x1; x <= x2; ++x) {
            double d = img[IDX(x, y, nx)] - mean;
            ss += d * d;
        }
    double sd = sqrt(ss / (double)(n - 1));
    for (size_t i = 0; i < (size_t)nx * ny; ++i) out[i] = (img[i] - mean) / sd;
    if (mean_std) { mean_std[0] = mean; mean_std[1] = sd; }
}

/* ------------------------------------------------------------------------------
 * a2  funImgGradient.m:53-78 -- correlation with [-1/60 3/20 -3/4 0 3/4 -3/20 1/60]
 *     along x (DfDx) and y (DfDy); valid for 4 <= x <= nx-3 and 4 <= y <= ny-3 (the
 *     reference crops 3 px, DfCropWidth = 3), times the reference-image mask.  Stored
 *     full-size with zeros outside the valid band: gx(x, y) here == DfDx(x-3, y-3).
 *     The symmetric difference form makes flat regions exact zeros (the mask branch
 *     of funICGN.m:54 counts exact zeros).
 * ---------------------------------------------------------------------------- */
void orc_gradient(const double *f, int nx, int ny, const uint8_t *mask, double *gx, double *gy)
{
    memset(gx, 0, sizeof(double) * (size_t)nx * ny);
    memset(gy, 0, sizeof(double) * (size_t)nx * ny);
#pragma omp parallel for schedule(static)
    for (int y = 4; y <= ny - 3; ++y)
        for (int x = 4; x <= nx - 3; ++x) {
            double dx = 0.75 * (f[IDX(x + 1, y, nx)] - f[IDX(x - 1, y, nx)])
                      - 0.15 * (f[IDX(x + 2, y, nx)] - f[IDX(x - 2, y, nx)])
                      + (1.0 / 60.0) * (f[IDX(x + 3, y, nx)] - f[IDX(x - 3, y, nx)]);
            double dy = 0.75 * (f[IDX(x, y + 1, nx)] - f[IDX(x, y - 1, nx)])
                      - 0.15 * (f[IDX(x, y + 2, nx)] - f[IDX(x, y - 2, nx)])
                      + (1.0 / 60.0) * (f[IDX(x, y + 3, nx)] - f[IDX(x, y - 3, nx)]);
            double m = mask ? (double)mask[IDX(x, y, nx)] : 1.0;
            gx[IDX(x, y, nx)] = dx * m;
            gy[IDX(x, y, nx)] = dy * m;
        }
}

/* ------------------------------------------------------------------------------
 * a11 ba_interp2.cpp (semantics of :60-66 access(), :68-130 indices_cubic(),
 *     :153-162 weights) -- Keys cubic convolution, a = -1/2, taps at floor-1 ..
 *     floor+2, clamp-to-edge.  Written from the Keys (1981) kernel.  x is the
 *     dim-1 (horizontal) coordinate, y the dim-2 one, both 1-based; funICGN.m:125
 *     calls ba_interp2(ImgDef, v22, u22) which samples at (x=u22, y=v22).
 * ---------------------------------------------------------------------------- */
static inline void keys_w(double d, double w[4])
{
    double d2 = d * d, d3 = d2 * d;
    w[0] = 0.5 * (-d + 2.0 * d2 - d3);
    w[1] = 0.5 * (2.0 - 5.0 * d2 + 3.0 * d3);
    w[2] = 0.5 * (d + 4.0 * d2 - 3.0 * d3);
    w[3] = 0.5 * (-d2 + d3);
}

static inline int clampi(int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); }

static inline double bicubic1(const double *img, int nx, int ny, double x, double y)
{
    double xf = floor(x), yf = floor(y);
    double wx[4], wy[4];
    keys_w(x - xf, wx);
    keys_w(y - yf, wy);
    int ix = (int)xf - 2, iy = (int)yf - 2; /* 0-based index of the first tap */
    double acc = 0.0;
    for (int a = 0; a < 4; ++a) {
        int cx = clampi(ix + a, 0, nx - 1);
        double col = 0.0;
        for (int b = 0; b < 4; ++b) {
            int cy = clampi(iy + b, 0, ny - 1);
            col += wy[b] * img[(size_t)cy * nx + cx];
        }
        acc += wx[a] * col;
    }
    return acc;
}

void orc_bicubic(const double *img, int nx, int ny, const double *X, const double *Y, long n,
                 double *Z)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) Z[i] = bicubic1(img, nx, ny, X[i], Y[i]);
}

/* ------------------------------------------------------------------------------
 * Small dense helpers.
 * ---------------------------------------------------------------------------- */

/* Solve A x = b for symmetric A (n <= 6) by Gaussian elimination with partial
 * pivoting (MATLAB's backslash on a dense square matrix; equal to its Cholesky path
 * to rounding).  Returns 0 on success, -1 if singular. */
static int dense_solve(int n, const double *A_in, const double *b_in, double *x)
{
    double A[36], b[6];
    memcpy(A, A_in, sizeof(double) * n * n);
    memcpy(b, b_in, sizeof(double) * n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = fabs(A[k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(A[i * n + k]) > best) { best = fabs(A[i * n + k]); p = i; }
        if (best == 0.0) return -1;
        if (p != k) {
            for (int j = 0; j < n; ++j) { double t = A[k * n + j]; A[k * n + j] = A[p * n + j]; A[p * n + j] = t; }
            double t = b[k]; b[k] = b[p]; b[p] = t;
        }
        for (int i = k + 1; i < n; ++i) {
            double m = A[i * n + k] / A[k * n + k];
            for (int j = k; j < n; ++j) A[i * n + j] -= m * A[k * n + j];
            b[i] -= m * b[k];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < n; ++j) s -= A[i * n + j] * x[j];
        x[i] = s / A[i * n + i];
    }
    return 0;
}

/* Inverse-compositional update, funICGN.m:268-286: P <- P o dP^-1 with
 * P = [u_x v_x u_y v_y u v].  Returns 0, or -1 when det(dP) == 0 (:269, :287-289). */
static int ic_compose(double P[6], const double D[6])
{
    double det = (1.0 + D[0]) * (1.0 + D[3]) - D[1] * D[2];
    if (det == 0.0) return -1;
    double t1 = (-D[0] - D[0] * D[3] + D[1] * D[2]) / det;
    double t2 = -D[1] / det;
    double t3 = -D[2] / det;
    double t4 = (-D[3] - D[0] * D[3] + D[1] * D[2]) / det;
    double t5 = (-D[4] - D[3] * D[4] + D[2] * D[5]) / det;
    double t6 = (-D[5] - D[0] * D[5] + D[1] * D[4]) / det;
    /* [1+P1 P3 P5; P2 1+P4 P6; 0 0 1] * [1+t1 t3 t5; t2 1+t4 t6; 0 0 1] */
    double a11 = 1.0 + P[0], a12 = P[2], a13 = P[4];
    double a21 = P[1], a22 = 1.0 + P[3], a23 = P[5];
    double m11 = a11 * (1.0 + t1) + a12 * t2;
    double m12 = a11 * t3 + a12 * (1.0 + t4);
    double m13 = a11 * t5 + a12 * t6 + a13;
    double m21 = a21 * (1.0 + t1) + a22 * t2;
    double m22 = a21 * t3 + a22 * (1.0 + t4);
    double m23 = a21 * t5 + a22 * t6 + a23;
    P[0] = m11 - 1.0; P[1] = m21; P[2] = m12; P[3] = m22 - 1.0; P[4] = m13; P[5] = m23;
    return 0;
}

void orc_ic_compose(double *P, const double *D, int *ok) { *ok = ic_compose(P, D); }

/* Lower triangle of a symmetric 6x6 in MATLAB column-major order, funICGN.m:320. */
static void h_pack21(const double H[36], double *h21)
{
    int k = 0;
    for (int c = 0; c < 6; ++c)
        for (int r = c; r < 6; ++r) h21[k++] = H[r * 6 + c];
}
static void h_unpack21(const double *h21, double H[36])
{
    int k = 0;
    for (int c = 0; c < 6; ++c)
        for (int r = c; r < 6; ++r) { H[r * 6 + c] = h21[k]; H[c * 6 + r] = h21[k]; ++k; }
}

/* Window geometry with the reference's indexing-error cases (caught as status -1 by
 * LocalICGN.m:101-106 / Subpb1.m:139-143).  Returns 0 and the 1-based corner on
 * success, -1 when x0 -+ w/2 is not an integer or the window (or its gradient,
 * which is cropped by 3 px) leaves the image. */
static int window_corner(double x0, double y0, int w, int nx, int ny, int *xa, int *ya)
{
    double xs = x0 - 0.5 * w, ys = y0 - 0.5 * w;
    if (xs != floor(xs) || ys != floor(ys)) return -1;
    int a = (int)xs, b = (int)ys;
    if (a < 4 || b < 4 || a + w > nx - 3 || b + w > ny - 3) return -1;
    *xa = a; *ya = b;
    return 0;
}

typedef struct {
    int w, n, xa, ya;       /* window size, pixel count, 1-based corner */
    double meanf, bottomf;
} window_t;

/* Window set-up plus the mask rule of funICGN.m:53-65 (n0 = exact zeros of DfDx):
 * returns 0 (ok), 102 (>= 50% masked) or -1 (indexing error). */
static int window_setup(const double *gx, int nx, int ny, double x0, double y0, int w,
                        window_t *W)
{
    int xa, ya;
    if (window_corner(x0, y0, w, nx, ny, &xa, &ya)) return -1;
    long n0 = 0;
    for (int j = 0; j <= w; ++j)
        for (int i = 0; i <= w; ++i) n0 += gx[IDX(xa + i, ya + j, nx)] == 0.0;
    double np = (double)(w + 1) * (w + 1);
    if (!((double)n0 < 0.5 * np)) return 102;
    if ((double)n0 > 0.1 * np) {
        w = (int)mround(sqrt(2.0) * w);
        if (window_corner(x0, y0, w, nx, ny, &xa, &ya)) return -1;
    }
    W->w = w; W->n = (w + 1) * (w + 1); W->xa = xa; W->ya = ya;
    return 0;
}

static void window_fstats(const double *f, int nx, window_t *W)
{
    double s = 0.0;
    for (int j = 0; j <= W->w; ++j)
        for (int i = 0; i <= W->w; ++i) s += f[IDX(W->xa + i, W->ya + j, nx)];
    W->meanf = s / W->n;
    double ss = 0.0;
    for (int j = 0; j <= W->w; ++j)
        for (int i = 0; i <= W->w; ++i) {
            double d = f[IDX(W->xa + i, W->ya + j, nx)] - W->meanf;
            ss += d * d;
        }
    W->bottomf = sqrt(ss);
}

/* Hessian of funICGN.m:67-79: H = sum J^T J with J = [fx dx, fy dx, fx dy, fy dy, fx, fy]. */
static void window_hessian(const double *gx, const double *gy, int nx, double x0, double y0,
                           const window_t *W, double H[36])
{
    memset(H, 0, sizeof(double) * 36);
    for (int j = 0; j <= W->w; ++j)
        for (int i = 0; i <= W->w; ++i) {
            int x = W->xa + i, y = W->ya + j;
            double fx = gx[IDX(x, y, nx)], fy = gy[IDX(x, y, nx)];
            double dx = x - x0, dy = y - y0;
            double J[6] = {fx * dx, fy * dx, fx * dy, fy * dy, fx, fy};
            for (int r = 0; r < 6; ++r)
                for (int c = r; c < 6; ++c) H[r * 6 + c] += J[r] * J[c];
        }
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < r; ++c) H[r * 6 + c] = H[c * 6 + r];
}

/* One residual evaluation (funICGN.m:113-156, :237-246).  Returns 1 if a warped
 * sample leaves [3, nx-2] x [3, ny-2] (:118-121), else 0 with b = bottomf*sum J r
 * (or only b5, b6 when full == 0) and zncc = sum f_n g_n. */
static int residual(const double *f, const double *g, const double *gx, const double *gy,
                    int nx, int ny, double x0, double y0, const window_t *W, const double P[6],
                    int full, double *b, double *zncc)
{
    int n = W->n, w = W->w;
    double *tg = (double *)malloc(sizeof(double) * n);
    int k = 0;
    for (int j = 0; j <= w; ++j)
        for (int i = 0; i <= w; ++i, ++k) {
            double dx = (W->xa + i) - x0, dy = (W->ya + j) - y0;
            double u = (1.0 + P[0]) * dx + P[2] * dy + (x0 + P[4]);
            double v = P[1] * dx + (1.0 + P[3]) * dy + (y0 + P[5]);
            if (u < 3.0 || u > nx - 2.0 || v < 3.0 || v > ny - 2.0) { free(tg); return 1; }
            if (isnan(u) || isnan(v)) { free(tg); return 2; } /* NaN warp -> NaN norm */
            tg[k] = bicubic1(g, nx, ny, u, v);
        }
    double s = 0.0;
    for (k = 0; k < n; ++k) s += tg[k];
    double meang = s / n, ss = 0.0;
    for (k = 0; k < n; ++k) { double d = tg[k] - meang; ss += d * d; }
    double bottomg = sqrt(ss);
    double b2[6] = {0, 0, 0, 0, 0, 0}, zc = 0.0;
    k = 0;
    for (int j = 0; j <= w; ++j)
        for (int i = 0; i <= w; ++i, ++k) {
            int x = W->xa + i, y = W->ya + j;
            double fn = (f[IDX(x, y, nx)] - W->meanf) / W->bottomf;
            double gn = (tg[k] - meang) / bottomg;
            double r = fn - gn;
            double fx = gx[IDX(x, y, nx)], fy = gy[IDX(x, y, nx)];
            double dx = x - x0, dy = y - y0;
            if (full) {
                b2[0] += dx * fx * r; b2[1] += dx * fy * r;
                b2[2] += dy * fx * r; b2[3] += dy * fy * r;
            }
            b2[4] += fx * r; b2[5] += fy * r;
            zc += fn * gn;
        }
    for (int q = 0; q < 6; ++q) b[q] = W->bottomf * b2[q];
    *zncc = zc;
    free(tg);
    return 0;
}

/* ------------------------------------------------------------------------------
 * a10 funICGN.m:1-322 -- 6-DOF IC-GN (Gauss-Newton branch; the LM branch calls
 *     ImgDef.eval on a plain matrix and is dead code, :199).  One subset.
 *     Returns the status: iterations n (1..100), 101 (not converged / out of the
 *     image / NaN), 102 (masked), -1 (indexing error, caught upstream).
 * ---------------------------------------------------------------------------- */
static int icgn6_one(const double *f, const double *g, const double *gx, const double *gy,
                     int nx, int ny, double x0, double y0, const double U0[2], int w0,
                     double tol, double P[6], double h21[21], double *zncc)
{
    double H[36];
    window_t W;
    P[0] = P[1] = P[2] = P[3] = 0.0;
    P[4] = U0[0]; P[5] = U0[1];
    memset(h21, 0, sizeof(double) * 21);
    *zncc = NAN;
    int st = window_setup(gx, nx, ny, x0, y0, w0, &W);
    if (st == -1) return -1;
    if (st == 102) return 102;
    window_hessian(gx, gy, nx, x0, y0, &W, H);
    h_pack21(H, h21);
    window_fstats(f, nx, &W);

    double nrmNew = 1.0, nrmAbs = 1.0, nrmInit = 0.0, zc = NAN;
    int it = 0;
    while (it <= 100 && nrmNew > tol && nrmAbs > tol) {
        ++it;
        double b[6];
        int rs = residual(f, g, gx, gy, nx, ny, x0, y0, &W, P, 1, b, &zc);
        if (rs == 1) { nrmNew = 1e6; break; }
        if (rs == 2) { nrmNew = NAN; break; }
        nrmAbs = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2] + b[3] * b[3] + b[4] * b[4] + b[5] * b[5]);
        nrmNew = nrmAbs;
        if (it == 1) nrmInit = nrmNew;
        nrmNew = nrmInit > tol ? nrmNew / nrmInit : 0.0;
        if (nrmNew < tol || nrmAbs < tol) break;
        double D[6];
        if (dense_solve(6, H, b, D)) { nrmNew = NAN; break; }
        for (int q = 0; q < 6; ++q) D[q] = -D[q];
        if (ic_compose(P, D)) break;
    }
    int conv = (nrmNew < tol) || (nrmAbs < tol);
    if (!conv || isnan(nrmNew)) it = 101;
    *zncc = zc;
    return it;
}

/* Batched driver: the parfor of LocalICGN.m:91-107.  ws[j] is the subset size of
 * node j (the reference passes one scalar, LocalICGN.m:40; per-node sizes are the
 * winsize argument funICGN_Subpb1 already takes).  Outputs: U (2S), F (4S, order
 * F11 F21 F12 F22), H21 (S x 21), iters (S), zncc (S).  Status -1 -> NaN U/F. */
void orc_icgn6(const double *f, const double *g, const double *gx, const double *gy, int nx,
               int ny, int S, const double *coords, const double *U0, const int *ws, double tol,
               double *U, double *F, double *H21, int *iters, double *zncc)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int j = 0; j < S; ++j) {
        double P[6], z;
        int st = icgn6_one(f, g, gx, gy, nx, ny, coords[2 * j], coords[2 * j + 1], U0 + 2 * j,
                           ws[j], tol, P, H21 + 21 * (size_t)j, &z);
        iters[j] = st;
        zncc[j] = z;
        if (st == -1) {
            U[2 * j] = U[2 * j + 1] = NAN;
            for (int q = 0; q < 4; ++q) F[4 * j + q] = NAN;
        } else {
            U[2 * j] = P[4]; U[2 * j + 1] = P[5];
            F[4 * j] = P[0]; F[4 * j + 1] = P[1]; F[4 * j + 2] = P[2]; F[4 * j + 3] = P[3];
        }
    }
}

/* ------------------------------------------------------------------------------
 * a13 funICGN_Subpb1.m:1-362 -- 2-DOF IC-GN with F frozen and the mu-penalty
 *     toward (UOld + vdual).  H is the node's 6x6 Hessian from LocalICGN.  beta
 *     and udual are unused by the reference (comments only).
 * ---------------------------------------------------------------------------- */
static int icgn2_one(const double *f, const double *g, const double *gx, const double *gy,
                     int nx, int ny, double x0, double y0, int w0, const double h21[21],
                     double mu, const double vdual[2], const double Uold[2],
                     const double Fold[4], double tol, double Uout[2])
{
    double P[6] = {Fold[0], Fold[1], Fold[2], Fold[3], Uold[0], Uold[1]};
    window_t W;
    Uout[0] = P[4]; Uout[1] = P[5];
    int st = window_setup(gx, nx, ny, x0, y0, w0, &W);
    if (st == -1) return -1;
    if (st == 102) return 102;
    window_fstats(f, nx, &W);
    double H[36];
    h_unpack21(h21, H);
    double bf2 = W.bottomf * W.bottomf;
    /* funICGN_Subpb1.m:281-283 */
    double H2[4] = {H[4 * 6 + 4] * 2.0 / bf2 + mu, H[4 * 6 + 5] * 2.0 / bf2,
                    H[5 * 6 + 4] * 2.0 / bf2, H[5 * 6 + 5] * 2.0 / bf2 + mu};
    double dmax = H2[0] > H2[3] ? H2[0] : H2[3];
    double tH[4] = {H2[0] + 1e-3 * dmax, H2[1], H2[2], H2[3] + 1e-3 * dmax};

    double nrmNew = 1.0, nrmAbs = 1.0, nrmInit = 0.0, zc;
    int it = 0;
    while (it <= 100 && nrmNew > tol && nrmAbs > tol) {
        ++it;
        double b[6];
        int rs = residual(f, g, gx, gy, nx, ny, x0, y0, &W, P, 0, b, &zc);
        if (rs == 1) { nrmNew = 1e6; break; }
        if (rs == 2) { nrmNew = NAN; break; }
        /* :269 */
        double tb[2] = {b[4] * 2.0 / bf2 + mu * (P[4] - Uold[0] - vdual[0]),
                        b[5] * 2.0 / bf2 + mu * (P[5] - Uold[1] - vdual[1])};
        double D[6] = {0, 0, 0, 0, 0, 0}, d2[2];
        int sing = dense_solve(2, tH, tb, d2);
        /* :294-308 */
        nrmAbs = sqrt(tb[0] * tb[0] + tb[1] * tb[1]);
        nrmNew = nrmAbs;
        if (it == 1) nrmInit = nrmNew;
        nrmNew = nrmInit > tol ? nrmNew / nrmInit : 0.0;
        if (nrmNew < tol || nrmAbs < tol) break;
        if (sing) { nrmNew = NAN; break; }
        D[4] = -d2[0]; D[5] = -d2[1];
        if (ic_compose(P, D)) break;
    }
    int conv = (nrmNew < tol) || (nrmAbs < tol);
    if (!conv || isnan(nrmNew)) it = 101;
    Uout[0] = P[4]; Uout[1] = P[5];
    return it;
}

/* Batched driver: the parfor of Subpb1.m:119-152 (without the post-processing).
 * udual (4S) is accepted for signature parity and unused, like the reference. */
void orc_icgn2(const double *f, const double *g, const double *gx, const double *gy, int nx,
               int ny, int S, const double *coords, const int *ws, const double *H21, double mu,
               const double *udual, const double *vdual, const double *Uold, const double *Fold,
               double tol, double *U, int *iters)
{
    (void)udual;
#pragma omp parallel for schedule(dynamic, 16)
    for (int j = 0; j < S; ++j) {
        double Uo[2];
        int st = icgn2_one(f, g, gx, gy, nx, ny, coords[2 * j], coords[2 * j + 1], ws[j],
                           H21 + 21 * (size_t)j, mu, vdual + 2 * j, Uold + 2 * j, Fold + 4 * j,
                           tol, Uo);
        iters[j] = st;
        if (st == -1) { U[2 * j] = U[2 * j + 1] = NAN; }
        else { U[2 * j] = Uo[0]; U[2 * j + 1] = Uo[1]; }
    }
}

/* ------------------------------------------------------------------------------
 * a3/a4/a5/a6  Integer search.
 * ---------------------------------------------------------------------------- */

/* funIntegerSearch.m:14-25 (trim) and :90-91 (corner lists).  grid = {x1, x2, y1, y2}.
 * Writes the corner counts; corners are x1' + k*step for k < Nx. */
void orc_search_grid(int nx, int ny, const int *grid, int w, int step, const int *s, int *out)
{
    int gx1 = grid[0], gx2 = grid[1], gy1 = grid[2], gy2 = grid[3];
    while (gx1 - s[0] - 0.5 * w < 1) ++gx1;
    while (gy1 - s[1] - 0.5 * w < 1) ++gy1;
    while (gx2 + 1.5 * w + s[0] > nx) --gx2;
    while (gy2 + 1.5 * w + s[1] > ny) --gy2;
    int Nx = gx2 - step >= gx1 ? (gx2 - step - gx1) / step + 1 : 0;
    int Ny = gy2 - step >= gy1 ? (gy2 - step - gy1) / step + 1 : 0;
    out[0] = gx1; out[1] = gy1; out[2] = Nx; out[3] = Ny;
}

/* Full normxcorr2 map (Lewis 1995 normalised cross-correlation, the definition the
 * reference's normxcorr2 call at funIntegerSearch.m:127 computes): template T
 * ((w+1)^2) against window D (Dx x Dy), zero-padded, output (w+Dx) x (w+Dy) in
 * x-fastest order.  Value = sum (T - mean T) D / sqrt(sum (T-mean T)^2 *
 * max(sum D^2 - (sum D)^2 / n, 0)) over the template footprint, n = (w+1)^2;
 * 0 where the window variance is zero.  Returns -1 for a constant template (the
 * reference's normxcorr2 raises, :153-157). */
/* NCC at one output position (cx, cy) of the full map (template corner at D offset
 * (cx - tw + 1, cy - tw + 1)); mt, st = template mean and sum of squared deviations. */
static double ncc_lag(const double *T, int tw, double mt, double st, const double *D, int Dx,
                      int Dy, int cx, int cy)
{
    int n = tw * tw;
    int ox = cx - (tw - 1), oy = cy - (tw - 1); /* template corner in D (0-based) */
    double num = 0.0, sd = 0.0, sd2 = 0.0;
    for (int j = 0; j < tw; ++j) {
        int dy = oy + j;
        if (dy < 0 || dy >= Dy) continue;
        for (int i = 0; i < tw; ++i) {
            int dx = ox + i;
            if (dx < 0 || dx >= Dx) continue;
            double d = D[(size_t)dy * Dx + dx];
            num += (T[j * tw + i] - mt) * d;
            sd += d;
            sd2 += d * d;
        }
    }
    double va = sd2 - sd * sd / n;
    if (va < 0) va = 0;
    double den = sqrt(st * va);
    return den > 0 ? num / den : 0.0;
}

/* Template mean and sum of squared deviations; -1 for a constant template. */
static int tstats(const double *T, int tw, double *mt_out, double *st_out)
{
    int n = tw * tw;
    double mt = 0.0;
    for (int i = 0; i < n; ++i) mt += T[i];
    mt /= n;
    double st = 0.0;
    for (int i = 0; i < n; ++i) st += (T[i] - mt) * (T[i] - mt);
    int allsame = 1;
    for (int i = 1; i < n; ++i) if (T[i] != T[0]) { allsame = 0; break; }
    *mt_out = mt;
    *st_out = st;
    return allsame ? -1 : 0;
}

static int ncc_full(const double *T, int tw, const double *D, int Dx, int Dy, double *C)
{
    double mt, st;
    if (tstats(T, tw, &mt, &st)) return -1;
    int Cx = tw - 1 + Dx, Cy = tw - 1 + Dy;
    for (int cy = 0; cy < Cy; ++cy)
        for (int cx = 0; cx < Cx; ++cx) C[(size_t)cy * Cx + cx] = ncc_lag(T, tw, mt, st, D, Dx, Dy, cx, cy);
    return 0;
}

void orc_ncc_full(const double *T, int tw, const double *D, int Dx, int Dy, double *C, int *ok)
{
    *ok = ncc_full(T, tw, D, Dx, Dy, C);
}

/* findpeak.m:1-68 on an (nr x nc) map stored with dim 1 (rows) fastest, i.e.
 * f[(c-1)*nr + (r-1)].  Returns col (xpeak), row (ypeak) and the fitted max,
 * 1-based, following the reference's [xpeak, ypeak, max_f] output order. */
static void findpeak(const double *f, int nr, int nc, double *xpeak, double *ypeak, double *maxf)
{
    /* :27-28 first maximum in column-major order (NaN ignored) */
    long imax = -1;
    double m = -INFINITY;
    for (long i = 0; i < (long)nr * nc; ++i)
        if (f[i] > m) { m = f[i]; imax = i; }
    if (imax < 0) { *xpeak = *ypeak = *maxf = NAN; return; }
    int r = (int)(imax % nr) + 1, c = (int)(imax / nr) + 1;
    *xpeak = c; *ypeak = r; *maxf = m;
    if (c == 1 || c == nc || r == 1 || r == nr) return; /* :30-35 */
    /* :37-46 least squares of [1 x y xy x^2 y^2] on the 3x3 neighbourhood
     * (x = column offset, y = row offset); closed form of the 6x9 pseudo-inverse. */
    double s0 = 0, sxu = 0, syu = 0, sxy = 0, sxx = 0, syy = 0;
    for (int dc = -1; dc <= 1; ++dc)
        for (int dr = -1; dr <= 1; ++dr) {
            double u = f[(long)(c - 1 + dc) * nr + (r - 1 + dr)];
            s0 += u; sxu += dc * u; syu += dr * u; sxy += dc * dr * u;
            sxx += dc * dc * u; syy += dr * dr * u;
        }
    double A1 = (5.0 * s0 - 3.0 * sxx - 3.0 * syy) / 9.0;
    double A2 = sxu / 6.0, A3 = syu / 6.0, A4 = sxy / 4.0;
    double A5 = sxx / 2.0 - s0 / 3.0, A6 = syy / 2.0 - s0 / 3.0;
    double den = A4 * A4 - 4.0 * A5 * A6;
    double xo = (-A3 * A4 + 2.0 * A6 * A2) / den;      /* :49 */
    double yo = -1.0 / den * (A4 * A2 - 2.0 * A5 * A3); /* :50 */
    if (fabs(xo) > 1.0 || fabs(yo) > 1.0) return;      /* :52-55 */
    if (isnan(xo) || isnan(yo)) return;
    xo = mround(1000.0 * xo) / 1000.0;                 /* :58-59 */
    yo = mround(1000.0 * yo) / 1000.0;
    *xpeak = c + xo;
    *ypeak = r + yo;
    *maxf = A1 + A2 * xo + A3 * yo + A4 * xo * yo + A5 * xo * xo + A6 * yo * yo; /* :65 */
}

void orc_findpeak(const double *f, int nr, int nc, double *out3)
{
    findpeak(f, nr, nc, out3, out3 + 1, out3 + 2);
}

/* q-factors of funIntegerSearch.m:211-264 on the full map: PCE = peak^2 /
 * mean(cc_min^2), PPE = 1 / entropy of a 30-bin histogram of cc_min over
 * [0, max] (equal bins, last closed). */
static void qfactors(const double *C, long n, double *pce, double *ppe)
{
    double mn = INFINITY, mx = -INFINITY;
    for (long i = 0; i < n; ++i) { if (C[i] < mn) mn = C[i]; }
    double s2 = 0.0;
    for (long i = 0; i < n; ++i) {
        double c = C[i] - mn;
        if (c > mx) mx = c;
        s2 += c * c;
    }
    *pce = mx * mx / (s2 / n);
    long hist[30] = {0};
    for (long i = 0; i < n; ++i) {
        double c = C[i] - mn;
        int b = mx > 0 ? (int)floor(c / mx * 30.0) : 0;
        if (b >= 30) b = 29;
        if (b < 0) b = 0;
        hist[b]++;
    }
    double e = 0.0;
    for (int b = 0; b < 30; ++b)
        if (hist[b]) { double p = (double)hist[b] / n; e += p * log(1.0 / p); }
    *ppe = 1.0 / e;
}

/* funIntegerSearchWholeField (funIntegerSearch.m:82-200).  Corners ii = x1 + k*step
 * (k < Nx), jj = y1 + l*step (l < Ny); subset (k, l) is output entry l*Nx + k
 * (x fastest == the reference's (Ny x Nx) arrays read row by row).  Outputs: centre
 * x, y; u, v; cc.max; raw q-factors (pce, ppe) before the per-column min-max
 * normalisation of :174-177 (done by the caller); status 0 ok / -1 constant template. */
void orc_search(const double *f, const double *g, int nx, int ny, int x1, int y1, int Nx, int Ny,
                int w, int step, const int *s, double *x, double *y, double *u, double *v,
                double *ccmax, double *pce, double *ppe, int *status)
{
    int tw = w + 1, Dx = tw + 2 * s[0], Dy = tw + 2 * s[1];
    int Cx = tw - 1 + Dx, Cy = tw - 1 + Dy;
    int nr = 2 * s[0] + 3, nc = 2 * s[1] + 3;
#pragma omp parallel for schedule(dynamic, 4)
    for (int q = 0; q < Nx * Ny; ++q) {
        int k = q % Nx, l = q / Nx;
        int ii = x1 + k * step, jj = y1 + l * step;
        double *T = (double *)malloc(sizeof(double) * tw * tw);
        double *D = (double *)malloc(sizeof(double) * Dx * Dy);
        double *C = (double *)malloc(sizeof(double) * Cx * Cy);
        double *crop = (double *)malloc(sizeof(double) * nr * nc);
        for (int j = 0; j < tw; ++j)
            for (int i = 0; i < tw; ++i) T[j * tw + i] = f[IDX(ii + i, jj + j, nx)];
        for (int j = 0; j < Dy; ++j)
            for (int i = 0; i < Dx; ++i) D[(size_t)j * Dx + i] = g[IDX(ii - s[0] + i, jj - s[1] + j, nx)];
        x[q] = (ii + ii + w) / 2.0; /* :158-159 */
        y[q] = (jj + jj + w) / 2.0;
        if (ncc_full(T, tw, D, Dx, Dy, C)) {
            u[q] = v[q] = ccmax[q] = NAN;
            pce[q] = ppe[q] = INFINITY;
            status[q] = -1;
        } else {
            qfactors(C, (long)Cx * Cy, pce + q, ppe + q);
            /* crop [w : end-w+1] in both dims (:134): rows (x-lags) a = w-1 .. w+2s+1
             * (0-based), stored row-fastest for findpeak */
            for (int c = 0; c < nc; ++c)
                for (int r = 0; r < nr; ++r) crop[c * nr + r] = C[(size_t)(w - 1 + c) * Cx + (w - 1 + r)];
            double xp, yp, mf;
            findpeak(crop, nr, nc, &xp, &yp, &mf);
            /* [v1, u1, max] = findpeak(...) (:134); zero_disp = ceil(size/2) (:148) */
            u[q] = yp - (s[0] + 2);
            v[q] = xp - (s[1] + 2);
            ccmax[q] = mf;
            status[q] = 0;
        }
        free(T); free(D); free(C); free(crop);
    }
}

/* orc_search restricted to a list of subsets (entries q = l*Nx + k of the whole-field
 * grid) and to the cropped lags (funIntegerSearch.m:134), without q-factors: the same
 * per-lag NCC as the full map, for sampled accuracy checks and CPU timing on large
 * frames.  Outputs are per list entry. */
void orc_search_sample(const double *f, const double *g, int nx, int ny, int x1, int y1, int Nx,
                       int w, int step, const int *s, int nq, const int *qlist, double *x,
                       double *y, double *u, double *v, double *ccmax, int *status)
{
    int tw = w + 1, Dx = tw + 2 * s[0], Dy = tw + 2 * s[1];
    int nr = 2 * s[0] + 3, nc = 2 * s[1] + 3;
#pragma omp parallel for schedule(dynamic, 4)
    for (int t = 0; t < nq; ++t) {
        int q = qlist[t];
        int k = q % Nx, l = q / Nx;
        int ii = x1 + k * step, jj = y1 + l * step;
        double *T = (double *)malloc(sizeof(double) * tw * tw);
        double *D = (double *)malloc(sizeof(double) * Dx * Dy);
        double *crop = (double *)malloc(sizeof(double) * nr * nc);
        for (int j = 0; j < tw; ++j)
            for (int i = 0; i < tw; ++i) T[j * tw + i] = f[IDX(ii + i, jj + j, nx)];
        for (int j = 0; j < Dy; ++j)
            for (int i = 0; i < Dx; ++i) D[(size_t)j * Dx + i] = g[IDX(ii - s[0] + i, jj - s[1] + j, nx)];
        x[t] = (ii + ii + w) / 2.0;
        y[t] = (jj + jj + w) / 2.0;
        double mt, st;
        if (tstats(T, tw, &mt, &st)) {
            u[t] = v[t] = ccmax[t] = NAN;
            status[t] = -1;
        } else {
            for (int c = 0; c < nc; ++c)
                for (int r = 0; r < nr; ++r)
                    crop[c * nr + r] = ncc_lag(T, tw, mt, st, D, Dx, Dy, w - 1 + r, w - 1 + c);
            double xp, yp, mf;
            findpeak(crop, nr, nc, &xp, &yp, &mf);
            u[t] = yp - (s[0] + 2);
            v[t] = xp - (s[1] + 2);
            ccmax[t] = mf;
            status[t] = 0;
        }
        free(T); free(D); free(crop);
    }
    (void)ny;
}

/* ------------------------------------------------------------------------------
 * a14  Subpb2.m:1-197 -- Q4 FEM global step.  elems are 0-based node indices
 * (E x 4, counter-clockwise as MeshSetUp.m:35-38 builds them).  The element matrix
 * and vector follow :100-151 with 3x3 Gauss points; the system is assembled into a
 * dense-free COO list (the MATLAB sparse() triplets of :155-163) that the caller
 * sums; orc_subpb2_solve solves it with FP64 Jacobi-PCG to a relative residual of
 * 1e-14 (the reference uses a direct sparse solve, :195).
 * ---------------------------------------------------------------------------- */
static void q4_dN(double ksi, double eta, double dNk[4], double dNe[4])
{
    dNk[0] = -(1 - eta) / 4; dNe[0] = -(1 - ksi) / 4;
    dNk[1] = (1 - eta) / 4;  dNe[1] = -(1 + ksi) / 4;
    dNk[2] = (1 + eta) / 4;  dNe[2] = (1 + ksi) / 4;
    dNk[3] = -(1 + eta) / 4; dNe[3] = (1 - ksi) / 4;
}

/* Element system, Subpb2.m:76-151: Ae (8x8 row-major, DOF order u1 v1 u2 v2 ...),
 * be (8). */
static void subpb2_element(const double *coords, const int *el, double beta, double mu,
                           double alpha, const double *U, const double *F, const double *udual,
                           const double *vdual, double Ae[64], double be[8])
{
    double px[4], py[4];
    for (int a = 0; a < 4; ++a) { px[a] = coords[2 * el[a]]; py[a] = coords[2 * el[a] + 1]; }
    /* FMinusW rows: DOF k's node [F11 F12 F21 F22] - udual (:88-94) */
    double FW[8][4], UV[8];
    for (int a = 0; a < 4; ++a) {
        int nd = el[a];
        double g4[4] = {F[4 * nd] - udual[4 * nd], F[4 * nd + 2] - udual[4 * nd + 2],
                        F[4 * nd + 1] - udual[4 * nd + 1], F[4 * nd + 3] - udual[4 * nd + 3]};
        for (int c = 0; c < 4; ++c) { FW[2 * a][c] = g4[c]; FW[2 * a + 1][c] = g4[c]; }
        UV[2 * a] = U[2 * nd] - vdual[2 * nd];
        UV[2 * a + 1] = U[2 * nd + 1] - vdual[2 * nd + 1];
    }
    static const double gp[3] = {-0.7745966692414834, 0.0, 0.7745966692414834};
    static const double gw[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    memset(Ae, 0, sizeof(double) * 64);
    memset(be, 0, sizeof(double) * 8);
    for (int ti = 0; ti < 3; ++ti)
        for (int tj = 0; tj < 3; ++tj) {
            double ksi = gp[ti], eta = gp[tj], wgt = gw[ti] * gw[tj];
            double N[4] = {(1 - ksi) * (1 - eta) * 0.25, (1 + ksi) * (1 - eta) * 0.25,
                           (1 + ksi) * (1 + eta) * 0.25, (1 - ksi) * (1 + eta) * 0.25};
            double dNk[4], dNe[4];
            q4_dN(ksi, eta, dNk, dNe);
            double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
            for (int a = 0; a < 4; ++a) {
                J11 += dNk[a] * px[a]; J12 += dNk[a] * py[a];
                J21 += dNe[a] * px[a]; J22 += dNe[a] * py[a];
            }
            double det = J11 * J22 - J12 * J21;
            double i11 = J22 / det, i12 = -J12 / det, i21 = -J21 / det, i22 = J11 / det;
            /* DN (4x8): rows u_x, u_y, v_x, v_y (:136-140) */
            double DN[4][8];
            memset(DN, 0, sizeof(DN));
            for (int a = 0; a < 4; ++a) {
                double ddx = i11 * dNk[a] + i12 * dNe[a];
                double ddy = i21 * dNk[a] + i22 * dNe[a];
                DN[0][2 * a] = ddx; DN[1][2 * a] = ddy;
                DN[2][2 * a + 1] = ddx; DN[3][2 * a + 1] = ddy;
            }
            double NN[2][8];
            memset(NN, 0, sizeof(NN));
            for (int a = 0; a < 4; ++a) { NN[0][2 * a] = N[a]; NN[1][2 * a + 1] = N[a]; }
            double wJ = det * wgt;
            for (int r = 0; r < 8; ++r)
                for (int c = 0; c < 8; ++c) {
                    double kd = 0, mm = 0;
                    for (int q = 0; q < 4; ++q) kd += DN[q][r] * DN[q][c];
                    for (int q = 0; q < 2; ++q) mm += NN[q][r] * NN[q][c];
                    Ae[r * 8 + c] += wJ * (beta * kd + mu * mm + alpha * mu * kd);
                }
            for (int r = 0; r < 8; ++r) {
                double dg = 0; /* diag(DN' * FMinusW')(r) */
                for (int q = 0; q < 4; ++q) dg += DN[q][r] * FW[r][q];
                double nuv = 0; /* (N'N (U - V))(r) */
                for (int c = 0; c < 8; ++c) {
                    double mm = 0;
                    for (int q = 0; q < 2; ++q) mm += NN[q][r] * NN[q][c];
                    nuv += mm * UV[c];
                }
                be[r] += wJ * (beta * dg + mu * nuv);
            }
        }
}

/* COO triplets for the whole mesh: rows/cols/vals (E*64) and b (2*nn, summed). */
void orc_subpb2_assemble(const double *coords, int nn, const int *elems, int E, double beta,
                         double mu, double alpha, const double *U, const double *F,
                         const double *udual, const double *vdual, int *ri, int *ci,
                         double *val, double *b)
{
    memset(b, 0, sizeof(double) * 2 * (size_t)nn);
    for (int e = 0; e < E; ++e) {
        double Ae[64], be[8];
        const int *el = elems + 4 * (size_t)e;
        subpb2_element(coords, el, beta, mu, alpha, U, F, udual, vdual, Ae, be);
        int dof[8];
        for (int a = 0; a < 4; ++a) { dof[2 * a] = 2 * el[a]; dof[2 * a + 1] = 2 * el[a] + 1; }
        for (int r = 0; r < 8; ++r) {
            for (int c = 0; c < 8; ++c) {
                size_t k = (size_t)e * 64 + r * 8 + c;
                ri[k] = dof[r]; ci[k] = dof[c]; val[k] = Ae[r * 8 + c];
            }
            b[dof[r]] += be[r];
        }
    }
}

/* CSR Jacobi-PCG in FP64 to relative residual rtol (x0 = 0). Returns iterations. */
int orc_pcg(int n, const int *rowptr, const int *col, const double *val, const double *b,
            double *x, double rtol, int maxit)
{
    double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
    double *p = malloc(sizeof(double) * n), *Ap = malloc(sizeof(double) * n);
    double *dinv = malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) {
        double d = 0;
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) if (col[k] == i) d += val[k];
        dinv[i] = d != 0 ? 1.0 / d : 1.0;
        x[i] = 0.0; r[i] = b[i]; z[i] = r[i] * dinv[i]; p[i] = z[i];
    }
    double bn = 0, rz = 0;
    for (int i = 0; i < n; ++i) { bn += b[i] * b[i]; rz += r[i] * z[i]; }
    bn = sqrt(bn);
    int it = 0;
    if (bn == 0) goto done;
    for (it = 1; it <= maxit; ++it) {
        double pAp = 0;
        for (int i = 0; i < n; ++i) {
            double s = 0;
            for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s += val[k] * p[col[k]];
            Ap[i] = s; pAp += p[i] * s;
        }
        if (!(pAp > 0) || rz == 0) break; /* breakdown (converged to round-off) */
        double a = rz / pAp, rn = 0;
        for (int i = 0; i < n; ++i) { x[i] += a * p[i]; r[i] -= a * Ap[i]; rn += r[i] * r[i]; }
        if (sqrt(rn) <= rtol * bn) break;
        double rz2 = 0;
        for (int i = 0; i < n; ++i) { z[i] = r[i] * dinv[i]; rz2 += r[i] * z[i]; }
        double bb = rz2 / rz;
        rz = rz2;
        for (int i = 0; i < n; ++i) p[i] = z[i] + bb * p[i];
    }
done:
    free(r); free(z); free(p); free(Ap); free(dinv);
    return it;
}

/* ------------------------------------------------------------------------------
 * a15  funGlobal_NodalStrainAvg.m:1-147 -- 2x2 Gauss-point gradients, extrapolated
 *     to the element's nodes (:105-119) and averaged over elements (:139-144).
 *     Written into the F slots as [u_x, u_y, v_x, v_y] (:123-137) exactly like the
 *     reference (its F convention elsewhere is [u_x v_x u_y v_y], so the
 *     off-diagonals come back swapped); swap_fix = 1 writes [u_x v_x u_y v_y].
 *     Nodes in no element get NaN (0/0 of :144).
 * ---------------------------------------------------------------------------- */
void orc_nodal_strain(const double *coords, int nn, const int *elems, int E, const double *U,
                      int swap_fix, double *Fout)
{
    double *acc = calloc(4 * (size_t)nn, sizeof(double));
    double *cnt = calloc(4 * (size_t)nn, sizeof(double));
    const double g = 1.0 / sqrt(3.0);
    /* Gauss points in node order (:55-64): (-,-) (+,-) (+,+) (-,+) */
    const double gk[4] = {-g, g, g, -g}, ge[4] = {-g, -g, g, g};
    const double s3 = sqrt(3.0);
    const double Mx[4][4] = {{1 + 0.5 * s3, -0.5, 1 - 0.5 * s3, -0.5},
                             {-0.5, 1 + 0.5 * s3, -0.5, 1 - 0.5 * s3},
                             {1 - 0.5 * s3, -0.5, 1 + 0.5 * s3, -0.5},
                             {-0.5, 1 - 0.5 * s3, -0.5, 1 + 0.5 * s3}};
    for (int e = 0; e < E; ++e) {
        const int *el = elems + 4 * (size_t)e;
        double px[4], py[4], ue[4], ve[4], Sg[4][4];
        for (int a = 0; a < 4; ++a) {
            px[a] = coords[2 * el[a]]; py[a] = coords[2 * el[a] + 1];
            ue[a] = U[2 * el[a]]; ve[a] = U[2 * el[a] + 1];
        }
        for (int p = 0; p < 4; ++p) {
            double dNk[4], dNe[4];
            q4_dN(gk[p], ge[p], dNk, dNe);
            double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
            for (int a = 0; a < 4; ++a) {
                J11 += dNk[a] * px[a]; J12 += dNk[a] * py[a];
                J21 += dNe[a] * px[a]; J22 += dNe[a] * py[a];
            }
            double det = J11 * J22 - J12 * J21;
            double i11 = J22 / det, i12 = -J12 / det, i21 = -J21 / det, i22 = J11 / det;
            double ux = 0, uy = 0, vx = 0, vy = 0;
            for (int a = 0; a < 4; ++a) {
                double ddx = i11 * dNk[a] + i12 * dNe[a];
                double ddy = i21 * dNk[a] + i22 * dNe[a];
                ux += ddx * ue[a]; uy += ddy * ue[a]; vx += ddx * ve[a]; vy += ddy * ve[a];
            }
            Sg[p][0] = ux; Sg[p][1] = uy; Sg[p][2] = vx; Sg[p][3] = vy;
        }
        for (int a = 0; a < 4; ++a) {
            double nv[4] = {0, 0, 0, 0};
            for (int p = 0; p < 4; ++p)
                for (int c = 0; c < 4; ++c) nv[c] += Mx[a][p] * Sg[p][c];
            int nd = el[a];
            if (swap_fix) { double t = nv[1]; nv[1] = nv[2]; nv[2] = t; }
            for (int c = 0; c < 4; ++c) { acc[4 * nd + c] += nv[c]; cnt[4 * nd + c] += 1.0; }
        }
    }
    for (size_t i = 0; i < 4 * (size_t)nn; ++i) Fout[i] = acc[i] / cnt[i];
    free(acc); free(cnt);
}
