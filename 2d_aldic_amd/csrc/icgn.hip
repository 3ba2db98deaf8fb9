// Batched inverse-compositional Gauss-Newton (IC-GN) for gfx950.
//
//   6-DOF: funICGN.m:1-322 (per subset of LocalICGN.m:91-107)
//   2-DOF: funICGN_Subpb1.m:1-362 (per subset of Subpb1.m:119-152)
//
// Layout: WPS wavefronts work on one subset (WPS = 1 for windows up to 33 x 33,
// 4 for larger ones); each lane owns PPL pixel slots of the (w+1)^2 window: pixel p
// (x fastest) sits in slot k = p / (64 * WPS) of thread p % (64 * WPS).  A per-thread
// LDS table holds each slot's offsets (dx, dy) and image index; every pass re-reads
// f, fx, fy at the slot's pixel (L1/L2 hits), which keeps the kernel at 128 VGPRs and
// four waves per SIMD to cover the gathers.  The Hessian, its inverse (FP64), mean/norm of f and the constant
// sums sum J_k and sum J_k f_n are formed once per subset.  Each iteration then makes
// ONE pass over the pixels: warp, Keys-cubic gather of g (16 loads through L1/L2),
// and the nine sums  S_g, S_gg, S_{g J_k} (k = 1..6), S_{g f}, from which
//   b_k  = bottomf * (S_{J_k f_n} - (S_{g J_k} - gbar S_{J_k}) / bottomg)
//   zncc = (S_{g f} - fbar S_g) / (bottomf bottomg)
// follow exactly (the identity sum J_k r with r = f_n - (g - gbar)/bottomg).  The
// convergence test, the 6x6 solve and the inverse-compositional update run in FP64
// on every lane with bitwise-identical inputs (butterfly all-reduce), so the loop
// is uniform across the wave.
#include <hip/hip_fp16.h>

#include <algorithm>

#include "aldic_internal.h"
#include "common.h"

namespace aldic {

constexpr int kMaxIter = 100;

// funICGN.m:34-65 window set-up with the reference's error cases and mask rule.
__device__ __forceinline__ int win_corner(double x0, double y0, int w, int nx, int ny, int &xa,
                                          int &ya)
{
    const double xs = x0 - 0.5 * w, ys = y0 - 0.5 * w;
    if (xs != floor(xs) || ys != floor(ys)) return -1;
    xa = (int)xs;
    ya = (int)ys;
    if (xa < 4 || ya < 4 || xa + w > nx - 3 || ya + w > ny - 3) return -1;
    return 0;
}

template <int WPS>
__device__ __forceinline__ float block_sum1(float v, float *red, int wid)
{
    v = wave_sum(v);
    if constexpr (WPS > 1) {
        if ((threadIdx.x & 63) == 0) red[wid] = v;
        __syncthreads();
        v = 0.f;
#pragma unroll
        for (int q = 0; q < WPS; ++q) v += red[q];
        __syncthreads();
    }
    return v;
}

template <int WPS>
__device__ __forceinline__ double block_sumd(double v, double *red, int wid)
{
    v = wave_sum(v);
    if constexpr (WPS > 1) {
        if ((threadIdx.x & 63) == 0) red[wid] = v;
        __syncthreads();
        v = 0.0;
#pragma unroll
        for (int q = 0; q < WPS; ++q) v += red[q];
        __syncthreads();
    }
    return v;
}

template <int WPS, int N>
__device__ __forceinline__ void block_sumN(float (&v)[N], float *red, int wid)
{
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = wave_sum(v[i]);
    if constexpr (WPS > 1) {
        if ((threadIdx.x & 63) == 0)
#pragma unroll
            for (int i = 0; i < N; ++i) red[i * WPS + wid] = v[i];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < N; ++i) {
            float s = 0.f;
#pragma unroll
            for (int q = 0; q < WPS; ++q) s += red[i * WPS + q];
            v[i] = s;
        }
        __syncthreads();
    }
}

// Inverse of a symmetric positive definite 6x6 by Cholesky, from the packed lower
// triangle h21 (MATLAB column-major order, as in H21), writing H^-1 (row-major) to hi.
// Returns false when a pivot is not positive.  Every lane of the caller runs it with
// identical inputs; the factor stays in registers, H and H^-1 live in LDS.  The six
// pivots are inverted once and every later division is a product.
__device__ __forceinline__ bool spd_inverse6_lds(const double *h21, double *hi)
{
    // column-major packed index of (r, c), r >= c
    auto hp = [&](int r, int c) { return h21[c * 6 - c * (c - 1) / 2 + (r - c)]; };
    double L[21], id[6];
#define LI(r, c) L[(r) * ((r) + 1) / 2 + (c)]
#pragma unroll
    for (int r = 0; r < 6; ++r) {
#pragma unroll
        for (int c = 0; c <= r; ++c) {
            double s = hp(r, c);
#pragma unroll
            for (int k = 0; k < c; ++k) s -= LI(r, k) * LI(c, k);
            if (r == c) {
                if (!(s > 0.0)) return false;
                LI(r, r) = sqrt(s);
                id[r] = 1.0 / LI(r, r);
            } else {
                LI(r, c) = s * id[c];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double y[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            double s = (r == j) ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < r; ++k) s -= LI(r, k) * y[k];
            y[r] = s * id[r];
        }
        double xcol[6];
#pragma unroll
        for (int r = 5; r >= 0; --r) {
            double s = y[r];
#pragma unroll
            for (int k = r + 1; k < 6; ++k) s -= LI(k, r) * xcol[k];
            xcol[r] = s * id[r];
        }
#pragma unroll
        for (int r = 0; r < 6; ++r) hi[r * 6 + j] = xcol[r];
    }
#undef LI
    return true;
}

// funICGN.m:268-286: P <- P o dP^-1.  Returns false when det(dP) == 0.
__device__ __forceinline__ bool ic_compose(double P[6], const double D[6])
{
    const double det = (1.0 + D[0]) * (1.0 + D[3]) - D[1] * D[2];
    if (det == 0.0) return false;
    const double id = 1.0 / det;  // one division, six products
    const double t1 = (-D[0] - D[0] * D[3] + D[1] * D[2]) * id;
    const double t2 = -D[1] * id;
    const double t3 = -D[2] * id;
    const double t4 = (-D[3] - D[0] * D[3] + D[1] * D[2]) * id;
    const double t5 = (-D[4] - D[3] * D[4] + D[2] * D[5]) * id;
    const double t6 = (-D[5] - D[0] * D[5] + D[1] * D[4]) * id;
    const double a11 = 1.0 + P[0], a12 = P[2], a13 = P[4];
    const double a21 = P[1], a22 = 1.0 + P[3], a23 = P[5];
    P[0] = a11 * (1.0 + t1) + a12 * t2 - 1.0;
    P[2] = a11 * t3 + a12 * (1.0 + t4);
    P[4] = a11 * t5 + a12 * t6 + a13;
    P[1] = a21 * (1.0 + t1) + a22 * t2;
    P[3] = a21 * t3 + a22 * (1.0 + t4) - 1.0;
    P[5] = a21 * t5 + a22 * t6 + a23;
    return true;
}

// Keys-cubic weights (a = -1/2, ba_interp2.cpp:153-162) of the fraction t, doubled
// (2 w, integer coefficients): w3 = t^3 - t^2, w0 = t^2 - t - w3, w1 = 2 - 2 t^2 + 3 w3,
// w2 = t + t^2 - 3 w3 (8 VALU).  A sample from doubled weights on both axes is 4 g.
__device__ __forceinline__ void keys_w2(float t, float (&w)[4])
{
    const float t2 = t * t;
    const float w3 = fmaf(t2, t, -t2);
    w[0] = (t2 - t) - w3;
    w[1] = fmaf(3.f, w3, fmaf(-2.f, t2, 2.f));
    w[2] = fmaf(-3.f, w3, t + t2);
    w[3] = w3;
}

// floor(x) as an integer in one instruction (x finite and well inside the int range)
__device__ __forceinline__ int floor_i(float x)
{
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

__device__ __forceinline__ float ld_byte(const float *p, unsigned off)
{
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(p) + off);
}

// One slot of an IC-GN pass, in stages for an explicit software pipeline.  The slot's
// table entry e holds its offsets (dx, dy) from the window centre as two halves and the
// byte offset of its pixel.  icgn_addr warps the offsets, u = p0 dx + p2 dy + p4,
// v = p1 dx + p3 dy + p5, and returns the byte offset of the sample's first tap row
// relative to rows[] (rows[b] = g - 4 (1 + nx) + 4 b nx: every tap row is one load at the
// same per-lane offset; the sample is known to lie in [3, nx-2] x [3, ny-2], 1-based, so
// every tap is inside the image) with the fractional parts; icgn_taps loads the 16 taps;
// icgn_accum interpolates and accumulates.
__device__ __forceinline__ unsigned icgn_addr(const uint2 e, int nx4, float p0, float p1, float p2,
                                              float p3, float p4, float p5, float &dx, float &dy,
                                              float &tx, float &ty)
{
    const __half2 h = *reinterpret_cast<const __half2 *>(&e.x);
    dx = __low2float(h);
    dy = __high2float(h);
    const float ux = fmaf(p0, dx, fmaf(p2, dy, p4));
    const float uy = fmaf(p1, dx, fmaf(p3, dy, p5));
    tx = __builtin_amdgcn_fractf(ux);
    ty = __builtin_amdgcn_fractf(uy);
    // images hold < 2^29 pixels and 4 nx < 2^23: byte offsets fit in 32 bits (arithmetic
    // mod 2^32), the row term in a 24-bit multiply; two instructions
    unsigned ob;
    asm("v_lshl_add_u32 %0, %1, 2, %2" : "=v"(ob) : "v"(floor_i(ux)), "v"(e.y));
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(ob) : "v"(floor_i(uy)), "s"(nx4), "v"(ob));
    return ob;
}

__device__ __forceinline__ void icgn_taps(const char *const (&rows)[4], unsigned ob, float (&v)[16])
{
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const float *row = reinterpret_cast<const float *>(rows[b] + ob);
#pragma unroll
        for (int i = 0; i < 4; ++i) v[4 * b + i] = row[i];
    }
}

// Accumulates, for a sample g (taken as 4 g from the doubled weights; the sums are
// scaled back by powers of two after the pass, exactly): S_g, S_gg, then S_{g J_k} with
// J = (fx dx, fy dx, fx dy, fy dy, fx, fy) and S_{g f} (6-DOF), or S_{g fx}, S_{g fy},
// S_{g f} (2-DOF).  `in` = false counts nothing.
template <bool SIX, int NS>
__device__ __forceinline__ void icgn_accum(const float (&v)[16], float tx, float ty, float dx,
                                           float dy, float f, float fx, float fy, bool in,
                                           float (&acc)[NS])
{
    float wx[4], wy[4];
    keys_w2(tx, wx);
    keys_w2(ty, wy);
    float r[4];
#pragma unroll
    for (int b = 0; b < 4; ++b)
        r[b] = wx[0] * v[4 * b] + wx[1] * v[4 * b + 1] + wx[2] * v[4 * b + 2] + wx[3] * v[4 * b + 3];
    float gv = wy[0] * r[0] + wy[1] * r[1] + wy[2] * r[2] + wy[3] * r[3];
    gv = in ? gv : 0.f;
    const float t1 = gv * fx, t2 = gv * fy;
    acc[0] += gv;
    acc[1] += gv * gv;
    if constexpr (SIX) {
        acc[2] += t1 * dx; acc[3] += t2 * dx;
        acc[4] += t1 * dy; acc[5] += t2 * dy;
        acc[6] += t1; acc[7] += t2;
        acc[8] += gv * f;
    } else {
        acc[2] += t1; acc[3] += t2;
        acc[4] += gv * f;
    }
}

// One slot without the pipeline.
template <bool SIX, int NS>
__device__ __forceinline__ void icgn_pixel(const uint2 e, const char *const (&rows)[4], int nx4,
                                           float p0, float p1, float p2, float p3, float p4,
                                           float p5, float f, float fx, float fy,
                                           float (&acc)[NS])
{
    float dx, dy, tx, ty, v[16];
    const unsigned ob = icgn_addr(e, nx4, p0, p1, p2, p3, p4, p5, dx, dy, tx, ty);
    icgn_taps(rows, ob, v);
    icgn_accum<SIX, NS>(v, tx, ty, dx, dy, f, fx, fy, true, acc);
}

// Shared body for both modes.  SIX = true: 6-DOF funICGN; false: 2-DOF funICGN_Subpb1.
template <int PPL, int WPS, bool SIX>
__device__ __forceinline__ void icgn_subset(const IcgnArgs &a, int slot)
{
    constexpr int SPB = 4 / WPS;          // subsets per block
    constexpr int T = 64 * WPS;           // threads per subset
    constexpr int NCS = 2 + 6 + 6 + (SIX ? 21 : 0);
    __shared__ double s_hinv[SPB][36];
    __shared__ double s_cst[SPB][12];
    // per subset: the warp P[0..5], then bottomf, bottomf^2, the 2-DOF system t00,
    // t01, t11, mean f, 1 / bottomf and 1 / n.  The Gauss-Newton loop re-reads them
    // after each pass, so no FP64 value is held in registers across the pass (which
    // would spill to scratch).
    __shared__ double s_k[SPB][14];
    __shared__ double s_cs[SPB][NCS];
    __shared__ float s_red[16 * WPS];
    __shared__ double s_redd[WPS];
    // per pixel slot of every thread: the offsets (dx, dy) from the window centre as
    // two halves (exact: multiples of 1/2 below 1024) and the pixel's byte offset.
    // Each thread writes and reads only its own entries.
    __shared__ uint2 s_tab[SPB * PPL * T];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int sub_in_block = wave / WPS;
    const int wid = wave % WPS;            // wave index inside the subset
    const int tid = wid * 64 + lane;       // thread index inside the subset
    const int j = a.list ? a.list[slot] : slot;
    const int nx = a.nx, ny = a.ny;
    const double x0 = a.coords[2 * j], y0 = a.coords[2 * j + 1];
    int w = a.ws ? a.ws[j] : a.ws_all;
    const double tol = a.tol;

    double P[6];
    if (SIX) {
        P[0] = P[1] = P[2] = P[3] = 0.0;
        if (a.F0)
            for (int q = 0; q < 4; ++q) P[q] = a.F0[4 * j + q];
        P[4] = a.U0[2 * j];
        P[5] = a.U0[2 * j + 1];
    } else {
        P[0] = a.Fold[4 * j]; P[1] = a.Fold[4 * j + 1];
        P[2] = a.Fold[4 * j + 2]; P[3] = a.Fold[4 * j + 3];
        P[4] = a.Uold[2 * j]; P[5] = a.Uold[2 * j + 1];
    }

    // H21 is written once the Hessian exists; the early exits (-1, 102) leave zeros.
    auto finish = [&](int status, bool zero_h, float zc) {
        if (tid != 0) return;
        if (SIX && a.cc_mode == 1 && zero_h) a.cc[8 * (size_t)j + 7] = (double)status;
        const bool failed = status == -1;
        if (SIX) {
            a.U[2 * j] = failed ? NAN : P[4];
            a.U[2 * j + 1] = failed ? NAN : P[5];
            for (int q = 0; q < 4; ++q) a.F[4 * j + q] = failed ? NAN : P[q];
            if (a.Ucopy) {
                a.Ucopy[2 * j] = failed ? NAN : P[4];
                a.Ucopy[2 * j + 1] = failed ? NAN : P[5];
            }
            if (a.Fcopy)
                for (int q = 0; q < 4; ++q) a.Fcopy[4 * j + q] = failed ? NAN : P[q];
            if (zero_h)
                for (int q = 0; q < 21; ++q) a.H21[21 * (size_t)j + q] = 0.0;
            if (a.zncc) a.zncc[j] = (double)zc;
        } else {
            a.U[2 * j] = failed ? NAN : P[4];
            a.U[2 * j + 1] = failed ? NAN : P[5];
        }
        a.iters[j] = status;
    };

    // ---- window, mask rule and per-subset constants (funICGN.m:34-94) -------------
    // One unrolled pass over the window's slots accumulates, per lane in FP32 (<= PPL
    // terms), the exact zeros of DfDx, sum f, sum f^2, sum J_k, sum J_k f and (6-DOF)
    // the 21 Hessian entries; lanes are combined in FP64.  Pixel p = tid + k*T sits at
    // column i, row r (p = r*w1 + i); the table's walk over k advances (i, r) by T
    // without divisions.
    uint2 *tab = s_tab + sub_in_block * (PPL * T);
    int xa = 0, ya = 0;
    int w1 = w + 1, n = w1 * w1;
    // The slot table of the current window: offsets from the centre and the pixel's byte
    // offset; an empty slot (past the window's last pixel) holds pixel (0, 0).  Every pass
    // over the window reads it back (f, fx, fy are re-read at each slot's pixel, L1/L2
    // hits, instead of holding 3 PPL registers: more waves in flight for the gathers).
    auto build_table = [&]() {
        const int di = T % w1, dr = T / w1;
        const float half = 0.5f * (float)w;
        int i = tid % w1, r = tid / w1;
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            const bool in = r < w1;
            const int ii = in ? i : 0, rr = in ? r : 0;
            const __half2 h = __floats2half2_rn((float)ii - half, (float)rr - half);
            uint2 e;
            e.x = *reinterpret_cast<const unsigned *>(&h);
            e.y = 4u * (unsigned)((ya - 1 + rr) * nx + (xa - 1 + ii));
            tab[k * T + tid] = e;
            i += di; r += dr;
            if (i >= w1) { i -= w1; ++r; }
        }
        // read back from LDS by every later pass, not held in registers
        asm volatile("" ::: "memory");
    };
    constexpr int NH = SIX ? 21 : 0;
    constexpr int NC = 2 + 6 + 6 + NH;   // sum f, sum f^2, sum J, sum J f, H
    double *cs = s_cs[sub_in_block];  // reduced sums, same value from every lane
    double *cst = s_cst[sub_in_block];  // [0..5] sum J_k, [6..11] sum J_k f_n
    double *hinv = s_hinv[sub_in_block];
    double meanf = 0.0, ssf = 0.0;
    bool singular = false;
    bool cached = false;
    if (!SIX && a.cc_mode == 2) {
        // constants of this subset kept by the 6-DOF run on the same window:
        // {mean f, sum (f - mean f)^2, S_J5, S_J6, S_J5fn, S_J6fn, w, status}
        const double *cc = a.cc + 8 * (size_t)j;
        const int st0 = (int)cc[7];
        if (st0 == -1 || st0 == 102) { finish(st0, true, NAN); return; }
        if (st0 == 0) {
            w = (int)cc[6];
            if (win_corner(x0, y0, w, nx, ny, xa, ya)) { finish(-1, true, NAN); return; }
            w1 = w + 1;
            n = w1 * w1;
            if (n > PPL * T) { if (tid == 0) a.iters[j] = -2; return; }
            build_table();
            meanf = cc[0];
            ssf = cc[1];
            cst[4] = cc[2]; cst[5] = cc[3]; cst[10] = cc[4]; cst[11] = cc[5];
            cached = true;
        }
    }
    if (!cached) {
    if (win_corner(x0, y0, w, nx, ny, xa, ya)) { finish(-1, true, NAN); return; }
    for (int pass = 0; pass < 2; ++pass) {
        if (n > PPL * T) {
            if (tid == 0) {
                a.iters[j] = -2;
                if (SIX && a.cc_mode == 1) a.cc[8 * (size_t)j + 7] = -2.0;
            }
            return;
        }
        build_table();
        float acc[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) acc[q] = 0.f;
        float zeros = 0.f;
        const int lim = n - tid;  // slot k holds a pixel iff k * T < lim
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            if (k * T < lim) {
                const uint2 e = tab[k * T + tid];
                const __half2 h = *reinterpret_cast<const __half2 *>(&e.x);
                const float dx = __low2float(h), dy = __high2float(h);
                const float fval = ld_byte(a.f, e.y), fx = ld_byte(a.gx, e.y), fy = ld_byte(a.gy, e.y);
                zeros += fx == 0.f ? 1.f : 0.f;
                const float J[6] = {fx * dx, fy * dx, fx * dy, fy * dy, fx, fy};
                acc[0] += fval;
                acc[1] += fval * fval;
#pragma unroll
                for (int q = 0; q < 6; ++q) { acc[2 + q] += J[q]; acc[8 + q] += J[q] * fval; }
                if (SIX) {
                    int q = 14;
#pragma unroll
                    for (int c = 0; c < 6; ++c)
#pragma unroll
                        for (int rr = c; rr < 6; ++rr) acc[q++] += J[rr] * J[c];
                }
            }
        }
        if (pass == 0) {  // mask rule on the original window
            const float n0 = block_sum1<WPS>(zeros, s_red, wid);
            if (!((double)n0 < 0.5 * n)) { finish(102, true, NAN); return; }
            if ((double)n0 > 0.1 * n) {
                w = (int)floor(sqrt(2.0) * w + 0.5);  // MATLAB round (funICGN.m:58)
                if (win_corner(x0, y0, w, nx, ny, xa, ya)) { finish(-1, true, NAN); return; }
                w1 = w + 1;
                n = w1 * w1;
                continue;  // reload the enlarged window
            }
        }
#pragma unroll
        for (int q = 0; q < NC; ++q) cs[q] = block_sumd<WPS>((double)acc[q], s_redd, wid);
        break;
    }
    meanf = cs[0] / n;
    ssf = cs[1] - cs[0] * meanf;  // sum (f - meanf)^2
    ssf = ssf > 0 ? ssf : 0.0;
    {
        const double bottomf = sqrt(ssf);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            cst[q] = cs[2 + q];  // every lane writes the same value and reads back its own
            cst[6 + q] = (cs[8 + q] - meanf * cs[2 + q]) / bottomf;
        }
    }
    if (SIX) {
        if (tid == 0) {
            for (int e = 0; e < 21; ++e) a.H21[21 * (size_t)j + e] = cs[14 + e];
            if (a.cc_mode == 1) {
                double *cc = a.cc + 8 * (size_t)j;
                cc[0] = meanf; cc[1] = ssf;
                cc[2] = cst[4]; cc[3] = cst[5]; cc[4] = cst[10]; cc[5] = cst[11];
                cc[6] = (double)w; cc[7] = 0.0;
            }
        }
        singular = !spd_inverse6_lds(cs + 14, hinv);
    }
    }  // !cached
    const double bottomf = sqrt(ssf);
    // 2-DOF system matrix
    double t00 = 0, t01 = 0, t11 = 0;
    const double bf2 = ssf;  // bottomf^2
    if (!SIX) {
        // tempH (funICGN_Subpb1.m:281-283)
        const double *hin = a.Hin + 21 * (size_t)j;
        const double mu = a.mu;
        const double h00 = hin[18] * 2.0 / bf2 + mu, h01 = hin[19] * 2.0 / bf2,
                     h11 = hin[20] * 2.0 / bf2 + mu;
        const double dmax = h00 > h11 ? h00 : h11;
        t00 = h00 + 1e-3 * dmax;
        t01 = h01;
        t11 = h11 + 1e-3 * dmax;
    }

    double *sk = s_k[sub_in_block];  // every lane writes the same values, reads its own
#pragma unroll
    for (int q = 0; q < 6; ++q) sk[q] = P[q];
    sk[6] = bottomf; sk[7] = bf2; sk[8] = t00; sk[9] = t01; sk[10] = t11; sk[11] = meanf;
    sk[12] = 1.0 / bottomf; sk[13] = 1.0 / n;

    // ---- Gauss-Newton loop (funICGN.m:107-294 / funICGN_Subpb1.m:94-337) ----------
    const float Xa = (float)xa, Ya = (float)ya;  // window pixel (i, r) is X = Xa + i, Y = Ya + r
    const float fhalf = 0.5f * (float)w;
    double nrmNew = 1.0, nrmAbs = 1.0, nrmInit = 0.0;
    float zc = NAN;
    int it = 0;
    while (it <= kMaxIter && nrmNew > tol && nrmAbs > tol) {
        ++it;
        asm volatile("" ::: "memory");
#pragma unroll
        for (int q = 0; q < 6; ++q) P[q] = sk[q];
        bool finite = true;
#pragma unroll
        for (int q = 0; q < 6; ++q) finite = finite && isfinite(P[q]);
        if (!finite) { nrmNew = NAN; break; }
        const float p0 = (float)P[0], p1 = (float)P[1], p2 = (float)P[2], p3 = (float)P[3],
                    p4 = (float)P[4], p5 = (float)P[5];
        // out-of-image test (funICGN.m:118-121).  The warp is affine, so over the
        // window's pixel box the extremes of u and v sit at its four corners; u = X + ux
        // with X the exact integer pixel, so u < 3 <=> ux < 3 - X.
        bool oob = false;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int ci = (c & 1) ? w : 0, cr = (c & 2) ? w : 0;
            const float dx = (float)ci - fhalf, dy = (float)cr - fhalf;
            const float ux = p0 * dx + p2 * dy + p4;
            const float uy = p1 * dx + p3 * dy + p5;
            const float X = Xa + (float)ci, Y = Ya + (float)cr;
            // written as "not inside" so a non-finite corner also counts as outside
            const bool in = (ux >= 3.f - X) && (ux <= (float)(nx - 2) - X) &&
                            (uy >= 3.f - Y) && (uy <= (float)(ny - 2) - Y);
            oob = oob || !in;
        }
        if (oob) { nrmNew = 1e6; break; }  // same on every lane

        constexpr int NS = SIX ? 9 : 5;
        float acc[NS];
#pragma unroll
        for (int q = 0; q < NS; ++q) acc[q] = 0.f;
        {
            // the table is re-read from LDS every pass (kept opaque to the compiler, which
            // would otherwise hold all of it in registers)
            asm volatile("" ::: "memory");
            const int lim = n - tid;  // slot k holds a pixel iff k * T < lim
            const int nx4 = 4 * nx;
            const char *gb = reinterpret_cast<const char *>(a.g) - 4 * (1 + (size_t)nx);
            const char *const rows[4] = {gb, gb + 4 * (size_t)nx, gb + 8 * (size_t)nx,
                                         gb + 12 * (size_t)nx};
            if (n > (PPL - 1) * T) {
                // every slot holds a pixel on some lane (a 33 x 33 window in 18 slots): one
                // branch-free run over the slots, software-pipelined by hand: the taps of
                // slot k + 1 are in flight while slot k is interpolated; only the last
                // slot's empty lanes are masked out
                float v[2][16], sdx[2], sdy[2], stx[2], sty[2], sf[2], sgx[2], sgy[2];
                {
                    const uint2 e = tab[tid];
                    const unsigned ob = icgn_addr(e, nx4, p0, p1, p2, p3, p4, p5, sdx[0], sdy[0],
                                                  stx[0], sty[0]);
                    icgn_taps(rows, ob, v[0]);
                    sf[0] = ld_byte(a.f, e.y); sgx[0] = ld_byte(a.gx, e.y);
                    sgy[0] = ld_byte(a.gy, e.y);
                }
#pragma unroll
                for (int k = 0; k < PPL; ++k) {
                    const int c = k & 1, m = c ^ 1;
                    if (k + 1 < PPL) {
                        const uint2 e = tab[(k + 1) * T + tid];
                        const unsigned ob = icgn_addr(e, nx4, p0, p1, p2, p3, p4, p5, sdx[m],
                                                      sdy[m], stx[m], sty[m]);
                        icgn_taps(rows, ob, v[m]);
                        sf[m] = ld_byte(a.f, e.y); sgx[m] = ld_byte(a.gx, e.y);
                        sgy[m] = ld_byte(a.gy, e.y);
                    }
                    icgn_accum<SIX, NS>(v[c], stx[c], sty[c], sdx[c], sdy[c], sf[c], sgx[c],
                                        sgy[c], k < PPL - 1 || k * T < lim, acc);
                }
            } else {
#pragma unroll
                for (int k = 0; k < PPL; ++k)
                    if (k * T < lim) {
                        const uint2 e = tab[k * T + tid];
                        icgn_pixel<SIX, NS>(e, rows, nx4, p0, p1, p2, p3, p4, p5, ld_byte(a.f, e.y),
                                            ld_byte(a.gx, e.y), ld_byte(a.gy, e.y), acc);
                    }
            }
        }
        block_sumN<WPS, NS>(acc, s_red, wid);
        asm volatile("" ::: "memory");
#pragma unroll
        for (int q = 0; q < 6; ++q) P[q] = sk[q];
        const double bottomf = sk[6], bf2 = sk[7], t00 = sk[8], t01 = sk[9], t11 = sk[10];
        // the pass summed 4 g: scale back (exact), S_{g f_n} = (S_{g f} - mean f S_g) / bottomf
        const double Sg = 0.25 * acc[0], Sgg = 0.0625 * acc[1];
        const double gbar = Sg * sk[13];
        const double vg = Sgg - Sg * gbar;
        const double bottomg = sqrt(vg > 0 ? vg : 0.0);
        const double ibg = 1.0 / bottomg;
        const double meanf_k = sk[11];
        double b[6];
        if (SIX) {
#pragma unroll
            for (int q = 0; q < 6; ++q)
                b[q] = bottomf * (cst[6 + q] - (0.25 * acc[2 + q] - gbar * cst[q]) * ibg);
            zc = (float)((0.25 * acc[8] - meanf_k * Sg) * sk[12] * ibg);
            nrmAbs = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2] + b[3] * b[3] + b[4] * b[4] +
                          b[5] * b[5]);
        } else {
            const double b5 = bottomf * (cst[10] - (0.25 * acc[2] - gbar * cst[4]) * ibg);
            const double b6 = bottomf * (cst[11] - (0.25 * acc[3] - gbar * cst[5]) * ibg);
            zc = (float)((0.25 * acc[4] - meanf_k * Sg) * sk[12] * ibg);
            // tempb (funICGN_Subpb1.m:269)
            const double two_bf2 = 2.0 / bf2;
            b[0] = b5 * two_bf2 + a.mu * (P[4] - a.Uold[2 * j] - a.vdual[2 * j]);
            b[1] = b6 * two_bf2 + a.mu * (P[5] - a.Uold[2 * j + 1] - a.vdual[2 * j + 1]);
            nrmAbs = sqrt(b[0] * b[0] + b[1] * b[1]);
        }
        nrmNew = nrmAbs;
        if (it == 1) nrmInit = nrmNew;
        nrmNew = nrmInit > tol ? nrmNew / nrmInit : 0.0;
        if (nrmNew < tol || nrmAbs < tol) break;

        double D[6];
        if (SIX) {
            if (singular) {  // MATLAB's backslash on a singular H gives non-finite steps
#pragma unroll
                for (int q = 0; q < 6; ++q) sk[q] = P[q] = NAN;
                continue;
            }
#pragma unroll
            for (int r = 0; r < 6; ++r) {
                double s = 0.0;
#pragma unroll
                for (int c = 0; c < 6; ++c) s += hinv[r * 6 + c] * b[c];
                D[r] = -s;
            }
        } else {
            const double det = t00 * t11 - t01 * t01;
            D[0] = D[1] = D[2] = D[3] = 0.0;
            if (det == 0.0) { sk[4] = P[4] = NAN; sk[5] = P[5] = NAN; continue; }
            const double id = 1.0 / det;
            D[4] = -(t11 * b[0] - t01 * b[1]) * id;
            D[5] = -(-t01 * b[0] + t00 * b[1]) * id;
        }
        if (!ic_compose(P, D)) break;
#pragma unroll
        for (int q = 0; q < 6; ++q) sk[q] = P[q];
    }
    const bool conv = (nrmNew < tol) || (nrmAbs < tol);
    if (!conv || isnan(nrmNew)) it = 101;
    finish(it, false, zc);
}

// Subset slot = blockIdx.x * SPB + wave / WPS.  LOOP: the grid is capped and steps over
// a list whose length is only known on the device.
template <int PPL, int WPS, bool SIX, bool LOOP>
__global__ __launch_bounds__(256, 4) void k_icgn(IcgnArgs a)
{
    constexpr int SPB = 4 / WPS;
    const int nlist = a.nlist_dev ? *a.nlist_dev : a.nlist;
    const int sub = (threadIdx.x >> 6) / WPS;
    if constexpr (!LOOP) {
        // workgroups are dealt to the 8 XCDs in turn; give each XCD a contiguous run of
        // subsets so that overlapping windows meet in the same L2
        const int nb = gridDim.x, b = blockIdx.x, xcd = b & 7, per = nb >> 3, rm = nb & 7;
        const int blk = xcd * per + min(xcd, rm) + (b >> 3);
        const int slot = blk * SPB + sub;
        if (slot < nlist) icgn_subset<PPL, WPS, SIX>(a, slot);  // uniform per subset
    } else {
        for (int base = blockIdx.x * SPB; base < nlist; base += gridDim.x * SPB) {
            if (base + sub < nlist) icgn_subset<PPL, WPS, SIX>(a, base + sub);
            if constexpr (WPS > 1) __syncthreads();  // the subset's LDS is free for the next
        }
    }
}

// Indices j with it[j] == code, in increasing order, and their count.  Two passes over
// runs of 4096 statuses per workgroup: k_list_cnt counts each run with ballots;
// k_list_put sums the counts of the runs before its own (every workgroup reads the
// same array) and writes its matches in order.
constexpr int kListRun = 4096;

__global__ __launch_bounds__(256) void k_list_cnt(const int *it, int S, int code, int *cnt)
{
    __shared__ int wc[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int b0 = blockIdx.x * kListRun, b1 = min(S, b0 + kListRun);
    int c = 0;
    for (int i = b0 + w * 64 + lane; i - lane < b1; i += 256)
        c += __popcll(__ballot(i < b1 && it[i] == code));
    if (lane == 0) wc[w] = c;
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

__global__ __launch_bounds__(256) void k_list_put(const int *it, int S, int code, const int *cnt,
                                                  int nrun, int *list, int *count)
{
    __shared__ int wc[4];
    __shared__ int base;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x < 64) {
        int sb = 0, st = 0;
        for (int k = threadIdx.x; k < nrun; k += 64) {
            const int v = cnt[k];
            if (k < (int)blockIdx.x) sb += v;
            st += v;
        }
        for (int m = 32; m >= 1; m >>= 1) {
            sb += __shfl_xor(sb, m, 64);
            st += __shfl_xor(st, m, 64);
        }
        if (threadIdx.x == 0) {
            base = sb;
            if (blockIdx.x == 0) *count = st;
        }
    }
    // each wave takes a contiguous quarter of the run
    const int b0 = blockIdx.x * kListRun, b1 = min(S, b0 + kListRun);
    const int q = kListRun / 4, w0 = min(b1, b0 + w * q), w1 = min(b1, w0 + q);
    int c = 0;
    for (int i = w0 + lane; i - lane < w1; i += 64) c += __popcll(__ballot(i < w1 && it[i] == code));
    if (lane == 0) wc[w] = c;
    __syncthreads();
    int o = base;
    for (int k = 0; k < w; ++k) o += wc[k];
    for (int i = w0 + lane; i - lane < w1; i += 64) {
        const bool m = i < w1 && it[i] == code;
        const unsigned long long bal = __ballot(m);
        if (m) list[o + __popcll(bal & ((1ull << lane) - 1ull))] = i;
        o += __popcll(bal);
    }
}

// scratch: int[(S + 4095) / 4096] of run counts
hipError_t launch_list_eq(const int *it, int S, int code, int *list, int *count, int *scratch,
                          hipStream_t st)
{
    const int nrun = (S + kListRun - 1) / kListRun;
    if (nrun <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_list_cnt, dim3(nrun), dim3(256), 0, st, it, S, code, scratch);
    hipLaunchKernelGGL(k_list_put, dim3(nrun), dim3(256), 0, st, it, S, code,
                       (const int *)scratch, nrun, list, count);
    return hipGetLastError();
}

int icgn_ppl_for(int ws)
{
    const int n = (ws + 1) * (ws + 1);
    if (n <= 5 * 64) return 5;      // windows up to 17 x 17
    if (n <= 8 * 64) return 8;
    if (n <= 18 * 64) return 18;
    if (n <= 18 * 256) return 72;   // 4 waves x 18 pixels
    return 0;
}

template <bool SIX>
static hipError_t launch_any(const IcgnArgs &a, int ppl, hipStream_t st)
{
    if (a.nlist <= 0) return hipSuccess;
    // a device-side length (a.nlist its upper bound) gets a capped, looping grid
    constexpr int cap = 512;
    const bool loop = a.nlist_dev != nullptr;
    if (ppl == 5) {
        if (loop)
            hipLaunchKernelGGL((k_icgn<5, 1, SIX, true>), dim3(std::min(cap, (a.nlist + 3) / 4)),
                               dim3(256), 0, st, a);
        else
            hipLaunchKernelGGL((k_icgn<5, 1, SIX, false>), dim3((a.nlist + 3) / 4), dim3(256), 0,
                               st, a);
    } else if (ppl == 8) {
        if (loop)
            hipLaunchKernelGGL((k_icgn<8, 1, SIX, true>), dim3(std::min(cap, (a.nlist + 3) / 4)),
                               dim3(256), 0, st, a);
        else
            hipLaunchKernelGGL((k_icgn<8, 1, SIX, false>), dim3((a.nlist + 3) / 4), dim3(256), 0,
                               st, a);
    } else if (ppl == 18) {
        if (loop)
            hipLaunchKernelGGL((k_icgn<18, 1, SIX, true>),
                               dim3(std::min(cap, (a.nlist + 3) / 4)), dim3(256), 0, st, a);
        else
            hipLaunchKernelGGL((k_icgn<18, 1, SIX, false>), dim3((a.nlist + 3) / 4), dim3(256), 0,
                               st, a);
    } else if (ppl == 72) {
        if (loop)
            hipLaunchKernelGGL((k_icgn<18, 4, SIX, true>), dim3(std::min(cap, a.nlist)),
                               dim3(256), 0, st, a);
        else
            hipLaunchKernelGGL((k_icgn<18, 4, SIX, false>), dim3(a.nlist), dim3(256), 0, st, a);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_icgn6(const IcgnArgs &a, int ppl, hipStream_t st) { return launch_any<true>(a, ppl, st); }
hipError_t launch_icgn2(const IcgnArgs &a, int ppl, hipStream_t st) { return launch_any<false>(a, ppl, st); }

// ---- ba_interp2 equivalent: Keys cubic with clamp-to-edge (ba_interp2.cpp:60-170) --
__global__ __launch_bounds__(256) void k_bicubic(const float *g, int nx, int ny, const double *X,
                                                 const double *Y, long n, double *Z)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = X[i], y = Y[i];
    if (!(isfinite(x) && isfinite(y))) { Z[i] = NAN; return; }
    const double xf = floor(x), yf = floor(y);
    float wx[4], wy[4];
    keys_weights((float)(x - xf), wx[0], wx[1], wx[2], wx[3]);
    keys_weights((float)(y - yf), wy[0], wy[1], wy[2], wy[3]);
    const long c0 = (long)xf - 2, r0 = (long)yf - 2;
    float acc = 0.f;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        long r = r0 + b;
        r = r < 0 ? 0 : (r >= ny ? ny - 1 : r);
        float s = 0.f;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            long cc = c0 + c;
            cc = cc < 0 ? 0 : (cc >= nx ? nx - 1 : cc);
            s += wx[c] * g[r * nx + cc];
        }
        acc += wy[b] * s;
    }
    Z[i] = (double)acc;
}

hipError_t launch_bicubic(const float *g, int nx, int ny, const double *X, const double *Y,
                          long n, double *Z, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bicubic, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, nx, ny,
                       X, Y, n, Z);
    return hipGetLastError();
}

}  // namespace aldic
