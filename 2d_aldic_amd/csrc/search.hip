// Whole-field integer search: normalised cross-correlation of every subset's
// template against its search window, argmax and the findpeak sub-pixel fit.
//
//   funIntegerSearch.m:82-200  subset loop (corner ii, jj; centre ii + w/2)
//   normxcorr2 (called :127)   Lewis-style NCC with zero-padded local sums
//   :134                       crop [w : end-w+1] -> (2s+3)^2 lags, lag -s-1 .. s+1
//   findpeak.m:1-68            first max in column-major order (x-lag fastest),
//                              3x3 quadratic fit off the crop edge, 1e-3 rounding
//
// One 256-thread workgroup per subset.  LDS holds the mean-free template T', the
// search window D with a one-pixel zero border (the partial-overlap lags of the crop
// edge), the per-band numerator partials and the window box sums.
//
// Numerator: a work item is (task group, band of template rows); each of its 64 lanes
// owns X neighbouring x-lags of one y-lag and slides an X+3 register window along the
// D row, so one D load feeds X FMAs and one T' float4 (the same address in every lane,
// an LDS broadcast) feeds 4X.  Lanes are numbered y-lag fastest (task = s nry + ly)
// and the D row stride is R = k (mod 64) with k = X nry^-1 (odd): the word lane L
// reads is then k L (mod 64), so the 64 lanes of a ds_read_b32 hit 64 different banks.
// Box sums of D and D^2 per lag run as FP64 running sums in wave 0, beside the
// numerator.  Bands are added in a fixed order, so the argmax is deterministic.
#include "aldic_internal.h"
#include "common.h"
#include "search_peak.h"

namespace aldic {

namespace {

struct SearchGeom {
    int tw, twp, nq, rem;         // template width, padded width (x4), full 4-col steps, rest
    int sx, sy, nrx, nry, nl;     // search radii, lag counts, lags
    int X, nstrip, ngroups, nb, rows_per, nitems;
    int R, DH, RW, VW;            // window row stride / rows / used columns; box-sum columns
    int off_T, off_D, off_P, off_V, off_S;  // float offsets into LDS (multiples of 4)
    int total;                    // floats
};

int round4(int v) { return (v + 3) & ~3; }

// Lays out one geometry for lag block width X (odd); returns false if it does not fit.
bool geom_for(int ws, int sx, int sy, int X, SearchGeom &G, long &cost)
{
    G.tw = ws + 1;
    G.twp = round4(G.tw);
    G.nq = G.tw / 4;
    G.rem = G.tw % 4;
    G.sx = sx;
    G.sy = sy;
    G.nrx = 2 * sx + 3;
    G.nry = 2 * sy + 3;
    G.nl = G.nrx * G.nry;
    G.X = X;
    G.nstrip = (G.nrx + X - 1) / X;
    const int ntask = G.nstrip * G.nry;
    G.ngroups = (ntask + 63) / 64;
    // bands: fewest (rounds of 4 waves) x (rows per band)
    long best = -1;
    int bestnb = 1;
    for (int nb = 1; nb <= G.tw && nb <= 32; ++nb) {
        const int rows = (G.tw + nb - 1) / nb;
        const int nbe = (G.tw + rows - 1) / rows;
        const long c = (long)((G.ngroups * nbe + 3) / 4) * (rows * G.tw + X) * X;
        if (best < 0 || c < best) { best = c; bestnb = nbe; }
    }
    G.rows_per = (G.tw + bestnb - 1) / bestnb;
    G.nb = (G.tw + G.rows_per - 1) / G.rows_per;
    G.nitems = G.ngroups * G.nb;
    cost = best;
    // window: rows ly + j, columns X*s + (0 .. X + tw - 2)
    G.DH = G.nry + G.tw - 1;
    G.RW = X * G.nstrip + G.tw - 1;
    G.VW = G.nrx + G.tw - 1;
    int inv = 1;
    while ((inv * G.nry) % 64 != 1) inv += 2;  // nry is odd
    const int c = (X * inv) % 64;
    G.R = G.RW + (((c - G.RW) % 64) + 64) % 64;
    G.off_T = 0;
    G.off_D = G.off_T + G.tw * G.twp;
    G.off_P = round4(G.off_D + G.DH * G.R);
    G.off_V = round4(G.off_P + G.nb * G.nl);
    G.off_S = round4(G.off_V + 2 * G.nry * G.VW);
    G.total = G.off_S + 2 * G.nl;
    return sizeof(float) * (size_t)G.total <= 160 * 1024 - 1024;
}

SearchGeom search_geom(int ws, int sx, int sy, bool &ok)
{
    SearchGeom best{};
    long bc = -1;
    ok = false;
    for (int X : {7, 5, 9, 3}) {
        SearchGeom G;
        long c;
        if (!geom_for(ws, sx, sy, X, G, c)) continue;
        if (bc < 0 || c < bc) { bc = c; best = G; }
        ok = true;
    }
    return best;
}

}  // namespace

size_t search_lds_bytes(int ws, int sx, int sy)
{
    bool ok;
    const SearchGeom G = search_geom(ws, sx, sy, ok);
    return ok ? sizeof(float) * (size_t)G.total + 64 : (size_t)1 << 30;
}

// One 4-column step of the register-window correlation.  w is a 12-entry ring of D
// values indexed by column mod 12 (the caller keeps every block start a multiple of 12
// columns); O is the step's column offset inside the block, NC its real columns.
template <int X, int O, int NC>
__device__ __forceinline__ void corr_step(float (&w)[12], float (&acc)[X], const float *dr,
                                          const float4 *t4)
{
#pragma unroll
    for (int k = 0; k < NC; ++k) w[(O + X - 1 + k) % 12] = dr[O + X - 1 + k];
    const float4 tv = t4[O / 4];
    const float tc[4] = {tv.x, tv.y, tv.z, tv.w};
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int x = 0; x < X; ++x) acc[x] = fmaf(tc[c], w[(O + x + c) % 12], acc[x]);
}

template <int X, int O>
__device__ __forceinline__ void corr_rest(int rem, float (&w)[12], float (&acc)[X],
                                          const float *dr, const float4 *t4)
{
    if (rem == 1) corr_step<X, O, 1>(w, acc, dr, t4);
    else if (rem == 2) corr_step<X, O, 2>(w, acc, dr, t4);
    else if (rem == 3) corr_step<X, O, 3>(w, acc, dr, t4);
}

// acc[x] += sum_i T'[i] * D[x + i] over one template row (tw = 4 nq + rem columns).
template <int X>
__device__ __forceinline__ void corr_row(float (&acc)[X], const float *dr, const float *tr,
                                         int nq, int rem)
{
    float w[12];
#pragma unroll
    for (int k = 0; k < X - 1; ++k) w[k] = dr[k];
    const float4 *t4 = reinterpret_cast<const float4 *>(tr);
    int q = 0;
    for (; q + 3 <= nq; q += 3) {
        corr_step<X, 0, 4>(w, acc, dr, t4);
        corr_step<X, 4, 4>(w, acc, dr, t4);
        corr_step<X, 8, 4>(w, acc, dr, t4);
        dr += 12;
        t4 += 3;
    }
    const int tail = nq - q;
    if (tail == 0) {
        corr_rest<X, 0>(rem, w, acc, dr, t4);
    } else if (tail == 1) {
        corr_step<X, 0, 4>(w, acc, dr, t4);
        corr_rest<X, 4>(rem, w, acc, dr, t4);
    } else {
        corr_step<X, 0, 4>(w, acc, dr, t4);
        corr_step<X, 4, 4>(w, acc, dr, t4);
        corr_rest<X, 8>(rem, w, acc, dr, t4);
    }
}

template <int X>
__global__ __launch_bounds__(256) void k_search(SearchArgs a, SearchGeom G)
{
    extern __shared__ __align__(16) float lds[];
    __shared__ double redd[4][2];
    __shared__ float redf[4][2];
    __shared__ float rv[4];
    __shared__ int ri[4];
    float *Tm = lds + G.off_T;
    float *Dp = lds + G.off_D;
    float *Pn = lds + G.off_P;
    float *V1 = lds + G.off_V;
    float *V2 = V1 + G.nry * G.VW;
    float *S1 = lds + G.off_S;
    float *S2 = S1 + G.nl;

    // subsets in contiguous runs per XCD (workgroups are dealt to the 8 XCDs in turn),
    // so neighbouring windows share an L2
    const int S = a.Nx * a.Ny;
    const int b = blockIdx.x, xcd = b & 7, per = S >> 3, rm = S & 7;
    int q = xcd * per + min(xcd, rm) + (b >> 3);
    if (a.list) {  // subset list: workgroup b takes entry b (whole workgroups leave)
        if (b >= *a.nlist_dev) return;
        q = a.list[b];
    }
    const int k = q % a.Nx, l = q / a.Nx;
    const int ii = a.x1 + k * a.step, jj = a.y1 + l * a.step;  // 1-based corner
    const int tw = G.tw, twp = G.twp, nx = a.nx;
    const int t = threadIdx.x, wv = t >> 6, lane = t & 63;

    // template (raw) and its statistics; window with a zero border:
    //   Dp[r][c] = g(ii - sx - 1 + c, jj - sy - 1 + r)
    double s1 = 0.0, s2 = 0.0;
    float mn = INFINITY, mx = -INFINITY;
    for (int r = wv; r < tw; r += 4) {
        const float *src = a.f + (size_t)(jj - 1 + r) * nx + (ii - 1);
        for (int c = lane; c < twp; c += 64) {
            float v = 0.f;
            if (c < tw) {
                v = src[c];
                s1 += v;
                s2 += (double)v * v;
                mn = fminf(mn, v);
                mx = fmaxf(mx, v);
            }
            Tm[r * twp + c] = v;
        }
    }
    const int Dx = tw + 2 * G.sx, Dy = tw + 2 * G.sy;
    for (int r = wv; r < G.DH; r += 4) {
        const bool rin = r >= 1 && r <= Dy;
        const float *src = a.g + (size_t)(jj - G.sy - 2 + r) * nx + (ii - G.sx - 2);
        for (int c = lane; c < G.RW; c += 64)
            Dp[r * G.R + c] = (rin && c >= 1 && c <= Dx) ? src[c] : 0.f;
    }
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        mn = fminf(mn, __shfl_xor(mn, m, 64));
        mx = fmaxf(mx, __shfl_xor(mx, m, 64));
    }
    if (lane == 0) {
        redd[wv][0] = s1;
        redd[wv][1] = s2;
        redf[wv][0] = mn;
        redf[wv][1] = mx;
    }
    __syncthreads();
    s1 = (redd[0][0] + redd[1][0]) + (redd[2][0] + redd[3][0]);
    s2 = (redd[0][1] + redd[1][1]) + (redd[2][1] + redd[3][1]);
    mn = fminf(fminf(redf[0][0], redf[1][0]), fminf(redf[2][0], redf[3][0]));
    mx = fmaxf(fmaxf(redf[0][1], redf[1][1]), fmaxf(redf[2][1], redf[3][1]));

    const double xc = (ii + ii + a.ws) / 2.0, yc = (jj + jj + a.ws) / 2.0;
    if (mn == mx) {  // constant template: normxcorr2 raises (funIntegerSearch.m:153-157)
        if (t == 0) {
            a.x[q] = xc; a.y[q] = yc;
            a.u[q] = a.v[q] = a.ccmax[q] = NAN;
            a.status[q] = -1;
            if (a.qf) { a.qf[2 * q] = INFINITY; a.qf[2 * q + 1] = INFINITY; }
            if (a.U0) { a.U0[2 * q] = NAN; a.U0[2 * q + 1] = NAN; }
        }
        return;
    }
    const double n = (double)tw * tw;
    const double mt = s1 / n;
    const double st = fmax(s2 - s1 * mt, 0.0);
    for (int r = wv; r < tw; r += 4)
        for (int c = lane; c < tw; c += 64) Tm[r * twp + c] = (float)((double)Tm[r * twp + c] - mt);
    __syncthreads();

    if (wv == 0) {
        // vertical running sums of D and D^2 over tw rows, one column per lane (FP64)
        for (int c = lane; c < G.VW; c += 64) {
            double v1 = 0.0, v2 = 0.0;
            for (int j = 0; j < tw; ++j) {
                const double d = Dp[j * G.R + c];
                v1 += d;
                v2 += d * d;
            }
            V1[c] = (float)v1;
            V2[c] = (float)v2;
            for (int ly = 1; ly < G.nry; ++ly) {
                const double dout = Dp[(ly - 1) * G.R + c], din = Dp[(ly - 1 + tw) * G.R + c];
                v1 += din - dout;
                v2 += din * din - dout * dout;
                V1[ly * G.VW + c] = (float)v1;
                V2[ly * G.VW + c] = (float)v2;
            }
        }
        wave_sync_lds();
        // horizontal running sums of the vertical sums -> box sums per lag
        for (int idx = lane; idx < 2 * G.nry; idx += 64) {
            const int qn = idx >= G.nry, ly = idx - qn * G.nry;
            const float *vr = (qn ? V2 : V1) + ly * G.VW;
            float *so = (qn ? S2 : S1) + ly * G.nrx;
            double h = 0.0;
            for (int i = 0; i < tw; ++i) h += vr[i];
            so[0] = (float)h;
            for (int lx = 1; lx < G.nrx; ++lx) {
                h += (double)vr[lx - 1 + tw] - (double)vr[lx - 1];
                so[lx] = (float)h;
            }
        }
    }

    // numerator items; wave 0 (box sums) takes the last, shortest band
    for (int it = G.nitems - 1 - wv; it >= 0; it -= 4) {
        const int grp = it % G.ngroups, band = it / G.ngroups;
        const int tau = grp * 64 + lane;
        int s = tau / G.nry;
        const int ly = tau - s * G.nry;
        const bool act = s < G.nstrip;
        if (!act) s = G.nstrip - 1;  // in-range reads, result dropped
        const int j0 = band * G.rows_per, j1 = min(tw, j0 + G.rows_per);
        float acc[X];
#pragma unroll
        for (int x = 0; x < X; ++x) acc[x] = 0.f;
        for (int j = j0; j < j1; ++j)
            corr_row<X>(acc, Dp + (ly + j) * G.R + X * s, Tm + j * twp, G.nq, G.rem);
        if (act) {
            float *pp = Pn + band * G.nl + ly * G.nrx;
#pragma unroll
            for (int x = 0; x < X; ++x)
                if (X * s + x < G.nrx) pp[X * s + x] = acc[x];
        }
    }
    __syncthreads();

    // NCC per lag, then argmax (ties -> smallest column-major index)
    const int nl = G.nl;
    float best = -INFINITY;
    int bi = 0x7fffffff;
    for (int p = t; p < nl; p += 256) {
        double num = 0.0;
        for (int bb = 0; bb < G.nb; ++bb) num += Pn[bb * nl + p];
        const double b1 = S1[p], b2 = S2[p];
        double var = b2 - b1 * b1 / n;
        var = var > 0.0 ? var : 0.0;
        const double den = sqrt(st * var);
        const float c = den > 0.0 ? (float)(num / den) : 0.f;
        if (c > best || (c == best && p < bi)) { best = c; bi = p; }
        Pn[p] = c;  // band 0 slot now holds the NCC value of lag p (only this thread read it)
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const float ob = __shfl_xor(best, m, 64);
        const int oi = __shfl_xor(bi, m, 64);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    if (lane == 0) { rv[wv] = best; ri[wv] = bi; }
    __syncthreads();
    if (t != 0) return;
    for (int w = 1; w < 4; ++w)
        if (rv[w] > best || (rv[w] == best && ri[w] < bi)) { best = rv[w]; bi = ri[w]; }

    double xp, yp, mf;
    findpeak_fit(Pn, G.nrx, G.nry, bi, best, xp, yp, mf);
    const double u = yp - (G.sx + 2), v = xp - (G.sy + 2);
    a.x[q] = xc;
    a.y[q] = yc;
    a.u[q] = u;
    a.v[q] = v;
    a.ccmax[q] = mf;
    a.status[q] = 0;
    if (a.qf) { a.qf[2 * q] = NAN; a.qf[2 * q + 1] = NAN; }
    if (a.U0) { a.U0[2 * q] = u; a.U0[2 * q + 1] = v; }
}

template <int X>
hipError_t launch_x(const SearchArgs &a, const SearchGeom &G, int S, hipStream_t st)
{
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)k_search<X>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
        attr_set = true;
    }
    hipLaunchKernelGGL(k_search<X>, dim3(S), dim3(256), sizeof(float) * (size_t)G.total, st, a,
                       G);
    return hipGetLastError();
}

hipError_t launch_search(const SearchArgs &a, hipStream_t st)
{
    const int S = a.Nx * a.Ny;
    if (S <= 0) return hipSuccess;
    if (a.ms && a.planes && !(a.f8 && a.g8 && search_i8_supported(a.ws, a.sx, a.sy)) &&
        ((a.f8 && a.g8 && search_mt_supported(a.ws, a.sx, a.sy, 8)) ||
         (a.f16 && a.g16 && search_mt_supported(a.ws, a.sx, a.sy, 16))))
        return launch_search_mt(a, st);  // records ev_kernel after its box-sum planes
    if (a.ev_kernel) {
        const hipError_t e = hipEventRecord(a.ev_kernel, st);
        if (e != hipSuccess) return e;
    }
    if (a.f8 && a.g8 && a.ms && search_i8_supported(a.ws, a.sx, a.sy))
        return launch_search_i8(a, st);
    bool ok;
    const SearchGeom G = search_geom(a.ws, a.sx, a.sy, ok);
    if (!ok) return hipErrorInvalidValue;
    switch (G.X) {
    case 3: return launch_x<3>(a, G, S, st);
    case 5: return launch_x<5>(a, G, S, st);
    case 7: return launch_x<7>(a, G, S, st);
    default: return launch_x<9>(a, G, S, st);
    }
}

hipError_t launch_search_list(const SearchArgs &a, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (!a.list || !a.nlist_dev || n > a.Nx * a.Ny) return hipErrorInvalidValue;
    bool ok;
    const SearchGeom G = search_geom(a.ws, a.sx, a.sy, ok);
    if (!ok) return hipErrorInvalidValue;
    switch (G.X) {
    case 3: return launch_x<3>(a, G, n, st);
    case 5: return launch_x<5>(a, G, n, st);
    case 7: return launch_x<7>(a, G, n, st);
    default: return launch_x<9>(a, G, n, st);
    }
}

}  // namespace aldic
