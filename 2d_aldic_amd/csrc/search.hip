// Whole-field integer search: normalised cross-correlation of every subset's
// template against its search window, argmax and the findpeak sub-pixel fit.
//
//   funIntegerSearch.m:82-200  subset loop (corner ii, jj; centre ii + w/2)
//   normxcorr2 (called :127)   Lewis-style NCC with zero-padded local sums
//   :134                       crop [w : end-w+1] -> (2s+3)^2 lags, lag -s-1 .. s+1
//   findpeak.m:1-68            first max in column-major order (x-lag fastest),
//                              3x3 quadratic fit off the crop edge, 1e-3 rounding
//
// One 256-thread workgroup per subset.  LDS holds the mean-free template T'
// ((w+1)^2), the search window D with a one-pixel zero border (the partial-overlap
// lags of the crop edge), vertical box sums of D and D^2 and the lag map.  Each
// work item accumulates 4 neighbouring x-lags over a band of template rows, sliding
// a 4-wide register window along the D row so every T' value feeds 4 FMAs; bands
// are summed in a fixed order (deterministic argmax).
#include "aldic_internal.h"
#include "common.h"

namespace aldic {

struct SearchGeom {
    int tw, twp, sx, sy, nrx, nry, nstrip, nyp, DW, DH, K, rows_per;
    int off_T, off_D, off_V, off_V2, off_P;  // float offsets into LDS (multiples of 4)
    int total;                               // floats
};

static int round4(int v) { return (v + 3) & ~3; }

static SearchGeom search_geom(int ws, int sx, int sy)
{
    SearchGeom G;
    G.tw = ws + 1;
    G.twp = round4(G.tw);          // template row stride, zero-padded to a float4 multiple
    G.sx = sx;
    G.sy = sy;
    G.nrx = 2 * sx + 3;
    G.nry = 2 * sy + 3;
    G.nstrip = (G.nrx + 3) / 4;    // 4 x-lags per work item
    G.nyp = (G.nry + 1) / 2;       // 2 y-lags per work item
    G.DW = G.nstrip * 4 + G.twp;   // window row stride: a multiple of 4 floats, and an
    if ((G.DW / 4) % 2 == 0) G.DW += 4;  // odd number of float4s so that neighbouring lanes'
                                         // rows fall on different LDS banks
    G.DH = G.tw + G.nry;           // one spare zero row for an odd lag count
    // Split the template rows into K bands so the 256 threads stay busy: pick the K
    // with the fewest (rounds of 256 items) x (rows per band + 2), keeping the band
    // partials within a 64 KB LDS budget when possible.
    const int base = G.nyp * G.nstrip;
    const int fixed = G.tw * G.twp + G.DH * G.DW + 2 * round4(G.nry * G.DW);
    int bestK = 1;
    long best = -1;
    for (int K = 1; K <= G.tw; ++K) {
        const int rows = (G.tw + K - 1) / K;
        const int Keff = (G.tw + rows - 1) / rows;
        const long lds = 4L * (fixed + (long)Keff * G.nrx * G.nry);
        if (K > 1 && lds > 64 * 1024) break;
        const long rounds = (base * (long)Keff + 255) / 256;
        const long cost = rounds * (rows + 2);
        if (best < 0 || cost < best) { best = cost; bestK = K; }
    }
    G.rows_per = (G.tw + bestK - 1) / bestK;
    G.K = (G.tw + G.rows_per - 1) / G.rows_per;
    G.off_T = 0;
    G.off_D = G.off_T + G.tw * G.twp;
    G.off_V = G.off_D + G.DH * G.DW;
    G.off_V2 = G.off_V + round4(G.nry * G.DW);
    G.off_P = G.off_V2 + round4(G.nry * G.DW);
    G.total = G.off_P + G.K * G.nrx * G.nry;
    return G;
}

size_t search_lds_bytes(int ws, int sx, int sy)
{
    return sizeof(float) * (size_t)search_geom(ws, sx, sy).total + 64;
}

__device__ __forceinline__ double block_sum256(double v, double *red)
{
    v = wave_sum(v);
    const int wid = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[wid] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void k_search(SearchArgs a, SearchGeom G)
{
    extern __shared__ __align__(16) float lds[];
    __shared__ double red[4];
    __shared__ float rv[4];
    __shared__ int ri[4];
    float *Tm = lds + G.off_T;
    float *Dp = lds + G.off_D;
    float *V = lds + G.off_V;
    float *V2 = lds + G.off_V2;
    float *Pn = lds + G.off_P;

    const int q = blockIdx.x;
    const int k = q % a.Nx, l = q / a.Nx;
    const int ii = a.x1 + k * a.step, jj = a.y1 + l * a.step;  // 1-based corner
    const int tw = G.tw, nx = a.nx;
    const int t = threadIdx.x;

    // template and its mean
    const int twp = G.twp;
    double s = 0.0;
    for (int p = t; p < tw * twp; p += 256) {
        const int r = p / twp, c = p % twp;
        const float v = c < tw ? a.f[(size_t)(jj - 1 + r) * nx + (ii - 1 + c)] : 0.f;
        Tm[p] = v;
        s += v;
    }
    s = block_sum256(s, red);
    const double mt = s / (tw * tw);
    double st = 0.0;
    for (int p = t; p < tw * twp; p += 256) {
        const float d = (p % twp) < tw ? (float)((double)Tm[p] - mt) : 0.f;  // padding stays 0
        Tm[p] = d;
        st += (double)d * d;
    }
    st = block_sum256(st, red);

    // window with a zero border: Dp[r][c] = g(ii - sx - 1 + c, jj - sy - 1 + r)
    const int Dx = tw + 2 * G.sx, Dy = tw + 2 * G.sy;
    for (int p = t; p < G.DH * G.DW; p += 256) {
        const int r = p / G.DW, c = p % G.DW;
        float v = 0.f;
        if (r >= 1 && r <= Dy && c >= 1 && c <= Dx)
            v = a.g[(size_t)(jj - G.sy - 2 + r) * nx + (ii - G.sx - 2 + c)];
        Dp[p] = v;
    }
    __syncthreads();

    const double xc = (ii + ii + a.ws) / 2.0, yc = (jj + jj + a.ws) / 2.0;
    if (st == 0.0) {  // constant template: normxcorr2 raises (funIntegerSearch.m:153-157)
        if (t == 0) {
            a.x[q] = xc; a.y[q] = yc;
            a.u[q] = a.v[q] = a.ccmax[q] = NAN;
            a.status[q] = -1;
            if (a.qf) { a.qf[2 * q] = INFINITY; a.qf[2 * q + 1] = INFINITY; }
            if (a.U0) { a.U0[2 * q] = NAN; a.U0[2 * q + 1] = NAN; }
        }
        return;
    }

    // vertical box sums over tw rows, for every lag row and column
    for (int p = t; p < G.nry * G.DW; p += 256) {
        const int ly = p / G.DW, c = p % G.DW;
        float s1 = 0.f, s2 = 0.f;
        for (int j = 0; j < tw; ++j) {
            const float d = Dp[(ly + j) * G.DW + c];
            s1 += d;
            s2 += d * d;
        }
        V[p] = s1;
        V2[p] = s2;
    }

    // numerator: item -> (y-lag pair fastest, strip of 4 x-lags, band of template rows).
    // Per 4 template columns one float4 of T' and one float4 from each of the two window
    // rows feed 32 FMAs (the 8-wide window slides in registers).
    const int nitems = G.nyp * G.nstrip * G.K;
    for (int it = t; it < nitems; it += 256) {
        const int lyp = it % G.nyp;
        const int rest = it / G.nyp;
        const int strip = rest % G.nstrip, band = rest / G.nstrip;
        const int lx0 = strip * 4, ly = lyp, ly2 = lyp + G.nyp;  // the item's two y-lags
        const int j0 = band * G.rows_per;
        const int j1 = min(tw, j0 + G.rows_per);
        float c0[4] = {0.f, 0.f, 0.f, 0.f}, c1[4] = {0.f, 0.f, 0.f, 0.f};
        for (int j = j0; j < j1; ++j) {
            const float4 *trow = reinterpret_cast<const float4 *>(Tm + j * twp);
            const float4 *d0 = reinterpret_cast<const float4 *>(Dp + (ly + j) * G.DW + lx0);
            const float4 *d1 = reinterpret_cast<const float4 *>(Dp + (ly2 + j) * G.DW + lx0);
            float4 a0 = d0[0], b0 = d1[0];
            for (int q = 0; q < twp / 4; ++q) {
                const float4 tv = trow[q];
                const float4 a1 = d0[q + 1], b1 = d1[q + 1];
                c0[0] += tv.x * a0.x + tv.y * a0.y + tv.z * a0.z + tv.w * a0.w;
                c0[1] += tv.x * a0.y + tv.y * a0.z + tv.z * a0.w + tv.w * a1.x;
                c0[2] += tv.x * a0.z + tv.y * a0.w + tv.z * a1.x + tv.w * a1.y;
                c0[3] += tv.x * a0.w + tv.y * a1.x + tv.z * a1.y + tv.w * a1.z;
                c1[0] += tv.x * b0.x + tv.y * b0.y + tv.z * b0.z + tv.w * b0.w;
                c1[1] += tv.x * b0.y + tv.y * b0.z + tv.z * b0.w + tv.w * b1.x;
                c1[2] += tv.x * b0.z + tv.y * b0.w + tv.z * b1.x + tv.w * b1.y;
                c1[3] += tv.x * b0.w + tv.y * b1.x + tv.z * b1.y + tv.w * b1.z;
                a0 = a1;
                b0 = b1;
            }
        }
        float *pp = Pn + band * G.nrx * G.nry;
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            if (lx0 + l >= G.nrx) continue;
            pp[ly * G.nrx + lx0 + l] = c0[l];
            if (ly2 < G.nry) pp[ly2 * G.nrx + lx0 + l] = c1[l];
        }
    }
    __syncthreads();

    // NCC per lag, then argmax (ties -> smallest column-major index)
    const int nl = G.nrx * G.nry;
    const float n = (float)(tw * tw);
    const float stf = (float)st;
    float best = -INFINITY;
    int bi = 0x7fffffff;
    for (int p = t; p < nl; p += 256) {
        const int ly = p / G.nrx, lx = p % G.nrx;
        float num = 0.f;
        for (int b = 0; b < G.K; ++b) num += Pn[b * nl + p];
        float s1 = 0.f, s2 = 0.f;
        const float *v1 = V + ly * G.DW + lx;
        const float *v2 = V2 + ly * G.DW + lx;
        for (int i = 0; i < tw; ++i) { s1 += v1[i]; s2 += v2[i]; }
        float var = s2 - s1 * s1 / n;
        var = var > 0.f ? var : 0.f;
        const float den = sqrtf(stf * var);
        const float c = den > 0.f ? num / den : 0.f;
        Pn[p] = c;  // band 0 slot now holds the NCC value of lag p
        if (c > best || (c == best && p < bi)) { best = c; bi = p; }
    }
    // wave argmax
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const float ob = __shfl_xor(best, m, 64);
        const int oi = __shfl_xor(bi, m, 64);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    if ((t & 63) == 0) { rv[t >> 6] = best; ri[t >> 6] = bi; }
    __syncthreads();
    if (t != 0) return;
    for (int w = 1; w < 4; ++w)
        if (rv[w] > best || (rv[w] == best && ri[w] < bi)) { best = rv[w]; bi = ri[w]; }

    // findpeak.m:30-66 (row = x-lag, column = y-lag, both 1-based)
    const int r = bi % G.nrx + 1, c = bi / G.nrx + 1;
    double xp = c, yp = r, mf = best;
    if (!(c == 1 || c == G.nry || r == 1 || r == G.nrx)) {
        double s0 = 0, sxu = 0, syu = 0, sxy = 0, sxx = 0, syy = 0;
        for (int dc = -1; dc <= 1; ++dc)
            for (int dr = -1; dr <= 1; ++dr) {
                const double u = Pn[(c - 1 + dc) * G.nrx + (r - 1 + dr)];
                s0 += u; sxu += dc * u; syu += dr * u; sxy += dc * dr * u;
                sxx += dc * dc * u; syy += dr * dr * u;
            }
        const double A1 = (5.0 * s0 - 3.0 * sxx - 3.0 * syy) / 9.0;
        const double A2 = sxu / 6.0, A3 = syu / 6.0, A4 = sxy / 4.0;
        const double A5 = sxx / 2.0 - s0 / 3.0, A6 = syy / 2.0 - s0 / 3.0;
        const double den = A4 * A4 - 4.0 * A5 * A6;
        double xo = (-A3 * A4 + 2.0 * A6 * A2) / den;
        double yo = -1.0 / den * (A4 * A2 - 2.0 * A5 * A3);
        if (!(fabs(xo) > 1.0 || fabs(yo) > 1.0) && !isnan(xo) && !isnan(yo)) {
            xo = xo < 0 ? -floor(-1000.0 * xo + 0.5) / 1000.0 : floor(1000.0 * xo + 0.5) / 1000.0;
            yo = yo < 0 ? -floor(-1000.0 * yo + 0.5) / 1000.0 : floor(1000.0 * yo + 0.5) / 1000.0;
            xp = c + xo;
            yp = r + yo;
            mf = A1 + A2 * xo + A3 * yo + A4 * xo * yo + A5 * xo * xo + A6 * yo * yo;
        }
    }
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

hipError_t launch_search(const SearchArgs &a, hipStream_t st)
{
    const int S = a.Nx * a.Ny;
    if (S <= 0) return hipSuccess;
    const SearchGeom G = search_geom(a.ws, a.sx, a.sy);
    const size_t lds = sizeof(float) * (size_t)G.total;
    if (lds > 160 * 1024 - 256) return hipErrorInvalidValue;
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)k_search,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 256);
        attr_set = true;
    }
    hipLaunchKernelGGL(k_search, dim3(S), dim3(256), lds, st, a, G);
    return hipGetLastError();
}

}  // namespace aldic
