// Multigrid integer search (funIntegerSearchMg.m:3-285, IntegerSearch.m:120): the device
// kernels; the level bookkeeping (splitting, border clamps) runs on the host in api.hip.
//
// Level 1 correlates the whole ROI of f with the same ROI of g: normxcorr2 in full, the
// numerator sum (T - mean T) D by FFT (FP64 radix-2, rows then columns, twiddle table),
// the window sums of D and D^2 by 2-D prefix sums, then the first maximum of the map in
// MATLAB's column-major order (x fastest) and findpeak's fit (search_peak.h).
// Later levels correlate each region of f with its g window (+-5 px) on the 13 x 14 lags
// findpeak reads (offsets -1..11 in x, -1..12 in y): tiles of 32 x 16 template pixels
// accumulate sum T D, sum D, sum D^2 per lag in FP64 (D zero outside the window), one
// workgroup per region sums its tiles in order and fits the peak.
// NCC = (sum T D - mean T sum D) / sqrt(st * max(sum D^2 - (sum D)^2 / n, 0)), 0 where the
// denominator is 0 (the oracle's definition of normxcorr2).
#include <algorithm>
#include <climits>
#include <cmath>

#include "aldic_internal.h"
#include "common.h"
#include "search_peak.h"

namespace aldic {

namespace {

constexpr int kFftThreads = 1024;
constexpr int kMgTX = 32, kMgTY = 16;          // template tile
constexpr int kMgLX = 13, kMgLY = 14;          // lags: offsets -1..11 (x), -1..12 (y)
constexpr int kMgLags = kMgLX * kMgLY;         // 182
constexpr int kMgDX = kMgTX + kMgLX - 1, kMgDY = kMgTY + kMgLY - 1;   // 44 x 29

__global__ void k_twiddles(double2 *tw, int N)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N / 2) return;
    double s, c;
    sincospi(-2.0 * (double)k / (double)N, &s, &c);
    tw[k] = make_double2(c, s);
}

template <bool INV>
__device__ __forceinline__ double2 tw_at(const double2 *__restrict__ tw, int i)
{
    const double2 w = tw[i];
    return INV ? make_double2(w.x, -w.y) : w;
}

// radix-2 butterfly: (a, b) <- (a + w b, a - w b)
__device__ __forceinline__ void bfly(double2 &a, double2 &b, double2 w)
{
    const double2 wb = make_double2(b.x * w.x - b.y * w.y, b.x * w.y + b.y * w.x);
    b = make_double2(a.x - wb.x, a.y - wb.y);
    a = make_double2(a.x + wb.x, a.y + wb.y);
}

// One FFT of length N = 2^logN per workgroup, in LDS: element i of line `line` at
// data[line * lstride + i * estride]; bit-reversed load, radix-2 decimation in time,
// twiddles exp(-2 pi i k / N) = tw[k * tmul] of the table of length L = N * tmul
// (conjugated for the inverse), no scaling.
template <bool INV>
__global__ __launch_bounds__(kFftThreads) void k_fft(double2 *data, const double2 *__restrict__ tw, int N,
                                                    int logN, int tmul, long lstride, long estride)
{
    extern __shared__ double2 s[];
    double2 *p = data + (long)blockIdx.x * lstride;
    for (int i = threadIdx.x; i < N; i += kFftThreads) {
        const int j = (int)(__brev((unsigned)i) >> (32 - logN));
        s[j] = p[(long)i * estride];
    }
    __syncthreads();
    // stages of half-length h: two at a time in registers (lengths 2h and 4h: each thread
    // owns elements i0, i0 + h, i0 + 2h, i0 + 3h), a single one when one is left
    for (int h = 1; h < N;) {
        if (2 * h < N) {
            const int sa = N / (2 * h) * tmul, sb = N / (4 * h) * tmul;
            for (int t = threadIdx.x; t < N / 4; t += kFftThreads) {
                const int k = t & (h - 1), i0 = (t / h) * 4 * h + k;
                double2 x0 = s[i0], x1 = s[i0 + h], x2 = s[i0 + 2 * h], x3 = s[i0 + 3 * h];
                const double2 wa = tw_at<INV>(tw, k * sa), w0 = tw_at<INV>(tw, k * sb),
                              w1 = tw_at<INV>(tw, (k + h) * sb);
                bfly(x0, x1, wa);
                bfly(x2, x3, wa);
                bfly(x0, x2, w0);
                bfly(x1, x3, w1);
                s[i0] = x0;
                s[i0 + h] = x1;
                s[i0 + 2 * h] = x2;
                s[i0 + 3 * h] = x3;
            }
            h *= 4;
        } else {
            const int sa = N / (2 * h) * tmul;
            for (int t = threadIdx.x; t < N / 2; t += kFftThreads) {
                const int k = t & (h - 1), i0 = (t / h) * 2 * h + k;
                double2 x0 = s[i0], x1 = s[i0 + h];
                bfly(x0, x1, tw_at<INV>(tw, k * sa));
                s[i0] = x0;
                s[i0 + h] = x1;
            }
            h *= 2;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < N; i += kFftThreads) p[(long)i * estride] = s[i];
}

// The level-1 operands packed in one complex array, Z = T' + i D: T' = f - mean (ROI of m x
// n pixels from (x1, y1), 1-based) and D = g on the same ROI, zero-padded to P x Q ([y][x],
// x fastest).  Both are real, so one transform gives both spectra.
__global__ void k_mg_operands(const float *__restrict__ f, const float *__restrict__ g, int nx, int x1,
                              int y1, int m, int n, const double *__restrict__ tmean, int P, int Q,
                              double2 *__restrict__ Z)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)P * Q) return;
    const int xx = (int)(i % P), yy = (int)(i / P);
    double t = 0.0, d = 0.0;
    if (xx < m && yy < n) {
        const long o = (long)(y1 - 1 + yy) * nx + (x1 - 1 + xx);
        t = (double)f[o] - *tmean;
        d = (double)g[o];
    }
    Z[i] = make_double2(t, d);
}

// From Z = FFT(T' + i D): FT(k) = (Z(k) + conj Z(-k)) / 2, FD(k) = (Z(k) - conj Z(-k)) / 2i;
// Z(k) <- conj(FT(k)) FD(k), and Z(-k) <- its conjugate (the correlation is real).  Z is
// A x B (B fastest, both powers of two); the thread of the smaller index of each pair
// does both.
__global__ void k_mg_cross(double2 *__restrict__ Z, int A, int B)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)A * B) return;
    const int kb = (int)(i % B), ka = (int)(i / B);
    const long j = (long)((A - ka) & (A - 1)) * B + ((B - kb) & (B - 1));
    if (j < i) return;
    const double2 a = Z[i], b = Z[j];
    // FT = ((a.x + b.x) + i (a.y - b.y)) / 2, FD = ((a.y + b.y) - i (a.x - b.x)) / 2
    const double tr = 0.5 * (a.x + b.x), ti = 0.5 * (a.y - b.y);
    const double dr = 0.5 * (a.y + b.y), di = -0.5 * (a.x - b.x);
    const double rr = tr * dr + ti * di, ri = tr * di - ti * dr;
    Z[i] = make_double2(rr, ri);
    if (j != i) Z[j] = make_double2(rr, -ri);
}

// out[c * rows + r] = in[r * cols + c], 32 x 32 tiles through LDS (both sides coalesced)
__global__ __launch_bounds__(256) void k_transpose(const double2 *__restrict__ in, double2 *__restrict__ out,
                                                   int rows, int cols)
{
    __shared__ double2 t[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) t[k][tx] = in[(long)r * cols + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) out[(long)c * rows + r] = t[tx][k];
    }
}

// sum, sum of squares and largest |v - first pixel| of the ROI of f (blocks of partials,
// then one block)
__global__ __launch_bounds__(256) void k_mg_tsum(const float *__restrict__ f, int nx, int x1, int y1, int m,
                                                 int n, double *__restrict__ part)
{
    __shared__ double r0[4], r1[4], r2[4];
    const double v0 = f[(long)(y1 - 1) * nx + (x1 - 1)];
    double s = 0.0, s2 = 0.0, dv = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (long)m * n; i += (long)gridDim.x * blockDim.x) {
        const double v = f[(long)(y1 - 1 + i / m) * nx + (x1 - 1 + i % m)];
        s += v;
        s2 += v * v;
        dv = fmax(dv, fabs(v - v0));
    }
    s = wave_sum(s);
    s2 = wave_sum(s2);
    for (int d = 32; d >= 1; d >>= 1) dv = fmax(dv, __shfl_xor(dv, d, kWave));
    if ((threadIdx.x & 63) == 0) { r0[threadIdx.x >> 6] = s; r1[threadIdx.x >> 6] = s2; r2[threadIdx.x >> 6] = dv; }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = r0[0] + r0[1] + r0[2] + r0[3];
        part[3 * blockIdx.x + 1] = r1[0] + r1[1] + r1[2] + r1[3];
        part[3 * blockIdx.x + 2] = fmax(fmax(r2[0], r2[1]), fmax(r2[2], r2[3]));
    }
}

// st = {mean, sum of squared deviations, largest |v - first pixel| (0: constant template)}
__global__ void k_mg_tsum_final(const double *__restrict__ part, int nb, long cnt, double *__restrict__ st)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0, s2 = 0.0, dv = 0.0;
    for (int b = 0; b < nb; ++b) { s += part[3 * b]; s2 += part[3 * b + 1]; dv = fmax(dv, part[3 * b + 2]); }
    st[0] = s / (double)cnt;
    st[1] = s2 - s * s / (double)cnt;
    st[2] = dv;
}

// Prefix sums S[y][x] = sum over x' < x, y' < y of D (and D^2) on the (m + 1) x (n + 1)
// grid.  Rows: one wave per row, 64 pixels a step (coalesced), wave inclusive scan plus
// the carry; row 0 is zero.
__global__ __launch_bounds__(256) void k_mg_prefix_rows(const float *__restrict__ g, int nx, int x1, int y1,
                                                        int m, int n, double *__restrict__ S1,
                                                        double *__restrict__ S2)
{
    const int yy = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (yy > n) return;
    const long row = (long)yy * (m + 1);
    if (lane == 0) { S1[row] = 0.0; S2[row] = 0.0; }
    double ca = 0.0, cb = 0.0;
    for (int x0 = 0; x0 < m; x0 += 64) {
        const int xx = x0 + lane;
        double a = 0.0;
        if (yy > 0 && xx < m) a = g[(long)(y1 - 1 + yy - 1) * nx + (x1 - 1 + xx)];
        double b = a * a;
        for (int d = 1; d < 64; d <<= 1) {
            const double ua = __shfl_up(a, d, kWave), ub = __shfl_up(b, d, kWave);
            if (lane >= d) { a += ua; b += ub; }
        }
        a += ca;
        b += cb;
        if (xx < m) { S1[row + xx + 1] = a; S2[row + xx + 1] = b; }
        ca = __shfl(a, 63, kWave);
        cb = __shfl(b, 63, kWave);
    }
}

// Columns in chunks of kMgCh rows: chunk totals per column, their exclusive scan down the
// chunks, then each chunk's inclusive scan plus its carry.
constexpr int kMgCh = 64;

__global__ void k_mg_col_chunks(int m, int n, const double *__restrict__ S1, const double *__restrict__ S2,
                                double *__restrict__ c1, double *__restrict__ c2)
{
    const int xx = blockIdx.x * blockDim.x + threadIdx.x, ch = blockIdx.y;
    if (xx > m) return;
    double a = 0.0, b = 0.0;
    for (int yy = ch * kMgCh; yy < min((ch + 1) * kMgCh, n + 1); ++yy) {
        a += S1[(long)yy * (m + 1) + xx];
        b += S2[(long)yy * (m + 1) + xx];
    }
    c1[(long)ch * (m + 1) + xx] = a;
    c2[(long)ch * (m + 1) + xx] = b;
}

__global__ void k_mg_col_carry(int m, int nch, double *__restrict__ c1, double *__restrict__ c2)
{
    const int xx = blockIdx.x * blockDim.x + threadIdx.x;
    if (xx > m) return;
    double a = 0.0, b = 0.0;
    for (int ch = 0; ch < nch; ++ch) {
        const long k = (long)ch * (m + 1) + xx;
        const double ta = c1[k], tb = c2[k];
        c1[k] = a;
        c2[k] = b;
        a += ta;
        b += tb;
    }
}

__global__ void k_mg_col_scan(int m, int n, double *__restrict__ S1, double *__restrict__ S2,
                              const double *__restrict__ c1, const double *__restrict__ c2)
{
    const int xx = blockIdx.x * blockDim.x + threadIdx.x, ch = blockIdx.y;
    if (xx > m) return;
    double a = c1[(long)ch * (m + 1) + xx], b = c2[(long)ch * (m + 1) + xx];
    for (int yy = ch * kMgCh; yy < min((ch + 1) * kMgCh, n + 1); ++yy) {
        const long i = (long)yy * (m + 1) + xx;
        a += S1[i];
        b += S2[i];
        S1[i] = a;
        S2[i] = b;
    }
}

// The full map (2m - 1) x (2n - 1), entry (r, c) at map[c * (2m - 1) + r] (x fastest as
// MATLAB stores it): template corner at D offset (r - m + 1, c - n + 1).
__global__ void k_mg_ncc_full(const double2 *__restrict__ corr, int P, int Q, int m, int n,
                              const double *__restrict__ S1, const double *__restrict__ S2,
                              const double *__restrict__ tst, double *__restrict__ map)
{
    const int R = 2 * m - 1, C = 2 * n - 1;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)R * C) return;
    const int r = (int)(i % R), c = (int)(i / R);
    const int kx = r - (m - 1), ky = c - (n - 1);
    const double num = corr[(long)((ky % Q + Q) % Q) * P + ((kx % P + P) % P)].x / ((double)P * (double)Q);
    const int xa = max(kx, 0), xb = min(kx + m, m), ya = max(ky, 0), yb = min(ky + n, n);
    const long w = m + 1;
    const double sd = S1[yb * w + xb] - S1[ya * w + xb] - S1[yb * w + xa] + S1[ya * w + xa];
    const double sd2 = S2[yb * w + xb] - S2[ya * w + xb] - S2[yb * w + xa] + S2[ya * w + xa];
    const double cnt = (double)m * (double)n;
    const double va = fmax(sd2 - sd * sd / cnt, 0.0);
    const double den = sqrt(tst[1] * va);
    map[i] = den > 0.0 ? num / den : 0.0;
}

// first maximum (lowest index among equal values): per-block candidates, then one block
__device__ __forceinline__ void amax_merge(double &v, long &k, double ov, long ok)
{
    if (ov > v || (ov == v && ok < k)) { v = ov; k = ok; }
}

__global__ __launch_bounds__(256) void k_mg_argmax(const double *__restrict__ map, long n, double *__restrict__ bv,
                                                   long *__restrict__ bk)
{
    __shared__ double sv[4];
    __shared__ long sk[4];
    double v = -INFINITY;
    long k = LONG_MAX;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        amax_merge(v, k, map[i], i);
    for (int d = 32; d >= 1; d >>= 1) {
        const double ov = __shfl_xor(v, d, kWave);
        const long ok = __shfl_xor(k, d, kWave);
        amax_merge(v, k, ov, ok);
    }
    if ((threadIdx.x & 63) == 0) { sv[threadIdx.x >> 6] = v; sk[threadIdx.x >> 6] = k; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) amax_merge(v, k, sv[w], sk[w]);
        bv[blockIdx.x] = v;
        bk[blockIdx.x] = k;
    }
}

// the winner over nb blocks and findpeak's fit on the map: out = {xp (column), yp (row), mf}
__global__ void k_mg_peak(const double *__restrict__ map, int R, int C, const double *__restrict__ bv,
                          const long *__restrict__ bk, int nb, double *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double v = -INFINITY;
    long k = LONG_MAX;
    for (int b = 0; b < nb; ++b) amax_merge(v, k, bv[b], bk[b]);
    if (k == LONG_MAX) {   // every entry NaN (NaN in the images)
        out[0] = out[1] = out[2] = NAN;
        return;
    }
    double xp, yp, mf;
    findpeak_fit(map, R, C, (int)k, v, xp, yp, mf);
    out[0] = xp;
    out[1] = yp;
    out[2] = mf;
}

// Later levels.  reg[q] = {xf1, xf2, yf1, yf2, xg1, yg1} (1-based, inclusive; the g window
// is xg1 - 5 .. xg1 + (xf2 - xf1) + 5); tile t = {region, x0, y0} (template pixels x0 ..
// x0 + 31, y0 .. y0 + 15, clipped to the region).  Partials per tile and lag: sum T D,
// sum D, sum D^2; per tile: sum T, sum T^2, min T, max T.
__global__ __launch_bounds__(256) void k_mg_tiles(const float *__restrict__ f, const float *__restrict__ g,
                                                  int nx, const int *__restrict__ reg, const int *__restrict__ tiles,
                                                  double *__restrict__ part, double *__restrict__ tstat)
{
    __shared__ double sT[kMgTY][kMgTX];
    __shared__ double sD[kMgDY][kMgDX + 1];
    const int t = blockIdx.x, q = tiles[3 * t], x0 = tiles[3 * t + 1], y0 = tiles[3 * t + 2];
    const int xf1 = reg[6 * q], xf2 = reg[6 * q + 1], yf1 = reg[6 * q + 2], yf2 = reg[6 * q + 3];
    const int xg1 = reg[6 * q + 4], yg1 = reg[6 * q + 5];
    const int wx = xf2 - xf1, wy = yf2 - yf1;
    // g window: x from xg1 - 5 to xg1 + wx + 5 (1-based)
    const int gxa = xg1 - 5, gxb = xg1 + wx + 5, gya = yg1 - 5, gyb = yg1 + wy + 5;
    for (int i = threadIdx.x; i < kMgTY * kMgTX; i += blockDim.x) {
        const int iy = i / kMgTX, ix = i % kMgTX, x = x0 + ix, y = y0 + iy;
        sT[iy][ix] = (x <= xf2 && y <= yf2) ? (double)f[(long)(y - 1) * nx + (x - 1)] : 0.0;
    }
    // D tile: template pixel (x, y) at lag (ox, oy) reads g at (x - xf1 + xg1 - 5 + ox, ...)
    for (int i = threadIdx.x; i < kMgDY * kMgDX; i += blockDim.x) {
        const int iy = i / kMgDX, ix = i % kMgDX;
        const int gx = x0 - xf1 + xg1 - 5 - 1 + ix, gy = y0 - yf1 + yg1 - 5 - 1 + iy;
        sD[iy][ix] = (gx >= gxa && gx <= gxb && gy >= gya && gy <= gyb) ? (double)g[(long)(gy - 1) * nx + (gx - 1)]
                                                                         : 0.0;
    }
    __syncthreads();
    const int nxt = min(kMgTX, xf2 - x0 + 1), nyt = min(kMgTY, yf2 - y0 + 1);
    const int l = threadIdx.x;
    if (l < kMgLags) {
        const int ox = l % kMgLX, oy = l / kMgLX;   // offset + 1
        double s = 0.0, sd = 0.0, sd2 = 0.0;
        for (int iy = 0; iy < nyt; ++iy)
            for (int ix = 0; ix < nxt; ++ix) {
                const double d = sD[iy + oy][ix + ox];
                s += sT[iy][ix] * d;
                sd += d;
                sd2 += d * d;
            }
        double *o = part + ((long)t * kMgLags + l) * 3;
        o[0] = s;
        o[1] = sd;
        o[2] = sd2;
    } else if (l == kMgLags) {
        double s = 0.0, s2 = 0.0, lo = INFINITY, hi = -INFINITY;
        for (int iy = 0; iy < nyt; ++iy)
            for (int ix = 0; ix < nxt; ++ix) {
                const double v = sT[iy][ix];
                s += v;
                s2 += v * v;
                lo = fmin(lo, v);
                hi = fmax(hi, v);
            }
        double *o = tstat + (long)t * 4;
        o[0] = s;
        o[1] = s2;
        o[2] = lo;
        o[3] = hi;
    }
}

// Coarse levels (many tiles per region): one wave per (region, lag) sums the region's tile
// partials (lanes stride over the tiles, then the shuffle tree: a fixed order), lag kMgLags
// being the template statistics; the result has the layout of one tile per region.
__global__ __launch_bounds__(256) void k_mg_reduce(const int *__restrict__ toff, const double *__restrict__ part,
                                                   const double *__restrict__ tstat, double *__restrict__ rpart,
                                                   double *__restrict__ rstat)
{
    const int q = blockIdx.x, l = blockIdx.y * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (l > kMgLags) return;
    const int t0 = toff[q], t1 = toff[q + 1];
    if (l < kMgLags) {
        double s = 0.0, sd = 0.0, sd2 = 0.0;
        for (int t = t0 + lane; t < t1; t += 64) {
            const double *o = part + ((long)t * kMgLags + l) * 3;
            s += o[0];
            sd += o[1];
            sd2 += o[2];
        }
        s = wave_sum(s);
        sd = wave_sum(sd);
        sd2 = wave_sum(sd2);
        if (lane == 0) {
            double *o = rpart + ((long)q * kMgLags + l) * 3;
            o[0] = s;
            o[1] = sd;
            o[2] = sd2;
        }
    } else {
        double s = 0.0, s2 = 0.0, lo = INFINITY, hi = -INFINITY;
        for (int t = t0 + lane; t < t1; t += 64) {
            const double *o = tstat + (long)t * 4;
            s += o[0];
            s2 += o[1];
            lo = fmin(lo, o[2]);
            hi = fmax(hi, o[3]);
        }
        s = wave_sum(s);
        s2 = wave_sum(s2);
        for (int d = 32; d >= 1; d >>= 1) {
            lo = fmin(lo, __shfl_xor(lo, d, kWave));
            hi = fmax(hi, __shfl_xor(hi, d, kWave));
        }
        if (lane == 0) {
            double *o = rstat + (long)q * 4;
            o[0] = s;
            o[1] = s2;
            o[2] = lo;
            o[3] = hi;
        }
    }
}

__global__ void k_iota_i(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) p[i] = i;
}

// One workgroup per region: its tiles' partials summed in order, the 13 x 14 NCC map,
// its first maximum and findpeak's fit.  out[q] = {du, dv, Phi} (displacement within the
// window: peak - 7), NaN, NaN, 0 for a constant template.
__global__ __launch_bounds__(256) void k_mg_finish(const int *__restrict__ reg, const int *__restrict__ toff,
                                                   const double *__restrict__ part, const double *__restrict__ tstat,
                                                   double *__restrict__ out)
{
    __shared__ double map[kMgLags];
    __shared__ double ts[4];
    const int q = blockIdx.x, l = threadIdx.x;
    const int t0 = toff[q], t1 = toff[q + 1];
    if (l == kMgLags) {
        double s = 0.0, s2 = 0.0, lo = INFINITY, hi = -INFINITY;
        for (int t = t0; t < t1; ++t) {
            const double *o = tstat + (long)t * 4;
            s += o[0];
            s2 += o[1];
            lo = fmin(lo, o[2]);
            hi = fmax(hi, o[3]);
        }
        ts[0] = s;
        ts[1] = s2;
        ts[2] = lo;
        ts[3] = hi;
    }
    __syncthreads();
    const double cnt = (double)(reg[6 * q + 1] - reg[6 * q] + 1) * (double)(reg[6 * q + 3] - reg[6 * q + 2] + 1);
    const double mean = ts[0] / cnt, stt = ts[1] - ts[0] * ts[0] / cnt;
    if (l < kMgLags) {
        double s = 0.0, sd = 0.0, sd2 = 0.0;
        for (int t = t0; t < t1; ++t) {
            const double *o = part + ((long)t * kMgLags + l) * 3;
            s += o[0];
            sd += o[1];
            sd2 += o[2];
        }
        const double num = s - mean * sd;
        const double va = fmax(sd2 - sd * sd / cnt, 0.0);
        const double den = sqrt(stt * va);
        map[l] = den > 0.0 ? num / den : 0.0;
    }
    __syncthreads();
    if (l != 0) return;
    if (!(ts[3] > ts[2])) {   // constant template: normxcorr2 raises, the reference catches
        out[3 * q] = NAN;
        out[3 * q + 1] = NAN;
        out[3 * q + 2] = 0.0;
        return;
    }
    double best = -INFINITY;
    int bi = 0;
    for (int k = 0; k < kMgLags; ++k)
        if (map[k] > best) { best = map[k]; bi = k; }
    double xp, yp, mf;
    findpeak_fit(map, kMgLX, kMgLY, bi, best, xp, yp, mf);
    out[3 * q] = yp - 7.0;
    out[3 * q + 1] = xp - 7.0;
    out[3 * q + 2] = mf;
}

int nblk(long n, int t = 256) { return (int)((n + t - 1) / t); }

}  // namespace

struct MgWork {
    DBuf zt, zy, zd, tw, s1, s2, map, part, red, peak, reg, tiles, toff, tpart, tstat, out;
    DBuf nodes, pts, fit, gstat, gw[3];
    GridfitDev gf;
    std::vector<double> gx_last, gy_last;   // the node lists gf was set up for
    ~MgWork()
    {
        for (DBuf *b : {&zt, &zy, &zd, &tw, &s1, &s2, &map, &part, &red, &peak, &reg, &tiles, &toff, &tpart, &tstat,
                        &out, &nodes, &pts, &fit, &gstat, &gw[0], &gw[1], &gw[2], &gf.w, &gf.a9, &gf.dinv})
            b->release();
    }
};

MgWork *mg_work_new() { return new (std::nothrow) MgWork(); }
void mg_work_free(MgWork *w) { delete w; }

namespace {

// nlines contiguous lines of length N = 2^logN from Z, transformed in place
hipError_t fft_rows(double2 *Z, const double2 *tw, int L, int N, int logN, int nlines, bool inv, hipStream_t st)
{
    static bool attr = false;
    if (!attr) {
        hipError_t e;
        if ((e = hipFuncSetAttribute((const void *)k_fft<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)sizeof(double2) * 8192)) ||
            (e = hipFuncSetAttribute((const void *)k_fft<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)sizeof(double2) * 8192)))
            return e;
        attr = true;
    }
    if (nlines <= 0) return hipSuccess;
    if (inv)
        hipLaunchKernelGGL(k_fft<true>, dim3(nlines), dim3(kFftThreads), sizeof(double2) * N, st, Z, tw, N, logN,
                           L / N, (long)N, 1L);
    else
        hipLaunchKernelGGL(k_fft<false>, dim3(nlines), dim3(kFftThreads), sizeof(double2) * N, st, Z, tw, N, logN,
                           L / N, (long)N, 1L);
    return hipGetLastError();
}

void transpose(const double2 *in, double2 *out, int rows, int cols, hipStream_t st)
{
    hipLaunchKernelGGL(k_transpose, dim3((cols + 31) / 32, (rows + 31) / 32), dim3(256), 0, st, in, out, rows, cols);
}

// Level 1 on the ROI x1..x1+m-1, y1..y1+n-1 (1-based) of the device images: out4 (host) =
// {xp (column = y), yp (row = x), max} of findpeak on the full (2m-1) x (2n-1) map and the
// template's largest |v - first pixel| (0: constant, normxcorr2 raises).
hipError_t mg_level1(MgWork &W, const float *f, const float *g, int nx, int x1, int y1, int m, int n,
                     double *out3, hipStream_t st)
{
    hipError_t e;
    int lp = 0, lq = 0;
    while ((1 << lp) < 2 * m - 1) ++lp;
    while ((1 << lq) < 2 * n - 1) ++lq;
    const int P = 1 << lp, Q = 1 << lq, L = P > Q ? P : Q;
    const long PQ = (long)P * Q, R = 2L * m - 1, C = 2L * n - 1;
    const int nbt = 512;
    if ((e = W.zt.ensure(sizeof(double2) * PQ)) || (e = W.zy.ensure(sizeof(double2) * PQ)) ||
        (e = W.zd.ensure(2 * sizeof(double) * (size_t)(m + 1) * ((n + 1 + kMgCh - 1) / kMgCh))) ||
        (e = W.tw.ensure(sizeof(double2) * (L / 2 + 1))) ||
        (e = W.s1.ensure(sizeof(double) * (m + 1) * (size_t)(n + 1))) ||
        (e = W.s2.ensure(sizeof(double) * (m + 1) * (size_t)(n + 1))) || (e = W.map.ensure(sizeof(double) * R * C)) ||
        (e = W.part.ensure(sizeof(double) * (3 * nbt + 3))) ||
        (e = W.red.ensure((sizeof(double) + sizeof(long)) * nbt)) || (e = W.peak.ensure(sizeof(double) * 4)))
        return e;
    double *tst = W.part.as<double>() + 3 * nbt;
    hipLaunchKernelGGL(k_twiddles, dim3(nblk(L / 2)), dim3(256), 0, st, W.tw.as<double2>(), L);
    hipLaunchKernelGGL(k_mg_tsum, dim3(nbt), dim3(256), 0, st, f, nx, x1, y1, m, n, W.part.as<double>());
    hipLaunchKernelGGL(k_mg_tsum_final, dim3(1), dim3(64), 0, st, (const double *)W.part.as<double>(), nbt,
                       (long)m * n, tst);
    hipLaunchKernelGGL(k_mg_operands, dim3(nblk(PQ)), dim3(256), 0, st, f, g, nx, x1, y1, m, n,
                       (const double *)tst, P, Q, W.zt.as<double2>());
    // 2-D transform as row transforms and transposes (every pass coalesced): Z is [y][x];
    // rows y >= n are zero and stay zero under the first pass; the spectrum is formed and
    // multiplied in the transposed layout [kx][ky]; the correlation returns to [y][x]
    double2 *Z = W.zt.as<double2>(), *Y = W.zy.as<double2>();
    const double2 *tw = W.tw.as<double2>();
    if ((e = fft_rows(Z, tw, L, P, lp, n, false, st))) return e;
    transpose(Z, Y, Q, P, st);
    if ((e = fft_rows(Y, tw, L, Q, lq, P, false, st))) return e;
    hipLaunchKernelGGL(k_mg_cross, dim3(nblk(PQ)), dim3(256), 0, st, Y, P, Q);
    if ((e = fft_rows(Y, tw, L, Q, lq, P, true, st))) return e;
    transpose(Y, Z, P, Q, st);
    if ((e = fft_rows(Z, tw, L, P, lp, Q, true, st))) return e;
    const int nch = (n + 1 + kMgCh - 1) / kMgCh;
    double *c1 = W.zd.as<double>(), *c2 = c1 + (size_t)nch * (m + 1);
    hipLaunchKernelGGL(k_mg_prefix_rows, dim3((n + 1 + 3) / 4), dim3(256), 0, st, g, nx, x1, y1, m, n,
                       W.s1.as<double>(), W.s2.as<double>());
    hipLaunchKernelGGL(k_mg_col_chunks, dim3(nblk(m + 1), nch), dim3(256), 0, st, m, n,
                       (const double *)W.s1.as<double>(), (const double *)W.s2.as<double>(), c1, c2);
    hipLaunchKernelGGL(k_mg_col_carry, dim3(nblk(m + 1)), dim3(256), 0, st, m, nch, c1, c2);
    hipLaunchKernelGGL(k_mg_col_scan, dim3(nblk(m + 1), nch), dim3(256), 0, st, m, n, W.s1.as<double>(),
                       W.s2.as<double>(), (const double *)c1, (const double *)c2);
    hipLaunchKernelGGL(k_mg_ncc_full, dim3(nblk(R * C)), dim3(256), 0, st, (const double2 *)W.zt.as<double2>(), P,
                       Q, m, n, (const double *)W.s1.as<double>(), (const double *)W.s2.as<double>(),
                       (const double *)tst, W.map.as<double>());
    double *bv = W.red.as<double>();
    long *bk = reinterpret_cast<long *>(bv + nbt);
    hipLaunchKernelGGL(k_mg_argmax, dim3(nbt), dim3(256), 0, st, (const double *)W.map.as<double>(), R * C, bv, bk);
    hipLaunchKernelGGL(k_mg_peak, dim3(1), dim3(64), 0, st, (const double *)W.map.as<double>(), (int)R, (int)C,
                       (const double *)bv, (const long *)bk, nbt, W.peak.as<double>());
    if ((e = hipGetLastError())) return e;
    if ((e = hipMemcpyAsync(out3, W.peak.p, sizeof(double) * 3, hipMemcpyDeviceToHost, st))) return e;
    if ((e = hipMemcpyAsync(out3 + 3, tst + 2, sizeof(double), hipMemcpyDeviceToHost, st))) return e;
    return hipStreamSynchronize(st);
}

// A later level: nr valid regions reg_host (6 ints each: xf1 xf2 yf1 yf2 xg1 yg1), out_host
// (3 per region) = {du, dv, Phi}.
hipError_t mg_regions(MgWork &W, const float *f, const float *g, int nx, const std::vector<int> &reg, int nr,
                      double *out_host, hipStream_t st)
{
    hipError_t e;
    if (nr <= 0) return hipSuccess;
    std::vector<int> tiles, toff(nr + 1, 0);
    for (int q = 0; q < nr; ++q) {
        const int *r = reg.data() + 6 * q;
        for (int y0 = r[2]; y0 <= r[3]; y0 += kMgTY)
            for (int x0 = r[0]; x0 <= r[1]; x0 += kMgTX) {
                tiles.push_back(q);
                tiles.push_back(x0);
                tiles.push_back(y0);
            }
        toff[q + 1] = (int)(tiles.size() / 3);
    }
    const int nt = toff[nr];
    int most = 0;
    for (int q = 0; q < nr; ++q) most = std::max(most, toff[q + 1] - toff[q]);
    // regions of many tiles: their partials reduced in parallel first (the finish kernel's
    // loop over the tiles would be long and serial)
    const bool coarse = most > 32;
    const size_t nrp = coarse ? (size_t)nr : 0;
    if ((e = W.reg.ensure(sizeof(int) * 6 * (size_t)nr)) || (e = W.tiles.ensure(sizeof(int) * tiles.size())) ||
        (e = W.toff.ensure(sizeof(int) * (2 * (size_t)nr + 2))) ||
        (e = W.tpart.ensure(sizeof(double) * 3 * kMgLags * ((size_t)nt + nrp))) ||
        (e = W.tstat.ensure(sizeof(double) * 4 * ((size_t)nt + nrp))) ||
        (e = W.out.ensure(sizeof(double) * 3 * (size_t)nr)))
        return e;
    if ((e = hipMemcpyAsync(W.reg.p, reg.data(), sizeof(int) * 6 * (size_t)nr, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(W.tiles.p, tiles.data(), sizeof(int) * tiles.size(), hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(W.toff.p, toff.data(), sizeof(int) * (nr + 1), hipMemcpyHostToDevice, st)))
        return e;
    hipLaunchKernelGGL(k_mg_tiles, dim3(nt), dim3(256), 0, st, f, g, nx, (const int *)W.reg.as<int>(),
                       (const int *)W.tiles.as<int>(), W.tpart.as<double>(), W.tstat.as<double>());
    const int *toff_d = W.toff.as<int>();
    const double *fpart = W.tpart.as<double>(), *fstat = W.tstat.as<double>();
    if (coarse) {
        int *ident = W.toff.as<int>() + nr + 1;
        double *rpart = W.tpart.as<double>() + 3 * (size_t)kMgLags * nt, *rstat = W.tstat.as<double>() + 4 * (size_t)nt;
        hipLaunchKernelGGL(k_mg_reduce, dim3(nr, (kMgLags + 1 + 3) / 4), dim3(256), 0, st, toff_d, fpart, fstat,
                           rpart, rstat);
        hipLaunchKernelGGL(k_iota_i, dim3(nblk(nr + 1)), dim3(256), 0, st, ident, nr);
        toff_d = ident;
        fpart = rpart;
        fstat = rstat;
    }
    hipLaunchKernelGGL(k_mg_finish, dim3(nr), dim3(256), 0, st, (const int *)W.reg.as<int>(), toff_d, fpart, fstat,
                       W.out.as<double>());
    if ((e = hipGetLastError())) return e;
    if ((e = hipMemcpyAsync(out_host, W.out.p, sizeof(double) * 3 * (size_t)nr, hipMemcpyDeviceToHost, st))) return e;
    return hipStreamSynchronize(st);   // the host vectors go out of scope
}

double mceil(double v) { return std::isfinite(v) ? std::ceil(v) : NAN; }

// floor-modulo (MATLAB mod) of a width by 2 is 1
bool odd_width(double w) { return std::isfinite(w) && w - 2.0 * std::floor(w / 2.0) == 1.0; }

// Border checks of funIntegerSearchMg.m:59-66 / :202-229 on one region [xf1 xf2 yf1 yf2
// xg1 xg2 yg1 yg2] and the odd-width fix of :68-69 / :231-239.
void mg_clamp(double *r, int nx, int ny, double gap)
{
    double &xf1 = r[0], &xf2 = r[1], &yf1 = r[2], &yf2 = r[3], &xg1 = r[4], &xg2 = r[5], &yg1 = r[6], &yg2 = r[7];
    if (xg1 < gap) { const double t = xg1; xg1 = gap; xf1 = xf1 + gap - t; }
    if (yg1 < gap) { const double t = yg1; yg1 = gap; yf1 = yf1 + gap - t; }
    if (xg2 > nx - gap) { const double t = xg2; xg2 = nx - gap; xf2 = xf2 + (nx - gap) - t; }
    if (yg2 > ny - gap) { const double t = yg2; yg2 = ny - gap; yf2 = yf2 + (ny - gap) - t; }
    if (odd_width(xf2 - xf1)) { xf2 -= 1; xg2 -= 1; }
    if (odd_width(yf2 - yf1)) { yf2 -= 1; yg2 -= 1; }
}

// funIntegerSearchMg.m:163-182's try: a region is searched when its bounds are finite
// integers inside the image and its g window (+-5 px) is too; otherwise NaN, NaN, 0.
bool mg_valid(const double *r, int nx, int ny)
{
    for (int k = 0; k < 8; ++k)
        if (!std::isfinite(r[k]) || r[k] != std::round(r[k])) return false;
    // (the g window has the template's size by construction; checked, as the kernel reads
    // g from xg1 - 5 to xg1 + (xf2 - xf1) + 5)
    return r[0] >= 1 && r[2] >= 1 && r[1] <= nx && r[3] <= ny && r[1] >= r[0] && r[3] >= r[2] && r[4] - 5 >= 1 &&
           r[6] - 5 >= 1 && r[5] + 5 <= nx && r[7] + 5 <= ny && r[5] - r[4] == r[1] - r[0] &&
           r[7] - r[6] == r[3] - r[2];
}

}  // namespace

// funIntegerSearchMg.m:3-285 (qfactors aside) on the device images.
int integer_search_mg(MgWork &W, const float *f, const float *g, int nx, int ny, const int *grid, int ws,
                      int step, MgResult &res, std::string &err, hipStream_t st)
{
    hipError_t e;
    double gridx[2] = {(double)grid[0], (double)grid[1]}, gridy[2] = {(double)grid[2], (double)grid[3]};
    const double gap = 1.0 + std::floor(0.75 * ws + 0.5);
    if (gridx[0] < gap) gridx[0] = gap;
    if (gridy[0] < gap) gridy[0] = gap;
    if (gridx[1] > nx - gap) gridx[1] = nx - gap;
    if (gridy[1] > ny - gap) gridy[1] = ny - gap;
    const int x1 = (int)gridx[0], y1 = (int)gridy[0];
    const int m = (int)gridx[1] - x1 + 1, n = (int)gridy[1] - y1 + 1;
    if (m < 2 || n < 2 || 2 * m - 1 > 8192 || 2 * n - 1 > 8192) {
        err = "integer_search_mg: ROI after the border gap must be 2..4096 px per side";
        return 1;
    }
    // level 1: normxcorr2 of the whole ROI (:38-57)
    double o1[4];
    if ((e = mg_level1(W, f, g, nx, x1, y1, m, n, o1, st))) { err = hipGetErrorString(e); return 2; }
    if (!(o1[3] > 0.0)) {
        err = "integer_search_mg: constant reference ROI (normxcorr2 raises)";
        return 1;
    }
    const double u = o1[1] - m, v = o1[0] - n;
    std::vector<double> regions(8), uc{mceil(u)}, vc{mceil(v)};
    {
        double r0[8] = {gridx[0], gridx[1], gridy[0], gridy[1], gridx[0] + uc[0], gridx[1] + uc[0],
                        gridy[0] + vc[0], gridy[1] + vc[0]};
        mg_clamp(r0, nx, ny, gap);
        regions.assign(r0, r0 + 8);
    }
    double wx = regions[1] - regions[0], wy = regions[3] - regions[2];
    int levels = 1;
    std::vector<double> kids, un, vn, phi, out;
    std::vector<int> regi, idx;
    for (;;) {
        ++levels;
        const double ratio = wx / wy;
        kids.clear();
        un.clear();
        vn.clear();
        const size_t nreg = regions.size() / 8;
        for (size_t q = 0; q < nreg; ++q) {
            const double *r = &regions[8 * q];
            const double xf1 = r[0], xf2 = r[1], yf1 = r[2], yf2 = r[3], xg1 = r[4], xg2 = r[5], yg1 = r[6],
                         yg2 = r[7];
            int nk;
            if (ratio > 2) {
                nk = 2;
                if (xf2 - xf1 > ws) {
                    const double mf = 0.5 * (xf1 + xf2), mg = 0.5 * (xg1 + xg2);
                    const double k[16] = {xf1, mf, yf1, yf2, xg1, mg, yg1, yg2, mf, xf2, yf1, yf2, mg, xg2, yg1, yg2};
                    kids.insert(kids.end(), k, k + 16);
                } else {
                    kids.insert(kids.end(), r, r + 8);
                    kids.insert(kids.end(), r, r + 8);
                }
            } else if (ratio < 0.5) {
                nk = 2;
                if (yf2 - yf1 > ws) {
                    const double mf = 0.5 * (yf1 + yf2), mg = 0.5 * (yg1 + yg2);
                    const double k[16] = {xf1, xf2, yf1, mf, xg1, xg2, yg1, mg, xf1, xf2, mf, yf2, xg1, xg2, mg, yg2};
                    kids.insert(kids.end(), k, k + 16);
                } else {
                    kids.insert(kids.end(), r, r + 8);
                    kids.insert(kids.end(), r, r + 8);
                }
            } else {
                nk = 4;
                if (xf2 - xf1 > ws && yf2 - yf1 > ws) {
                    const double mfx = 0.5 * (xf1 + xf2), mgx = 0.5 * (xg1 + xg2);
                    const double mfy = 0.5 * (yf1 + yf2), mgy = 0.5 * (yg1 + yg2);
                    const double k[32] = {xf1, mfx, yf1, mfy, xg1, mgx, yg1, mgy, mfx, xf2, yf1, mfy, mgx, xg2, yg1, mgy,
                                          xf1, mfx, mfy, yf2, xg1, mgx, mgy, yg2, mfx, xf2, mfy, yf2, mgx, xg2, mgy, yg2};
                    kids.insert(kids.end(), k, k + 32);
                } else {
                    for (int c = 0; c < 4; ++c) kids.insert(kids.end(), r, r + 8);
                }
            }
            for (int c = 0; c < nk; ++c) {
                un.push_back(uc[q]);
                vn.push_back(vc[q]);
            }
        }
        double wxn, wyn;
        if (ratio > 2) { wxn = wx / 2; wyn = wy; }
        else if (ratio < 0.5) { wxn = wx; wyn = wy / 2; }
        else { wxn = wx / 2; wyn = wy / 2; }
        // :163-182 on the device for the regions the try would not catch
        const int nk = (int)(kids.size() / 8);
        regi.clear();
        idx.clear();
        for (int q = 0; q < nk; ++q) {
            const double *r = &kids[8 * q];
            if (!mg_valid(r, nx, ny)) continue;
            idx.push_back(q);
            const int ri[6] = {(int)r[0], (int)r[1], (int)r[2], (int)r[3], (int)r[4], (int)r[6]};
            regi.insert(regi.end(), ri, ri + 6);
        }
        out.assign(3 * idx.size(), 0.0);
        if ((e = mg_regions(W, f, g, nx, regi, (int)idx.size(), out.data(), st))) {
            err = hipGetErrorString(e);
            return 2;
        }
        phi.assign(nk, 0.0);
        std::vector<double> du(nk, NAN), dv(nk, NAN);
        for (size_t t = 0; t < idx.size(); ++t) {
            du[idx[t]] = out[3 * t];
            dv[idx[t]] = out[3 * t + 1];
            phi[idx[t]] = out[3 * t + 2];
        }
        for (int q = 0; q < nk; ++q) {
            un[q] += du[q];
            vn[q] += dv[q];
        }
        if (wxn < ws && wyn < ws) break;
        regions.assign(kids.size(), 0.0);
        uc.assign(nk, 0.0);
        vc.assign(nk, 0.0);
        for (int q = 0; q < nk; ++q) {
            const double *r = &kids[8 * q];
            const double cu = mceil(un[q]), cv = mceil(vn[q]);
            double rr[8] = {r[0], r[1], r[2], r[3], r[0] + cu, r[1] + cu, r[2] + cv, r[3] + cv};
            mg_clamp(rr, nx, ny, gap);
            std::copy(rr, rr + 8, &regions[8 * q]);
            uc[q] = cu;
            vc[q] = cv;
        }
        wx = wxn;
        wy = wyn;
    }
    const int nk = (int)(kids.size() / 8);
    res.levels = levels;
    res.tempx.resize(nk);
    res.tempy.resize(nk);
    for (int q = 0; q < nk; ++q) {
        res.tempx[q] = 0.5 * (kids[8 * q] + kids[8 * q + 1]);
        res.tempy[q] = 0.5 * (kids[8 * q + 2] + kids[8 * q + 3]);
    }
    res.tempu = un;
    res.tempv = vn;
    res.tempphi = phi;
    // :262-270 the grid away from the borders the final displacements need (first minimum
    // and first maximum, as MATLAB's min / max)
    int i1 = 0, i2 = 0, j1 = 0, j2 = 0;
    for (int q = 1; q < nk; ++q) {
        if (res.tempx[q] < res.tempx[i1]) i1 = q;
        if (res.tempx[q] > res.tempx[i2]) i2 = q;
        if (res.tempy[q] < res.tempy[j1]) j1 = q;
        if (res.tempy[q] > res.tempy[j2]) j2 = q;
    }
    const double gx1 = mceil(1 + 1.25 * ws + un[i1]), gx2 = mceil(1 + 1.25 * ws + un[i2]);
    const double gy1 = mceil(1 + 1.25 * ws + vn[j1]), gy2 = mceil(1 + 1.25 * ws + vn[j2]);
    if (gridx[0] < gx1) gridx[0] = gx1;
    if (gridy[0] < gy1) gridy[0] = gy1;
    if (gridx[1] > nx - gx2 + 1) gridx[1] = nx - gx2 + 1;
    if (gridy[1] > ny - gy2 + 1) gridy[1] = ny - gy2 + 1;
    if (!(gridx[1] >= gridx[0] + step) || !(gridy[1] >= gridy[0] + step)) {
        err = "integer_search_mg: fewer than two grid nodes per axis after the final border trim";
        return 1;
    }
    res.xlist.clear();
    res.ylist.clear();
    for (double t = gridx[0]; t <= gridx[1] + 1e-9; t += step) res.xlist.push_back(t);
    for (double t = gridy[0]; t <= gridy[1] + 1e-9; t += step) res.ylist.push_back(t);
    const int Nx = (int)res.xlist.size(), Ny = (int)res.ylist.size();
    const double xa = res.xlist.front(), xb = res.xlist.back(), ya = res.ylist.front(), yb = res.ylist.back();
    // :272-276 gridfit of u, v (points with a value: gridfit.m:332 drops NaN) and Phi
    std::vector<double> pu, pc;   // x, y, u, v of the points with a displacement; x, y, Phi, 0
    for (int q = 0; q < nk; ++q) {
        const double x = res.tempx[q], y = res.tempy[q];
        if (!(x > xa && x < xb && y > ya && y < yb)) continue;
        if (std::isfinite(un[q]) && std::isfinite(vn[q])) {
            const double p[4] = {x, y, un[q], vn[q]};
            pu.insert(pu.end(), p, p + 4);
        }
        if (std::isfinite(phi[q])) {
            const double p[4] = {x, y, phi[q], 0.0};
            pc.insert(pc.end(), p, p + 4);
        }
    }
    if (pu.empty() || pc.empty()) {
        err = "integer_search_mg: no region centre inside the final grid";
        return 1;
    }
    const int ng = Nx * Ny;
    // the grid's springs depend on the node lists only: set up again when they change
    if (W.gx_last != res.xlist || W.gy_last != res.ylist) {
        W.gx_last.clear();
        if ((e = gridfit_setup(W.gf, res.xlist.data(), Nx, res.ylist.data(), Ny, st))) {
            err = hipGetErrorString(e);
            return 2;
        }
        W.gx_last = res.xlist;
        W.gy_last = res.ylist;
    }
    if ((e = W.nodes.ensure(sizeof(double) * (Nx + Ny))) || (e = W.fit.ensure(sizeof(double2) * 2 * (size_t)ng)) ||
        (e = W.gstat.ensure(sizeof(double) * 6)) ||
        (e = W.pts.ensure(sizeof(double) * 4 * (pu.size() / 4 + pc.size() / 4))))
        { err = hipGetErrorString(e); return 2; }
    double *nodes = W.nodes.as<double>(), *pts = W.pts.as<double>();
    if ((e = hipMemcpyAsync(nodes, res.xlist.data(), sizeof(double) * Nx, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(nodes + Nx, res.ylist.data(), sizeof(double) * Ny, hipMemcpyHostToDevice, st)))
        { err = hipGetErrorString(e); return 2; }
    // point arrays: x | y | (value pairs), per set
    // point arrays x | y | value pairs; u, v and Phi share one CG (two pairs per node) when
    // every point has a displacement, else u, v and Phi are fitted apart
    auto stage = [&](const std::vector<double> &p, int npair, double *dst) -> hipError_t {
        const int w = 2 + 2 * npair, np = (int)(p.size() / w);
        std::vector<double> soa(p.size());
        for (int i = 0; i < np; ++i) {
            soa[i] = p[(size_t)w * i];
            soa[np + i] = p[(size_t)w * i + 1];
            for (int k = 0; k < 2 * npair; ++k) soa[2 * (size_t)np + 2 * npair * (size_t)i + k] = p[(size_t)w * i + 2 + k];
        }
        hipError_t r = hipMemcpyAsync(dst, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice, st);
        if (r == hipSuccess) r = hipStreamSynchronize(st);
        return r;
    };
    const int npu = (int)(pu.size() / 4), npc = (int)(pc.size() / 4);
    double *fit = W.fit.as<double>();
    int su, sc;   // node i: u = uv[su * i], v = uv[su * i + 1], Phi = cc[sc * i] (below)
    if (npu == npc) {
        std::vector<double> pj(6 * (size_t)npu);
        for (int i = 0; i < npu; ++i) {
            const double q[6] = {pu[4 * i], pu[4 * i + 1], pu[4 * i + 2], pu[4 * i + 3], pc[4 * i + 2], 0.0};
            std::copy(q, q + 6, &pj[6 * (size_t)i]);
        }
        if ((e = stage(pj, 2, pts)) ||
            (e = gridfit_scattered(W.gf, nodes, nodes + Nx, npu, pts, pts + npu,
                                   reinterpret_cast<const double2 *>(pts + 2 * npu), 2, fit, 1e-12, 4 * ng + 100,
                                   W.gw, W.gstat.as<double>(), st)))
            { err = hipGetErrorString(e); return 2; }
        su = sc = 4;
    } else {
        double *ppu = pts, *ppc = pts + 4 * (size_t)npu;
        if ((e = stage(pu, 1, ppu)) || (e = stage(pc, 1, ppc)) ||
            (e = gridfit_scattered(W.gf, nodes, nodes + Nx, npu, ppu, ppu + npu,
                                   reinterpret_cast<const double2 *>(ppu + 2 * npu), 1, fit, 1e-12, 4 * ng + 100,
                                   W.gw, W.gstat.as<double>(), st)) ||
            (e = gridfit_scattered(W.gf, nodes, nodes + Nx, npc, ppc, ppc + npc,
                                   reinterpret_cast<const double2 *>(ppc + 2 * npc), 1, fit + 2 * (size_t)ng,
                                   1e-12, 4 * ng + 100, W.gw, W.gstat.as<double>() + 3, st)))
            { err = hipGetErrorString(e); return 2; }
        su = sc = 2;
    }
    std::vector<double> hfit(4 * (size_t)ng);
    double hst[6] = {0, 0, 0, 0, 0, 0};
    if ((e = hipMemcpyAsync(hfit.data(), fit, sizeof(double) * hfit.size(), hipMemcpyDeviceToHost, st)) ||
        (e = hipMemcpyAsync(hst, W.gstat.p, sizeof(double) * (npu == npc ? 3 : 6), hipMemcpyDeviceToHost, st)) ||
        (e = hipStreamSynchronize(st)))
        { err = hipGetErrorString(e); return 2; }
    const double *uv = hfit.data(), *cc = npu == npc ? hfit.data() + 2 : hfit.data() + 2 * (size_t)ng;
    res.u.resize(ng);
    res.v.resize(ng);
    res.cc.resize(ng);
    for (int i = 0; i < ng; ++i) {
        res.u[i] = uv[(size_t)su * i];
        res.v[i] = uv[(size_t)su * i + 1];
        res.cc[i] = cc[(size_t)sc * i];
    }
    res.fit_iters[0] = (int)hst[1];
    res.fit_iters[1] = (int)hst[4];
    return 0;
}

}  // namespace aldic
