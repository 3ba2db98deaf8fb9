// Outlier removal and NaN in-painting on a 2-D field of node values:
//   funRemoveOutliers.m:45-86 with Thr0 given (no prompts), its median test
//   funMedRemoveOutliers / medFilt2 (:371-445), and inpaint_nans.m methods 3 (the del^4
//   plate) and 4 (springs), as IntegerSearch.m:114 and Init.m:13 apply them to the
//   integer search's field.
//
// A field is a strided view: element (i, j) at p[i si + j sj].  The unknowns (NaN or
// flagged elements) are numbered in memory order (the faster axis first) by a
// deterministic compaction, and each in-painting is the least squares of its stencil rows
// solved through the normal equations: the matrix is assembled by gathering over the
// stencil rows (no atomics, fixed order), banded in that numbering (bandwidth b from the
// unknowns' neighbourhoods), and factored by a banded Cholesky in one workgroup (n b^2
// flops), then two triangular solves.  Everything is deterministic.
//
// The median test runs in single precision as the reference's does (medFilt2 stacks the
// neighbours in a single array, and MATLAB's mixed single/double arithmetic stays single):
// median of V(i,j), V(i+1,j), V(i+1,j+1) (one mirrored row and column after the array),
// fluct = u - med, res = the same median of |fluct|, nf = |fluct / (res + 0.1)|,
// flagged where sqrt(nf_u^2 + nf_v^2) > Thr0; the float operations are the
// correctly rounded intrinsics, never contracted.
#include <algorithm>
#include <cmath>
#include <vector>

#include "aldic_internal.h"
#include "common.h"

namespace aldic {

namespace {

constexpr int kChunk = 1024;  // order indices per compaction block (4 per thread)

struct View {
    double *p;
    int rows, cols;
    long si, sj;
};

// memory-order index of (i, j) and back: the axis with the smaller stride is the faster
__device__ __forceinline__ long ord_of(const View &v, int i, int j)
{
    return v.sj <= v.si ? (long)i * v.cols + j : (long)j * v.rows + i;
}
__device__ __forceinline__ void ij_of(const View &v, long o, int &i, int &j)
{
    if (v.sj <= v.si) { i = (int)(o / v.cols); j = (int)(o % v.cols); }
    else { j = (int)(o / v.rows); i = (int)(o % v.rows); }
}
__device__ __forceinline__ double &at(const View &v, int i, int j) { return v.p[i * v.si + j * v.sj]; }

// ---- deterministic compaction of the unknowns ----------------------------------------

// unk[o] (int, memory order) -> per-block counts
__global__ __launch_bounds__(256) void k_unk_count(const int *unk, long N, int *bcnt)
{
    __shared__ int red[4];
    const long o0 = (long)blockIdx.x * kChunk;
    int c = 0;
    for (int t = threadIdx.x; t < kChunk; t += 256) {
        const long o = o0 + t;
        if (o < N && unk[o]) ++c;
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m, kWave);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) bcnt[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// exclusive scan of the block counts (one workgroup); total into *n
__global__ __launch_bounds__(1024) void k_unk_scan(int *bcnt, int nb, int *n)
{
    __shared__ int part[1024];
    const int per = (nb + 1023) / 1024, b0 = threadIdx.x * per;
    int s = 0;
    for (int k = 0; k < per && b0 + k < nb; ++k) s += bcnt[b0 + k];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int k = 0; k < 1024; ++k) { const int t = part[k]; part[k] = acc; acc += t; }
        *n = acc;
    }
    __syncthreads();
    int acc = part[threadIdx.x];
    for (int k = 0; k < per && b0 + k < nb; ++k) { const int t = bcnt[b0 + k]; bcnt[b0 + k] = acc; acc += t; }
}

// ordered writes: list[k] = memory-order index of unknown k, slot[o] = k (or -1)
__global__ __launch_bounds__(256) void k_unk_put(const int *unk, long N, const int *boff, long *list,
                                                 int *slot)
{
    __shared__ int wsum[4];
    const long o0 = (long)blockIdx.x * kChunk + 4L * threadIdx.x;
    int f[4], c = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const long o = o0 + q;
        f[q] = (o < N && unk[o]) ? 1 : 0;
        c += f[q];
    }
    // exclusive prefix over the 256 threads (in memory order: thread t holds o0 + 4t ..)
    int x = c;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, kWave);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int base = boff[blockIdx.x] + x - c;
    for (int k = 0; k < w; ++k) base += wsum[k];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const long o = o0 + q;
        if (o >= N) break;
        if (f[q]) { list[base] = o; slot[o] = base; ++base; }
        else slot[o] = -1;
    }
}

// ---- stencil rows ------------------------------------------------------------------

// inpaint_nans.m method 3: weight of offset (di, dj) in the row centred at (ri, rj)
// (0-based) of an n x m array: the del^4 template inside, the 5-point Laplacian on the
// second ring, second differences along the first/last row and column, no row at the
// corners.
__device__ __forceinline__ int plate_w(int ri, int rj, int n, int m, int di, int dj)
{
    const int r = ri + 1, c = rj + 1;
    if (r >= 3 && r <= n - 2 && c >= 3 && c <= m - 2) {
        const int ad = abs(di) + abs(dj);
        if (di == 0 && dj == 0) return 20;
        if (ad == 1) return -8;
        if (abs(di) == 1 && abs(dj) == 1) return 2;
        if ((abs(di) == 2 && dj == 0) || (abs(dj) == 2 && di == 0)) return 1;
        return 0;
    }
    if (((r == 2 || r == n - 1) && c >= 2 && c <= m - 1) ||
        ((c == 2 || c == m - 1) && r >= 2 && r <= n - 1)) {
        if (di == 0 && dj == 0) return -4;
        return (abs(di) + abs(dj) == 1) ? 1 : 0;
    }
    if ((r == 1 || r == n) && c >= 2 && c <= m - 1) {
        if (di != 0) return 0;
        return dj == 0 ? -2 : (abs(dj) == 1 ? 1 : 0);
    }
    if ((c == 1 || c == m) && r >= 2 && r <= n - 1) {
        if (dj != 0) return 0;
        return di == 0 ? -2 : (abs(di) == 1 ? 1 : 0);
    }
    return 0;
}

// band of unknown k: k minus the smallest unknown number within +-R in both axes
__global__ __launch_bounds__(256) void k_bandwidth(View v, const long *list, const int *slot, int n,
                                                   int R, int *band)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int i, j;
    ij_of(v, list[k], i, j);
    int kmin = k;
    for (int di = -R; di <= R; ++di)
        for (int dj = -R; dj <= R; ++dj) {
            const int a = i + di, b = j + dj;
            if (a < 0 || a >= v.rows || b < 0 || b >= v.cols) continue;
            const int s = slot[ord_of(v, a, b)];
            if (s >= 0 && s < kmin) kmin = s;
        }
    atomicMax(band, k - kmin);
}

// Normal equations in band storage B[k (b + 1) + d] = A[k][k - d] and the right-hand
// sides (nrhs views sharing the unknowns).  method 4: A = the springs' Laplacian (in-array
// 4-neighbour count on the diagonal, -1 between unknown neighbours), rhs = the known
// neighbours' sum.  method 3: A = M^T M, rhs = -M^T (known part), gathered over the
// plate rows that touch unknown k.
__global__ __launch_bounds__(256) void k_assemble_band(View v0, View v1, int nrhs, const long *list,
                                                       const int *slot, int n, int b, int method,
                                                       double *B, double *rhs)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)n * (b + 1);
    if (t >= total) return;
    const int k = (int)(t / (b + 1)), d = (int)(t % (b + 1));
    const int c = k - d;
    int i, j;
    ij_of(v0, list[k], i, j);
    const int n_ = v0.rows, m_ = v0.cols;
    double a = 0.0;
    if (c >= 0) {
        int ci, cj;
        ij_of(v0, list[c], ci, cj);
        if (method == 4) {
            if (d == 0) {
                a = (i > 0) + (i < n_ - 1) + (j > 0) + (j < m_ - 1);
            } else if (abs(ci - i) + abs(cj - j) == 1) {
                a = -1.0;
            }
        } else {
            // rows centred at (i - di, j - dj) for every template offset (di, dj)
            for (int di = -2; di <= 2; ++di)
                for (int dj = -2; dj <= 2; ++dj) {
                    if (abs(di) + abs(dj) > 2) continue;
                    const int ri = i - di, rj = j - dj;
                    if (ri < 0 || ri >= n_ || rj < 0 || rj >= m_) continue;
                    const int wk = plate_w(ri, rj, n_, m_, di, dj);
                    if (!wk) continue;
                    const int ei = ci - ri, ej = cj - rj;
                    if (abs(ei) + abs(ej) > 2) continue;
                    const int wc = plate_w(ri, rj, n_, m_, ei, ej);
                    a += (double)wk * wc;
                }
        }
    }
    B[t] = a;
    if (d != 0) return;
    for (int q = 0; q < nrhs; ++q) {
        const View &v = q ? v1 : v0;
        double r = 0.0;
        if (method == 4) {
            const int nb[4][2] = {{i - 1, j}, {i + 1, j}, {i, j - 1}, {i, j + 1}};
            for (int e = 0; e < 4; ++e) {
                const int a2 = nb[e][0], b2 = nb[e][1];
                if (a2 < 0 || a2 >= n_ || b2 < 0 || b2 >= m_) continue;
                if (slot[ord_of(v, a2, b2)] >= 0) continue;
                r += at(v, a2, b2);
            }
        } else {
            for (int di = -2; di <= 2; ++di)
                for (int dj = -2; dj <= 2; ++dj) {
                    if (abs(di) + abs(dj) > 2) continue;
                    const int ri = i - di, rj = j - dj;
                    if (ri < 0 || ri >= n_ || rj < 0 || rj >= m_) continue;
                    const int wk = plate_w(ri, rj, n_, m_, di, dj);
                    if (!wk) continue;
                    double known = 0.0;  // the row's known part
                    for (int ei = -2; ei <= 2; ++ei)
                        for (int ej = -2; ej <= 2; ++ej) {
                            if (abs(ei) + abs(ej) > 2) continue;
                            const int pi = ri + ei, pj = rj + ej;
                            if (pi < 0 || pi >= n_ || pj < 0 || pj >= m_) continue;
                            const int w = plate_w(ri, rj, n_, m_, ei, ej);
                            if (!w || slot[ord_of(v, pi, pj)] >= 0) continue;
                            known += (double)w * at(v, pi, pj);
                        }
                    r -= (double)wk * known;
                }
        }
        rhs[q * (long)n + k] = r;
    }
}

// Banded Cholesky A = L L^T in place (band storage) and the solves L L^T x = rhs for
// nrhs right-hand sides, one workgroup; *err = 1 on a non-positive pivot.
__global__ __launch_bounds__(1024) void k_band_solve(double *B, int n, int b, double *rhs, int nrhs,
                                                     int *err)
{
    const int tid = threadIdx.x, W = b + 1;
    __shared__ int bad;
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        if (tid == 0) {
            const double p = B[(long)k * W];
            if (!(p > 0.0)) bad = 1;
            B[(long)k * W] = sqrt(p > 0.0 ? p : 1.0);
        }
        __syncthreads();
        const double dk = B[(long)k * W];
        const int m = min(b, n - 1 - k);
        for (int r = 1 + tid; r <= m; r += blockDim.x) B[(long)(k + r) * W + r] /= dk;
        __syncthreads();
        // trailing update: A[k+r][k+s] -= L[k+r][k] L[k+s][k], 1 <= s <= r <= m
        const long pairs = (long)m * (m + 1) / 2;
        for (long t = tid; t < pairs; t += blockDim.x) {
            // t -> (r, s) with r >= s >= 1: r = floor((sqrt(8t + 1) - 1) / 2) + 1
            int r = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
            while ((long)(r + 1) * (r + 2) / 2 <= t) ++r;
            while ((long)r * (r + 1) / 2 > t) --r;
            const int s = (int)(t - (long)r * (r + 1) / 2) + 1;
            r += 1;
            const double lr = B[(long)(k + r) * W + r], ls = B[(long)(k + s) * W + s];
            B[(long)(k + r) * W + (r - s)] -= lr * ls;
        }
        __syncthreads();
    }
    for (int q = 0; q < nrhs; ++q) {
        double *y = rhs + (long)q * n;
        for (int k = 0; k < n; ++k) {  // forward: L y = rhs
            if (tid == 0) y[k] /= B[(long)k * W];
            __syncthreads();
            const double yk = y[k];
            const int m = min(b, n - 1 - k);
            for (int r = 1 + tid; r <= m; r += blockDim.x) y[k + r] -= B[(long)(k + r) * W + r] * yk;
            __syncthreads();
        }
        for (int k = n - 1; k >= 0; --k) {  // backward: L^T x = y
            if (tid == 0) y[k] /= B[(long)k * W];
            __syncthreads();
            const double xk = y[k];
            const int m = min(b, k);
            for (int r = 1 + tid; r <= m; r += blockDim.x) y[k - r] -= B[(long)k * W + r] * xk;
            __syncthreads();
        }
    }
    if (tid == 0 && bad) *err = 1;
}

__global__ __launch_bounds__(256) void k_scatter_sol(View v0, View v1, int nrhs, const long *list,
                                                     int n, const double *x)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int i, j;
    ij_of(v0, list[k], i, j);
    at(v0, i, j) = x[k];
    if (nrhs > 1) at(v1, i, j) = x[(long)n + k];
}

// ---- NaN marks, the median test ------------------------------------------------------

__global__ __launch_bounds__(256) void k_mark_nan(View v, int *unk)
{
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (long)v.rows * v.cols) return;
    int i, j;
    ij_of(v, o, i, j);
    unk[o] = isnan(at(v, i, j)) ? 1 : 0;
}

__device__ __forceinline__ float med3(float a, float b, float c)
{
    return fmaxf(fminf(a, b), fminf(fmaxf(a, b), c));
}

// fl = single(u) - med3(single(u)) at each element, for u and v (filled copies)
__global__ __launch_bounds__(256) void k_fluct(View u, View v, float *flu, float *flv)
{
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (long)u.rows * u.cols) return;
    int i, j;
    ij_of(u, o, i, j);
    const int i1 = min(i + 1, u.rows - 1), j1 = min(j + 1, u.cols - 1);
    const float a = (float)at(u, i, j), mu = med3(a, (float)at(u, i1, j), (float)at(u, i1, j1));
    const float c = (float)at(v, i, j), mv = med3(c, (float)at(v, i1, j), (float)at(v, i1, j1));
    flu[o] = __fsub_rn(a, mu);
    flv[o] = __fsub_rn(c, mv);
}

// the normalised fluctuation magnitude against thr; unk = flagged or NaN in u0
__global__ __launch_bounds__(256) void k_flag(View u0, const float *flu, const float *flv, float thr,
                                              int *unk, int *flagged, long fsi, long fsj)
{
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (long)u0.rows * u0.cols) return;
    int i, j;
    ij_of(u0, o, i, j);
    const int i1 = min(i + 1, u0.rows - 1), j1 = min(j + 1, u0.cols - 1);
    const long o10 = ord_of(u0, i1, j), o11 = ord_of(u0, i1, j1);
    const float ru = med3(fabsf(flu[o]), fabsf(flu[o10]), fabsf(flu[o11]));
    const float rv = med3(fabsf(flv[o]), fabsf(flv[o10]), fabsf(flv[o11]));
    const float nu = fabsf(__fdiv_rn(flu[o], __fadd_rn(ru, 0.1f)));
    const float nv = fabsf(__fdiv_rn(flv[o], __fadd_rn(rv, 0.1f)));
    const float mag = __fsqrt_rn(__fadd_rn(__fmul_rn(nu, nu), __fmul_rn(nv, nv)));
    const int f = mag > thr ? 1 : 0;
    unk[o] = (f || isnan(at(u0, i, j))) ? 1 : 0;
    if (flagged) flagged[i * fsi + j * fsj] = f;
}

__global__ __launch_bounds__(256) void k_copy_view(View src, View dst)
{
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (long)src.rows * src.cols) return;
    int i, j;
    ij_of(src, o, i, j);
    at(dst, i, j) = at(src, i, j);
}

__global__ __launch_bounds__(256) void k_restore_nan(View v, const int *nanmask)
{
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (long)v.rows * v.cols) return;
    int i, j;
    ij_of(v, o, i, j);
    if (nanmask[o]) at(v, i, j) = NAN;
}

// the frame's mask holes: a node whose centre pixel is masked gets NaN (IntegerSearch.m:68-76;
// a non-integer centre leaves the field unmasked, as the reference's try/catch does)
__global__ __launch_bounds__(256) void k_mask_nodes(double *U, const double *coords, int S,
                                                    const uint8_t *mask, int nx, int ny)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= S) return;
    const double x = coords[2 * n], y = coords[2 * n + 1];
    if (x != floor(x) || y != floor(y)) return;
    const int xi = (int)x - 1, yi = (int)y - 1;
    if (xi < 0 || xi >= nx || yi < 0 || yi >= ny) return;
    if (!mask[(size_t)yi * nx + xi]) { U[2 * n] = NAN; U[2 * n + 1] = NAN; }
}

int nblk(long n, int t = 256) { return (int)((n + t - 1) / t); }

}  // namespace

// Work buffers of one field operation (grown on demand).
struct OutlierWork {
    DBuf unk, bcnt, nint, list, slot, band, B, rhs, err, flu, flv, copy, nanm;
};

namespace {

hipError_t sync_read(int *dst, const void *src, hipStream_t st)
{
    hipError_t e = hipMemcpyAsync(dst, src, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

// Numbers the marked unknowns of a field of N elements; returns n on the host.
hipError_t number_unknowns(OutlierWork &w, long N, int &n, hipStream_t st)
{
    const int nb = (int)((N + kChunk - 1) / kChunk);
    hipError_t e;
    if ((e = w.bcnt.ensure(sizeof(int) * (nb + 1))) || (e = w.nint.ensure(2 * sizeof(int))) ||
        (e = w.slot.ensure(sizeof(int) * N)))
        return e;
    hipLaunchKernelGGL(k_unk_count, dim3(nb), dim3(256), 0, st, w.unk.as<int>(), N, w.bcnt.as<int>());
    hipLaunchKernelGGL(k_unk_scan, dim3(1), dim3(1024), 0, st, w.bcnt.as<int>(), nb, w.nint.as<int>());
    if ((e = sync_read(&n, w.nint.p, st))) return e;
    if ((e = w.list.ensure(sizeof(long) * (n > 0 ? n : 1)))) return e;
    hipLaunchKernelGGL(k_unk_put, dim3(nb), dim3(256), 0, st, w.unk.as<int>(), N,
                       (const int *)w.bcnt.as<int>(), w.list.as<long>(), w.slot.as<int>());
    return hipGetLastError();
}

// The least squares in-painting (method 3 or 4) of the numbered unknowns of one or two
// views sharing them.  status: 0, or ALDIC_ERR_NUMERIC / ALDIC_ERR_ARG with msg set.
int solve_unknowns(OutlierWork &w, View v0, View v1, int nrhs, int n, int method, hipStream_t st,
                   const char *&msg)
{
    if (n == 0) return 0;
    hipError_t e;
    if ((e = w.band.ensure(sizeof(int)))) return -1;
    if ((e = launch_iset(w.band.as<int>(), 1, 0, st))) return -1;
    hipLaunchKernelGGL(k_bandwidth, dim3(nblk(n)), dim3(256), 0, st, v0, (const long *)w.list.as<long>(),
                       (const int *)w.slot.as<int>(), n, method == 3 ? 4 : 1, w.band.as<int>());
    int b = 0;
    if ((e = sync_read(&b, w.band.p, st))) return -1;
    const double bytes = 8.0 * n * (b + 1.0);
    if (bytes > 512.0 * 1024 * 1024) {
        msg = "inpaint: the band of the in-painting system exceeds 512 MiB";
        return ALDIC_ERR_NUMERIC;
    }
    if ((e = w.B.ensure((size_t)bytes)) || (e = w.rhs.ensure(sizeof(double) * 2 * (size_t)n)) ||
        (e = w.err.ensure(sizeof(int))))
        return -1;
    if ((e = launch_iset(w.err.as<int>(), 1, 0, st))) return -1;
    const long tot = (long)n * (b + 1);
    hipLaunchKernelGGL(k_assemble_band, dim3(nblk(tot)), dim3(256), 0, st, v0, v1, nrhs,
                       (const long *)w.list.as<long>(), (const int *)w.slot.as<int>(), n, b, method,
                       w.B.as<double>(), w.rhs.as<double>());
    hipLaunchKernelGGL(k_band_solve, dim3(1), dim3(1024), 0, st, w.B.as<double>(), n, b,
                       w.rhs.as<double>(), nrhs, w.err.as<int>());
    int bad = 0;
    if ((e = sync_read(&bad, w.err.p, st))) return -1;
    if (bad) {
        msg = "inpaint: the in-painting system is singular (no known value reaches the NaNs)";
        return ALDIC_ERR_NUMERIC;
    }
    hipLaunchKernelGGL(k_scatter_sol, dim3(nblk(n)), dim3(256), 0, st, v0, v1, nrhs,
                       (const long *)w.list.as<long>(), n, (const double *)w.rhs.as<double>());
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

// ---- entry points used by api.hip ------------------------------------------------------

// inpaint_nans(A, method) in place on a device view; returns 0, -1 (HIP error) or an
// ALDIC error code with msg.
int inpaint_view(OutlierWork &w, double *p, int rows, int cols, long si, long sj, int method,
                 hipStream_t st, const char *&msg)
{
    const View v{p, rows, cols, si, sj};
    const long N = (long)rows * cols;
    if (w.unk.ensure(sizeof(int) * N)) return -1;
    hipLaunchKernelGGL(k_mark_nan, dim3(nblk(N)), dim3(256), 0, st, v, w.unk.as<int>());
    int n = 0;
    if (number_unknowns(w, N, n, st)) return -1;
    if (n == N) {
        msg = "inpaint: every element is NaN";
        return ALDIC_ERR_NUMERIC;
    }
    return solve_unknowns(w, v, v, 1, n, method, st, msg);
}

// funRemoveOutliers(u, v, [], 0.5, thr) without prompts on two device views of one shape:
// the median test (NaNs pre-filled by method 3 in copies), the flagged and the NaN
// elements in-painted together by springs; keep_nan: put the NaNs back (the reference's
// return value) or keep them filled (what Init.m:13 then does, in the same solve).
int remove_outliers_views(OutlierWork &w, double *up, double *vp, int rows, int cols, long si,
                          long sj, float thr, bool keep_nan, int *flagged, long fsi, long fsj,
                          int *n_flagged, hipStream_t st, const char *&msg)
{
    const View u{up, rows, cols, si, sj}, v{vp, rows, cols, si, sj};
    const long N = (long)rows * cols;
    if (w.unk.ensure(sizeof(int) * N) || w.copy.ensure(sizeof(double) * 2 * N) ||
        w.flu.ensure(sizeof(float) * N) || w.flv.ensure(sizeof(float) * N) ||
        w.nanm.ensure(sizeof(int) * N))
        return -1;
    hipLaunchKernelGGL(k_mark_nan, dim3(nblk(N)), dim3(256), 0, st, u, w.nanm.as<int>());
    hipLaunchKernelGGL(k_mark_nan, dim3(nblk(N)), dim3(256), 0, st, u, w.unk.as<int>());
    int n = 0;
    if (number_unknowns(w, N, n, st)) return -1;
    if (n == N) {
        msg = "remove_outliers: every element is NaN";
        return ALDIC_ERR_NUMERIC;
    }
    int r;
    View uf = u, vf = v;
    if (n > 0) {
        // the NaNs pre-filled by method 3 in copies (memory order, faster axis first)
        const bool jfast = sj <= si;
        uf = View{w.copy.as<double>(), rows, cols, jfast ? (long)cols : 1L, jfast ? 1L : (long)rows};
        vf = View{w.copy.as<double>() + N, rows, cols, uf.si, uf.sj};
        hipLaunchKernelGGL(k_copy_view, dim3(nblk(N)), dim3(256), 0, st, u, uf);
        hipLaunchKernelGGL(k_copy_view, dim3(nblk(N)), dim3(256), 0, st, v, vf);
        if ((r = solve_unknowns(w, uf, vf, 2, n, 3, st, msg))) return r;
    }
    hipLaunchKernelGGL(k_fluct, dim3(nblk(N)), dim3(256), 0, st, uf, vf, w.flu.as<float>(),
                       w.flv.as<float>());
    hipLaunchKernelGGL(k_flag, dim3(nblk(N)), dim3(256), 0, st, u, (const float *)w.flu.as<float>(),
                       (const float *)w.flv.as<float>(), thr, w.unk.as<int>(), flagged, fsi, fsj);
    int m = 0;
    if (number_unknowns(w, N, m, st)) return -1;
    if (n_flagged) *n_flagged = m - n;
    if (m > n && (r = solve_unknowns(w, u, v, 2, m, 4, st, msg))) return r;
    if (m == n && n > 0 && !keep_nan && (r = solve_unknowns(w, u, v, 2, m, 4, st, msg))) return r;
    if (keep_nan) {
        hipLaunchKernelGGL(k_restore_nan, dim3(nblk(N)), dim3(256), 0, st, u, (const int *)w.nanm.as<int>());
        hipLaunchKernelGGL(k_restore_nan, dim3(nblk(N)), dim3(256), 0, st, v, (const int *)w.nanm.as<int>());
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

hipError_t launch_mask_nodes(double *U, const double *coords, int S, const uint8_t *mask, int nx,
                             int ny, hipStream_t st)
{
    hipLaunchKernelGGL(k_mask_nodes, dim3(nblk(S)), dim3(256), 0, st, U, coords, S, mask, nx, ny);
    return hipGetLastError();
}

OutlierWork *outlier_work_new() { return new OutlierWork(); }
void outlier_work_free(OutlierWork *w)
{
    if (!w) return;
    for (DBuf *b : {&w->unk, &w->bcnt, &w->nint, &w->list, &w->slot, &w->band, &w->B, &w->rhs,
                    &w->err, &w->flu, &w->flv, &w->copy, &w->nanm})
        b->release();
    delete w;
}

}  // namespace aldic
