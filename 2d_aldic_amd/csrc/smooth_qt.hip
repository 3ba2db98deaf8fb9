// funSmoothStrainQuadtree (func_quadtree/funSmoothStrainQuadtree.m:27-104): two passes of
// regularizeNd (func/regularizeNd.m, 2-D, 'linear', 'normal' solver) of the four F
// components over the grid min:h:max, a 3x3 Gaussian nanconv ('edge') and a read-back at
// the nodes, which lie on the grid.
//
// With every node on a grid point, regularizeNd's fidelity rows select single grid points
// (bilinear weights 1 and 0), so its normal equations are
//     (D + s_x^2 Lx^T Lx + s_y^2 Ly^T Ly) z = b,
// D = nodes per grid point, b = the sum of their values, L_d the second difference
// (1, -2, 1) along d at the interior points and s_d = smoothness sqrt(n_nodes / n_eq_d)
// (n_d - 1)^2 (regularizeNd.m:666-672 with secondDerivativeWeights on a uniform grid:
// (range_d / h)^2 = (n_d - 1)^2).  The operator is applied matrix-free (a 13-point
// stencil) and the four components are solved together by Jacobi-preconditioned CG in
// FP64 (one shared recurrence over the stacked vector).
#include <cmath>
#include <vector>

#include "aldic_internal.h"
#include "common.h"

namespace aldic {

namespace {

constexpr int kSqT = 256;
constexpr int kSqMaxB = 256;

struct SqOp {
    int nx, ny;
    double cx, cy;           // s_x^2, s_y^2
    const double *D;         // nodes per grid point
};

// (L^T L v)_j along one axis of n points for the value at offset k (v(k) loads), row
// i = j - 1, j, j + 1 contributing c_(j - i + 1) (L v)_i, c = (1, -2, 1), rows 1..n-2
template <class V>
__device__ __forceinline__ double4 ltl(int j, int n, V v)
{
    double4 y = make_double4(0, 0, 0, 0);
    const double c[3] = {1.0, -2.0, 1.0};
#pragma unroll
    for (int a = -1; a <= 1; ++a) {
        const int i = j + a;   // row centre
        if (i < 1 || i > n - 2) continue;
        const double4 l = v(i - 1 - j), m = v(i - j), r = v(i + 1 - j);
        const double w = c[j - i + 1];
        y.x += w * (l.x - 2.0 * m.x + r.x);
        y.y += w * (l.y - 2.0 * m.y + r.y);
        y.z += w * (l.z - 2.0 * m.z + r.z);
        y.w += w * (l.w - 2.0 * m.w + r.w);
    }
    return y;
}

__device__ __forceinline__ double ltl_diag(int j, int n)
{
    double d = 0.0;
    const double c2[3] = {1.0, 4.0, 1.0};
    for (int a = -1; a <= 1; ++a) {
        const int i = j + a;
        if (i >= 1 && i <= n - 2) d += c2[j - i + 1];
    }
    return d;
}

__device__ __forceinline__ double4 sq_apply(const SqOp &o, const double4 *__restrict__ x, int g)
{
    const int ix = g % o.nx, iy = g / o.nx;
    const double4 xc = x[g];
    const double4 ax = ltl(ix, o.nx, [&](int k) { return x[g + k]; });
    const double4 ay = ltl(iy, o.ny, [&](int k) { return x[g + k * o.nx]; });
    const double d = o.D[g];
    return make_double4(d * xc.x + o.cx * ax.x + o.cy * ay.x, d * xc.y + o.cx * ax.y + o.cy * ay.y,
                        d * xc.z + o.cx * ax.z + o.cy * ay.z, d * xc.w + o.cx * ax.w + o.cy * ay.w);
}

__device__ __forceinline__ double dot4(double4 a, double4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }

// block sums of up to three values, thread 0 writes them
__device__ __forceinline__ void sq_block(double a, double b, double c, double *oa, double *ob, double *oc)
{
    __shared__ double red[3][kSqT / 64];
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][w] = a; red[1][w] = b; red[2][w] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0, s1 = 0, s2 = 0;
        for (int k = 0; k < kSqT / 64; ++k) { s0 += red[0][k]; s1 += red[1][k]; s2 += red[2][k]; }
        if (oa) oa[blockIdx.x] = s0;
        if (ob) ob[blockIdx.x] = s1;
        if (oc) oc[blockIdx.x] = s2;
    }
}

// sum of nb partials, the same in every block (fixed order)
__device__ __forceinline__ double sq_total(const double *p, int nb)
{
    __shared__ double sh[kSqT / 64];
    double v = 0.0;
    for (int k = threadIdx.x; k < nb; k += kSqT) v += p[k];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < kSqT / 64; ++k) s += sh[k];
    __syncthreads();
    return s;
}

// nodes -> grid: counts and sums (the fidelity part of the normal equations)
__global__ void k_sq_scatter(int nn, const int *__restrict__ gidx, const double4 *__restrict__ F,
                             double *__restrict__ D, double4 *__restrict__ b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int g = gidx[i];
    const double4 f = F[i];
    atomicAdd(&D[g], 1.0);
    atomicAdd(&b[g].x, f.x);
    atomicAdd(&b[g].y, f.y);
    atomicAdd(&b[g].z, f.z);
    atomicAdd(&b[g].w, f.w);
}

struct SqVec {
    double4 *x, *r, *z, *p, *q;
    const double4 *b;
    double *dinv;
    double *prz[2], *ppq, *prr, *pbb;   // per-block partials
    double *stat;                        // {rr, bb} after the last update
    int n, nb;
};

// from the starting x (zero in the first pass, the first pass's solution in the second):
// r = b - N x, z = M^-1 r, p = z; partials of r.z and b.b
__global__ __launch_bounds__(kSqT) void k_sq_init(SqOp o, SqVec v)
{
    double rz = 0.0, bb = 0.0;
    for (int g = blockIdx.x * kSqT + threadIdx.x; g < v.n; g += gridDim.x * kSqT) {
        const int ix = g % o.nx, iy = g / o.nx;
        const double dg = o.D[g] + o.cx * ltl_diag(ix, o.nx) + o.cy * ltl_diag(iy, o.ny);
        const double di = 1.0 / dg;
        v.dinv[g] = di;
        const double4 b = v.b[g], nx = sq_apply(o, v.x, g);
        const double4 r = make_double4(b.x - nx.x, b.y - nx.y, b.z - nx.z, b.w - nx.w);
        const double4 z = make_double4(r.x * di, r.y * di, r.z * di, r.w * di);
        v.r[g] = r;
        v.z[g] = z;
        v.p[g] = z;
        rz += dot4(r, z);
        bb += dot4(b, b);
    }
    sq_block(rz, bb, 0.0, v.prz[0], v.pbb, nullptr);
}

// q = N p, partials of p.q
__global__ __launch_bounds__(kSqT) void k_sq_apply(SqOp o, SqVec v)
{
    double pq = 0.0;
    for (int g = blockIdx.x * kSqT + threadIdx.x; g < v.n; g += gridDim.x * kSqT) {
        const double4 q = sq_apply(o, v.p, g);
        v.q[g] = q;
        pq += dot4(v.p[g], q);
    }
    sq_block(pq, 0.0, 0.0, v.ppq, nullptr, nullptr);
}

// alpha = r.z / p.q; x += alpha p, r -= alpha q, z = M^-1 r; partials of the new r.z, r.r
__global__ __launch_bounds__(kSqT) void k_sq_update(SqVec v, int par)
{
    const double rz = sq_total(v.prz[par], v.nb), pq = sq_total(v.ppq, v.nb);
    const double alpha = (pq > 0.0 && isfinite(pq)) ? rz / pq : 0.0;
    double rzn = 0.0, rr = 0.0;
    for (int g = blockIdx.x * kSqT + threadIdx.x; g < v.n; g += gridDim.x * kSqT) {
        const double4 p = v.p[g], q = v.q[g];
        double4 x = v.x[g], r = v.r[g];
        x.x += alpha * p.x; x.y += alpha * p.y; x.z += alpha * p.z; x.w += alpha * p.w;
        r.x -= alpha * q.x; r.y -= alpha * q.y; r.z -= alpha * q.z; r.w -= alpha * q.w;
        const double di = v.dinv[g];
        const double4 z = make_double4(r.x * di, r.y * di, r.z * di, r.w * di);
        v.x[g] = x;
        v.r[g] = r;
        v.z[g] = z;
        rzn += dot4(r, z);
        rr += dot4(r, r);
    }
    sq_block(rzn, rr, 0.0, v.prz[par ^ 1], v.prr, nullptr);
}

// beta = r.z (new) / r.z (old); p = z + beta p; block 0 records {r.r, b.b}
__global__ __launch_bounds__(kSqT) void k_sq_dir(SqVec v, int par)
{
    const double rzo = sq_total(v.prz[par], v.nb), rzn = sq_total(v.prz[par ^ 1], v.nb);
    const double rr = sq_total(v.prr, v.nb), bb = sq_total(v.pbb, v.nb);
    const double beta = rzo > 0.0 ? rzn / rzo : 0.0;
    for (int g = blockIdx.x * kSqT + threadIdx.x; g < v.n; g += gridDim.x * kSqT) {
        const double4 z = v.z[g];
        double4 p = v.p[g];
        p.x = z.x + beta * p.x; p.y = z.y + beta * p.y; p.z = z.z + beta * p.z; p.w = z.w + beta * p.w;
        v.p[g] = p;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        v.stat[0] = rr;
        v.stat[1] = bb;
    }
}

// grid -> nodes
__global__ void k_sq_gather(int nn, const int *__restrict__ gidx, const double4 *__restrict__ G,
                            double4 *__restrict__ F)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nn) F[i] = G[gidx[i]];
}

}  // namespace

// F (4 nn, in/out, on the device): one call = both passes.  iters[0..1] = CG iterations.
hipError_t smooth_strain_qt(int nn, const int *gidx_dev, int nx, int ny, double smoothness, double sigma, double rtol,
                            int maxit, double *F, DBuf *w, int *iters, hipStream_t st)
{
    const int n = nx * ny;
    hipError_t e;
    const int nb = std::min(kSqMaxB, (n + kSqT - 1) / kSqT);
    const size_t v4 = sizeof(double4) * (size_t)n;
    if ((e = w[0].ensure(7 * v4)) || (e = w[1].ensure(sizeof(double) * (2 * (size_t)n + 6 * kSqMaxB + 4))))
        return e;
    double4 *base = w[0].as<double4>();
    double4 *bv = base + 5 * (size_t)n, *blur = base + 6 * (size_t)n;
    double *D = w[1].as<double>(), *dinv = D + n, *parts = dinv + n;
    SqOp o{nx, ny, 0.0, 0.0, D};
    const double neq_x = (double)(nx - 2) * ny, neq_y = (double)(ny - 2) * nx;
    const double sx = smoothness * std::sqrt(nn / neq_x) * (double)(nx - 1) * (nx - 1);
    const double sy = smoothness * std::sqrt(nn / neq_y) * (double)(ny - 1) * (ny - 1);
    o.cx = sx * sx;
    o.cy = sy * sy;
    SqVec v{};
    v.x = base;
    v.r = base + n;
    v.z = base + 2 * (size_t)n;
    v.p = base + 3 * (size_t)n;
    v.q = base + 4 * (size_t)n;
    v.b = bv;
    v.dinv = dinv;
    v.prz[0] = parts;
    v.prz[1] = parts + kSqMaxB;
    v.ppq = parts + 2 * kSqMaxB;
    v.prr = parts + 3 * kSqMaxB;
    v.pbb = parts + 4 * kSqMaxB;
    v.stat = parts + 5 * kSqMaxB;
    v.n = n;
    v.nb = nb;
    double hs[2];
    if ((e = hipMemsetAsync(v.x, 0, v4, st))) return e;
    for (int pass = 0; pass < 2; ++pass) {
        if ((e = hipMemsetAsync(D, 0, sizeof(double) * n, st)) || (e = hipMemsetAsync(bv, 0, v4, st))) return e;
        hipLaunchKernelGGL(k_sq_scatter, dim3((nn + 255) / 256), dim3(256), 0, st, nn, gidx_dev,
                           (const double4 *)F, D, bv);
        hipLaunchKernelGGL(k_sq_init, dim3(nb), dim3(kSqT), 0, st, o, v);
        int it = 0;
        bool done = false;
        while (!done && it < maxit) {
            const int chunk = std::min(16, maxit - it);
            for (int k = 0; k < chunk; ++k, ++it) {
                hipLaunchKernelGGL(k_sq_apply, dim3(nb), dim3(kSqT), 0, st, o, v);
                hipLaunchKernelGGL(k_sq_update, dim3(nb), dim3(kSqT), 0, st, v, it & 1);
                hipLaunchKernelGGL(k_sq_dir, dim3(nb), dim3(kSqT), 0, st, v, it & 1);
            }
            if ((e = hipGetLastError())) return e;
            if ((e = hipMemcpyAsync(hs, v.stat, sizeof(hs), hipMemcpyDeviceToHost, st)) ||
                (e = hipStreamSynchronize(st)))
                return e;
            done = !(hs[0] > rtol * rtol * hs[1]);
        }
        if (iters) iters[pass] = it;
        if (!std::isfinite(hs[0])) return hipErrorUnknown;
        if ((e = launch_blur3(reinterpret_cast<const double *>(v.x), reinterpret_cast<double *>(blur), nx, ny, 4,
                              sigma, st)))
            return e;
        hipLaunchKernelGGL(k_sq_gather, dim3((nn + 255) / 256), dim3(256), 0, st, nn, gidx_dev,
                           (const double4 *)blur, reinterpret_cast<double4 *>(F));
    }
    return hipGetLastError();
}

}  // namespace aldic
