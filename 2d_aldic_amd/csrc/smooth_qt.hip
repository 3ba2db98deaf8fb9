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
// stencil) and the four components are solved together by preconditioned CG in FP64 (one
// shared recurrence over the stacked vector).  The preconditioner is the operator with D
// replaced by its mean d: M = d I + s_x^2 (T_x (+) I) + s_y^2 (I (+) T_y), T = L^T L,
// inverted exactly by fast diagonalisation (T's eigenvectors from a host Jacobi
// eigensolver, cached per grid size; four 16 x 16-tiled FP64 products per application).
#include <cmath>
#include <map>
#include <mutex>
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
    double *prz[2], *ppq, *prr, *pbb;   // per-block partials
    double *stat;                        // {rr, bb} after the last update
    int n, nb;
};

// from the starting x (zero in the first pass, the first pass's solution in the second):
// r = b - N x; partials of b.b
__global__ __launch_bounds__(kSqT) void k_sq_init(SqOp o, SqVec v)
{
    double bb = 0.0;
    for (int g = blockIdx.x * kSqT + threadIdx.x; g < v.n; g += gridDim.x * kSqT) {
        const double4 b = v.b[g], nx = sq_apply(o, v.x, g);
        v.r[g] = make_double4(b.x - nx.x, b.y - nx.y, b.z - nx.z, b.w - nx.w);
        bb += dot4(b, b);
    }
    sq_block(bb, 0.0, 0.0, v.pbb, nullptr, nullptr);
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

// alpha = r.z / p.q; x += alpha p, r -= alpha q; partials of r.r
__global__ __launch_bounds__(kSqT) void k_sq_update(SqVec v, int par)
{
    const double rz = sq_total(v.prz[par], v.nb), pq = sq_total(v.ppq, v.nb);
    const double alpha = (pq > 0.0 && isfinite(pq)) ? rz / pq : 0.0;
    double rr = 0.0;
    for (int g = blockIdx.x * kSqT + threadIdx.x; g < v.n; g += gridDim.x * kSqT) {
        const double4 p = v.p[g], q = v.q[g];
        double4 x = v.x[g], r = v.r[g];
        x.x += alpha * p.x; x.y += alpha * p.y; x.z += alpha * p.z; x.w += alpha * p.w;
        r.x -= alpha * q.x; r.y -= alpha * q.y; r.z -= alpha * q.z; r.w -= alpha * q.w;
        v.x[g] = x;
        v.r[g] = r;
        rr += dot4(r, r);
    }
    sq_block(rr, 0.0, 0.0, v.prr, nullptr, nullptr);
}

// partials of r.z (z = M^-1 r, just formed) into prz[par]; INIT: p = z
template <bool INIT>
__global__ __launch_bounds__(kSqT) void k_sq_rz(SqVec v, int par)
{
    double rz = 0.0;
    for (int g = blockIdx.x * kSqT + threadIdx.x; g < v.n; g += gridDim.x * kSqT) {
        const double4 z = v.z[g];
        rz += dot4(v.r[g], z);
        if (INIT) v.p[g] = z;
    }
    sq_block(rz, 0.0, 0.0, v.prz[par], nullptr, nullptr);
}

// Fast-diagonalisation products on the ny x nx grid of double4 values (x fastest), 16 x 16
// tiles through LDS.  Right: Y[r][c] = sum_k X[r][k] V'(k, c), V' = V or V^T (V n x n,
// row-major), X rows x n.  Left: Y[r][c] = sum_k V'(r, k) X[k][c], X n x cols; SCALE
// divides by d + cx lx[c] + cy ly[r] (the eigenvalues of M).
template <bool TR>
__global__ __launch_bounds__(256) void k_fd_right(const double4 *__restrict__ X, const double *__restrict__ V,
                                                  double4 *__restrict__ Y, int rows, int n)
{
    __shared__ double4 xs[16][17];
    __shared__ double vs[16][17];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int r = blockIdx.y * 16 + ty, c = blockIdx.x * 16 + tx;
    double4 acc = make_double4(0, 0, 0, 0);
    for (int k0 = 0; k0 < n; k0 += 16) {
        xs[ty][tx] = (r < rows && k0 + tx < n) ? X[(size_t)r * n + k0 + tx] : make_double4(0, 0, 0, 0);
        const int kk = k0 + ty, cc = blockIdx.x * 16 + tx;
        vs[ty][tx] = (kk < n && cc < n) ? (TR ? V[(size_t)cc * n + kk] : V[(size_t)kk * n + cc]) : 0.0;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const double4 a = xs[ty][j];
            const double w = vs[j][tx];
            acc.x += a.x * w; acc.y += a.y * w; acc.z += a.z * w; acc.w += a.w * w;
        }
        __syncthreads();
    }
    if (r < rows && c < n) Y[(size_t)r * n + c] = acc;
}

template <bool TR, bool SCALE>
__global__ __launch_bounds__(256) void k_fd_left(const double *__restrict__ V, const double4 *__restrict__ X,
                                                 double4 *__restrict__ Y, int n, int cols, const double *lx,
                                                 const double *ly, double d, double cx, double cy)
{
    __shared__ double vs[16][17];
    __shared__ double4 xs[16][17];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int r = blockIdx.y * 16 + ty, c = blockIdx.x * 16 + tx;
    double4 acc = make_double4(0, 0, 0, 0);
    for (int k0 = 0; k0 < n; k0 += 16) {
        const int k = k0 + tx;
        vs[ty][tx] = (r < n && k < n) ? (TR ? V[(size_t)k * n + r] : V[(size_t)r * n + k]) : 0.0;
        xs[ty][tx] = (k0 + ty < n && c < cols) ? X[(size_t)(k0 + ty) * cols + c] : make_double4(0, 0, 0, 0);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const double w = vs[ty][j];
            const double4 a = xs[j][tx];
            acc.x += w * a.x; acc.y += w * a.y; acc.z += w * a.z; acc.w += w * a.w;
        }
        __syncthreads();
    }
    if (r < n && c < cols) {
        if (SCALE) {
            const double iv = 1.0 / (d + cx * lx[c] + cy * ly[r]);
            acc.x *= iv; acc.y *= iv; acc.z *= iv; acc.w *= iv;
        }
        Y[(size_t)r * cols + c] = acc;
    }
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

__global__ void k_blend_f(double *__restrict__ F2, const double *__restrict__ F1, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) F2[i] = 0.1 * F2[i] + 0.9 * F1[i];
}

// T = L^T L for n points (L: second differences (1, -2, 1) at the interior points) and
// its eigen-decomposition by cyclic Jacobi rotations: V (row-major, column j the
// eigenvector of lam[j]).  Cached per n (host).
struct FdAxis {
    std::vector<double> V, lam;
};

const FdAxis &fd_axis(int n)
{
    static std::mutex mu;
    static std::map<int, FdAxis> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(n);
    if (it != cache.end()) return it->second;
    std::vector<double> A((size_t)n * n, 0.0), V((size_t)n * n, 0.0);
    const double c[3] = {1.0, -2.0, 1.0};
    for (int i = 1; i + 1 < n; ++i)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) A[(size_t)(i - 1 + a) * n + (i - 1 + b)] += c[a] * c[b];
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    double norm = 0.0;
    for (double x : A) norm += x * x;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += A[(size_t)p * n + q] * A[(size_t)p * n + q];
        if (off <= 1e-30 * norm) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[(size_t)p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double th = (A[(size_t)q * n + q] - A[(size_t)p * n + p]) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < n; ++k) {   // columns p, q
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = cs * akp - sn * akq;
                    A[(size_t)k * n + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; ++k) {   // rows p, q
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = cs * apk - sn * aqk;
                    A[(size_t)q * n + k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = cs * vkp - sn * vkq;
                    V[(size_t)k * n + q] = sn * vkp + cs * vkq;
                }
            }
    }
    FdAxis f;
    f.V = std::move(V);
    f.lam.resize(n);
    for (int i = 0; i < n; ++i) f.lam[i] = std::max(A[(size_t)i * n + i], 0.0);
    return cache.emplace(n, std::move(f)).first->second;
}

}  // namespace

hipError_t launch_blend_f(double *F2, const double *F1, long n, hipStream_t st)
{
    hipLaunchKernelGGL(k_blend_f, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, F2, F1, n);
    return hipGetLastError();
}

bool smooth_grid(int nn, const double *coords, double h, int &nx, int &ny, std::vector<int> &gidx)
{
    if (nn < 1 || !(h > 0)) return false;
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    for (int i = 0; i < nn; ++i) {
        x0 = std::min(x0, coords[2 * i]); x1 = std::max(x1, coords[2 * i]);
        y0 = std::min(y0, coords[2 * i + 1]); y1 = std::max(y1, coords[2 * i + 1]);
    }
    if (!std::isfinite(x0 + x1 + y0 + y1)) return false;
    const long nxl = (long)std::floor((x1 - x0) / h + 1e-9) + 1, nyl = (long)std::floor((y1 - y0) / h + 1e-9) + 1;
    if (nxl < 3 || nyl < 3 || nxl * nyl > (1L << 26)) return false;
    nx = (int)nxl;
    ny = (int)nyl;
    gidx.resize(nn);
    for (int i = 0; i < nn; ++i) {
        const long ix = std::lround((coords[2 * i] - x0) / h), iy = std::lround((coords[2 * i + 1] - y0) / h);
        if (ix < 0 || ix >= nx || iy < 0 || iy >= ny || x0 + ix * h != coords[2 * i] || y0 + iy * h != coords[2 * i + 1])
            return false;
        gidx[i] = (int)(iy * nx + ix);
    }
    return true;
}

// F (4 nn, in/out, on the device): one call = both passes.  iters[0..1] = CG iterations.
hipError_t smooth_strain_qt(int nn, const int *gidx_dev, int nx, int ny, double smoothness, double sigma, double rtol,
                            int maxit, double *F, DBuf *w, int *iters, hipStream_t st)
{
    const int n = nx * ny;
    hipError_t e;
    const int nb = std::min(kSqMaxB, (n + kSqT - 1) / kSqT);
    const size_t v4 = sizeof(double4) * (size_t)n;
    const FdAxis &ax = fd_axis(nx), &ay = fd_axis(ny);
    const size_t nfd = (size_t)nx * nx + (size_t)ny * ny + nx + ny;
    if ((e = w[0].ensure(7 * v4)) ||
        (e = w[1].ensure(sizeof(double) * ((size_t)n + 6 * kSqMaxB + 4 + nfd))))
        return e;
    double4 *base = w[0].as<double4>();
    double4 *bv = base + 5 * (size_t)n, *blur = base + 6 * (size_t)n;
    double *D = w[1].as<double>(), *parts = D + n, *Vx = parts + 6 * kSqMaxB + 4;
    double *Vy = Vx + (size_t)nx * nx, *lx = Vy + (size_t)ny * ny, *ly = lx + nx;
    if ((e = hipMemcpyAsync(Vx, ax.V.data(), sizeof(double) * ax.V.size(), hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(Vy, ay.V.data(), sizeof(double) * ay.V.size(), hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(lx, ax.lam.data(), sizeof(double) * nx, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(ly, ay.lam.data(), sizeof(double) * ny, hipMemcpyHostToDevice, st)))
        return e;
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
    v.prz[0] = parts;
    v.prz[1] = parts + kSqMaxB;
    v.ppq = parts + 2 * kSqMaxB;
    v.prr = parts + 3 * kSqMaxB;
    v.pbb = parts + 4 * kSqMaxB;
    v.stat = parts + 5 * kSqMaxB;
    v.n = n;
    v.nb = nb;
    // z = M^-1 r: R Vx, Vy^T (.) scaled, Vy (.), (.) Vx^T; q and blur as temporaries
    const double dmean = (double)nn / (double)n;
    const dim3 gfd((nx + 15) / 16, (ny + 15) / 16);
    auto precond = [&]() {
        hipLaunchKernelGGL(k_fd_right<false>, gfd, dim3(256), 0, st, (const double4 *)v.r, (const double *)Vx, v.q,
                           ny, nx);
        hipLaunchKernelGGL((k_fd_left<true, true>), gfd, dim3(256), 0, st, (const double *)Vy, (const double4 *)v.q,
                           blur, ny, nx, (const double *)lx, (const double *)ly, dmean, o.cx, o.cy);
        hipLaunchKernelGGL((k_fd_left<false, false>), gfd, dim3(256), 0, st, (const double *)Vy,
                           (const double4 *)blur, v.q, ny, nx, (const double *)lx, (const double *)ly, dmean, o.cx,
                           o.cy);
        hipLaunchKernelGGL(k_fd_right<true>, gfd, dim3(256), 0, st, (const double4 *)v.q, (const double *)Vx, v.z,
                           ny, nx);
    };
    double hs[2];
    if ((e = hipMemsetAsync(v.x, 0, v4, st))) return e;
    for (int pass = 0; pass < 2; ++pass) {
        if ((e = hipMemsetAsync(D, 0, sizeof(double) * n, st)) || (e = hipMemsetAsync(bv, 0, v4, st))) return e;
        hipLaunchKernelGGL(k_sq_scatter, dim3((nn + 255) / 256), dim3(256), 0, st, nn, gidx_dev,
                           (const double4 *)F, D, bv);
        hipLaunchKernelGGL(k_sq_init, dim3(nb), dim3(kSqT), 0, st, o, v);
        precond();
        hipLaunchKernelGGL(k_sq_rz<true>, dim3(nb), dim3(kSqT), 0, st, v, 0);
        int it = 0;
        bool done = false;
        while (!done && it < maxit) {
            const int chunk = std::min(8, maxit - it);
            for (int k = 0; k < chunk; ++k, ++it) {
                hipLaunchKernelGGL(k_sq_apply, dim3(nb), dim3(kSqT), 0, st, o, v);
                hipLaunchKernelGGL(k_sq_update, dim3(nb), dim3(kSqT), 0, st, v, it & 1);
                precond();
                hipLaunchKernelGGL(k_sq_rz<false>, dim3(nb), dim3(kSqT), 0, st, v, (it & 1) ^ 1);
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
