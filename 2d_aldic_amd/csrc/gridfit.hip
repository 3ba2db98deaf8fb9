// gridfit smoothing of nodal data on a rectilinear node grid (Subpb1.m:185-199, which
// calls gridfit.m with 'regularizer','springs' and otherwise the defaults: smoothness 1,
// 'triangle' interpolation, autoscale, backslash).
//
// The data sit exactly on the nodes (Subpb1 fits the nodes' own U over xnodes = unique
// node x, ynodes = unique node y), so gridfit's interpolation matrix A selects the
// nodes that carry data (gridfit.m:460-511 with tx, ty in {0, 1}) and |A|_1 = 1.  The
// least-squares fit of [A; lam Areg] z = [d; 0] (gridfit.m:651-666, lam = 1 / |Areg|_1)
// is the solution of the normal equations
//     (D + lam^2 L) z = D d,
// D = diag(node has data), L = Areg^T Areg the weighted graph Laplacian of the springs:
// axis springs of stiffness xscale/dx, yscale/dy and both diagonals of every cell with
// 1/sqrt((dx/xscale)^2 + (dy/yscale)^2) (gridfit.m:540-569; L weights = stiffness^2).
//
// The FP64 9-point stencil of N is assembled on the device from the data mask and
// solved by Jacobi-preconditioned conjugate gradients with one kernel per iteration.
#include <algorithm>
#include <cmath>
#include <vector>

#include "aldic_internal.h"
#include "common.h"

namespace aldic {

namespace {

constexpr int kGfThreads = 256;
constexpr int kGfMaxBlocks = 512;

// block sums of three values, written by thread 0 to o0[b], o1[b], o2[b]
__device__ __forceinline__ void gf_block3(double a, double b, double c, double *o0, double *o1,
                                          double *o2)
{
    __shared__ double red[3][kGfThreads / 64];
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][w] = a; red[1][w] = b; red[2][w] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int q = 0; q < kGfThreads / 64; ++q) { s0 += red[0][q]; s1 += red[1][q]; s2 += red[2][q]; }
        o0[blockIdx.x] = s0;
        if (o1) o1[blockIdx.x] = s1;
        if (o2) o2[blockIdx.x] = s2;
    }
}

// Conjugate gradients with one global reduction per iteration (Chronopoulos & Gear
// 1989), Jacobi preconditioner M = diag(N), on the pair of systems N z = b (the u and
// v components, one shared scalar recurrence), one kernel per iteration.  Kernel i
// forms u_{i+1} = M^-1 r_{i+1} at its neighbours from the previous kernel's r_i, w_i,
// s_{i-1} (r_{i+1} = r_i - alpha_i (w_i + beta_i s_{i-1})), so w_{i+1} = N u_{i+1}
// needs no second grid-wide step.  N has the 9-point stencil of the node grid: its
// values sit slot-major, a[k * n + node] for the neighbour (dx, dy) = (k % 3 - 1,
// k / 3 - 1), so every load of an iteration is independent of the others.
struct GfCg {
    int n, gx, gy;  // nodes (NP value pairs each), grid
    double *a9, *dinv;         // stencil values of N (0 outside the grid), 1 / diag(N)
    double2 *x;                // in: starting guess, out: solution (interleaved u, v)
    double2 *r[2], *w[2], *s[2], *p;
    double *G[2], *D[2], *RR[2], *BB;  // per-workgroup partials (gamma, delta, r.r, b.b)
    double *hist;              // [parity][2] = {alpha_i, gamma_i}, written by workgroup 0
    double *stat;              // {done, iterations, relative residual}
    double *rec;               // optional copy of stat, written alongside it
    int nb;
    double rtol2;
};

__device__ __forceinline__ bool gf_in(int ix, int iy, int k, int gx, int gy)
{
    const int jx = ix + k % 3 - 1, jy = iy + k / 3 - 1;
    return jx >= 0 && jx < gx && jy >= 0 && jy < gy;
}

// K0: stencil of N from the data mask (squared spring stiffnesses w), b = D d with the
// data d = x at the data nodes, r_0 = b - N x_0; p and s start at 0.
template <int NP>
__global__ __launch_bounds__(kGfThreads) void k_gf_cg0(GfCg c, const double *w, double lam2,
                                                       const int *bad)
{
    const int gx = c.gx, gy = c.gy;
    const double *wx = w, *wy = wx + (gx - 1), *axs = wy + (gy - 1), *ays = axs + (gx - 1);
    double bb = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < c.n; i += gridDim.x * blockDim.x) {
        const int ix = i % gx, iy = i / gx;
        double dsum = 0.0;
        double2 ax[NP];
#pragma unroll
        for (int k = 0; k < NP; ++k) ax[k] = make_double2(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            if (k == 4) continue;
            double v = 0.0;
            if (gf_in(ix, iy, k, gx, gy)) {
                const int dx = k % 3 - 1, dy = k / 3 - 1;
                double wt;
                if (dy == 0) wt = wx[min(ix, ix + dx)];
                else if (dx == 0) wt = wy[min(iy, iy + dy)];
                else wt = 1.0 / (axs[min(ix, ix + dx)] + ays[min(iy, iy + dy)]);
                v = -lam2 * wt;
                dsum += wt;
                const int j = i + dy * gx + dx;
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    const double2 xj = c.x[j * NP + q];
                    ax[q].x += v * xj.x;
                    ax[q].y += v * xj.y;
                }
            }
            c.a9[k * c.n + i] = v;
        }
        const bool gd = !bad || !bad[i];
        const double d = (gd ? 1.0 : 0.0) + lam2 * dsum;
        c.a9[4 * c.n + i] = d;
        c.dinv[i] = 1.0 / d;
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const double2 xi = c.x[i * NP + q];
            const double2 b = gd ? xi : make_double2(0.0, 0.0);
            c.r[0][i * NP + q] = make_double2(b.x - (ax[q].x + d * xi.x), b.y - (ax[q].y + d * xi.y));
            c.p[i * NP + q] = make_double2(0.0, 0.0);
            c.s[0][i * NP + q] = make_double2(0.0, 0.0);
            bb += b.x * b.x + b.y * b.y;
        }
    }
    gf_block3(bb, 0.0, 0.0, c.BB, nullptr, nullptr);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c.stat[0] = 0.0; c.stat[1] = 0.0; c.stat[2] = 1.0;
        if (c.rec) { c.rec[0] = 0.0; c.rec[1] = 0.0; c.rec[2] = 1.0; }
    }
}

// K1: w_0 = N u_0 with u_0 = M^-1 r_0; partials gamma_0, delta_0, r_0.r_0
template <int NP>
__global__ __launch_bounds__(kGfThreads) void k_gf_cg1(GfCg c)
{
    const int gx = c.gx, gy = c.gy;
    double g = 0.0, dl = 0.0, rr = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < c.n; i += gridDim.x * blockDim.x) {
        const int ix = i % gx, iy = i / gx;
        double2 wa[NP];
#pragma unroll
        for (int q = 0; q < NP; ++q) wa[q] = make_double2(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            if (!gf_in(ix, iy, k, gx, gy)) continue;
            const int j = i + (k / 3 - 1) * gx + (k % 3 - 1);
            const double av = c.a9[k * c.n + i] * c.dinv[j];
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const double2 rj = c.r[0][j * NP + q];
                wa[q].x += av * rj.x;
                wa[q].y += av * rj.y;
            }
        }
        const double dinv = c.dinv[i];
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const double2 ri = c.r[0][i * NP + q];
            const double uu = ri.x * dinv, uv = ri.y * dinv;
            c.w[0][i * NP + q] = wa[q];
            g += ri.x * uu + ri.y * uv;
            dl += wa[q].x * uu + wa[q].y * uv;
            rr += ri.x * ri.x + ri.y * ri.y;
        }
    }
    gf_block3(g, dl, rr, c.G[0], c.D[0], c.RR[0]);
}

// Operands of one node's iteration: the 9 stencil values scaled by M^-1 at the neighbour
// (0 outside the grid, where the neighbour index is clamped into it), the neighbours'
// r_i, w_i, s_{i-1}, and the node's p_{i-1}, x_i.  None depends on the iteration's
// scalars, so they are all loaded before the global reduction.
template <int NP>
struct GfOps {
    double av[9], dq;
    double2 r[9][NP], w[9][NP], s[9][NP], p[NP], x[NP];
};

template <int NP>
__device__ __forceinline__ void gf_load(const GfCg &c, int cur, int q, GfOps<NP> &o)
{
    const int gx = c.gx, gy = c.gy;
    const int ix = q % gx, iy = q / gx;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int jx = min(max(ix + k % 3 - 1, 0), gx - 1);
        const int jy = min(max(iy + k / 3 - 1, 0), gy - 1);
        const int j = jy * gx + jx;
        o.av[k] = c.a9[k * c.n + q] * c.dinv[j];
#pragma unroll
        for (int m = 0; m < NP; ++m) {
            o.r[k][m] = c.r[cur][j * NP + m];
            o.w[k][m] = c.w[cur][j * NP + m];
            o.s[k][m] = c.s[cur][j * NP + m];
        }
    }
    o.dq = c.dinv[q];
#pragma unroll
    for (int m = 0; m < NP; ++m) {
        o.p[m] = c.p[q * NP + m];
        o.x[m] = c.x[q * NP + m];
    }
}

__device__ __forceinline__ void gf_pin(double v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void gf_pin(double2 v) { gf_pin(v.x); gf_pin(v.y); }

// Holds every operand in registers at this point, so the loads are issued (and waited
// for) before the reduction instead of being sunk past it.
template <int NP>
__device__ __forceinline__ void gf_pin(const GfOps<NP> &o)
{
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        gf_pin(o.av[k]);
#pragma unroll
        for (int m = 0; m < NP; ++m) { gf_pin(o.r[k][m]); gf_pin(o.w[k][m]); gf_pin(o.s[k][m]); }
    }
    gf_pin(o.dq);
#pragma unroll
    for (int m = 0; m < NP; ++m) { gf_pin(o.p[m]); gf_pin(o.x[m]); }
}

// One node of iteration i: r_{i+1} = r_i - alpha (w_i + beta s_{i-1}) at the node and its
// neighbours, w_{i+1} = N M^-1 r_{i+1}, p_i, x_{i+1}; adds the node's gamma, delta, r.r.
template <int NP>
__device__ __forceinline__ void gf_node(const GfCg &c, int nxt, int q, const GfOps<NP> &o,
                                        double alpha, double beta, double &g, double &dl,
                                        double &rn)
{
    double2 wa[NP], rnew[NP], snew[NP];
#pragma unroll
    for (int m = 0; m < NP; ++m) {
        wa[m] = make_double2(0.0, 0.0);
        rnew[m] = wa[m];
        snew[m] = wa[m];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
#pragma unroll
        for (int m = 0; m < NP; ++m) {
            const double2 rj = o.r[k][m], wj = o.w[k][m], sj = o.s[k][m];
            const double2 sn = make_double2(wj.x + beta * sj.x, wj.y + beta * sj.y);
            const double2 rnj = make_double2(rj.x - alpha * sn.x, rj.y - alpha * sn.y);
            wa[m].x += o.av[k] * rnj.x;
            wa[m].y += o.av[k] * rnj.y;
            if (k == 4) { rnew[m] = rnj; snew[m] = sn; }
        }
    }
#pragma unroll
    for (int m = 0; m < NP; ++m) {
        const int e = q * NP + m;
        const double2 ri = o.r[4][m];
        // p_i = u_i + beta p_{i-1}, x_{i+1} = x_i + alpha p_i
        double2 pi = o.p[m];
        pi.x = ri.x * o.dq + beta * pi.x;
        pi.y = ri.y * o.dq + beta * pi.y;
        c.p[e] = pi;
        double2 xi = o.x[m];
        xi.x += alpha * pi.x;
        xi.y += alpha * pi.y;
        c.x[e] = xi;
        c.r[nxt][e] = rnew[m];
        c.s[nxt][e] = snew[m];
        c.w[nxt][e] = wa[m];
        const double uu = rnew[m].x * o.dq, uv = rnew[m].y * o.dq;
        g += rnew[m].x * uu + rnew[m].y * uv;
        dl += wa[m].x * uu + wa[m].y * uv;
        rn += rnew[m].x * rnew[m].x + rnew[m].y * rnew[m].y;
    }
}

// iteration i (reads parity i & 1, writes (i + 1) & 1).  With one value pair per node
// and one node per thread (the launch covers the grid), every load of the iteration --
// stop flag, partials, history and the node's operands -- is issued at once: one memory
// round trip per kernel.  Otherwise the node loads follow the reduction.
template <int NP, bool PRE, int NPL>
__global__ __launch_bounds__(kGfThreads) void k_gf_cgit(GfCg c, int i)
{
    const int cur = i & 1, nxt = cur ^ 1;
    const double stop = c.stat[0];
    // wave wv sums partial array wv (G, D, RR of this parity, then BB); the arrays are
    // consecutive runs of nb words from G[0] (G0 G1 D0 D1 RR0 RR1 BB), so the offset is
    // arithmetic, not a pointer select
    const int wv = threadIdx.x >> 6;
    const double *ps = c.G[0] + (size_t)(wv < 3 ? 2 * wv + cur : 6) * c.nb;
    // nb <= 64 NPL: NPL independent (clamped) loads per lane, summed in the order
    // k = lane, lane + 64, ...
    const int lane = threadIdx.x & 63;
    double pv[NPL];
#pragma unroll
    for (int u = 0; u < NPL; ++u) pv[u] = ps[min(lane + 64 * u, c.nb - 1)];
    const double h_alpha = c.hist[2 * nxt], h_gamma = c.hist[2 * nxt + 1];  // i - 1
    const int q0 = blockIdx.x * blockDim.x + threadIdx.x;
    GfOps<NP> o;
    if (PRE) gf_load(c, cur, min(q0, c.n - 1), o);
    double part = 0.0;
#pragma unroll
    for (int u = 0; u < NPL; ++u)
        if (lane + 64 * u < c.nb) part += pv[u];
    if (PRE) gf_pin(o);
    // sticky stop: once set, later partials are stale and the kernel does nothing (the
    // flag is uniform, so every thread leaves before the barrier)
    if (stop != 0.0) return;
    __shared__ double sh[4];
    part = wave_sum(part);
    if ((threadIdx.x & 63) == 0) sh[wv] = part;
    __syncthreads();
    const double gam = sh[0], del = sh[1], rr = sh[2], bb = sh[3];
    double alpha, beta;
    if (i == 0) {
        beta = 0.0;
        alpha = gam / del;
    } else {
        beta = gam / h_gamma;
        alpha = gam / (del - beta * gam / h_alpha);
    }
    // stop at the tolerance, on b == 0, or on a breakdown (non-finite or non-positive
    // step: the system is symmetric positive definite)
    const bool done = !(rr > c.rtol2 * bb) || !(alpha > 0.0) || !isfinite(alpha) ||
                      !isfinite(beta);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c.stat[0] = done ? 1.0 : 0.0;
        c.stat[1] = (double)i;
        c.stat[2] = bb > 0 ? sqrt(rr / bb) : 0.0;
        if (c.rec) { c.rec[0] = c.stat[0]; c.rec[1] = c.stat[1]; c.rec[2] = c.stat[2]; }
        c.hist[2 * cur] = alpha;
        c.hist[2 * cur + 1] = gam;
    }
    if (done) return;
    double g = 0.0, dl = 0.0, rn = 0.0;
    if (PRE) {
        if (q0 < c.n) gf_node(c, nxt, q0, o, alpha, beta, g, dl, rn);
    } else {
        for (int q = q0; q < c.n; q += gridDim.x * blockDim.x) {
            gf_load(c, cur, q, o);
            gf_node(c, nxt, q, o, alpha, beta, g, dl, rn);
        }
    }
    gf_block3(g, dl, rn, c.G[nxt], c.D[nxt], c.RR[nxt]);
}

// Scattered data (gridfit.m:460-511, 'triangle'): point p lies in cell (ix, iy) of the node
// grid (histc: the last node counts in the last cell) at (tx, ty) in [0, 1]^2; it
// interpolates from node (ix, iy) with weight 1 - max(tx, ty), (ix + 1, iy + 1) with
// min(tx, ty), and (ix + 1, iy) (tx > ty) or (ix, iy + 1) with |tx - ty|.  Adds the point's
// rows of A to A^T A (stencil slots), A^T d and the column sums of |A| (atomics: the
// order of the FP64 sums varies between runs in the last bits).
__device__ __forceinline__ int gf_cell(const double *nodes, int n, double v)
{
    int lo = 0, hi = n;  // count of nodes <= v
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (nodes[mid] <= v) lo = mid + 1; else hi = mid;
    }
    return min(max(lo, 1), n - 1) - 1;
}

template <int NP>
__global__ __launch_bounds__(256) void k_gf_scatter(int np, const double *px, const double *py,
                                                    const double2 *pd, const double *xn, int gx,
                                                    const double *yn, int gy, double *a9,
                                                    double2 *b, double *colsum)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int ix = gf_cell(xn, gx, px[p]), iy = gf_cell(yn, gy, py[p]);
    const double tx = fmin(fmax((px[p] - xn[ix]) / (xn[ix + 1] - xn[ix]), 0.0), 1.0);
    const double ty = fmin(fmax((py[p] - yn[iy]) / (yn[iy + 1] - yn[iy]), 0.0), 1.0);
    const double t1 = fmin(tx, ty), t2 = fmax(tx, ty);
    const int n = gx * gy;
    const int nd[3] = {iy * gx + ix, (iy + 1) * gx + ix + 1, tx > ty ? iy * gx + ix + 1 : (iy + 1) * gx + ix};
    const int ox[3] = {0, 1, tx > ty ? 1 : 0}, oy[3] = {0, 1, tx > ty ? 0 : 1};
    const double w[3] = {1.0 - t2, t1, t2 - t1};
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const double2 d = pd[(size_t)p * NP + q];
            atomicAdd(&b[(size_t)nd[a] * NP + q].x, w[a] * d.x);
            atomicAdd(&b[(size_t)nd[a] * NP + q].y, w[a] * d.y);
        }
        atomicAdd(&colsum[nd[a]], w[a]);
        for (int c = 0; c < 3; ++c) {
            const int k = (oy[c] - oy[a] + 1) * 3 + (ox[c] - ox[a] + 1);
            atomicAdd(&a9[(size_t)k * n + nd[a]], w[a] * w[c]);
        }
    }
}

// lam^2 = (|A|_1 / |Areg|_1)^2 (smoothness 1, gridfit.m:651-654): the largest column sum
// per block, then out = reg2 * max^2 from one block
__device__ __forceinline__ double gf_block_max(double m)
{
    __shared__ double red[4];
    for (int d = 32; d >= 1; d >>= 1) m = fmax(m, __shfl_xor(m, d, kWave));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

constexpr int kGfMaxParts = 256;

__global__ __launch_bounds__(256) void k_gf_colmax(const double *colsum, int n, double *part)
{
    double m = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        m = fmax(m, colsum[i]);
    m = gf_block_max(m);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ __launch_bounds__(256) void k_gf_lam2(const double *part, int nb, double reg2, double *out)
{
    const double m = gf_block_max(threadIdx.x < nb ? part[threadIdx.x] : 0.0);
    if (threadIdx.x == 0) out[0] = reg2 * m * m;
}

// K0 of the scattered fit: N = A^T A (already in a9) + lam^2 L, r_0 = A^T d (x_0 = 0)
template <int NP>
__global__ __launch_bounds__(kGfThreads) void k_gf_cg0s(GfCg c, const double *w, const double *lam2p,
                                                        const double2 *bvec)
{
    const int gx = c.gx, gy = c.gy;
    const double lam2 = *lam2p;
    const double *wx = w, *wy = wx + (gx - 1), *axs = wy + (gy - 1), *ays = axs + (gx - 1);
    double bb = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < c.n; i += gridDim.x * blockDim.x) {
        const int ix = i % gx, iy = i / gx;
        double dsum = 0.0;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            if (k == 4 || !gf_in(ix, iy, k, gx, gy)) continue;
            const int dx = k % 3 - 1, dy = k / 3 - 1;
            double wt;
            if (dy == 0) wt = wx[min(ix, ix + dx)];
            else if (dx == 0) wt = wy[min(iy, iy + dy)];
            else wt = 1.0 / (axs[min(ix, ix + dx)] + ays[min(iy, iy + dy)]);
            c.a9[k * c.n + i] -= lam2 * wt;
            dsum += wt;
        }
        const double d = c.a9[4 * c.n + i] + lam2 * dsum;
        c.a9[4 * c.n + i] = d;
        c.dinv[i] = 1.0 / d;
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const size_t e = (size_t)i * NP + q;
            const double2 b = bvec[e];
            c.x[e] = make_double2(0.0, 0.0);
            c.r[0][e] = b;
            c.p[e] = make_double2(0.0, 0.0);
            c.s[0][e] = make_double2(0.0, 0.0);
            bb += b.x * b.x + b.y * b.y;
        }
    }
    gf_block3(bb, 0.0, 0.0, c.BB, nullptr, nullptr);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c.stat[0] = 0.0; c.stat[1] = 0.0; c.stat[2] = 1.0;
        if (c.rec) { c.rec[0] = 0.0; c.rec[1] = 0.0; c.rec[2] = 1.0; }
    }
}

}  // namespace

hipError_t gridfit_setup(GridfitDev &g, const double *xnodes, int gx, const double *ynodes,
                         int gy, hipStream_t st)
{
    if (gx < 2 || gy < 2) return hipErrorInvalidValue;
    std::vector<double> dx(gx - 1), dy(gy - 1);
    double xs = 0.0, ys = 0.0;
    for (int i = 0; i + 1 < gx; ++i) { dx[i] = xnodes[i + 1] - xnodes[i]; xs += dx[i]; }
    for (int j = 0; j + 1 < gy; ++j) { dy[j] = ynodes[j + 1] - ynodes[j]; ys += dy[j]; }
    for (double v : dx) if (!(v > 0.0)) return hipErrorInvalidValue;
    for (double v : dy) if (!(v > 0.0)) return hipErrorInvalidValue;
    xs /= (gx - 1);  // autoscale: mean spacing (gridfit.m:361-365)
    ys /= (gy - 1);
    // springs' stiffnesses and the squares the normal equations use
    std::vector<double> w(2 * (gx - 1) + 2 * (gy - 1));
    double *wx = w.data(), *wy = wx + (gx - 1), *axs = wy + (gy - 1), *ays = axs + (gx - 1);
    for (int i = 0; i + 1 < gx; ++i) {
        const double s = 1.0 / (dx[i] / xs);
        wx[i] = s * s;
        axs[i] = (dx[i] / xs) * (dx[i] / xs);
    }
    for (int j = 0; j + 1 < gy; ++j) {
        const double s = 1.0 / (dy[j] / ys);
        wy[j] = s * s;
        ays[j] = (dy[j] / ys) * (dy[j] / ys);
    }
    // |Areg|_1: largest sum of the stiffnesses of the springs at one node
    double nr = 0.0;
    for (int iy = 0; iy < gy; ++iy)
        for (int ix = 0; ix < gx; ++ix) {
            double sum = 0.0;
            if (ix > 0) sum += std::sqrt(wx[ix - 1]);
            if (ix + 1 < gx) sum += std::sqrt(wx[ix]);
            if (iy > 0) sum += std::sqrt(wy[iy - 1]);
            if (iy + 1 < gy) sum += std::sqrt(wy[iy]);
            for (int cx = ix - 1; cx <= ix; ++cx)
                for (int cy = iy - 1; cy <= iy; ++cy)
                    if (cx >= 0 && cx + 1 < gx && cy >= 0 && cy + 1 < gy)
                        sum += 1.0 / std::sqrt(axs[cx] + ays[cy]);
            nr = std::max(nr, sum);
        }
    g.gx = gx;
    g.gy = gy;
    g.lam2 = 1.0 / (nr * nr);  // smoothness 1, |A|_1 = 1
    hipError_t e;
    if ((e = g.w.ensure(sizeof(double) * w.size())) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(g.w.p, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice,
                            st)) != hipSuccess)
        return e;
    const size_t n = (size_t)gx * gy;
    if ((e = g.a9.ensure(9 * sizeof(double) * n)) != hipSuccess) return e;
    if ((e = g.dinv.ensure(sizeof(double) * n)) != hipSuccess) return e;
    // the host vector above is copied before it goes out of scope
    return hipStreamSynchronize(st);
}

// scat (nullable): the scattered fit's {lam^2 (device), A^T d (device, NP pairs per node)}
struct GfScat {
    const double *lam2;
    const double2 *b;
};

template <int NP>
static hipError_t solve_np(GridfitDev &g, const int *bad, double *U, double rtol, int maxit,
                           DBuf *work, double *stats_dev, hipStream_t st, int poll,
                           const GfScat *scat = nullptr)
{
    const int n = g.gx * g.gy;
    // one node per thread when the grid allows it (the operand prefetch needs it); at
    // most kGfMaxBlocks partials per array
    const int nfull = (n + kGfThreads - 1) / kGfThreads;
    const bool pre = NP == 1 && nfull <= kGfMaxBlocks;
    const int nb = std::min(kGfMaxBlocks, nfull);
    hipError_t e;
    const size_t vb = sizeof(double2) * (size_t)n * NP;
    if ((e = work[0].ensure(7 * vb)) != hipSuccess) return e;
    if ((e = work[1].ensure(sizeof(double) * (7 * (size_t)nb + 8))) != hipSuccess) return e;
    GfCg c{};
    c.n = n;
    c.gx = g.gx;
    c.gy = g.gy;
    c.a9 = g.a9.as<double>();
    c.dinv = g.dinv.as<double>();
    c.x = reinterpret_cast<double2 *>(U);
    double2 *v = work[0].as<double2>();
    const size_t m = (size_t)n * NP;
    c.r[0] = v; c.r[1] = v + m; c.w[0] = v + 2 * m; c.w[1] = v + 3 * m;
    c.s[0] = v + 4 * m; c.s[1] = v + 5 * m; c.p = v + 6 * m;
    double *ps = work[1].as<double>();
    c.G[0] = ps; c.G[1] = ps + nb; c.D[0] = ps + 2 * nb; c.D[1] = ps + 3 * nb;
    c.RR[0] = ps + 4 * nb; c.RR[1] = ps + 5 * nb; c.BB = ps + 6 * nb;
    c.hist = ps + 7 * nb;
    c.stat = ps + 7 * nb + 4;
    c.rec = stats_dev;
    c.nb = nb;
    c.rtol2 = rtol * rtol;
    if (scat)
        hipLaunchKernelGGL(k_gf_cg0s<NP>, dim3(nb), dim3(kGfThreads), 0, st, c, (const double *)g.w.p,
                           scat->lam2, scat->b);
    else
        hipLaunchKernelGGL(k_gf_cg0<NP>, dim3(nb), dim3(kGfThreads), 0, st, c,
                           (const double *)g.w.p, g.lam2, bad);
    hipLaunchKernelGGL(k_gf_cg1<NP>, dim3(nb), dim3(kGfThreads), 0, st, c);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // poll <= 0: maxit iteration kernels and no host wait (the caller reads stat)
    const int chunk = poll > 0 ? poll : maxit;
    for (int it = 0; it < maxit; it += chunk) {
        for (int q = 0; q < chunk && it + q < maxit; ++q) {
            if (pre && nb <= 64)
                hipLaunchKernelGGL((k_gf_cgit<NP, NP == 1, 1>), dim3(nb), dim3(kGfThreads), 0, st,
                                   c, it + q);
            else if (pre)
                hipLaunchKernelGGL((k_gf_cgit<NP, NP == 1, kGfMaxBlocks / 64>), dim3(nb),
                                   dim3(kGfThreads), 0, st, c, it + q);
            else
                hipLaunchKernelGGL((k_gf_cgit<NP, false, kGfMaxBlocks / 64>), dim3(nb),
                                   dim3(kGfThreads), 0, st, c, it + q);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (poll <= 0) break;
        double done = 0.0;
        if ((e = hipMemcpyAsync(&done, c.stat, sizeof(double), hipMemcpyDeviceToHost, st)) !=
            hipSuccess)
            return e;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
        if (done != 0.0) break;
    }
    // stats_dev ({done, iterations, relative residual} of the last iteration kernel that
    // ran) is written by the kernels themselves
    return hipSuccess;
}

hipError_t gridfit_solve(GridfitDev &g, const int *bad, double *U, int np, double rtol, int maxit,
                         DBuf *work, double *stats_dev, hipStream_t st, int poll)
{
    if (np == 1) return solve_np<1>(g, bad, U, rtol, maxit, work, stats_dev, st, poll);
    if (np == 2) return solve_np<2>(g, bad, U, rtol, maxit, work, stats_dev, st, poll);
    return hipErrorInvalidValue;
}

hipError_t gridfit_scattered(GridfitDev &g, const double *xnodes_dev, const double *ynodes_dev,
                             int np, const double *px, const double *py, const double2 *pd, int npair,
                             double *U, double rtol, int maxit, DBuf *work, double *stats_dev,
                             hipStream_t st)
{
    const int n = g.gx * g.gy;
    hipError_t e;
    if (np <= 0 || npair < 1 || npair > 2) return hipErrorInvalidValue;
    const size_t nb2 = (size_t)n * npair;
    if ((e = work[2].ensure(sizeof(double2) * nb2 + sizeof(double) * ((size_t)n + 1 + kGfMaxParts))) != hipSuccess)
        return e;
    double2 *b = work[2].as<double2>();
    double *colsum = reinterpret_cast<double *>(b + nb2), *lam2 = colsum + n, *parts = lam2 + 1;
    const int nbm = std::min(kGfMaxParts, (n + 255) / 256);
    if ((e = hipMemsetAsync(g.a9.p, 0, 9 * sizeof(double) * (size_t)n, st)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(work[2].p, 0, sizeof(double2) * nb2 + sizeof(double) * n, st)) != hipSuccess)
        return e;
    if (npair == 1)
        hipLaunchKernelGGL(k_gf_scatter<1>, dim3((np + 255) / 256), dim3(256), 0, st, np, px, py, pd, xnodes_dev,
                           g.gx, ynodes_dev, g.gy, g.a9.as<double>(), b, colsum);
    else
        hipLaunchKernelGGL(k_gf_scatter<2>, dim3((np + 255) / 256), dim3(256), 0, st, np, px, py, pd, xnodes_dev,
                           g.gx, ynodes_dev, g.gy, g.a9.as<double>(), b, colsum);
    hipLaunchKernelGGL(k_gf_colmax, dim3(nbm), dim3(256), 0, st, (const double *)colsum, n, parts);
    hipLaunchKernelGGL(k_gf_lam2, dim3(1), dim3(256), 0, st, (const double *)parts, nbm, g.lam2, lam2);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const GfScat sc{lam2, b};
    if (npair == 2) return solve_np<2>(g, nullptr, U, rtol, maxit, work, stats_dev, st, 32, &sc);
    return solve_np<1>(g, nullptr, U, rtol, maxit, work, stats_dev, st, 32, &sc);
}

namespace {

// nanconv(Z, fspecial('gaussian', 3, sigma), 'edge') on the node grid (no NaNs): the 3x3
// Gaussian is separable, w = [e, 1, e] / (1 + 2 e) with e = exp(-1 / (2 sigma^2)) per
// axis, and the edge correction divides by the weight inside the grid, wx(ix) wy(iy).
__global__ __launch_bounds__(256) void k_blur3(const double *in, double *out, int gx, int gy,
                                               int nc, double e)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= gx * gy * nc) return;
    const int c = t % nc, i = t / nc, ix = i % gx, iy = i / gx;
    const double w1 = 1.0 / (1.0 + 2.0 * e), w0 = e * w1;
    const double wl[3] = {w0, w1, w0};
    double sx = 0.0, sy = 0.0, acc = 0.0;
    for (int d = -1; d <= 1; ++d) {
        if (ix + d >= 0 && ix + d < gx) sx += wl[d + 1];
        if (iy + d >= 0 && iy + d < gy) sy += wl[d + 1];
    }
    for (int dy = -1; dy <= 1; ++dy) {
        if (iy + dy < 0 || iy + dy >= gy) continue;
        double row = 0.0;
        for (int dx = -1; dx <= 1; ++dx) {
            if (ix + dx < 0 || ix + dx >= gx) continue;
            row += wl[dx + 1] * in[(size_t)(i + dy * gx + dx) * nc + c];
        }
        acc += wl[dy + 1] * row;
    }
    out[t] = acc / (sx * sy);
}

}  // namespace

hipError_t launch_blur3(const double *in, double *out, int gx, int gy, int nc, double sigma,
                        hipStream_t st)
{
    const int n = gx * gy * nc;
    hipLaunchKernelGGL(k_blur3, dim3((n + 255) / 256), dim3(256), 0, st, in, out, gx, gy, nc,
                       std::exp(-1.0 / (2.0 * sigma * sigma)));
    return hipGetLastError();
}

}  // namespace aldic
