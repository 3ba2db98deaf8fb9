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
// The FP64 CSR matrix of N (9-point pattern) is assembled on the device from the data
// mask and solved by the Jacobi-PCG of the FEM step (two kernels per iteration, no host
// round trip between convergence polls).
#include <algorithm>
#include <cmath>
#include <vector>

#include "aldic_internal.h"
#include "common.h"

namespace aldic {

namespace {

// CSR values of N = D + lam2 L for general weights; pattern: node itself then its
// existing neighbours in the order the host built them (rowptr, col)
__global__ __launch_bounds__(256) void k_gridfit_csr(int gx, int gy, const double *w,
                                                     double lam2, const int *bad,
                                                     const int *rowptr, const int *col,
                                                     double *aval, double *diag, int ncomp,
                                                     const double *U, double2 *b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= gx * gy) return;
    const double *wx = w, *wy = wx + (gx - 1), *axs = wy + (gy - 1), *ays = axs + (gx - 1);
    const int ix = i % gx, iy = i / gx;
    const bool gd = !bad[i];
    double dsum = 0.0;
    for (int k = rowptr[i] + 1; k < rowptr[i + 1]; ++k) {
        const int j = col[k], jx = j % gx, jy = j / gx;
        double wt;
        if (jy == iy) wt = wx[min(ix, jx)];
        else if (jx == ix) wt = wy[min(iy, jy)];
        else wt = 1.0 / (axs[min(ix, jx)] + ays[min(iy, jy)]);
        aval[k] = -lam2 * wt;
        dsum += wt;
    }
    const double d = (gd ? 1.0 : 0.0) + lam2 * dsum;
    aval[rowptr[i]] = d;
    diag[i] = d;
    const double u = gd ? U[(size_t)i * ncomp] : 0.0;
    const double v = gd && ncomp > 1 ? U[(size_t)i * ncomp + 1] : 0.0;
    b[i] = make_double2(u, v);
}

__global__ __launch_bounds__(256) void k_gridfit_csr_out(int n, int ncomp, const double2 *x,
                                                         double *U)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    U[(size_t)i * ncomp] = x[i].x;
    if (ncomp > 1) U[(size_t)i * ncomp + 1] = x[i].y;
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
    {
        // 9-point pattern: the node first, then its neighbours
        std::vector<int> rp(gx * gy + 1), cl;
        cl.reserve(9 * (size_t)gx * gy);
        for (int iy = 0; iy < gy; ++iy)
            for (int ix = 0; ix < gx; ++ix) {
                const int i = iy * gx + ix;
                rp[i] = (int)cl.size();
                cl.push_back(i);
                for (int dyy = -1; dyy <= 1; ++dyy)
                    for (int dxx = -1; dxx <= 1; ++dxx) {
                        const int jx = ix + dxx, jy = iy + dyy;
                        if ((dxx || dyy) && jx >= 0 && jx < gx && jy >= 0 && jy < gy)
                            cl.push_back(jy * gx + jx);
                    }
            }
        rp[gx * gy] = (int)cl.size();
        MeshDev &m = g.csr;
        m.nn = gx * gy;
        m.nnz = (int)cl.size();
        if ((e = m.rowptr.ensure(sizeof(int) * rp.size())) != hipSuccess) return e;
        if ((e = m.col.ensure(sizeof(int) * cl.size())) != hipSuccess) return e;
        if ((e = m.aval.ensure(sizeof(double) * cl.size())) != hipSuccess) return e;
        if ((e = m.diag.ensure(sizeof(double) * (size_t)m.nn)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(m.rowptr.p, rp.data(), sizeof(int) * rp.size(),
                                hipMemcpyHostToDevice, st)) != hipSuccess)
            return e;
        if ((e = hipMemcpyAsync(m.col.p, cl.data(), sizeof(int) * cl.size(),
                                hipMemcpyHostToDevice, st)) != hipSuccess)
            return e;
    }
    // the host vectors above are copied before they go out of scope
    return hipStreamSynchronize(st);
}

hipError_t gridfit_solve(GridfitDev &g, const int *bad, double *U, int ncomp, double rtol,
                         int maxit, DBuf *work, double *stats_dev, hipStream_t st, int poll)
{
    if (ncomp < 1 || ncomp > 2) return hipErrorInvalidValue;
    const int n = g.gx * g.gy, nb = (n + 255) / 256;
    hipError_t e;
    if ((e = work[7].ensure(2 * sizeof(double2) * (size_t)n)) != hipSuccess) return e;
    double2 *b = work[7].as<double2>(), *x = b + n;
    MeshDev &m = g.csr;
    hipLaunchKernelGGL(k_gridfit_csr, dim3(nb), dim3(256), 0, st, g.gx, g.gy,
                       (const double *)g.w.p, g.lam2, bad, m.rowptr.as<int>(), m.col.as<int>(),
                       m.aval.as<double>(), m.diag.as<double>(), ncomp, (const double *)U, b);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    double hs[2] = {0, 0};
    if ((e = fem_pcg(m, (const double *)b, (double *)x, rtol, maxit, work, hs, st, poll)) !=
        hipSuccess)
        return e;
    hipLaunchKernelGGL(k_gridfit_csr_out, dim3(nb), dim3(256), 0, st, n, ncomp,
                       (const double2 *)x, U);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (stats_dev) {
        const double s4[4] = {hs[0], hs[1], hs[0], hs[1]};
        e = hipMemcpyAsync(stats_dev, s4, sizeof s4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    return e;
}

}  // namespace aldic
