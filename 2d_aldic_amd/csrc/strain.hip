// Strain post-processing on a uniform node grid (ComputeStrain.m:11-137): the
// deformation gradient from the displacements by finite differences or plane fits,
// the strain type, and the trimmed, world-oriented output.  Node n = i + M j (i along
// x, fastest), U interleaved (u, v) per node, F per node (F11, F21, F12, F22) =
// (du/dx, dv/dx, du/dy, dv/dy).  One thread per node; every kernel is a single pass over
// its inputs (HBM-bound, 16 + 32 bytes per node).
#include "aldic_internal.h"

namespace aldic {

namespace {

// funDerivativeOp.m:28-67: central differences (u(i+1) - u(i-1)) / 2h inside, one-sided
// (u(i+1) - u(i)) / h or (u(i) - u(i-1)) / h on the first and last node of a line (the
// operator's border rows take the node itself as the missing neighbour, with weights
// -+2 / 2h); a line of one node has no entries (derivative 0).
__global__ __launch_bounds__(256) void k_fd(int M, int N, double h, const double *__restrict__ U,
                                            double *__restrict__ F)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= M * N) return;
    const int i = n % M, j = n / M;
    double dux = 0, dvx = 0, duy = 0, dvy = 0;
    if (M > 1) {
        const int il = i > 0 ? i - 1 : i, ir = i < M - 1 ? i + 1 : i;
        const double c = (i == 0 || i == M - 1) ? 2.0 : 1.0;
        const int a = il + M * j, b = ir + M * j;
        dux = c * (U[2 * b] - U[2 * a]) / (2.0 * h);
        dvx = c * (U[2 * b + 1] - U[2 * a + 1]) / (2.0 * h);
    }
    if (N > 1) {
        const int jl = j > 0 ? j - 1 : j, jr = j < N - 1 ? j + 1 : j;
        const double c = (j == 0 || j == N - 1) ? 2.0 : 1.0;
        const int a = i + M * jl, b = i + M * jr;
        duy = c * (U[2 * b] - U[2 * a]) / (2.0 * h);
        dvy = c * (U[2 * b + 1] - U[2 * a + 1]) / (2.0 * h);
    }
    F[4 * n] = dux;
    F[4 * n + 1] = dvx;
    F[4 * n + 2] = duy;
    F[4 * n + 3] = dvy;
}

// PlaneFit2.m: least squares of [1, ti h, tj h] over the (2R+1)^2 nodes around each node
// at least R from the border.  On this symmetric stencil the normal matrix is diagonal,
// so the slopes are sum(ti u) / (h (2R+1) S2) and sum(tj u) / (h (2R+1) S2) with
// S2 = sum_t t^2 = R (R+1) (2R+1) / 3; the other nodes keep their finite differences.
__global__ __launch_bounds__(256) void k_planefit(int M, int N, int R, double h,
                                                  const double *__restrict__ U,
                                                  double *__restrict__ F)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= M * N) return;
    const int i = n % M, j = n / M;
    if (i < R || i >= M - R || j < R || j >= N - R) return;
    double sux = 0, svx = 0, suy = 0, svy = 0;
    for (int tj = -R; tj <= R; ++tj)
        for (int ti = -R; ti <= R; ++ti) {
            const int m = (i + ti) + M * (j + tj);
            const double u = U[2 * m], v = U[2 * m + 1];
            sux += ti * u;
            svx += ti * v;
            suy += tj * u;
            svy += tj * v;
        }
    const double den = h * (2.0 * R + 1.0) * (R * (R + 1.0) * (2.0 * R + 1.0) / 3.0);
    F[4 * n] = sux / den;
    F[4 * n + 1] = svx / den;
    F[4 * n + 2] = suy / den;
    F[4 * n + 3] = svy / den;
}

// ComputeStrain.m:113-137: strain type per node (0 infinitesimal: unchanged; 1 Eulerian;
// 2 Green-Lagrangian as the reference writes it, E = (C - I) / 2 with the gradient taken
// as I - grad u), in place.
__global__ __launch_bounds__(256) void k_strain_type(int nn, int type, double *__restrict__ F)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn || type == 0) return;
    const double dudx = F[4 * n], dvdx = F[4 * n + 1], dudy = F[4 * n + 2], dvdy = F[4 * n + 3];
    double e11, e21, e12, e22;
    if (type == 1) {
        e11 = 1.0 / (1.0 - dudx) - 1.0;
        e22 = 1.0 / (1.0 - dvdy) - 1.0;
        e12 = dudy / (1.0 - dvdy);
        e21 = dvdx / (1.0 - dudx);
    } else {
        e11 = 0.5 * (dudx * 2.0 - dudx * dudx - dvdx * dvdx);
        e22 = 0.5 * (dvdy * 2.0 - dudy * dudy - dvdy * dvdy);
        e12 = 0.5 * (dudy + dvdx - dudx * dudy - dvdx * dvdy);
        e21 = 0.5 * (dvdx + dudy - dudy * dudx - dvdy * dvdx);
    }
    F[4 * n] = e11;
    F[4 * n + 1] = e21;
    F[4 * n + 2] = e12;
    F[4 * n + 3] = e22;
}

// ComputeStrain.m:18-24,138-139: the nodes at least r from the border (x fastest) with
// F21 and F12 negated (y pointing up in the world frame).
__global__ __launch_bounds__(256) void k_trim_world(int M, int N, int r,
                                                    const double *__restrict__ F,
                                                    double *__restrict__ W)
{
    const int Mi = M - 2 * r, Ni = N - 2 * r;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= Mi * Ni) return;
    const int n = (r + k % Mi) + M * (r + k / Mi);
    W[4 * k] = F[4 * n];
    W[4 * k + 1] = -F[4 * n + 1];
    W[4 * k + 2] = -F[4 * n + 2];
    W[4 * k + 3] = F[4 * n + 3];
}

int nblk(long n) { return (int)((n + 255) / 256); }

}  // namespace

hipError_t launch_strain_grad(int M, int N, double h, int method, int R, const double *U,
                              double *F, hipStream_t st)
{
    const int nn = M * N;
    if (method == 1 || method == 2)
        hipLaunchKernelGGL(k_fd, dim3(nblk(nn)), dim3(256), 0, st, M, N, h, U, F);
    if (method == 2 && R > 0)
        hipLaunchKernelGGL(k_planefit, dim3(nblk(nn)), dim3(256), 0, st, M, N, R, h, U, F);
    return hipGetLastError();
}

hipError_t launch_strain_finish(int M, int N, int type, int r, double *F, double *W,
                                hipStream_t st)
{
    const int nn = M * N, nw = (M - 2 * r) * (N - 2 * r);
    hipLaunchKernelGGL(k_strain_type, dim3(nblk(nn)), dim3(256), 0, st, nn, type, F);
    if (nw > 0) hipLaunchKernelGGL(k_trim_world, dim3(nblk(nw)), dim3(256), 0, st, M, N, r, F, W);
    return hipGetLastError();
}

}  // namespace aldic
