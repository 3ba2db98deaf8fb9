// Quadtree global step (Subpb2Quadtree.m:1-346 as main_ALDIC_Quadtree.m:181-212 calls
// it: alpha = 0, 2x2 Gauss points, duals multiplied by 0).
//
// Elements have four corners and up to four midside nodes (5..8 on the right, top,
// left, bottom edges; absent = -1), shape functions of Subpb2Quadtree.m:126-133:
//   N5..N8 = d/2 (1 +- ksi or eta)(1 - |other|),  corners = bilinear - (neighbouring
//   midside functions)/2.
// The system has the Q4 step's structure (A = beta K + mu M in both displacement blocks,
// rhs beta G_i . F_i + mu (M U)_i with each node's own F), so it is assembled into the
// same scalar CSR over nodes, node-gather and deterministic, and solved by the same
// Jacobi-PCG.  Fixed nodes (Dirichlet: U kept; nodes outside every element, and the
// node Subpb2Quadtree.m:165-168 drops when no element lacks a midside node: 0) become
// identity rows with zero right-hand side, their columns move into a coupling array
// that the free rows' right-hand side subtracts (times U at Dirichlet nodes), so the
// system stays symmetric positive definite and the fixed values are written exactly.
#include "aldic_internal.h"
#include "common.h"

namespace aldic {

namespace {

__device__ __forceinline__ double sgn(double v) { return v > 0.0 ? 1.0 : (v < 0.0 ? -1.0 : 0.0); }

// N, dN/dksi, dN/deta of the 8 functions; d[k] = 1 if midside node 5 + k exists
__device__ __forceinline__ void q8_shape(double ksi, double eta, const double d[4], double N[8],
                                         double dk[8], double de[8])
{
    const double ae = fabs(eta), ak = fabs(ksi), se = sgn(-eta), sk = sgn(-ksi);
    N[4] = d[0] * 0.5 * (1 + ksi) * (1 - ae);
    N[5] = d[1] * 0.5 * (1 + eta) * (1 - ak);
    N[6] = d[2] * 0.5 * (1 - ksi) * (1 - ae);
    N[7] = d[3] * 0.5 * (1 - eta) * (1 - ak);
    N[0] = (1 - ksi) * (1 - eta) * 0.25 - 0.5 * (N[6] + N[7]);
    N[1] = (1 + ksi) * (1 - eta) * 0.25 - 0.5 * (N[7] + N[4]);
    N[2] = (1 + ksi) * (1 + eta) * 0.25 - 0.5 * (N[4] + N[5]);
    N[3] = (1 - ksi) * (1 + eta) * 0.25 - 0.5 * (N[5] + N[6]);
    dk[4] = d[0] * 0.5 * (1 - ae);          de[4] = d[0] * 0.5 * (1 + ksi) * se;
    dk[5] = d[1] * 0.5 * (1 + eta) * sk;    de[5] = d[1] * 0.5 * (1 - ak);
    dk[6] = -d[2] * 0.5 * (1 - ae);         de[6] = d[2] * 0.5 * (1 - ksi) * se;
    dk[7] = d[3] * 0.5 * (1 - eta) * sk;    de[7] = -d[3] * 0.5 * (1 - ak);
    dk[0] = -0.25 * (1 - eta) - 0.5 * (dk[6] + dk[7]);
    de[0] = -0.25 * (1 - ksi) - 0.5 * (de[6] + de[7]);
    dk[1] = 0.25 * (1 - eta) - 0.5 * (dk[7] + dk[4]);
    de[1] = -0.25 * (1 + ksi) - 0.5 * (de[7] + de[4]);
    dk[2] = 0.25 * (1 + eta) - 0.5 * (dk[4] + dk[5]);
    de[2] = 0.25 * (1 + ksi) - 0.5 * (de[4] + de[5]);
    dk[3] = -0.25 * (1 + eta) - 0.5 * (dk[5] + dk[6]);
    de[3] = 0.25 * (1 - ksi) - 0.5 * (de[5] + de[6]);
}

// one thread per node: its row of K and M from the elements around it (fixed order)
__global__ __launch_bounds__(256) void k_assemble8(int nn, const double2 *coords, const int *elems,
                                                   const int *n2e_ptr, const N2E8 *n2e, double beta,
                                                   double mu, const int *rowptr, double *aval,
                                                   double *mval, double2 *gvec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) { aval[k] = 0.0; mval[k] = 0.0; }
    const double g = 0.57735026918962576;  // 1 / sqrt(3)
    double gxs = 0.0, gys = 0.0;
    for (int q = n2e_ptr[i]; q < n2e_ptr[i + 1]; ++q) {
        const N2E8 r = n2e[q];
        const int *nd = elems + 8 * r.e;
        double px[8], py[8], d[4];
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool on = nd[b] >= 0;
            px[b] = on ? coords[nd[b]].x : 0.0;
            py[b] = on ? coords[nd[b]].y : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) d[k] = nd[4 + k] >= 0 ? 1.0 : 0.0;
        double Kr[8], Mr[8];
#pragma unroll
        for (int b = 0; b < 8; ++b) { Kr[b] = 0.0; Mr[b] = 0.0; }
        const int a = r.a;
        for (int gi = 0; gi < 4; ++gi) {
            const double ksi = (gi & 1) ? g : -g, eta = (gi & 2) ? g : -g;
            double N[8], dk[8], de[8];
            q8_shape(ksi, eta, d, N, dk, de);
            double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                J11 += dk[b] * px[b]; J12 += dk[b] * py[b];
                J21 += de[b] * px[b]; J22 += de[b] * py[b];
            }
            const double det = J11 * J22 - J12 * J21;
            const double i11 = J22 / det, i12 = -J12 / det, i21 = -J21 / det, i22 = J11 / det;
            const double ax = i11 * dk[a] + i12 * de[a], ay = i21 * dk[a] + i22 * de[a];
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const double bx = i11 * dk[b] + i12 * de[b], by = i21 * dk[b] + i22 * de[b];
                Kr[b] += det * (ax * bx + ay * by);
                Mr[b] += det * N[a] * N[b];
            }
            gxs += det * ax;
            gys += det * ay;
        }
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            if (nd[b] < 0) continue;
            aval[r.slot[b]] += beta * Kr[b] + mu * Mr[b];
            mval[r.slot[b]] += Mr[b];
        }
    }
    gvec[i] = make_double2(gxs, gys);
}

// fixed nodes (code != 0) -> identity rows; columns of fixed nodes -> cval
__global__ __launch_bounds__(256) void k_bc_rows(int nn, const int *rowptr, const int *col,
                                                 const int *code, double *aval, double *cval,
                                                 double *diag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    double dg = 0.0;
    const bool fixed = code[i] != 0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
        const int c = col[k];
        double v = aval[k], cv = 0.0;
        if (fixed) {
            v = c == i ? 1.0 : 0.0;
        } else if (code[c] != 0) {
            cv = v;
            v = 0.0;
        }
        aval[k] = v;
        cval[k] = cv;
        if (c == i) dg = v;
    }
    diag[i] = fixed ? 1.0 : dg;
}

// b = 0 at fixed nodes (their identity rows then keep x = 0 through the whole PCG, the
// values go in afterwards); free rows lose the Dirichlet columns' share
__global__ __launch_bounds__(256) void k_rhs_bc(int nn, const int *rowptr, const int *col,
                                                const int *code, const double *cval,
                                                const double2 *U, double2 *b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    if (code[i] != 0) { b[i] = make_double2(0.0, 0.0); return; }
    double2 bi = b[i];
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
        const int c = col[k];
        if (code[c] != 1) continue;
        const double2 u = U[c];
        bi.x -= cval[k] * u.x;
        bi.y -= cval[k] * u.y;
    }
    b[i] = bi;
}

__global__ __launch_bounds__(256) void k_fix(int nn, const int *code, const double2 *U, double2 *x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int c = code[i];
    if (c == 1) x[i] = U[i];
    else if (c != 0) x[i] = make_double2(0.0, 0.0);
}

}  // namespace

hipError_t quad_assemble(QuadDev &q, double beta, double mu, hipStream_t st)
{
    MeshDev &m = q.m;
    const int nn = m.nn;
    hipLaunchKernelGGL(k_assemble8, dim3((nn + 255) / 256), dim3(256), 0, st, nn,
                       m.coords.as<double2>(), q.elems.as<int>(), m.n2e_ptr.as<int>(),
                       q.n2e.as<N2E8>(), beta, mu, m.rowptr.as<int>(), m.aval.as<double>(),
                       m.mval.as<double>(), m.gvec.as<double2>());
    hipLaunchKernelGGL(k_bc_rows, dim3((nn + 255) / 256), dim3(256), 0, st, nn, m.rowptr.as<int>(),
                       m.col.as<int>(), q.code.as<int>(), m.aval.as<double>(),
                       q.cval.as<double>(), m.diag.as<double>());
    m.beta = beta;
    m.mu = mu;
    m.alpha = 0.0;
    return hipGetLastError();
}

hipError_t quad_rhs(QuadDev &q, double beta, double mu, const double *U, const double *F,
                    double *b, hipStream_t st)
{
    hipError_t e;
    if ((e = fem_rhs(q.m, beta, mu, U, F, nullptr, nullptr, b, st)) != hipSuccess) return e;
    const int nn = q.m.nn;
    hipLaunchKernelGGL(k_rhs_bc, dim3((nn + 255) / 256), dim3(256), 0, st, nn,
                       q.m.rowptr.as<int>(), q.m.col.as<int>(), q.code.as<int>(),
                       q.cval.as<double>(), reinterpret_cast<const double2 *>(U),
                       reinterpret_cast<double2 *>(b));
    return hipGetLastError();
}

hipError_t quad_finish(QuadDev &q, const double *U, double *x, hipStream_t st)
{
    const int nn = q.m.nn;
    hipLaunchKernelGGL(k_fix, dim3((nn + 255) / 256), dim3(256), 0, st, nn, q.code.as<int>(),
                       reinterpret_cast<const double2 *>(U), reinterpret_cast<double2 *>(x));
    return hipGetLastError();
}

}  // namespace aldic
