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
#include <utility>

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

// ---- nodal strain (funGlobalNodalStrainQuadtree.m:31-255) ---------------------------
// The midside nodes are dropped (:33), so each element is a bilinear Q4 over its corners.
// Gauss point (i, j) of element e sits at x_i = g_i (x3 - x1)/2 + (x3 + x1)/2,
// y_j likewise (corners 1 and 3, :116-121), stored at 4 e + 2 i + j (:201-205); its
// (ksi, eta) come from the element's map ksi = l1 x y + l2 x + l3 y + l4 (eta with m),
// l = [x y, x, y, 1]^-1 [-1 1 1 -1]' over the corners (:102-107).  The strain there is
// [u_x, u_y, v_x, v_y] (:184-197).

// 4x4 solve with partial pivoting, two right-hand sides (overwritten with the solutions)
__device__ __forceinline__ void solve4(double A[4][4], double b0[4], double b1[4])
{
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int p = k;
        double best = fabs(A[k][k]);
#pragma unroll
        for (int r = k + 1; r < 4; ++r)
            if (fabs(A[r][k]) > best) { best = fabs(A[r][k]); p = r; }
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (r == p && p != k) {
#pragma unroll
                for (int c = 0; c < 4; ++c) { const double t = A[k][c]; A[k][c] = A[r][c]; A[r][c] = t; }
                double t = b0[k]; b0[k] = b0[r]; b0[r] = t;
                t = b1[k]; b1[k] = b1[r]; b1[r] = t;
            }
#pragma unroll
        for (int r = k + 1; r < 4; ++r) {
            const double f = A[r][k] / A[k][k];
#pragma unroll
            for (int c = k; c < 4; ++c) A[r][c] -= f * A[k][c];
            b0[r] -= f * b0[k];
            b1[r] -= f * b1[k];
        }
    }
#pragma unroll
    for (int k = 3; k >= 0; --k) {
#pragma unroll
        for (int c = k + 1; c < 4; ++c) { b0[k] -= A[k][c] * b0[c]; b1[k] -= A[k][c] * b1[c]; }
        b0[k] /= A[k][k];
        b1[k] /= A[k][k];
    }
}

__global__ __launch_bounds__(256) void k_gp_strain(int E, const double2 *coords, const int *elems,
                                                   const double2 *U, double4 *Fgp, double2 *gpxy)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int *nd = elems + 8 * e;
    double2 p[4];
    double2 u[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) { p[a] = coords[nd[a]]; u[a] = U[nd[a]]; }
    double A[4][4], l[4] = {-1, 1, 1, -1}, m[4] = {-1, -1, 1, 1};
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        A[a][0] = p[a].x * p[a].y; A[a][1] = p[a].x; A[a][2] = p[a].y; A[a][3] = 1.0;
    }
    solve4(A, l, m);
    const double g = 0.57735026918962576;  // 1 / sqrt(3)
    const double hx = 0.5 * (p[2].x - p[0].x), cx = 0.5 * (p[2].x + p[0].x);
    const double hy = 0.5 * (p[2].y - p[0].y), cy = 0.5 * (p[2].y + p[0].y);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const double x = (i ? g : -g) * hx + cx, y = (j ? g : -g) * hy + cy;
            const double ksi = l[0] * x * y + l[1] * x + l[2] * y + l[3];
            const double eta = m[0] * x * y + m[1] * x + m[2] * y + m[3];
            const double dk[4] = {-0.25 * (1 - eta), 0.25 * (1 - eta), 0.25 * (1 + eta), -0.25 * (1 + eta)};
            const double de[4] = {-0.25 * (1 - ksi), -0.25 * (1 + ksi), 0.25 * (1 + ksi), 0.25 * (1 - ksi)};
            double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                J11 += dk[a] * p[a].x; J12 += dk[a] * p[a].y;
                J21 += de[a] * p[a].x; J22 += de[a] * p[a].y;
            }
            const double det = J11 * J22 - J12 * J21;
            double ux = 0, uy = 0, vx = 0, vy = 0;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const double dx = (J22 * dk[a] - J12 * de[a]) / det;
                const double dy = (-J21 * dk[a] + J11 * de[a]) / det;
                ux += dx * u[a].x; uy += dy * u[a].x;
                vx += dx * u[a].y; vy += dy * u[a].y;
            }
            const int q = 4 * e + 2 * i + j;
            Fgp[q] = make_double4(ux, uy, vx, vy);
            if (gpxy) gpxy[q] = make_double2(x, y);
        }
}

// F at node i = sum_k w[3i + k] Fgp[tri[3i + k]] (linear interpolation weights of the
// Gauss points, given by the caller)
__global__ __launch_bounds__(256) void k_gp_gather(int nn, const int *tri, const double *w,
                                                   const double4 *Fgp, double4 *F)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    double4 s = make_double4(0.0, 0.0, 0.0, 0.0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double wk = w[3 * i + k];
        const double4 f = Fgp[tri[3 * i + k]];
        s.x += wk * f.x; s.y += wk * f.y; s.z += wk * f.z; s.w += wk * f.w;
    }
    F[i] = s;
}

// ---- bad-node refill on the mesh -----------------------------------------------------
// One sweep: a node not yet known takes the inverse-distance average of its known
// neighbours (nodes sharing an element) and becomes known for the next sweep; nodes it
// reads were known before this sweep, so no node reads a value written in the same sweep.
__global__ __launch_bounds__(256) void k_fill_init(int nn, const int *bad, int *known)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nn) known[i] = bad[i] ? 0 : 1;
}

__global__ __launch_bounds__(256) void k_fill_sweep(int nn, const int *rowptr, const int *col,
                                                    const double2 *xy, const int *kin, int *kout,
                                                    double2 *U, double4 *F)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    if (kin[i]) { kout[i] = 1; return; }
    const double2 pi = xy[i];
    double ws = 0.0;
    double2 su = make_double2(0.0, 0.0);
    double4 sf = make_double4(0.0, 0.0, 0.0, 0.0);
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
        const int c = col[k];
        if (c == i || !kin[c]) continue;
        const double dx = xy[c].x - pi.x, dy = xy[c].y - pi.y;
        const double w = 1.0 / sqrt(dx * dx + dy * dy);
        ws += w;
        su.x += w * U[c].x; su.y += w * U[c].y;
        if (F) {
            const double4 f = F[c];
            sf.x += w * f.x; sf.y += w * f.y; sf.z += w * f.z; sf.w += w * f.w;
        }
    }
    if (ws > 0.0) {
        U[i] = make_double2(su.x / ws, su.y / ws);
        if (F) F[i] = make_double4(sf.x / ws, sf.y / ws, sf.z / ws, sf.w / ws);
        kout[i] = 1;
    } else {
        kout[i] = 0;
    }
}

__global__ __launch_bounds__(256) void k_fill_rest(int nn, const int *known, double2 *U, double4 *F)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn || known[i]) return;
    U[i] = make_double2(0.0, 0.0);
    if (F) F[i] = make_double4(0.0, 0.0, 0.0, 0.0);
}

__global__ __launch_bounds__(256) void k_start(int nn, const int *code, const double2 *U, double2 *x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nn) x[i] = code[i] == 0 ? U[i] : make_double2(0.0, 0.0);
}

__global__ __launch_bounds__(256) void k_copy_dir(int nn, const int *code, const double4 *src,
                                                  double4 *x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nn && code[i] == 1) x[i] = src[i];
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

hipError_t quad_gp_strain(const QuadDev &q, const double *U, double *Fgp, double *gpxy,
                          hipStream_t st)
{
    const int E = q.m.E;
    if (E <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gp_strain, dim3((E + 255) / 256), dim3(256), 0, st, E,
                       q.m.coords.as<double2>(), q.elems.as<int>(),
                       reinterpret_cast<const double2 *>(U), reinterpret_cast<double4 *>(Fgp),
                       reinterpret_cast<double2 *>(gpxy));
    return hipGetLastError();
}

hipError_t quad_gp_gather(int nn, const int *tri, const double *w, const double *Fgp, double *F,
                          hipStream_t st)
{
    hipLaunchKernelGGL(k_gp_gather, dim3((nn + 255) / 256), dim3(256), 0, st, nn, tri, w,
                       reinterpret_cast<const double4 *>(Fgp), reinterpret_cast<double4 *>(F));
    return hipGetLastError();
}

hipError_t launch_meshfill(const MeshDev &m, const int *bad, double *U, double *F, int *known,
                           int sweeps, hipStream_t st)
{
    const int nn = m.nn, nb = (nn + 255) / 256;
    int *k0 = known, *k1 = known + nn;
    hipLaunchKernelGGL(k_fill_init, dim3(nb), dim3(256), 0, st, nn, bad, k0);
    for (int s = 0; s < sweeps; ++s) {
        hipLaunchKernelGGL(k_fill_sweep, dim3(nb), dim3(256), 0, st, nn, m.rowptr.as<int>(),
                           m.col.as<int>(), m.coords.as<double2>(), (const int *)k0, k1,
                           reinterpret_cast<double2 *>(U), reinterpret_cast<double4 *>(F));
        std::swap(k0, k1);
    }
    hipLaunchKernelGGL(k_fill_rest, dim3(nb), dim3(256), 0, st, nn, (const int *)k0,
                       reinterpret_cast<double2 *>(U), reinterpret_cast<double4 *>(F));
    return hipGetLastError();
}

hipError_t quad_start(const QuadDev &q, const double *U, double *x, hipStream_t st)
{
    const int nn = q.m.nn;
    hipLaunchKernelGGL(k_start, dim3((nn + 255) / 256), dim3(256), 0, st, nn, q.code.as<int>(),
                       reinterpret_cast<const double2 *>(U), reinterpret_cast<double2 *>(x));
    return hipGetLastError();
}

hipError_t quad_copy_dirichlet(const QuadDev &q, const double *src, double *x, hipStream_t st)
{
    const int nn = q.m.nn;
    hipLaunchKernelGGL(k_copy_dir, dim3((nn + 255) / 256), dim3(256), 0, st, nn, q.code.as<int>(),
                       reinterpret_cast<const double4 *>(src), reinterpret_cast<double4 *>(x));
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
