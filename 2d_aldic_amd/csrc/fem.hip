// Q4 finite-element global step and nodal strain on the device.
//
//   Subpb2.m:76-195                  element matrices with 3x3 Gauss points, solve
//   funGlobal_NodalStrainAvg.m:32-144 2x2 Gauss-point gradients, extrapolation,
//                                     element averaging
//
// The global matrix of Subpb2 couples u with u and v with v only (DN'DN and N'N are
// block-diagonal in the displacement component), with the same scalar node matrix
// A = beta K + mu M + alpha mu K in both blocks.  It is assembled once per
// (mesh, beta, mu, alpha) into a scalar CSR over nodes: one thread per node gathers
// the elements around it (node -> element list) in a fixed order, so the assembly is
// atomic-free and deterministic.  The right-hand side is
//   b_u(i) = beta (G_i . (F11 - W11, F12 - W12)_i) + mu (M (U - V)_u)_i
//   b_v(i) = beta (G_i . (F21 - W21, F22 - W22)_i) + mu (M (U - V)_v)_i
// with G_i = sum_e int_e grad N_i dA (the reference's diag(DN' (F-W)') uses the
// node's own F, Subpb2.m:88-94,149).  The solve is Jacobi-preconditioned CG in FP64
// on interleaved (u, v) vectors, two kernels per iteration; scalars stay on the
// device and the dot products finish in the last block (fixed order).
#include <algorithm>

#include "aldic_internal.h"
#include "common.h"

namespace aldic {

__device__ __forceinline__ void q4_dN(double ksi, double eta, double dNk[4], double dNe[4])
{
    dNk[0] = -(1 - eta) * 0.25; dNe[0] = -(1 - ksi) * 0.25;
    dNk[1] = (1 - eta) * 0.25;  dNe[1] = -(1 + ksi) * 0.25;
    dNk[2] = (1 + eta) * 0.25;  dNe[2] = (1 + ksi) * 0.25;
    dNk[3] = -(1 + eta) * 0.25; dNe[3] = (1 - ksi) * 0.25;
}

// x-derivative and y-derivative of the 4 shape functions at (ksi, eta); returns det J.
__device__ __forceinline__ double q4_grad(const double px[4], const double py[4], double ksi,
                                          double eta, double ddx[4], double ddy[4])
{
    double dNk[4], dNe[4];
    q4_dN(ksi, eta, dNk, dNe);
    double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        J11 += dNk[b] * px[b]; J12 += dNk[b] * py[b];
        J21 += dNe[b] * px[b]; J22 += dNe[b] * py[b];
    }
    const double det = J11 * J22 - J12 * J21;
    const double i11 = J22 / det, i12 = -J12 / det, i21 = -J21 / det, i22 = J11 / det;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        ddx[b] = i11 * dNk[b] + i12 * dNe[b];
        ddy[b] = i21 * dNk[b] + i22 * dNe[b];
    }
    return det;
}

__global__ __launch_bounds__(256) void k_assemble(int nn, const double2 *coords, const int4 *elems,
                                                  const int *n2e_ptr, const N2E *n2e,
                                                  double beta, double mu, double alpha,
                                                  const int *rowptr, const int *col, double *aval,
                                                  double *mval, double *diag, double2 *gvec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) { aval[k] = 0.0; mval[k] = 0.0; }
    const double gp[3] = {-0.7745966692414834, 0.0, 0.7745966692414834};
    const double gw[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    double gx = 0.0, gy = 0.0;
    for (int q = n2e_ptr[i]; q < n2e_ptr[i + 1]; ++q) {
        const N2E r = n2e[q];
        const int4 el = elems[r.e];
        const int nd[4] = {el.x, el.y, el.z, el.w};
        double px[4], py[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) { px[b] = coords[nd[b]].x; py[b] = coords[nd[b]].y; }
        double Kr[4] = {0, 0, 0, 0}, Mr[4] = {0, 0, 0, 0};
        for (int ti = 0; ti < 3; ++ti)
            for (int tj = 0; tj < 3; ++tj) {
                const double ksi = gp[ti], eta = gp[tj];
                double ddx[4], ddy[4];
                const double det = q4_grad(px, py, ksi, eta, ddx, ddy);
                const double wJ = det * gw[ti] * gw[tj];
                const double N[4] = {(1 - ksi) * (1 - eta) * 0.25, (1 + ksi) * (1 - eta) * 0.25,
                                     (1 + ksi) * (1 + eta) * 0.25, (1 - ksi) * (1 + eta) * 0.25};
                const int a = r.a;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    Kr[b] += wJ * (ddx[a] * ddx[b] + ddy[a] * ddy[b]);
                    Mr[b] += wJ * (N[a] * N[b]);
                }
                gx += wJ * ddx[a];
                gy += wJ * ddy[a];
            }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            aval[r.slot[b]] += beta * Kr[b] + mu * Mr[b] + alpha * mu * Kr[b];
            mval[r.slot[b]] += Mr[b];
        }
    }
    double d = 0.0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k)
        if (col[k] == i) d = aval[k];
    diag[i] = d;
    gvec[i] = make_double2(gx, gy);
}

__global__ __launch_bounds__(256) void k_rhs(int nn, const int *rowptr, const int *col,
                                             const double *mval, const double2 *gvec, double beta,
                                             double mu, const double2 *U, const double4 *F,
                                             const double4 *udual, const double2 *vdual,
                                             double2 *b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    double su = 0.0, sv = 0.0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
        const int c = col[k];
        const double m = mval[k];
        const double2 u = U[c], v = vdual ? vdual[c] : make_double2(0.0, 0.0);
        su += m * (u.x - v.x);
        sv += m * (u.y - v.y);
    }
    const double4 f = F[i];  // [F11 F21 F12 F22]
    const double4 w = udual ? udual[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    const double2 g = gvec[i];
    b[i] = make_double2(beta * (g.x * (f.x - w.x) + g.y * (f.z - w.z)) + mu * su,
                        beta * (g.x * (f.y - w.y) + g.y * (f.w - w.w)) + mu * sv);
}

// ---- Jacobi-PCG, one reduction and one kernel per iteration -------------------------
// Chronopoulos-Gear form: with u = D^-1 r and w = A u,
//   gamma = r.u, delta = w.u;  beta = gamma / gamma_prev,
//   alpha = gamma / (delta - beta gamma / alpha_prev);
//   p = u + beta p, s = w + beta s (= A p), x += alpha p, r -= alpha s.
// Iteration i reads parity i & 1 and writes the other.  Each row forms its neighbours'
// next residual r_j - alpha (w_j + beta s_j) from the previous kernel's vectors, so the
// product w = A u of the next u needs no second grid-wide step; the per-block partials
// of gamma, delta and r.r are reduced by the next kernel in every block in the same
// order (identical scalars, deterministic).  A sticky stop flag (block 0 of the deciding
// kernel) turns later kernels into no-ops.
struct CgPart {
    double *G[2], *D[2], *RR[2], *BB;  // per-block partials
    double *hist;                      // {alpha, gamma} per parity
    double *stat;                      // [0] done, [1] iterations, [2] relative residual
    double *rec;                       // optional copy of stat, written alongside it
    int nb;
    double rtol2;
};

struct CgVec {
    double2 *r[2], *w[2], *s[2], *p[2];
};

constexpr int kPcgThreads = 256;
constexpr int kPcgMaxBlocks = 512;
constexpr int kCgChunk = 8;  // row entries whose loads are issued together

// Sum of one partial array, identical in every block (fixed order).
// nb <= 64 NPL: NPL independent (clamped) loads per lane, summed in the order
// k = lane, lane + 64, ...
template <int NPL>
__device__ __forceinline__ double part_sum(const double *p, int nb)
{
    const int lane = threadIdx.x & 63;
    double v[NPL];
#pragma unroll
    for (int u = 0; u < NPL; ++u) v[u] = p[min(lane + 64 * u, nb - 1)];
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < NPL; ++u)
        if (lane + 64 * u < nb) s += v[u];
    return wave_sum(s);
}

__device__ __forceinline__ void block_reduce3(double a, double b, double c, double *pa, double *pb,
                                              double *pc)
{
    __shared__ double red[3][4];
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][w] = a; red[1][w] = b; red[2][w] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        pa[blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        pb[blockIdx.x] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        pc[blockIdx.x] = red[2][0] + red[2][1] + red[2][2] + red[2][3];
    }
}

__device__ __forceinline__ double2 csr_row(int i, const int *rowptr, const int *col,
                                           const double *aval, const double2 *v)
{
    double au = 0.0, av = 0.0;
    for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) {
        const double2 vc = v[col[q]];
        au += aval[q] * vc.x;
        av += aval[q] * vc.y;
    }
    return make_double2(au, av);
}

// r0 = b - A x0 (x0 = 0 unless warm), s = p = 0; partial b.b
__global__ __launch_bounds__(kPcgThreads) void k_cg_init(int nn, const int *rowptr, const int *col,
                                                         const double *aval, const double2 *b,
                                                         double2 *x, int warm, CgVec v, CgPart P)
{
    double bb = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
        const double2 bi = b[i];
        double2 ri = bi;
        if (warm) {
            const double2 ax = csr_row(i, rowptr, col, aval, x);
            ri = make_double2(bi.x - ax.x, bi.y - ax.y);
        } else {
            x[i] = make_double2(0.0, 0.0);
        }
        v.r[0][i] = ri;
        v.s[0][i] = make_double2(0.0, 0.0);
        v.p[0][i] = make_double2(0.0, 0.0);
        bb += bi.x * bi.x + bi.y * bi.y;
    }
    block_reduce3(bb, 0.0, 0.0, P.BB, P.D[1], P.RR[1]);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        P.stat[0] = 0.0; P.stat[1] = 0.0; P.stat[2] = 1.0;
        if (P.rec) { P.rec[0] = 0.0; P.rec[1] = 0.0; P.rec[2] = 1.0; }
    }
}

// w0 = A D^-1 r0; partials gamma0, delta0, r0.r0
__global__ __launch_bounds__(kPcgThreads) void k_cg_w0(int nn, const int *rowptr, const int *col,
                                                       const double *aval, const double *diag,
                                                       CgVec v, CgPart P)
{
    double g = 0.0, dl = 0.0, rr = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
        double au = 0.0, av = 0.0;
        for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) {
            const int c = col[q];
            const double dc = diag[c] != 0.0 ? 1.0 / diag[c] : 1.0;
            const double2 rc = v.r[0][c];
            au += aval[q] * (rc.x * dc);
            av += aval[q] * (rc.y * dc);
        }
        const double di = diag[i] != 0.0 ? 1.0 / diag[i] : 1.0;
        const double2 ri = v.r[0][i];
        v.w[0][i] = make_double2(au, av);
        g += ri.x * ri.x * di + ri.y * ri.y * di;
        dl += au * ri.x * di + av * ri.y * di;
        rr += ri.x * ri.x + ri.y * ri.y;
    }
    block_reduce3(g, dl, rr, P.G[0], P.D[0], P.RR[0]);
}

template <int NPL>
__global__ __launch_bounds__(kPcgThreads) void k_cg_it(int nn, const int *rowptr, const int *col,
                                                       const double *aval, const double *diag,
                                                       double2 *x, CgVec v, CgPart P, int i)
{
    const int cur = i & 1, nxt = cur ^ 1;
    // stop flag, partials and history are loaded together.  Wave w sums partial array w
    // (G, D, RR of this parity, then BB); the arrays are consecutive runs of nb words from
    // G[0] (G0 G1 D0 D1 RR0 RR1 BB), so the offset is arithmetic, not a pointer select.
    const double stop = P.stat[0];
    const double h_alpha = P.hist[2 * nxt], h_gamma = P.hist[2 * nxt + 1];  // i - 1
    const int w = threadIdx.x >> 6;
    const double *ps = P.G[0] + (size_t)(w < 3 ? 2 * w + cur : 6) * P.nb;
    const double val = part_sum<NPL>(ps, P.nb);
    if (stop != 0.0) return;  // sticky stop (uniform: every thread leaves)
    __shared__ double sh[4];
    if ((threadIdx.x & 63) == 0) sh[w] = val;
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
    // stop at the tolerance, on b == 0, on a NaN residual, or on a breakdown (the
    // system is symmetric positive definite: a non-positive or non-finite step)
    const bool done = !(rr > P.rtol2 * bb) || !(alpha > 0.0) || !isfinite(alpha) ||
                      !isfinite(beta);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        P.stat[0] = done ? 1.0 : 0.0;
        P.stat[1] = (double)i;
        P.stat[2] = bb > 0 ? sqrt(rr / bb) : 0.0;
        if (P.rec) { P.rec[0] = P.stat[0]; P.rec[1] = P.stat[1]; P.rec[2] = P.stat[2]; }
        P.hist[2 * cur] = alpha;
        P.hist[2 * cur + 1] = gam;
    }
    if (done) return;
    const double2 *rc = v.r[cur], *wc = v.w[cur], *sc = v.s[cur], *pc = v.p[cur];
    double g = 0.0, dl = 0.0, rn2 = 0.0;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nn; q += gridDim.x * blockDim.x) {
        double2 wa = make_double2(0.0, 0.0), rnew = wa, snew = wa, unew = wa;
        // the row in chunks of kCgChunk entries: the chunk's column indices and values
        // are loaded together, then all its neighbours' operands, then the arithmetic
        // in entry order (entries past the row's end point at the row's own node with
        // value 0 and are skipped)
        const int rs = rowptr[q], re = rowptr[q + 1];
        for (int k0 = rs; k0 < re; k0 += kCgChunk) {
            int cj[kCgChunk];
            double av[kCgChunk];
#pragma unroll
            for (int u = 0; u < kCgChunk; ++u) {
                const bool in = k0 + u < re;
                const int kc = min(k0 + u, re - 1);  // in the row: loads stay in bounds
                cj[u] = in ? col[kc] : q;
                av[u] = in ? aval[kc] : 0.0;
            }
            double2 wj[kCgChunk], sj[kCgChunk], rj[kCgChunk];
            double dd[kCgChunk];
#pragma unroll
            for (int u = 0; u < kCgChunk; ++u) {
                wj[u] = wc[cj[u]];
                sj[u] = sc[cj[u]];
                rj[u] = rc[cj[u]];
                dd[u] = diag[cj[u]];
            }
#pragma unroll
            for (int u = 0; u < kCgChunk; ++u) {
                if (k0 + u >= re) break;
                const int j = cj[u];
                const double dj = dd[u] != 0.0 ? 1.0 / dd[u] : 1.0;
                const double2 snj = make_double2(wj[u].x + beta * sj[u].x, wj[u].y + beta * sj[u].y);
                const double2 rnj = make_double2(rj[u].x - alpha * snj.x, rj[u].y - alpha * snj.y);
                const double2 unj = make_double2(rnj.x * dj, rnj.y * dj);
                wa.x += av[u] * unj.x;
                wa.y += av[u] * unj.y;
                if (j == q) { rnew = rnj; snew = snj; unew = unj; }
            }
        }
        const double dq = diag[q] != 0.0 ? 1.0 / diag[q] : 1.0;
        const double2 rq = rc[q], pq = pc[q];
        const double2 pn = make_double2(rq.x * dq + beta * pq.x, rq.y * dq + beta * pq.y);
        double2 xq = x[q];
        xq.x += alpha * pn.x;
        xq.y += alpha * pn.y;
        x[q] = xq;
        v.p[nxt][q] = pn;
        v.s[nxt][q] = snew;
        v.r[nxt][q] = rnew;
        v.w[nxt][q] = wa;
        g += rnew.x * unew.x + rnew.y * unew.y;
        dl += wa.x * unew.x + wa.y * unew.y;
        rn2 += rnew.x * rnew.x + rnew.y * rnew.y;
    }
    block_reduce3(g, dl, rn2, P.G[nxt], P.D[nxt], P.RR[nxt]);
}

// ---- fast diagonalisation for structured grids ------------------------------------
// On a uniform gx x gy lattice with bilinear elements (x-fastest numbering) the Q4
// matrices are Kronecker forms of the 1-D FE matrices K1 = 1/h tridiag(-1, 2, -1) and
// M1 = h/6 tridiag(1, 4, 1) (end diagonals halved):  M = M1y (x) M1x,
// K = K1y (x) M1x + M1y (x) K1x.  With V^T M1 V = I and V^T K1 V = diag(lambda) (the
// DCT-I basis, computed in closed form on the host), the system
// (cK K + mu M) X = B, with X, B the gy x gx nodal arrays, has the exact solution
//   X = Vy [ (Vy^T B Vx) ./ (mu + cK (lambda_y,i + lambda_x,j)) ] Vx^T
// (Lynch, Rice & Thomas 1964).  Four FP64 GEMM passes replace the whole CG loop; u and
// v ride along as the interleaved column pairs of the gy x 2gx nodal array.
struct GemmArgs {
    int M, N, K;
    const double *A; long sam, sak, boA;   // A(m,k) = A[m*sam + k*sak + batch*boA]
    const double *B; long sbk, sbn, boB;   // B(k,n)
    double *C; long scm, scn, boC;         // C(m,n)
    const double *lamy, *lamx;             // scale C(m,n) /= mu + cK (lamy[m] + lamx[n])
    double mu, cK;
};

static hipError_t fem_fd_solve4(const MeshDev &m, const double *b, double *x, DBuf *work,
                                hipStream_t st);

// C = A B (then ./ (mu + cK (lamy[m] + lamx[n])) when SCALE) on the FP64 matrix cores, for
// the four-pass solve: one wave per TM x 32 tile of C (TM / 16 x 2 tiles of
// v_mfma_f64_16x16x4f64: A 16 x 4 with lane l holding (row l % 16, k l / 16), B 4 x 16 with
// lane l holding (k l / 16, col l % 16), D 16 x 16 with lane l, element v holding
// (row l / 16 + 4 v, col l % 16)).  Every lane loads its own fragments from global
// memory (no LDS, no barrier; the matrices hold < 2^31 bytes, so offsets are 32-bit).
// K runs in chunks of 16 steps of 4: all of a chunk's fragments are loaded into their own
// registers and waited for, then one MFMA chain per output tile starts from zero and
// runs the chunk back to back, and the chunk's result is added to an FP64 register sum
// before the next chunk's loads: no fragment register is reloaded while an MFMA that
// reads it may be in flight, and no MFMA accumulator is carried between chunks.  Addresses are clamped into the matrices; A's
// word is multiplied by 0 past K (finite data), and rows or columns past M, N only feed
// results that are not stored.  Tiles run along N fastest; the batch is blockIdx.y.
typedef double gm_d4 __attribute__((ext_vector_type(4)));

constexpr int kGemmChunk = 16;  // k steps of 4 per MFMA chain

template <int TM, bool SCALE>
__global__ __launch_bounds__(256) void k_gemm_mfma(GemmArgs g)
{
    constexpr int MT = TM / 16;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int tilesN = (g.N + 31) / 32, tilesM = (g.M + TM - 1) / TM;
    const int tile = blockIdx.x * 4 + wv;
    if (tile >= tilesM * tilesN) return;  // whole wave
    const int m0 = (tile / tilesN) * TM, n0 = (tile % tilesN) * 32;
    const int bz = blockIdx.y;
    const double *A = g.A + (long)bz * g.boA;
    const double *B = g.B + (long)bz * g.boB;
    double *C = g.C + (long)bz * g.boC;
    const int li = lane & 15, lk = lane >> 4;
    const int sak = (int)g.sak, sbk = (int)g.sbk;
    int ar[MT], bc[2];
#pragma unroll
    for (int i = 0; i < MT; ++i) ar[i] = min(m0 + 16 * i + li, g.M - 1) * (int)g.sam;
#pragma unroll
    for (int j = 0; j < 2; ++j) bc[j] = min(n0 + 16 * j + li, g.N - 1) * (int)g.sbn;
    double sum[MT][2][4];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) sum[i][j][v] = 0.0;
    const int ks = (g.K + 3) / 4;
    for (int s0 = 0; s0 < ks; s0 += kGemmChunk) {
        double a[kGemmChunk][MT], b[kGemmChunk][2];
#pragma unroll
        for (int u = 0; u < kGemmChunk; ++u) {
            const int k = 4 * (s0 + u) + lk, kc = min(k, g.K - 1);
            const double km = k < g.K ? 1.0 : 0.0;
#pragma unroll
            for (int i = 0; i < MT; ++i) a[u][i] = A[ar[i] + kc * sak] * km;
#pragma unroll
            for (int j = 0; j < 2; ++j) b[u][j] = B[kc * sbk + bc[j]];
        }
        // every fragment of the chunk is loaded before its first MFMA (no instruction is
        // scheduled across the barrier), and no load of the next chunk is issued before
        // this chunk's results are summed
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < MT; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                gm_d4 part = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int u = 0; u < kGemmChunk; ++u)
                    part = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][i], b[u][j], part, 0, 0, 0);
#pragma unroll
                for (int v = 0; v < 4; ++v) sum[i][j][v] += part[v];
            }
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int c = n0 + 16 * j + li;
            if (c >= g.N) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = m0 + 16 * i + lk + 4 * v;
                if (r >= g.M) continue;
                double o = sum[i][j][v];
                if (SCALE) o /= g.mu + g.cK * (g.lamy[r] + g.lamx[c]);
                C[(long)r * g.scm + (long)c * g.scn] = o;
            }
        }
}

// Wave tiles of 32 rows when they already give one wave per SIMD (1024), else 16.
static hipError_t gemm_mfma(const GemmArgs &g, int batch, bool scale, hipStream_t st)
{
    const long tn = (g.N + 31) / 32;
    const bool big = ((g.M + 31) / 32) * tn * batch >= 1024;
    const long tiles = (big ? (g.M + 31) / 32 : (g.M + 15) / 16) * tn;
    const dim3 grid((unsigned)((tiles + 3) / 4), (unsigned)batch);
    if (big) {
        if (scale) hipLaunchKernelGGL((k_gemm_mfma<32, true>), grid, dim3(256), 0, st, g);
        else hipLaunchKernelGGL((k_gemm_mfma<32, false>), grid, dim3(256), 0, st, g);
    } else {
        if (scale) hipLaunchKernelGGL((k_gemm_mfma<16, true>), grid, dim3(256), 0, st, g);
        else hipLaunchKernelGGL((k_gemm_mfma<16, false>), grid, dim3(256), 0, st, g);
    }
    return hipGetLastError();
}

// Fused pass of the fast-diagonalisation solve for gx, gy <= 128: workgroup (row strip
// of 16 rows, component) computes
//   out[m, :] = ((L B)[m, :] R) (./ (mu + cK (lamy[m] + lamx[n])) when SCALE)
// with L(m, k) = Lsrc[k * gy + m] and R(j, n) = Rsrc[j * gx + n].  B and out are the
// interleaved gy x 2gx nodal arrays.  Both products run on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64: A 16 x 4 with lane l holding (l % 16, l / 16), B 4 x 16 with
// lane l holding (l / 16, l % 16), D 16 x 16 with lane l, element v holding
// (l / 16 + 4 v, l % 16)).  16 waves = 8 column tiles of 16 x 2 halves of the inner
// index; every wave loads its 16 A, B and R fragments from global memory up front (one
// round trip), indices clamped in range and the words past gy or gx zeroed.  The strip
// (L B) goes through LDS once; the halves are summed in a fixed order.
constexpr int kFdRows = 16;
constexpr int kFdMax = 128;                 // gx, gy limit of the fused path
constexpr int kFdThreads = 1024;
constexpr int kFdTiles = kFdMax / 16;       // column tiles
constexpr int kFdSteps = kFdMax / 4 / 2;    // inner steps of 4 per half

typedef double fd_d4 __attribute__((ext_vector_type(4)));

template <bool SCALE>
__global__ __launch_bounds__(kFdThreads) void k_fd_pass(int gy, int gx,
                                                        const double *__restrict__ Lsrc,
                                                        const double *__restrict__ B,
                                                        const double *__restrict__ Rsrc,
                                                        double *__restrict__ out,
                                                        const double *__restrict__ lamy,
                                                        const double *__restrict__ lamx,
                                                        double mu, double cK)
{
    __shared__ double Tt[kFdMax * kFdRows];        // Tt[j * 16 + i] = (L B)(m0 + i, j)
    __shared__ double Pp[kFdTiles][64 * 4];        // second half's tiles
    const int comp = blockIdx.y, m0 = blockIdx.x * kFdRows;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int ct = wv % kFdTiles, h = wv / kFdTiles;
    const int li = lane & 15, lk = lane >> 4;
    const int w2 = 2 * gx;
    const int kh = ((gy + 3) / 4 + 1) / 2, jh = ((gx + 3) / 4 + 1) / 2;  // steps per half
    const int rowc = min(m0 + li, gy - 1);    // A row (rows past gy are not stored)
    const int col = ct * 16 + li, colc = min(col, gx - 1);
    double fa[kFdSteps], fb[kFdSteps], fr[kFdSteps];
#pragma unroll
    for (int s = 0; s < kFdSteps; ++s) {
        const int k = (h * kh + s) * 4 + lk, kc = min(k, gy - 1);
        const bool kin = s < kh && k < gy;
        const double av = Lsrc[(long)kc * gy + rowc];
        const double bv = B[(long)kc * w2 + 2 * colc + comp];
        fa[s] = kin ? av : 0.0;
        fb[s] = kin && col < gx ? bv : 0.0;
        const int j = (h * jh + s) * 4 + lk, jc = min(j, gx - 1);
        const double rv = Rsrc[(long)jc * gx + colc];
        fr[s] = s < jh && j < gx && col < gx ? rv : 0.0;
    }
    // Tt = L_strip B
    fd_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < kFdSteps; ++s)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[s], fb[s], acc, 0, 0, 0);
    if (h == 1) {
#pragma unroll
        for (int v = 0; v < 4; ++v) Pp[ct][lane * 4 + v] = acc[v];
    }
    __syncthreads();
    if (h == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) Tt[col * kFdRows + lk + 4 * v] = acc[v] + Pp[ct][lane * 4 + v];
    }
    __syncthreads();
    // out strip = Tt R
    acc = fd_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < kFdSteps; ++s) {
        const double ta = Tt[((h * jh + s) * 4 + lk) * kFdRows + li];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ta, fr[s], acc, 0, 0, 0);
    }
    if (h == 1) {
#pragma unroll
        for (int v = 0; v < 4; ++v) Pp[ct][lane * 4 + v] = acc[v];
    }
    __syncthreads();
    if (h == 0 && col < gx) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = m0 + lk + 4 * v;
            if (m < gy) {
                double o = acc[v] + Pp[ct][lane * 4 + v];
                if (SCALE) o /= mu + cK * (lamy[m] + lamx[col]);
                out[(long)m * w2 + 2 * col + comp] = o;
            }
        }
    }
}

static hipError_t fem_fd_solve(const MeshDev &m, const double *b, double *x, DBuf *work,
                               hipStream_t st)
{
    if (m.gx <= kFdMax && m.gy <= kFdMax) {
        const int gx = m.gx, gy = m.gy;
        hipError_t e;
        if ((e = work[0].ensure(sizeof(double) * 2 * (size_t)gx * gy)) != hipSuccess) return e;
        double *T2 = work[0].as<double>();
        const double *Vx = m.fdV.as<double>(), *Vy = Vx + (size_t)gx * gx;
        const double *VxT = Vy + (size_t)gy * gy, *VyT = VxT + (size_t)gx * gx;
        const double *lx = m.fdlam.as<double>(), *ly = lx + gx;
        const dim3 grid((gy + kFdRows - 1) / kFdRows, 2);
        // T2 = (Vy^T B Vx) ./ D:  L(m, k) = Vy[k][m],  R(j, n) = Vx[j][n]
        hipLaunchKernelGGL(k_fd_pass<true>, grid, dim3(kFdThreads), 0, st, gy, gx, Vy, b, Vx, T2, ly,
                           lx, m.mu, m.beta + m.alpha * m.mu);
        // X = Vy T2 Vx^T:  L(m, k) = Vy[m][k] = VyT[k][m],  R(j, n) = Vx[n][j] = VxT[j][n]
        hipLaunchKernelGGL(k_fd_pass<false>, grid, dim3(kFdThreads), 0, st, gy, gx, VyT,
                           (const double *)T2, VxT, x, nullptr, nullptr, 0.0, 0.0);
        return hipGetLastError();
    }
    return fem_fd_solve4(m, b, x, work, st);
}

// Four separate GEMM passes (grids too large for the fused path).
static hipError_t fem_fd_solve4(const MeshDev &m, const double *b, double *x, DBuf *work,
                                hipStream_t st)
{
    const int gx = m.gx, gy = m.gy;
    const size_t nv = 2 * (size_t)gx * gy;
    hipError_t e;
    if ((e = work[0].ensure(sizeof(double) * nv)) != hipSuccess) return e;
    if ((e = work[1].ensure(sizeof(double) * nv)) != hipSuccess) return e;
    double *T1 = work[0].as<double>(), *T2 = work[1].as<double>();
    const double *Vx = m.fdV.as<double>(), *Vy = Vx + (size_t)gx * gx;
    const double *lx = m.fdlam.as<double>(), *ly = lx + gx;
    const long w2 = 2L * gx;
    GemmArgs g{};
    // T1 = Vy^T B
    g = GemmArgs{gy, 2 * gx, gy, Vy, 1, gy, 0, b, w2, 1, 0, T1, w2, 1, 0, nullptr, nullptr, 0, 0};
    if ((e = gemm_mfma(g, 1, false, st)) != hipSuccess) return e;
    // T2 = (T1 Vx) ./ D, per component
    g = GemmArgs{gy, gx, gx, T1, w2, 2, 1, Vx, gx, 1, 0, T2, w2, 2, 1, ly, lx, m.mu,
                 m.beta + m.alpha * m.mu};
    if ((e = gemm_mfma(g, 2, true, st)) != hipSuccess) return e;
    // T1 = T2 Vx^T, per component
    g = GemmArgs{gy, gx, gx, T2, w2, 2, 1, Vx, 1, gx, 0, T1, w2, 2, 1, nullptr, nullptr, 0, 0};
    if ((e = gemm_mfma(g, 2, false, st)) != hipSuccess) return e;
    // X = Vy T1
    g = GemmArgs{gy, 2 * gx, gy, Vy, gy, 1, 0, T1, w2, 1, 0, x, w2, 1, 0, nullptr, nullptr, 0, 0};
    return gemm_mfma(g, 1, false, st);
}

hipError_t fem_assemble(MeshDev &m, double beta, double mu, double alpha, hipStream_t st)
{
    const int nb = (m.nn + 255) / 256;
    hipLaunchKernelGGL(k_assemble, dim3(nb), dim3(256), 0, st, m.nn, m.coords.as<double2>(),
                       m.elems.as<int4>(), m.n2e_ptr.as<int>(), m.n2e.as<N2E>(), beta, mu, alpha,
                       m.rowptr.as<int>(), m.col.as<int>(), m.aval.as<double>(),
                       m.mval.as<double>(), m.diag.as<double>(), m.gvec.as<double2>());
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) { m.beta = beta; m.mu = mu; m.alpha = alpha; }
    return e;
}

hipError_t fem_rhs(const MeshDev &m, double beta, double mu, const double *U, const double *F,
                   const double *udual, const double *vdual, double *b, hipStream_t st)
{
    const int nb = (m.nn + 255) / 256;
    hipLaunchKernelGGL(k_rhs, dim3(nb), dim3(256), 0, st, m.nn, m.rowptr.as<int>(),
                       m.col.as<int>(), m.mval.as<double>(), m.gvec.as<double2>(), beta, mu,
                       (const double2 *)U, (const double4 *)F, (const double4 *)udual,
                       (const double2 *)vdual, (double2 *)b);
    return hipGetLastError();
}

int pcg_blocks(int nn)
{
    int nb = (nn + kPcgThreads - 1) / kPcgThreads;
    return nb < 1 ? 1 : (nb > kPcgMaxBlocks ? kPcgMaxBlocks : nb);
}

// work[0..4]: r, z, pA, pB, Ap (double2 each); work[5]: partials + status.
// poll > 0: check convergence on the host every `poll` iterations (synchronises);
// poll == 0: enqueue exactly maxit iterations (kernels after convergence return at
// once), no host synchronisation -- usable inside stream capture.
hipError_t fem_pcg(const MeshDev &m, const double *b, double *x, double rtol, int maxit,
                   DBuf *work, double *stats_host, hipStream_t st, int poll, double *stats_dev,
                   bool warm)
{
    const int nn = m.nn;
    const int nb = pcg_blocks(nn);
    const size_t vb = sizeof(double2) * (size_t)nn;
    hipError_t e;
    // work[0]: r, w, s, p in two parities; work[5]: partials, history, status
    if ((e = work[0].ensure(8 * vb)) != hipSuccess) return e;
    if ((e = work[5].ensure(sizeof(double) * (7 * (size_t)nb + 8))) != hipSuccess) return e;
    double *pp = work[5].as<double>();
    CgPart P;
    P.G[0] = pp; P.G[1] = pp + nb; P.D[0] = pp + 2 * nb; P.D[1] = pp + 3 * nb;
    P.RR[0] = pp + 4 * nb; P.RR[1] = pp + 5 * nb; P.BB = pp + 6 * nb;
    P.hist = pp + 7 * nb; P.stat = P.hist + 4; P.nb = nb; P.rtol2 = rtol * rtol;
    P.rec = stats_dev;
    double2 *vv = work[0].as<double2>();
    CgVec v;
    for (int k = 0; k < 2; ++k) {
        v.r[k] = vv + (0 + k) * (size_t)nn;
        v.w[k] = vv + (2 + k) * (size_t)nn;
        v.s[k] = vv + (4 + k) * (size_t)nn;
        v.p[k] = vv + (6 + k) * (size_t)nn;
    }
    const int *rp = m.rowptr.as<int>(), *cl = m.col.as<int>();
    const double *av = m.aval.as<double>(), *dg = m.diag.as<double>();
    hipLaunchKernelGGL(k_cg_init, dim3(nb), dim3(kPcgThreads), 0, st, nn, rp, cl, av,
                       (const double2 *)b, (double2 *)x, warm ? 1 : 0, v, P);
    hipLaunchKernelGGL(k_cg_w0, dim3(nb), dim3(kPcgThreads), 0, st, nn, rp, cl, av, dg, v, P);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int chunk = poll > 0 ? poll : maxit;
    for (int it = 0; it < maxit; it += chunk) {
        for (int q = 0; q < chunk && it + q < maxit; ++q)
            if (nb <= 64)
                hipLaunchKernelGGL(k_cg_it<1>, dim3(nb), dim3(kPcgThreads), 0, st, nn, rp, cl,
                                   av, dg, (double2 *)x, v, P, it + q);
            else
                hipLaunchKernelGGL(k_cg_it<kPcgMaxBlocks / 64>, dim3(nb), dim3(kPcgThreads), 0,
                                   st, nn, rp, cl, av, dg, (double2 *)x, v, P, it + q);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (poll <= 0) break;
        double done = 0.0;
        if ((e = hipMemcpyAsync(&done, P.stat, sizeof(double), hipMemcpyDeviceToHost, st)) !=
            hipSuccess)
            return e;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
        if (done != 0.0) break;
    }
    // stats_dev, when given, is written by the kernels alongside P.stat
    if (stats_host) {
        // status of the last iteration kernel: {done, iterations, relative residual}
        double hs[3];
        if ((e = hipMemcpyAsync(hs, P.stat, sizeof hs, hipMemcpyDeviceToHost, st)) != hipSuccess)
            return e;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
        stats_host[0] = hs[1];
        stats_host[1] = hs[2];
    }
    return hipSuccess;
}

hipError_t fem_solve(const MeshDev &m, const double *b, double *x, double rtol, int maxit,
                     DBuf *work, double *stats_host, hipStream_t st, int poll)
{
    if (m.structured && m.mu > 0.0) {
        hipError_t e = fem_fd_solve(m, b, x, work + 6, st);
        if (stats_host) { stats_host[0] = 0.0; stats_host[1] = 0.0; }
        return e;
    }
    return fem_pcg(m, b, x, rtol, maxit, work, stats_host, st, poll);
}

// ---- nodal strain (funGlobal_NodalStrainAvg.m) ------------------------------------
// With udual set, also the ADMM dual update of the FEM branch (main_ALDIC.m:340,440):
// udual = F - F1, vdual = U - U1 at the same node.
__global__ __launch_bounds__(256) void k_strain(int nn, const double2 *coords, const int4 *elems,
                                                const int *n2e_ptr, const N2E *n2e,
                                                const double2 *U, int swap_fix, double4 *F,
                                                const double4 *F1, const double2 *U1,
                                                double4 *udual, double2 *vdual)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const double g = 0.57735026918962573;  // 1/sqrt(3)
    const double gk[4] = {-g, g, g, -g}, ge[4] = {-g, -g, g, g};
    const double s3h = 0.86602540378443865;  // sqrt(3)/2
    // extrapolation matrix of :105-108 (row a = node, column p = Gauss point)
    const double Mx[4][4] = {{1 + s3h, -0.5, 1 - s3h, -0.5},
                             {-0.5, 1 + s3h, -0.5, 1 - s3h},
                             {1 - s3h, -0.5, 1 + s3h, -0.5},
                             {-0.5, 1 - s3h, -0.5, 1 + s3h}};
    double acc[4] = {0, 0, 0, 0};
    int cnt = 0;
    for (int q = n2e_ptr[i]; q < n2e_ptr[i + 1]; ++q) {
        const N2E r = n2e[q];
        const int4 el = elems[r.e];
        const int nd[4] = {el.x, el.y, el.z, el.w};
        double px[4], py[4], ue[4], ve[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            px[b] = coords[nd[b]].x; py[b] = coords[nd[b]].y;
            ue[b] = U[nd[b]].x; ve[b] = U[nd[b]].y;
        }
        double nv[4] = {0, 0, 0, 0};
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            double ddx[4], ddy[4];
            q4_grad(px, py, gk[p], ge[p], ddx, ddy);
            double ux = 0, uy = 0, vx = 0, vy = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                ux += ddx[b] * ue[b]; uy += ddy[b] * ue[b];
                vx += ddx[b] * ve[b]; vy += ddy[b] * ve[b];
            }
            const double m = Mx[r.a][p];
            nv[0] += m * ux; nv[1] += m * uy; nv[2] += m * vx; nv[3] += m * vy;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] += nv[c];
        ++cnt;
    }
    double4 o;
    const double d = (double)cnt;  // 0 -> NaN like the reference's 0/0
    o.x = acc[0] / d;
    o.y = (swap_fix ? acc[2] : acc[1]) / d;
    o.z = (swap_fix ? acc[1] : acc[2]) / d;
    o.w = acc[3] / d;
    F[i] = o;
    if (udual) {
        const double4 f1 = F1[i];
        udual[i] = make_double4(o.x - f1.x, o.y - f1.y, o.z - f1.z, o.w - f1.w);
        const double2 u = U[i], u1 = U1[i];
        vdual[i] = make_double2(u.x - u1.x, u.y - u1.y);
    }
}

hipError_t fem_strain(const MeshDev &m, const double *U, int swap_fix, double *F, hipStream_t st,
                      const double *F1, const double *U1, double *udual, double *vdual)
{
    const int nb = (m.nn + 255) / 256;
    hipLaunchKernelGGL(k_strain, dim3(nb), dim3(256), 0, st, m.nn, m.coords.as<double2>(),
                       m.elems.as<int4>(), m.n2e_ptr.as<int>(), m.n2e.as<N2E>(),
                       (const double2 *)U, swap_fix, (double4 *)F, (const double4 *)F1,
                       (const double2 *)U1, (double4 *)udual, (double2 *)vdual);
    return hipGetLastError();
}

// ---- ADMM helpers --------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sub(const double *a, const double *b, double *o, long n)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) o[i] = a[i] - b[i];
}

// out[0] = sum (a - b)^2 over n values: one workgroup, fixed order (deterministic)
__global__ __launch_bounds__(1024) void k_sqdiff(const double *a, const double *b, long n,
                                                 double *out)
{
    __shared__ double red[16];
    double s = 0.0;
    for (long i = threadIdx.x; i < n; i += 1024) {
        const double d = a[i] - b[i];
        s = fma(d, d, s);
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < 16; ++k) t += red[k];
        out[0] = t;
    }
}

hipError_t launch_sqdiff(const double *a, const double *b, long n, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_sqdiff, dim3(1), dim3(1024), 0, st, a, b, n, out);
    return hipGetLastError();
}

hipError_t launch_sub(const double *a, const double *b, double *out, long n, hipStream_t st)
{
    hipLaunchKernelGGL(k_sub, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b, out, n);
    return hipGetLastError();
}

// Bad-point rule and refill (three grid launches, no host round trip).
//   LocalICGN.m:119-139 (mean_rule = 1) / Subpb1.m:167-182 (mean_rule = 0): a subset is
//   bad when its status is < 0 or > 99, and with the mean rule also when it exceeds
//   mean + 0.15 std of the good ones (std with N-1); bad U (and F) become NaN.
//   Refill (the reference interpolates the bad nodes from the good ones with
//   scatteredInterpolant 'natural', LocalICGN.m:144-161): every bad node takes the
//   inverse-distance average of the nearest good node along each of the four grid
//   directions (left, right, down, up, summed in that order); an isolated bad node gets
//   the mean of its four neighbours.  A node with no good node in its row or column
//   gets 0.
// k_itstats: per-block partials {count, sum, sum of squares} of the good statuses.
__global__ __launch_bounds__(256) void k_itstats(const int *it, int S, double *part)
{
    __shared__ double red[3][4];
    double c = 0, s = 0, s2 = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < S; i += gridDim.x * blockDim.x) {
        const int v = it[i];
        if (v >= 0 && v <= 99) { c += 1; s += v; s2 += (double)v * v; }
    }
    c = wave_sum(c); s = wave_sum(s); s2 = wave_sum(s2);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][w] = c; red[1][w] = s; red[2][w] = s2; }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) part[3 * blockIdx.x + q] = red[q][0] + red[q][1] + red[q][2] + red[q][3];
}

// k_mark: threshold from the partials (same order in every block), mark, copy, NaN.
// mean_rule 2 is the quadtree Subpb1's max(mean + 0.15 std, 20).
__global__ __launch_bounds__(256) void k_mark(const int *it, int S, const double *part, int np,
                                              int mean_rule, const double *Usrc,
                                              const double *Fsrc, double *U, double *F, int *bad,
                                              int *nbad)
{
    __shared__ double tot[3];
    if (threadIdx.x < 64) {
        double a[3] = {0, 0, 0};
        for (int k = threadIdx.x; k < np; k += 64)
            for (int q = 0; q < 3; ++q) a[q] += part[3 * k + q];
        for (int q = 0; q < 3; ++q) a[q] = wave_sum(a[q]);
        if (threadIdx.x == 0) { tot[0] = a[0]; tot[1] = a[1]; tot[2] = a[2]; }
    }
    __syncthreads();
    const double C = tot[0], Sm = tot[1], S2 = tot[2];
    const double mean = C > 0 ? Sm / C : 0.0;
    const double var = C > 1 ? (S2 - Sm * mean) / (C - 1) : 0.0;
    double thr = mean + 0.15 * sqrt(var > 0 ? var : 0.0);
    if (mean_rule == 2) thr = thr > 20.0 ? thr : 20.0;  // Subpb1Quadtree.m:180
    // statuses < 0 or > 99 = all but the counted good ones (Subpb1.m:166-169)
    if (nbad && blockIdx.x == 0 && threadIdx.x == 0) *nbad = S - (int)C;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < S; i += gridDim.x * blockDim.x) {
        const int v = it[i];
        const bool b = v < 0 || v > 99 || (mean_rule && C > 0 && (double)v > thr);
        bad[i] = b ? 1 : 0;
        double2 u = Usrc ? reinterpret_cast<const double2 *>(Usrc)[i]
                         : reinterpret_cast<const double2 *>(U)[i];
        if (b) u = make_double2(NAN, NAN);
        reinterpret_cast<double2 *>(U)[i] = u;
        if (F) {
            double4 f = Fsrc ? reinterpret_cast<const double4 *>(Fsrc)[i]
                             : reinterpret_cast<const double4 *>(F)[i];
            if (b) f = make_double4(NAN, NAN, NAN, NAN);
            reinterpret_cast<double4 *>(F)[i] = f;
        }
    }
}

// k_markfill: the threshold from k_itstats' partials (same order in every block), then
// per node the bad flag and, for a bad node, the axis refill: along each of the four grid
// directions the nearest good node (its flag recomputed from its status), weighted by
// inverse distance.  U (and F) are refilled in place: bad nodes are written, good nodes
// only read, so one pass suffices.
__device__ __forceinline__ bool is_bad(int v, int rule, double C, double thr)
{
    return v < 0 || v > 99 || (rule && C > 0 && (double)v > thr);
}

__global__ __launch_bounds__(256) void k_markfill(const int *it, int Nx, int Ny, const double *part,
                                                  int np, int rule, double *U, double *F, int *bad,
                                                  int *nbad)
{
    __shared__ double tot[3];
    if (threadIdx.x < 64) {
        double a[3] = {0, 0, 0};
        for (int k = threadIdx.x; k < np; k += 64)
            for (int q = 0; q < 3; ++q) a[q] += part[3 * k + q];
        for (int q = 0; q < 3; ++q) a[q] = wave_sum(a[q]);
        if (threadIdx.x == 0) { tot[0] = a[0]; tot[1] = a[1]; tot[2] = a[2]; }
    }
    __syncthreads();
    const double C = tot[0], Sm = tot[1], S2 = tot[2];
    const double mean = C > 0 ? Sm / C : 0.0;
    const double var = C > 1 ? (S2 - Sm * mean) / (C - 1) : 0.0;
    const double thr = mean + 0.15 * sqrt(var > 0 ? var : 0.0);
    if (nbad && blockIdx.x == 0 && threadIdx.x == 0) *nbad = Nx * Ny - (int)C;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nx * Ny) return;
    const bool b = is_bad(it[i], rule, C, thr);
    bad[i] = b ? 1 : 0;
    if (!b) return;
    const int x = i % Nx, y = i / Nx;
    int src[4];
    double w[4];
    int q = x - 1;
    while (q >= 0 && is_bad(it[y * Nx + q], rule, C, thr)) --q;
    src[0] = q >= 0 ? y * Nx + q : -1; w[0] = q >= 0 ? 1.0 / (x - q) : 0.0;
    q = x + 1;
    while (q < Nx && is_bad(it[y * Nx + q], rule, C, thr)) ++q;
    src[1] = q < Nx ? y * Nx + q : -1; w[1] = q < Nx ? 1.0 / (q - x) : 0.0;
    q = y - 1;
    while (q >= 0 && is_bad(it[q * Nx + x], rule, C, thr)) --q;
    src[2] = q >= 0 ? q * Nx + x : -1; w[2] = q >= 0 ? 1.0 / (y - q) : 0.0;
    q = y + 1;
    while (q < Ny && is_bad(it[q * Nx + x], rule, C, thr)) ++q;
    src[3] = q < Ny ? q * Nx + x : -1; w[3] = q < Ny ? 1.0 / (q - y) : 0.0;
    double ws = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) ws += w[k];
    for (int c = 0; c < 2; ++c) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (src[k] >= 0) s += w[k] * U[2 * (size_t)src[k] + c];
        U[2 * (size_t)i + c] = ws > 0 ? s / ws : 0.0;
    }
    if (F)
        for (int c = 0; c < 4; ++c) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (src[k] >= 0) s += w[k] * F[4 * (size_t)src[k] + c];
            F[4 * (size_t)i + c] = ws > 0 ? s / ws : 0.0;
        }
}

// In place (U, F hold the values to refill); Usrc / Fsrc must be null.
hipError_t launch_badfill(const int *iters, int S, int Nx, int Ny, int use_mean_rule,
                          const double *Usrc, const double *Fsrc, double *U, double *F, int *bad,
                          double *part, hipStream_t st, int *nbad)
{
    if (S <= 0) return hipSuccess;
    if (Usrc || Fsrc) return hipErrorInvalidValue;
    int nb = (S + 255) / 256;
    if (nb > 256) nb = 256;
    const int nbf = (S + 255) / 256;
    hipLaunchKernelGGL(k_itstats, dim3(nb), dim3(256), 0, st, iters, S, part);
    hipLaunchKernelGGL(k_markfill, dim3(nbf), dim3(256), 0, st, iters, Nx, Ny,
                       (const double *)part, nb, use_mean_rule, U, F, bad, nbad);
    return hipGetLastError();
}

__global__ void k_iset(int *p, int n, int v)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

hipError_t launch_iset(int *p, int n, int v, hipStream_t st)
{
    hipLaunchKernelGGL(k_iset, dim3(1), dim3(64), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_badmark(const int *iters, int S, int rule, const double *Usrc,
                          const double *Fsrc, double *U, double *F, int *bad, double *part,
                          hipStream_t st, int *nbad)
{
    if (S <= 0) return hipSuccess;
    int nb = (S + 255) / 256;
    if (nb > 256) nb = 256;
    const int nbf = (S + 255) / 256;
    hipLaunchKernelGGL(k_itstats, dim3(nb), dim3(256), 0, st, iters, S, part);
    hipLaunchKernelGGL(k_mark, dim3(nbf), dim3(256), 0, st, iters, S, (const double *)part, nb,
                       rule, Usrc, Fsrc, U, F, bad, nbad);
    return hipGetLastError();
}

}  // namespace aldic
