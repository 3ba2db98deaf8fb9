// Learned initial guess for IC-GN: a siamese correlation network on 4 x 4 pixel cells maps
// a node's reference and deformed patches (the normalised images f, g, 64 x 64 centred on
// the node, plus a 4-pixel context ring: 72 x 72) to the six IC-GN parameters (u_x, v_x,
// u_y, v_y, u, v), in place of the integer search.
//
//   E1  4 x 4 stride 4 patchify of one image, 1 -> 16 channels, bias, ReLU   18 x 18 cells
//   E2  3 x 3 'valid' convolution, 16 -> 16, bias                            16 x 16 cells
//   N   features / (norm + 1e-3) per cell
//   C   quadrant cost volumes over the central 6 x 6 cells of f and displacements of -5..5
//       cells: Q[q][d] = 1/9 sum_{p in q} <nf(p), ng(p + d)>, and their mean G
//   S   soft-argmax of G and the four Q (learned sharpness beta): 10 displacements (px)
//   L   dense 10 -> 6: the scaled outputs (F / 0.05, U / 4)
//
// E1 and E2 are GEMMs on the BF16 matrix cores (v_mfma_f32_16x16x16bf16_1k, FP32
// accumulation) whose operands are gathered on the fly (implicit im2col) and rounded to
// BF16 (round to nearest even) as they are staged in LDS; the rest runs in FP32.
// Inference on a node grid whose patches share one 4-pixel cell lattice (any step that
// is a multiple of 4) computes E1 and E2 once per image over that lattice and reads each
// node's cells from the two feature maps: the 72-pixel input makes those features
// identical to the per-patch ones.  Other node sets, and training, run per patch.
// Training (mean squared error on the scaled outputs, Adam) keeps the patch pairs of its
// set in device memory; its backward GEMMs reuse the same kernel with the transposed
// loaders, split over K when K is long.  Master weights and optimiser moments stay FP32.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <vector>

#include "aldic_internal.h"
#include "common.h"

namespace aldic {

namespace {

constexpr int kC = 16;                  // channels
constexpr int kR = 5, kW = 2 * kR + 1;  // displacement radius (cells)
constexpr int kND = kW * kW;            // displacements per volume
constexpr int kQD = 4 * kND;            // quadrant volumes per node
constexpr int kPatch = 72, kN1 = 18, kN2 = 16, kP0 = 5;
constexpr int kCell1 = kN1 * kN1, kCell2 = kN2 * kN2;
constexpr int kOut = 6, kNF = 10;
constexpr float kEps = 1e-3f;
// flat parameter vector: W1 b1 W2 b2 beta Wl bl, W row-major (N x K)
constexpr long kOffW1 = 0, kOffB1 = kOffW1 + kC * 16;
constexpr long kOffW2 = kOffB1 + kC, kOffB2 = kOffW2 + kC * 9 * kC;
constexpr long kOffBeta = kOffB2 + kC, kOffWl = kOffBeta + 1, kOffBl = kOffWl + kOut * kNF;
constexpr long kNParams = kOffBl + kOut;
__constant__ float c_scale[kOut] = {0.05f, 0.05f, 0.05f, 0.05f, 4.f, 4.f};

// ---- operand loaders: element (row, col) of a GEMM operand -----------------------------
struct LdDense {  // X[r * ld + c]
    const float *p;
    long ld;
    __device__ float operator()(long r, long c) const { return p[r * ld + c]; }
};
struct LdTrans {  // X[c * ld + r]
    const float *p;
    long ld;
    __device__ float operator()(long r, long c) const { return p[c * ld + r]; }
};
// E1 operand from stored pairs (72 x 72 x 2 per sample, NHWC): rows (m, cy, cx) over
// 18 x 18 cells, m < B the f patch of sample idx[m], m >= B the g patch of idx[m - B];
// cols (ky, kx).
struct LdSetCells {
    const float *set;
    const int *idx;
    int B;
    __device__ float operator()(long r, long c) const
    {
        const int m = (int)(r / kCell1), cell = (int)(r % kCell1);
        const int cy = cell / kN1, cx = cell % kN1, ky = (int)c >> 2, kx = (int)c & 3;
        const int ch = m >= B;
        const long s = idx[ch ? m - B : m];
        return set[((s * kPatch + 4 * cy + ky) * kPatch + 4 * cx + kx) * 2 + ch];
    }
};
// 3 x 3 'valid' convolution operand of maps (M maps of H x W cells x kC, NHWC): rows (m,
// oy, ox) over (H - 2) x (W - 2), cols (ky, kx, ci).
struct LdValid {
    const float *x;
    int H, W;
    __device__ float operator()(long r, long c) const
    {
        const int Wo = W - 2;
        const long HWo = (long)(H - 2) * Wo;
        const long m = r / HWo;
        const int o = (int)(r % HWo), oy = o / Wo, ox = o % Wo;
        const int t = (int)c / kC, ci = (int)c % kC, ky = t / 3, kx = t % 3;
        return x[((m * H + oy + ky) * W + ox + kx) * kC + ci];
    }
};

__device__ __forceinline__ unsigned short bf16_rne(float v)
{
    const unsigned u = __float_as_uint(v);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

typedef short cnn_s4 __attribute__((ext_vector_type(4)));
typedef float cnn_f4 __attribute__((ext_vector_type(4)));

constexpr int kGT = 64;   // C tile
constexpr int kGK = 16;   // k step (one MFMA)
constexpr int kGPad = 4;  // LDS row padding (bf16 elements)

// C(i, j) = sum_r A(i, r) B(r, j) for i < M, j < N, r in this workgroup's K range
// [z * kchunk, min(K, (z + 1) kchunk)), z = blockIdx.z; then + bias[j] and ReLU (MODE 1),
// + bias[j] (MODE 2), or plain (MODE 0; with split K each z writes its own slice of C,
// C + z M ldc, summed afterwards).  256 threads, a 64 x 64 tile of C, each wave 32 x 32
// (2 x 2 MFMA tiles of 16 x 16: A with lane l holding (row l % 16, k 4 (l / 16) + e), B
// with lane l holding (k 4 (l / 16) + e, col l % 16), D with lane l, element e holding
// (row 4 (l / 16) + e, col l % 16)).  Operands are staged per k step as BF16 in LDS, A
// by rows and B by columns (so each lane's four k are one 8-byte read).
template <class LA, class LB, int MODE>
__global__ __launch_bounds__(256) void k_cnn_gemm(LA la, LB lb, long M, int N, long K, long kchunk,
                                                  const float *__restrict__ bias,
                                                  float *__restrict__ C, long ldc)
{
    __shared__ unsigned short As[kGT][kGK + kGPad];
    __shared__ unsigned short Bs[kGT][kGK + kGPad];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const long m0 = (long)blockIdx.x * kGT;
    const int n0 = blockIdx.y * kGT;
    const long kb = (long)blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    C += (long)blockIdx.z * M * ldc;
    const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
    cnn_f4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = cnn_f4{0.f, 0.f, 0.f, 0.f};
    // staging: thread t loads row t / 4, k (t % 4) * 4 .. + 3 of both tiles
    const int sr = t >> 2, sk = (t & 3) * 4;
    for (long k0 = kb; k0 < ke; k0 += kGK) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const long k = k0 + sk + e;
            const long gm = m0 + sr;
            const int gn = n0 + sr;
            const float av = (gm < M && k < ke) ? la(gm, k) : 0.f;
            const float bv = (gn < N && k < ke) ? lb(k, gn) : 0.f;
            As[sr][sk + e] = bf16_rne(av);
            Bs[sr][sk + e] = bf16_rne(bv);
        }
        __syncthreads();
        const int li = lane & 15, lk = (lane >> 4) * 4;
        cnn_s4 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
            a[i] = *reinterpret_cast<const cnn_s4 *>(&As[wm + 16 * i + li][lk]);
#pragma unroll
        for (int j = 0; j < 2; ++j)
            b[j] = *reinterpret_cast<const cnn_s4 *>(&Bs[wn + 16 * j + li][lk]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    const int li = lane & 15, lr = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn + 16 * j + li;
            if (col >= N) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const long row = m0 + wm + 16 * i + lr + e;
                if (row >= M) continue;
                float v = acc[i][j][e];
                if (MODE == 1) {
                    v += bias[col];
                    v = v > 0.f ? v : 0.f;
                } else if (MODE == 2) {
                    v += bias[col];
                }
                C[row * ldc + col] = v;
            }
        }
}

// out[i] = sum_z part[z * n + i] in z order (split-K partial products)
__global__ void k_sum_split(const float *__restrict__ part, int nz, long n, float *__restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int z = 0; z < nz; ++z) s += part[z * n + i];
    out[i] = s;
}

int nblk(long n) { return (int)((n + 255) / 256); }

template <int MODE, class LA, class LB>
hipError_t gemm(LA la, LB lb, long M, int N, long K, const float *bias, float *C, long ldc,
                hipStream_t st)
{
    if (M <= 0 || N <= 0) return hipSuccess;
    const dim3 grid((unsigned)((M + kGT - 1) / kGT), (unsigned)((N + kGT - 1) / kGT));
    hipLaunchKernelGGL((k_cnn_gemm<LA, LB, MODE>), grid, dim3(256), 0, st, la, lb, M, N, K, K, bias,
                       C, ldc);
    return hipGetLastError();
}

// Plain product with a long K on a small C (weight gradients): K split over up to 256
// workgroups of at least 1024 steps each, slices summed in a fixed order into C (dense,
// ldc == N).  part must hold splits * M * N floats (kSplitMax * M * N suffices).
constexpr int kSplitMax = 256;
template <class LA, class LB>
hipError_t gemm_splitk(LA la, LB lb, long M, int N, long K, float *C, float *part, hipStream_t st)
{
    if (M <= 0 || N <= 0) return hipSuccess;
    long chunk = std::max<long>(1024, (K + kSplitMax - 1) / kSplitMax);
    chunk = (chunk + kGK - 1) / kGK * kGK;
    const int nz = (int)((K + chunk - 1) / chunk);
    if (nz <= 1) return gemm<0>(la, lb, M, N, K, nullptr, C, N, st);
    const dim3 grid((unsigned)((M + kGT - 1) / kGT), (unsigned)((N + kGT - 1) / kGT), (unsigned)nz);
    hipLaunchKernelGGL((k_cnn_gemm<LA, LB, 0>), grid, dim3(256), 0, st, la, lb, M, N, K, chunk, nullptr,
                       part, (long)N);
    hipLaunchKernelGGL(k_sum_split, dim3(nblk(M * N)), dim3(256), 0, st, part, nz, M * N, C);
    return hipGetLastError();
}

// ---- elementwise and reductions -------------------------------------------------------
// dPre = dOut where the (post-ReLU) output is positive, else 0 (in place)
__global__ void k_relu_bwd(const float *__restrict__ out, float *__restrict__ d, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(out[i] > 0.f)) d[i] = 0.f;
}

// db[j] = sum_i d[i * N + j], one workgroup per column, fixed order
__global__ __launch_bounds__(256) void k_colsum(const float *__restrict__ d, long M, int N,
                                                float *__restrict__ db)
{
    __shared__ float red[4];
    const int j = blockIdx.x;
    float s = 0.f;
    for (long i = threadIdx.x; i < M; i += blockDim.x) s += d[i * N + j];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) db[j] = red[0] + red[1] + red[2] + red[3];
}

// Gradient of the input of a 3 x 3 'valid' convolution (maps of H x W cells x kC) from
// the gradient of its gathered operand dA (rows (m, oy, ox), cols (ky, kx, ci)).
__global__ void k_col2im_valid(const float *__restrict__ dA, float *__restrict__ dX, long nmaps, int H,
                               int W)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nmaps * H * W * kC) return;
    const int ci = (int)(i % kC);
    const long pix = i / kC;
    const int ix = (int)(pix % W), iy = (int)((pix / W) % H);
    const long m = pix / ((long)H * W);
    const int Ho = H - 2, Wo = W - 2;
    float s = 0.f;
#pragma unroll
    for (int ky = 0; ky < 3; ++ky) {
        const int oy = iy - ky;
        if (oy < 0 || oy >= Ho) continue;
#pragma unroll
        for (int kx = 0; kx < 3; ++kx) {
            const int ox = ix - kx;
            if (ox < 0 || ox >= Wo) continue;
            const long row = (m * Ho + oy) * Wo + ox;
            s += dA[row * (9 * kC) + (ky * 3 + kx) * kC + ci];
        }
    }
    dX[i] = s;
}

// n = e / (|e| + eps) per cell of kC features; r = |e| kept for the backward pass
__global__ void k_norm_fwd(const float *__restrict__ e, float *__restrict__ n, float *__restrict__ r,
                           long ncell)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncell) return;
    const float4 *p = reinterpret_cast<const float4 *>(e + i * kC);
    float4 v[4];
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = p[k];
        s += v[k].x * v[k].x + v[k].y * v[k].y + v[k].z * v[k].z + v[k].w * v[k].w;
    }
    const float rr = sqrtf(s), inv = 1.f / (rr + kEps);
    float4 *q = reinterpret_cast<float4 *>(n + i * kC);
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = make_float4(v[k].x * inv, v[k].y * inv, v[k].z * inv, v[k].w * inv);
    if (r) r[i] = rr;
}

// de = dn / (r + eps) - e (e . dn) / (r (r + eps)^2), the second term 0 when r == 0
__global__ void k_norm_bwd(const float *__restrict__ e, const float *__restrict__ r,
                           const float *__restrict__ dn, float *__restrict__ de, long ncell)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncell) return;
    const float rr = r[i], a = 1.f / (rr + kEps);
    float dot = 0.f;
#pragma unroll
    for (int k = 0; k < kC; ++k) dot += e[i * kC + k] * dn[i * kC + k];
    const float b = rr > 0.f ? dot / (rr * (rr + kEps) * (rr + kEps)) : 0.f;
#pragma unroll
    for (int k = 0; k < kC; ++k) de[i * kC + k] = dn[i * kC + k] * a - e[i * kC + k] * b;
}

// Quadrant cost volumes of node b: its 16 x 16 cells start at nf + off_b and ng + off_b
// (per-node offsets, row stride rs cells) or, without offsets, at nf + b * 4096 and
// ng + b * 4096 (rs 16).  One workgroup per node: the f central cells and all g cells in
// LDS (rows padded to 17 floats), one thread per (quadrant, displacement) summing its
// nine cells' dot products; Q[b][q][d] = sum / 9.
__global__ __launch_bounds__(256) void k_corr(const float *__restrict__ nf, const float *__restrict__ ng,
                                              const long *__restrict__ off, int rs,
                                              float *__restrict__ Q)
{
    __shared__ float sf[36][kC + 1];
    __shared__ float sg[kCell2][kC + 1];
    const int b = blockIdx.x, t = threadIdx.x;
    const long o = off ? off[b] : (long)b * kCell2 * kC;
    for (int i = t; i < kCell2 * kC; i += 256) {
        const int cell = i / kC, ch = i % kC, cy = cell / kN2, cx = cell % kN2;
        sg[cell][ch] = ng[o + ((long)cy * rs + cx) * kC + ch];
    }
    for (int i = t; i < 36 * kC; i += 256) {
        const int p = i / kC, ch = i % kC, cy = kP0 + p / 6, cx = kP0 + p % 6;
        sf[p][ch] = nf[o + ((long)cy * rs + cx) * kC + ch];
    }
    __syncthreads();
    for (int task = t; task < kQD; task += 256) {
        const int q = task / kND, d = task % kND, dy = d / kW - kR, dx = d % kW - kR;
        const int qy = q >> 1, qx = q & 1;
        float s = 0.f;
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            const int py = 3 * qy + j / 3, px = 3 * qx + j % 3;
            const int gc = (kP0 + py + dy) * kN2 + kP0 + px + dx;
            const int fc = py * 6 + px;
#pragma unroll
            for (int ch = 0; ch < kC; ++ch) s += sf[fc][ch] * sg[gc][ch];
        }
        Q[(long)b * kQD + task] = s * (1.f / 9.f);
    }
}

// Soft-argmax of the five volumes (G = mean of Q, then Q0..Q3) with sharpness beta and
// the dense head: one wave per node; lanes own displacements d and d + 64.
__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, kWave));
    return v;
}

// Quadrant volume q at displacement d of node b: dense per-node volumes (training) ...
struct QDense {
    const float *Q;
    __device__ float operator()(int b, int q, int d) const { return Q[(long)b * kQD + q * kND + d]; }
};
// ... or the lattice's cost map: B3[P][d] = sum over the 3 x 3 f cells from P of
// <nf, ng(+ d)>, so quadrant (qy, qx) of the node whose first E2 cell is cell0 is B3 at
// cell0 + (kP0 + 3 qy) rows + (kP0 + 3 qx), over 9
struct QLattice {
    const float *B3;
    const long *cell0;
    int Wo;
    __device__ float operator()(int b, int q, int d) const
    {
        const long P = cell0[b] + (long)(kP0 + 3 * (q >> 1)) * Wo + kP0 + 3 * (q & 1);
        return B3[P * kND + d] * (1.f / 9.f);
    }
};

// volume v (0: the mean of the four quadrants, 1..4: quadrant v - 1) at this lane's two
// displacements (lane, lane + 64; 0 past kND)
// the five volumes at this lane's two displacements (lane, lane + 64; 0 past kND):
// a[0] = the mean of the four quadrants, a[1 + q] = quadrant q
template <class QS>
__device__ __forceinline__ void volumes5(const QS &qs, int b, int lane, float a0[5], float a1[5])
{
    const int d1 = lane + 64;
    const bool has1 = d1 < kND;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        a0[1 + q] = qs(b, q, lane);
        a1[1 + q] = has1 ? qs(b, q, d1) : 0.f;
    }
    a0[0] = 0.25f * (a0[1] + a0[2] + a0[3] + a0[4]);
    a1[0] = 0.25f * (a1[1] + a1[2] + a1[3] + a1[4]);
}

// softmax weights of one volume for this lane's two displacements (w1 = 0 past kND)
__device__ __forceinline__ void soft_weights(float a0, float a1, float beta, int lane, float &w0, float &w1)
{
    const bool has1 = lane + 64 < kND;
    const float z0 = beta * a0, z1 = has1 ? beta * a1 : -INFINITY;
    const float mx = wave_max(fmaxf(z0, z1));
    const float e0 = expf(z0 - mx), e1 = has1 ? expf(z1 - mx) : 0.f;
    const float inv = 1.f / wave_sum(e0 + e1);
    w0 = e0 * inv;
    w1 = e1 * inv;
}

template <class QS>
__global__ __launch_bounds__(256) void k_head_fwd(QS qs, int S, const float *__restrict__ w,
                                                  float *__restrict__ feats, float *__restrict__ y)
{
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (b >= S) return;
    const float beta = w[kOffBeta];
    const int d1 = lane + 64;
    const float dx0 = 4.f * (lane % kW - kR), dy0 = 4.f * (lane / kW - kR);
    const float dx1 = 4.f * (d1 % kW - kR), dy1 = 4.f * (d1 / kW - kR);
    float m[kNF], A0[5], A1[5];
    volumes5(qs, b, lane, A0, A1);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const float a0 = A0[v], a1 = A1[v];
        float w0, w1;
        soft_weights(a0, a1, beta, lane, w0, w1);
        m[2 * v] = wave_sum(w0 * dx0 + w1 * dx1);
        m[2 * v + 1] = wave_sum(w0 * dy0 + w1 * dy1);
    }
    if (lane < kNF) {
        float mv = m[0];
#pragma unroll
        for (int k = 1; k < kNF; ++k) mv = lane == k ? m[k] : mv;
        feats[(long)b * kNF + lane] = mv;
    }
    if (lane < kOut) {
        float s = w[kOffBl + lane];
#pragma unroll
        for (int k = 0; k < kNF; ++k) s += w[kOffWl + lane * kNF + k] * m[k];
        y[(long)b * kOut + lane] = s;
    }
}

// Backward of the soft-argmax and the head's input: dm = Wl^T dy, then per volume dz =
// w (s - sum w s) with s_d = a_d . dm, dV = beta dz, dbeta_b = sum dz V; dQ = dV(Q) +
// dV(G) / 4.  One wave per node.
__global__ __launch_bounds__(256) void k_head_bwd(const float *__restrict__ Q, int S,
                                                  const float *__restrict__ w,
                                                  const float *__restrict__ dy, float *__restrict__ dQ,
                                                  float *__restrict__ dbeta)
{
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (b >= S) return;
    const QDense qs{Q};
    const float beta = w[kOffBeta];
    float dm[kNF];
#pragma unroll
    for (int k = 0; k < kNF; ++k) {
        float s = 0.f;
#pragma unroll
        for (int o = 0; o < kOut; ++o) s += w[kOffWl + o * kNF + k] * dy[(long)b * kOut + o];
        dm[k] = s;
    }
    const int d1 = lane + 64;
    const bool has1 = d1 < kND;
    const float dx0 = 4.f * (lane % kW - kR), dy0 = 4.f * (lane / kW - kR);
    const float dx1 = 4.f * (d1 % kW - kR), dy1 = 4.f * (d1 / kW - kR);
    float dg0 = 0.f, dg1 = 0.f, db = 0.f;
    float *dQb = dQ + (long)b * kQD;
    float A0[5], A1[5];
    volumes5(qs, b, lane, A0, A1);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const float a0 = A0[v], a1 = A1[v];
        float w0, w1;
        soft_weights(a0, a1, beta, lane, w0, w1);
        const float s0 = dx0 * dm[2 * v] + dy0 * dm[2 * v + 1];
        const float s1 = dx1 * dm[2 * v] + dy1 * dm[2 * v + 1];
        const float sbar = wave_sum(w0 * s0 + w1 * s1);
        const float z0 = w0 * (s0 - sbar), z1 = w1 * (s1 - sbar);
        db += z0 * a0 + (has1 ? z1 * a1 : 0.f);
        if (v == 0) {
            dg0 = beta * z0 * 0.25f;
            dg1 = beta * z1 * 0.25f;
        } else {
            dQb[(v - 1) * kND + lane] = beta * z0 + dg0;
            if (has1) dQb[(v - 1) * kND + d1] = beta * z1 + dg1;
        }
    }
    db = wave_sum(db);
    if (lane == 0) dbeta[b] = db;
}

// Adjoint of k_corr for training (no offsets): dn of node b's f cells (zero outside the
// central 6 x 6) and g cells, at dn + b * 4096 and dn + (B + b) * 4096.
__global__ __launch_bounds__(256) void k_corr_bwd(const float *__restrict__ n, const float *__restrict__ dQ,
                                                  int B, float *__restrict__ dn)
{
    __shared__ float sf[36][kC + 1];
    __shared__ float sg[kCell2][kC + 1];
    __shared__ float sq[kQD];
    const int b = blockIdx.x, t = threadIdx.x;
    const float *nf = n + (long)b * kCell2 * kC, *ng = n + (long)(B + b) * kCell2 * kC;
    for (int i = t; i < kCell2 * kC; i += 256) sg[i / kC][i % kC] = ng[i];
    for (int i = t; i < 36 * kC; i += 256) {
        const int p = i / kC, ch = i % kC, cy = kP0 + p / 6, cx = kP0 + p % 6;
        sf[p][ch] = nf[(cy * kN2 + cx) * kC + ch];
    }
    for (int i = t; i < kQD; i += 256) sq[i] = dQ[(long)b * kQD + i] * (1.f / 9.f);
    __syncthreads();
    float *dnf = dn + (long)b * kCell2 * kC, *dng = dn + (long)(B + b) * kCell2 * kC;
    for (int i = t; i < kCell2 * kC; i += 256) {
        const int cell = i / kC, ch = i % kC, cy = cell / kN2, cx = cell % kN2;
        // f: central cells only
        float s = 0.f;
        const int py = cy - kP0, px = cx - kP0;
        if (py >= 0 && py < 6 && px >= 0 && px < 6) {
            const float *qv = sq + ((py / 3) * 2 + px / 3) * kND;
            for (int d = 0; d < kND; ++d) {
                const int dy = d / kW - kR, dx = d % kW - kR;
                s += qv[d] * sg[(cy + dy) * kN2 + cx + dx][ch];
            }
        }
        dnf[i] = s;
        // g: every central f cell p with cell - p inside the displacement window
        float sgv = 0.f;
        for (int p = 0; p < 36; ++p) {
            const int fy = kP0 + p / 6, fx = kP0 + p % 6;
            const int dy = cy - fy, dx = cx - fx;
            if (dy < -kR || dy > kR || dx < -kR || dx > kR) continue;
            const int q = ((p / 6) / 3) * 2 + (p % 6) / 3;
            sgv += sq[q * kND + (dy + kR) * kW + dx + kR] * sf[p][ch];
        }
        dng[i] = sgv;
    }
}

// Encoder over a cell lattice (inference), one workgroup per 16 x 16 E2 cells of image
// blockIdx.z (0: f, 1: g): the 72 x 72 pixels of its 18 x 18 E1 cells (lattice cell
// (cy, cx) covers pixels px0 + 4 cx .., py0 + 4 cy .., clamped) in LDS; E1 and E2 on the
// BF16 matrix cores with the GEMM path's operand order, K chunks and roundings (E1 one
// 16-wide K step, E2 nine: one per 3 x 3 tap); the E2 outputs normalised per cell as
// k_norm_fwd does, into n[img][Ho][Wo][kC].
__global__ __launch_bounds__(256) void k_feat_lat(const float *__restrict__ f, const float *__restrict__ g,
                                                  int nx, int ny, int px0, int py0, int Ho, int Wo,
                                                  const float *__restrict__ w, float *__restrict__ n)
{
    __shared__ float pix[kPatch][kPatch + 1];
    __shared__ float h1[kCell1][kC + 1];
    __shared__ float e2[kCell2][kC + 1];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int ty0 = blockIdx.y * kN2, tx0 = blockIdx.x * kN2, img = blockIdx.z;
    const float *src = img ? g : f;
    for (int i = t; i < kPatch * kPatch; i += 256) {
        const int r = i / kPatch, c = i % kPatch;
        const int x = min(max(px0 + 4 * tx0 + c, 0), nx - 1), y = min(max(py0 + 4 * ty0 + r, 0), ny - 1);
        pix[r][c] = src[(long)y * nx + x];
    }
    const int li = lane & 15, lk = lane >> 4;
    // E1: rows = the tile's 324 E1 cells in 21 tiles of 16 (past 324: zero), K = (ky, kx)
    cnn_s4 b1f;
    for (int e = 0; e < 4; ++e) b1f[e] = (short)bf16_rne(w[kOffW1 + li * 16 + 4 * lk + e]);
    __syncthreads();
    for (int i = wv; i < (kCell1 + 15) / 16; i += 4) {
        const int cell = 16 * i + li, cy = cell / kN1, cx = cell % kN1;
        cnn_s4 a;
        for (int e = 0; e < 4; ++e)
            a[e] = cell < kCell1 ? (short)bf16_rne(pix[4 * cy + lk][4 * cx + e]) : (short)0;
        cnn_f4 acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b1f, cnn_f4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        for (int e = 0; e < 4; ++e) {
            const int oc = 16 * i + 4 * lk + e;
            if (oc < kCell1) h1[oc][li] = fmaxf(acc[e] + w[kOffB1 + li], 0.f);
        }
    }
    __syncthreads();
    // E2: row tile i = E2 row oy, A rows ox = li; K = (tap, ci) in nine 16-wide steps
    cnn_s4 b2f[9];
    for (int tp = 0; tp < 9; ++tp)
        for (int e = 0; e < 4; ++e) b2f[tp][e] = (short)bf16_rne(w[kOffW2 + li * 9 * kC + tp * kC + 4 * lk + e]);
    for (int oy = wv; oy < kN2; oy += 4) {
        cnn_f4 acc = cnn_f4{0.f, 0.f, 0.f, 0.f};
        for (int tp = 0; tp < 9; ++tp) {
            const int ky = tp / 3, kx = tp % 3;
            const float *hr = h1[(oy + ky) * kN1 + li + kx];
            cnn_s4 a;
            for (int e = 0; e < 4; ++e) a[e] = (short)bf16_rne(hr[4 * lk + e]);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b2f[tp], acc, 0, 0, 0);
        }
        for (int e = 0; e < 4; ++e) e2[oy * kN2 + 4 * lk + e][li] = acc[e] + w[kOffB2 + li];
    }
    __syncthreads();
    // normalise: thread t = E2 cell (t / 16, t % 16) of the tile
    const int oy = t >> 4, ox = t & 15, gy = ty0 + oy, gx = tx0 + ox;
    if (gy >= Ho || gx >= Wo) return;
    float v[kC];
    for (int k = 0; k < kC; ++k) v[k] = e2[t][k];
    float s = 0.f;
    for (int k = 0; k < 4; ++k)
        s += v[4 * k] * v[4 * k] + v[4 * k + 1] * v[4 * k + 1] + v[4 * k + 2] * v[4 * k + 2] +
             v[4 * k + 3] * v[4 * k + 3];
    const float inv = 1.f / (sqrtf(s) + kEps);
    float4 *q = reinterpret_cast<float4 *>(n + (((long)img * Ho + gy) * Wo + gx) * kC);
    for (int k = 0; k < 4; ++k) q[k] = make_float4(v[4 * k] * inv, v[4 * k + 1] * inv, v[4 * k + 2] * inv, v[4 * k + 3] * inv);
}

// Cost map of the lattice: B3[P][d] = sum over the 3 x 3 f cells from P of <nf(c),
// ng(c + d)> (cells outside the map count as zero), for every cell P of an 8 x 8 tile per
// workgroup: the tile's 10 x 10 f cells and 20 x 20 g cells in LDS; the 100 x 121 cell
// products register-blocked (a work item = one f-cell row, one y-displacement and five f
// cells: 5 x 11 products from 5 + 15 cells, four channels per LDS read), then the 3 x 3
// box sums.
__global__ __launch_bounds__(256) void k_cost_b3(const float *__restrict__ nf, const float *__restrict__ ng,
                                                 int Ho, int Wo, float *__restrict__ B3)
{
    constexpr int T = 8, TF = T + 2, TG = TF + 2 * kR, XB = 5;
    __shared__ float4 sf[TF * TF][kC / 4];
    __shared__ float4 sg[TG * TG][kC / 4];
    __shared__ float cp[TF * TF][kND];
    const int t = threadIdx.x, py0 = blockIdx.y * T, px0 = blockIdx.x * T;
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int i = t; i < TF * TF * (kC / 4); i += 256) {
        const int c = i >> 2, k = i & 3, y = py0 + c / TF, x = px0 + c % TF;
        sf[c][k] = (y < Ho && x < Wo) ? reinterpret_cast<const float4 *>(nf + ((long)y * Wo + x) * kC)[k] : z4;
    }
    for (int i = t; i < TG * TG * (kC / 4); i += 256) {
        const int c = i >> 2, k = i & 3, y = py0 - kR + c / TG, x = px0 - kR + c % TG;
        sg[c][k] = (y >= 0 && y < Ho && x >= 0 && x < Wo)
                       ? reinterpret_cast<const float4 *>(ng + ((long)y * Wo + x) * kC)[k]
                       : z4;
    }
    __syncthreads();
    // work items: (f row cy 0..9, dy 0..10, half h 0..1) = 220
    if (t < TF * kW * 2) {
        const int h = t & 1, r = t >> 1, cy = r / kW, dyi = r % kW, cx0 = h * XB;
        float acc[XB][kW];
#pragma unroll
        for (int i = 0; i < XB; ++i)
#pragma unroll
            for (int j = 0; j < kW; ++j) acc[i][j] = 0.f;
        const int grow = (cy + dyi) * TG + cx0;   // g cell of (f cell cx0, dx = -5)
#pragma unroll
        for (int k = 0; k < kC / 4; ++k) {
            float4 fv[XB], gv[XB + kW - 1];
#pragma unroll
            for (int i = 0; i < XB; ++i) fv[i] = sf[cy * TF + cx0 + i][k];
#pragma unroll
            for (int j = 0; j < XB + kW - 1; ++j) gv[j] = sg[grow + j][k];
#pragma unroll
            for (int i = 0; i < XB; ++i)
#pragma unroll
                for (int j = 0; j < kW; ++j) {
                    const float4 a = fv[i], b = gv[i + j];
                    acc[i][j] += a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
                }
        }
#pragma unroll
        for (int i = 0; i < XB; ++i)
#pragma unroll
            for (int j = 0; j < kW; ++j) cp[cy * TF + cx0 + i][dyi * kW + j] = acc[i][j];
    }
    __syncthreads();
    for (int i = t; i < T * T * kND; i += 256) {
        const int c = i / kND, d = i % kND, cy = c / T, cx = c % T;
        const int y = py0 + cy, x = px0 + cx;
        if (y >= Ho || x >= Wo) continue;
        float s = 0.f;
#pragma unroll
        for (int j = 0; j < 9; ++j) s += cp[(cy + j / 3) * TF + cx + j % 3][d];
        B3[((long)y * Wo + x) * kND + d] = s;
    }
}

// outputs scaled back; loss and its gradient on the scaled targets
__global__ void k_out_scale(const float *__restrict__ y, int n, double *__restrict__ P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n * kOut) P[i] = (double)y[i] * (double)c_scale[i % kOut];
}

__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ y,
                                             const float *__restrict__ tgt, const int *idx,
                                             int B, float *__restrict__ dy, double *loss)
{
    __shared__ double red[4];
    double s = 0.0;
    const float inv = 2.f / (float)(B * kOut);
    for (int i = threadIdx.x; i < B * kOut; i += blockDim.x) {
        const int b = i / kOut, o = i % kOut;
        const float e = y[i] - tgt[(long)idx[b] * kOut + o];
        dy[i] = e * inv;
        s += (double)e * e;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *loss = (red[0] + red[1] + red[2] + red[3]) / (B * kOut);
}

__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, long n, float lr, float c1, float c2)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi, vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
}

// 72 x 72 pair tensors of nodes (NHWC from f, g), clamped at the image border
__global__ void k_extract(const float *__restrict__ f, const float *__restrict__ g, int nx,
                          int ny, const int *__restrict__ cx0, const int *__restrict__ cy0,
                          int S, float *__restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)S * kPatch * kPatch) return;
    const long p = i / (kPatch * kPatch);
    const int r = (int)(i % (kPatch * kPatch)), y = r / kPatch, x = r % kPatch;
    const int gx = min(max(cx0[p] + x, 0), nx - 1), gy = min(max(cy0[p] + y, 0), ny - 1);
    out[2 * i] = f[(long)gy * nx + gx];
    out[2 * i + 1] = g[(long)gy * nx + gx];
}

// first pixel (0-based) of node i's 72-pixel input: the 64-pixel patch of the 1-based
// node x0 starts at pixel x0 - 32 (0-based x0 - 33), its context ring 4 pixels before
__global__ void k_corner(const double *__restrict__ xy, int S, int *__restrict__ cx0,
                         int *__restrict__ cy0)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    cx0[i] = (int)floor(xy[2 * i]) - 37;
    cy0[i] = (int)floor(xy[2 * i + 1]) - 37;
}

// node i's first E2 cell in the lattice's E2 map (Wo cells per row)
__global__ void k_node_off(const double *__restrict__ xy, int S, int px0, int py0, int Wo,
                           long *__restrict__ cell0)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    const int gx = ((int)floor(xy[2 * i]) - 37 - px0) >> 2;
    const int gy = ((int)floor(xy[2 * i + 1]) - 37 - py0) >> 2;
    cell0[i] = (long)gy * Wo + gx;
}

__global__ void k_scale_tgt(const double *__restrict__ P, int n, float *__restrict__ t)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n * kOut) t[i] = (float)(P[i] / (double)c_scale[i % kOut]);
}

__global__ void k_iota(int *__restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

}  // namespace

// ---- network passes -------------------------------------------------------------------
struct CnnWork {
    DBuf w, m, v, grad;                     // parameters, Adam moments, gradient
    DBuf h1, e, r, n, Q, feats, y;          // activations of a batch
    DBuf dy, dQ, dbeta, dn, de, dcol, dh1;  // backward
    DBuf part;                              // split-K partial products
    DBuf set, tgt;                          // training set: pair tensors, scaled targets
    long n_set = 0, cap_set = 0;
    DBuf corner, idx, loss, off, pset, iota, b3;
    long adam_t = 0;
    bool have_w = false;
    ~CnnWork()
    {
        for (DBuf *b : {&w, &m, &v, &grad, &h1, &e, &r, &n, &Q, &feats, &y, &dy, &dQ, &dbeta, &dn, &de, &dcol,
                        &dh1, &part, &set, &tgt, &corner, &idx, &loss, &off, &pset, &iota, &b3})
            b->release();
    }
};

CnnWork *cnn_work_new() { return new (std::nothrow) CnnWork(); }
void cnn_work_free(CnnWork *c) { delete c; }
long cnn_nparams() { return kNParams; }

// activations for B pairs on the per-patch path
static hipError_t ensure_acts(CnnWork &c, long B)
{
    hipError_t e;
    const long m = 2 * B;
    if ((e = c.h1.ensure(sizeof(float) * m * kCell1 * kC)) || (e = c.e.ensure(sizeof(float) * m * kCell2 * kC)) ||
        (e = c.r.ensure(sizeof(float) * m * kCell2)) || (e = c.n.ensure(sizeof(float) * m * kCell2 * kC)) ||
        (e = c.Q.ensure(sizeof(float) * B * kQD)) || (e = c.feats.ensure(sizeof(float) * B * kNF)) ||
        (e = c.y.ensure(sizeof(float) * B * kOut)))
        return e;
    return hipSuccess;
}

// forward of B pairs of the set tensor `set` (rows idx): outputs in c.y, features kept
static hipError_t forward_set(CnnWork &c, const float *set, const int *idx, long B, hipStream_t st)
{
    hipError_t e;
    const float *w = c.w.as<float>();
    const long m = 2 * B;
    LdSetCells l1{set, idx, (int)B};
    if ((e = gemm<1>(l1, LdTrans{w + kOffW1, 16}, m * kCell1, kC, 16, w + kOffB1, c.h1.as<float>(), kC, st)))
        return e;
    if ((e = gemm<2>(LdValid{c.h1.as<float>(), kN1, kN1}, LdTrans{w + kOffW2, 9 * kC}, m * kCell2, kC, 9 * kC,
                     w + kOffB2, c.e.as<float>(), kC, st)))
        return e;
    hipLaunchKernelGGL(k_norm_fwd, dim3(nblk(m * kCell2)), dim3(256), 0, st, c.e.as<float>(), c.n.as<float>(),
                       c.r.as<float>(), m * kCell2);
    hipLaunchKernelGGL(k_corr, dim3((unsigned)B), dim3(256), 0, st, c.n.as<float>(),
                       c.n.as<float>() + B * kCell2 * kC, (const long *)nullptr, kN2, c.Q.as<float>());
    hipLaunchKernelGGL(k_head_fwd<QDense>, dim3((unsigned)((B + 3) / 4)), dim3(256), 0, st, QDense{c.Q.as<float>()},
                       (int)B, w, c.feats.as<float>(), c.y.as<float>());
    return hipGetLastError();
}

static hipError_t ensure_back(CnnWork &c, long B)
{
    hipError_t e;
    const long m = 2 * B;
    if ((e = c.dy.ensure(sizeof(float) * B * kOut)) || (e = c.dQ.ensure(sizeof(float) * B * kQD)) ||
        (e = c.dbeta.ensure(sizeof(float) * B)) || (e = c.dn.ensure(sizeof(float) * m * kCell2 * kC)) ||
        (e = c.de.ensure(sizeof(float) * m * kCell2 * kC)) ||
        (e = c.dcol.ensure(sizeof(float) * m * kCell2 * 9 * kC)) ||
        (e = c.dh1.ensure(sizeof(float) * m * kCell1 * kC)) ||
        (e = c.part.ensure(sizeof(float) * kSplitMax * kC * 9 * kC)) ||
        (e = c.grad.ensure(sizeof(float) * kNParams)))
        return e;
    return hipSuccess;
}

// gradient of the mean squared error of a batch (the forward pass must have run on it,
// c.dy holding dLoss/dy) into c.grad
static hipError_t backward(CnnWork &c, const float *set, const int *idx, long B, hipStream_t st)
{
    hipError_t e;
    const float *w = c.w.as<float>();
    float *g = c.grad.as<float>();
    const long m = 2 * B;
    // head: dWl = dy^T feats, dbl, then the soft-argmax into dQ and dbeta
    if ((e = gemm<0>(LdTrans{c.dy.as<float>(), kOut}, LdDense{c.feats.as<float>(), kNF}, kOut, kNF, B, nullptr,
                     g + kOffWl, kNF, st)))
        return e;
    hipLaunchKernelGGL(k_colsum, dim3(kOut), dim3(256), 0, st, c.dy.as<float>(), B, kOut, g + kOffBl);
    hipLaunchKernelGGL(k_head_bwd, dim3((unsigned)((B + 3) / 4)), dim3(256), 0, st, c.Q.as<float>(), (int)B, w,
                       c.dy.as<float>(), c.dQ.as<float>(), c.dbeta.as<float>());
    hipLaunchKernelGGL(k_colsum, dim3(1), dim3(256), 0, st, c.dbeta.as<float>(), B, 1, g + kOffBeta);
    // correlation and normalisation
    hipLaunchKernelGGL(k_corr_bwd, dim3((unsigned)B), dim3(256), 0, st, c.n.as<float>(), c.dQ.as<float>(), (int)B,
                       c.dn.as<float>());
    hipLaunchKernelGGL(k_norm_bwd, dim3(nblk(m * kCell2)), dim3(256), 0, st, c.e.as<float>(), c.r.as<float>(),
                       c.dn.as<float>(), c.de.as<float>(), m * kCell2);
    // E2: dW2 = de^T col(h1), db2, dcol = de W2, col2im
    if ((e = gemm_splitk(LdTrans{c.de.as<float>(), kC}, LdValid{c.h1.as<float>(), kN1, kN1}, kC, 9 * kC,
                         m * kCell2, g + kOffW2, c.part.as<float>(), st)))
        return e;
    hipLaunchKernelGGL(k_colsum, dim3(kC), dim3(256), 0, st, c.de.as<float>(), m * kCell2, kC, g + kOffB2);
    if ((e = gemm<0>(LdDense{c.de.as<float>(), kC}, LdDense{w + kOffW2, 9 * kC}, m * kCell2, 9 * kC, kC, nullptr,
                     c.dcol.as<float>(), 9 * kC, st)))
        return e;
    hipLaunchKernelGGL(k_col2im_valid, dim3(nblk(m * kCell1 * kC)), dim3(256), 0, st, c.dcol.as<float>(),
                       c.dh1.as<float>(), m, kN1, kN1);
    // E1: ReLU, dW1 = dh1^T patches, db1
    hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(m * kCell1 * kC)), dim3(256), 0, st, c.h1.as<float>(),
                       c.dh1.as<float>(), m * kCell1 * kC);
    LdSetCells l1{set, idx, (int)B};
    if ((e = gemm_splitk(LdTrans{c.dh1.as<float>(), kC}, l1, kC, 16, m * kCell1, g + kOffW1, c.part.as<float>(),
                         st)))
        return e;
    hipLaunchKernelGGL(k_colsum, dim3(kC), dim3(256), 0, st, c.dh1.as<float>(), m * kCell1, kC, g + kOffB1);
    return hipGetLastError();
}

hipError_t cnn_set_weights(CnnWork &c, const float *w_host, hipStream_t st)
{
    hipError_t e;
    if ((e = c.w.ensure(sizeof(float) * kNParams)) || (e = c.m.ensure(sizeof(float) * kNParams)) ||
        (e = c.v.ensure(sizeof(float) * kNParams)))
        return e;
    if ((e = hipMemcpyAsync(c.w.p, w_host, sizeof(float) * kNParams, hipMemcpyHostToDevice, st))) return e;
    if ((e = hipMemsetAsync(c.m.p, 0, sizeof(float) * kNParams, st))) return e;
    if ((e = hipMemsetAsync(c.v.p, 0, sizeof(float) * kNParams, st))) return e;
    c.adam_t = 0;
    c.have_w = true;
    return hipStreamSynchronize(st);
}

hipError_t cnn_get_weights(CnnWork &c, float *w_host, hipStream_t st)
{
    hipError_t e = hipMemcpyAsync(w_host, c.w.p, sizeof(float) * kNParams, hipMemcpyDeviceToHost, st);
    if (e) return e;
    return hipStreamSynchronize(st);
}

// Whether the nodes' 72-pixel inputs (from floor(x) - 37, floor(y) - 37) share one
// 4-pixel lattice, and its origin and extent (host arithmetic only)
CnnPlan cnn_plan(const double *xy_host, int S)
{
    CnnPlan p;
    if (S <= 0) return p;
    auto fl = [](double v) { return (long)std::floor(v); };
    auto md = [](long v) { return (int)(((v % 4) + 4) % 4); };
    const int ox = md(fl(xy_host[0]) - 37), oy = md(fl(xy_host[1]) - 37);
    long xlo = LONG_MAX, xhi = LONG_MIN, ylo = LONG_MAX, yhi = LONG_MIN;
    for (int i = 0; i < S; ++i) {
        const long sx = fl(xy_host[2 * i]) - 37, sy = fl(xy_host[2 * i + 1]) - 37;
        if (md(sx) != ox || md(sy) != oy) return p;
        xlo = std::min(xlo, sx);
        xhi = std::max(xhi, sx);
        ylo = std::min(ylo, sy);
        yhi = std::max(yhi, sy);
    }
    p.nx1 = (xhi - xlo) / 4 + kN1;
    p.ny1 = (yhi - ylo) / 4 + kN1;
    p.px0 = (int)xlo;
    p.py0 = (int)ylo;
    // the lattice pays off when the maps are not much larger than the per-node cells
    p.lattice = p.nx1 * p.ny1 <= 4L * S * kCell1 + 4096;
    return p;
}

// P (S x 6, device) for nodes xy (device, 2 S) planned by cnn_plan
hipError_t cnn_forward(CnnWork &c, const float *f, const float *g, int nx, int ny, const double *xy,
                       const CnnPlan &plan, int S, double *P, hipStream_t st)
{
    hipError_t e;
    if (S <= 0) return hipSuccess;
    const bool lattice = plan.lattice;
    const long nx1 = plan.nx1, ny1 = plan.ny1;
    const long xlo = plan.px0, ylo = plan.py0;
    if (lattice) {
        const long Wo = nx1 - 2, Ho = ny1 - 2, ncell2 = Wo * Ho;
        const int px0 = (int)xlo, py0 = (int)ylo;
        if ((e = c.n.ensure(sizeof(float) * 2 * ncell2 * kC)) || (e = c.b3.ensure(sizeof(float) * ncell2 * kND)) ||
            (e = c.off.ensure(sizeof(long) * (size_t)S)) || (e = c.feats.ensure(sizeof(float) * (size_t)S * kNF)) ||
            (e = c.y.ensure(sizeof(float) * (size_t)S * kOut)))
            return e;
        const float *w = c.w.as<float>();
        float *nm = c.n.as<float>();
        hipLaunchKernelGGL(k_feat_lat, dim3((unsigned)((Wo + kN2 - 1) / kN2), (unsigned)((Ho + kN2 - 1) / kN2), 2),
                           dim3(256), 0, st, f, g, nx, ny, px0, py0, (int)Ho, (int)Wo, w, nm);
        hipLaunchKernelGGL(k_cost_b3, dim3((unsigned)((Wo + 7) / 8), (unsigned)((Ho + 7) / 8)), dim3(256), 0, st,
                           nm, nm + ncell2 * kC, (int)Ho, (int)Wo, c.b3.as<float>());
        hipLaunchKernelGGL(k_node_off, dim3(nblk(S)), dim3(256), 0, st, xy, S, px0, py0, (int)Wo,
                           c.off.as<long>());
        const QLattice ql{c.b3.as<float>(), c.off.as<long>(), (int)Wo};
        hipLaunchKernelGGL(k_head_fwd<QLattice>, dim3((unsigned)((S + 3) / 4)), dim3(256), 0, st, ql, S, w,
                           c.feats.as<float>(), c.y.as<float>());
        hipLaunchKernelGGL(k_out_scale, dim3(nblk((long)S * kOut)), dim3(256), 0, st, c.y.as<float>(), S, P);
        return hipGetLastError();
    }
    // per patch, in batches
    const long Bmax = std::min<long>(4096, S);
    if ((e = c.corner.ensure(sizeof(int) * 2 * (size_t)S)) ||
        (e = c.pset.ensure(sizeof(float) * Bmax * kPatch * kPatch * 2)) || (e = c.iota.ensure(sizeof(int) * Bmax)) ||
        (e = ensure_acts(c, Bmax)))
        return e;
    int *cx0 = c.corner.as<int>(), *cy0 = cx0 + S;
    hipLaunchKernelGGL(k_corner, dim3(nblk(S)), dim3(256), 0, st, xy, S, cx0, cy0);
    hipLaunchKernelGGL(k_iota, dim3(nblk(Bmax)), dim3(256), 0, st, c.iota.as<int>(), (int)Bmax);
    for (long b0 = 0; b0 < S; b0 += Bmax) {
        const long B = std::min<long>(Bmax, S - b0);
        hipLaunchKernelGGL(k_extract, dim3(nblk(B * kPatch * kPatch)), dim3(256), 0, st, f, g, nx, ny, cx0 + b0,
                           cy0 + b0, (int)B, c.pset.as<float>());
        if ((e = forward_set(c, c.pset.as<float>(), c.iota.as<int>(), B, st))) return e;
        hipLaunchKernelGGL(k_out_scale, dim3(nblk(B * kOut)), dim3(256), 0, st, c.y.as<float>(), (int)B,
                           P + kOut * b0);
    }
    return hipGetLastError();
}

// appends S pair tensors (from f, g at nodes xy, device) and their targets P (device,
// S x 6) to the training set
hipError_t cnn_add_samples(CnnWork &c, const float *f, const float *g, int nx, int ny,
                           const double *xy, const double *P, int S, hipStream_t st)
{
    hipError_t e;
    const long need = c.n_set + S;
    if (need > c.cap_set) {
        const long cap = std::max<long>(need, 2 * c.cap_set);
        DBuf ns, nt;
        if ((e = ns.ensure(sizeof(float) * cap * kPatch * kPatch * 2)) || (e = nt.ensure(sizeof(float) * cap * kOut)))
            return e;
        if (c.n_set > 0) {
            if ((e = hipMemcpyAsync(ns.p, c.set.p, sizeof(float) * c.n_set * kPatch * kPatch * 2,
                                    hipMemcpyDeviceToDevice, st)) ||
                (e = hipMemcpyAsync(nt.p, c.tgt.p, sizeof(float) * c.n_set * kOut, hipMemcpyDeviceToDevice, st)))
                return e;
        }
        if ((e = hipStreamSynchronize(st))) return e;
        std::swap(c.set, ns);
        std::swap(c.tgt, nt);
        ns.release();
        nt.release();
        c.cap_set = cap;
    }
    if ((e = c.corner.ensure(sizeof(int) * 2 * (size_t)S))) return e;
    int *cx0 = c.corner.as<int>(), *cy0 = cx0 + S;
    hipLaunchKernelGGL(k_corner, dim3(nblk(S)), dim3(256), 0, st, xy, S, cx0, cy0);
    hipLaunchKernelGGL(k_extract, dim3(nblk((long)S * kPatch * kPatch)), dim3(256), 0, st, f, g, nx, ny, cx0,
                       cy0, S, c.set.as<float>() + c.n_set * kPatch * kPatch * 2);
    hipLaunchKernelGGL(k_scale_tgt, dim3(nblk((long)S * kOut)), dim3(256), 0, st, P, S,
                       c.tgt.as<float>() + c.n_set * kOut);
    c.n_set = need;
    return hipGetLastError();
}

void cnn_clear_samples(CnnWork &c) { c.n_set = 0; }
long cnn_n_samples(const CnnWork &c) { return c.n_set; }

// one batch (rows idx_host of the training set): loss into *loss_dev, gradient into c.grad;
// then, when lr > 0, one Adam step
hipError_t cnn_step(CnnWork &c, const int *idx_host, int B, float lr, double *loss_dev,
                    hipStream_t st)
{
    hipError_t e;
    if ((e = ensure_acts(c, B)) || (e = ensure_back(c, B)) || (e = c.idx.ensure(sizeof(int) * B))) return e;
    if ((e = hipMemcpyAsync(c.idx.p, idx_host, sizeof(int) * B, hipMemcpyHostToDevice, st))) return e;
    if ((e = forward_set(c, c.set.as<float>(), c.idx.as<int>(), B, st))) return e;
    hipLaunchKernelGGL(k_mse, dim3(1), dim3(256), 0, st, c.y.as<float>(), c.tgt.as<float>(), c.idx.as<int>(), B,
                       c.dy.as<float>(), loss_dev);
    if ((e = backward(c, c.set.as<float>(), c.idx.as<int>(), B, st))) return e;
    if (lr > 0.f) {
        ++c.adam_t;
        const float c1 = 1.f / (1.f - std::pow(0.9f, (float)c.adam_t));
        const float c2 = 1.f / (1.f - std::pow(0.999f, (float)c.adam_t));
        hipLaunchKernelGGL(k_adam, dim3(nblk(kNParams)), dim3(256), 0, st, c.w.as<float>(), c.grad.as<float>(),
                           c.m.as<float>(), c.v.as<float>(), kNParams, lr, c1, c2);
    }
    return hipGetLastError();
}

// ---- LocalICGN with the learned start: small helpers -------------------------------------
namespace {
// node centres of the search grid (funIntegerSearch.m: x = (ii + ii + ws) / 2), x fastest
__global__ void k_grid_coords(int x1, int y1, int Nx, int Ny, int step, int ws, double *__restrict__ xy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nx * Ny) return;
    const int ii = x1 + (i % Nx) * step, jj = y1 + (i / Nx) * step;
    xy[2 * i] = (ii + ii + ws) / 2.0;
    xy[2 * i + 1] = (jj + jj + ws) / 2.0;
}
// P (u_x, v_x, u_y, v_y, u, v per node) -> U0 (u, v) and F0 (F11 F21 F12 F22 = u_x v_x u_y v_y)
__global__ void k_split_start(const double *__restrict__ P, int S, double *__restrict__ U0,
                              double *__restrict__ F0)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    for (int k = 0; k < 4; ++k) F0[4 * i + k] = P[6 * i + k];
    U0[2 * i] = P[6 * i + 4];
    U0[2 * i + 1] = P[6 * i + 5];
}
// 1 where IC-GN from the learned start did not converge within max_it residual
// evaluations or ended below min_zncc (or with a non-finite ZNCC)
__global__ void k_flag_fallback(const int *__restrict__ it, const double *__restrict__ z, int S, int max_it,
                                double min_z, int *__restrict__ flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    flag[i] = !(it[i] >= 1 && it[i] <= max_it && z[i] >= min_z);
}
}  // namespace

hipError_t launch_grid_coords(int x1, int y1, int Nx, int Ny, int step, int ws, double *xy, hipStream_t st)
{
    hipLaunchKernelGGL(k_grid_coords, dim3(nblk((long)Nx * Ny)), dim3(256), 0, st, x1, y1, Nx, Ny, step, ws, xy);
    return hipGetLastError();
}

hipError_t launch_split_start(const double *P, int S, double *U0, double *F0, hipStream_t st)
{
    hipLaunchKernelGGL(k_split_start, dim3(nblk(S)), dim3(256), 0, st, P, S, U0, F0);
    return hipGetLastError();
}

hipError_t launch_flag_fallback(const int *it, const double *z, int S, int max_it, double min_z, int *flag,
                                hipStream_t st)
{
    hipLaunchKernelGGL(k_flag_fallback, dim3(nblk(S)), dim3(256), 0, st, it, z, S, max_it, min_z, flag);
    return hipGetLastError();
}

const float *cnn_grad(const CnnWork &c) { return c.grad.as<float>(); }
bool cnn_have_weights(const CnnWork &c) { return c.have_w; }
const float *cnn_outputs(const CnnWork &c) { return c.y.as<float>(); }

}  // namespace aldic
