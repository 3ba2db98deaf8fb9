// Learned initial guess for IC-GN: a small convolutional network maps a subset's 64 x 64
// reference and deformed patches (the normalised images f, g, both centred on the node)
// to the six IC-GN parameters (u_x, v_x, u_y, v_y, u, v), in place of the integer search.
//
//   L1  4 x 4 stride 4 patchify, 2 -> 32 channels, bias, ReLU      16 x 16 x 32
//   L2  3 x 3 stride 2 pad 1,   32 -> 64, bias, ReLU               8 x 8 x 64
//   L3  3 x 3 stride 2 pad 1,   64 -> 64, bias, ReLU               4 x 4 x 64
//   L4  dense 1024 -> 128, bias, ReLU
//   L5  dense 128 -> 6, bias                                        outputs x (0.05 x 4, 16 x 2)
//
// Activations are NHWC FP32 in device memory; every layer is one GEMM on the BF16 matrix
// cores (v_mfma_f32_16x16x16bf16_1k, FP32 accumulation) whose operands are gathered on
// the fly (implicit im2col) and rounded to BF16 (round to nearest even) as they are
// staged in LDS.  Training (mean squared error on the scaled outputs, Adam) runs on a
// device-held set of patch tensors; its backward GEMMs reuse the same kernel with the
// transposed loaders.  Master weights and optimiser moments stay FP32.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "aldic_internal.h"
#include "common.h"

namespace aldic {

namespace {

constexpr int kPatch = 64;
constexpr int kOut = 6;
// layer geometry: (rows per patch, K, N)
constexpr int kR1 = 256, kK1 = 32, kN1 = 32;    // 16 x 16 cells, 4 x 4 x 2 taps
constexpr int kR2 = 64, kK2 = 288, kN2 = 64;    // 8 x 8, 3 x 3 x 32
constexpr int kR3 = 16, kK3 = 576, kN3 = 64;    // 4 x 4, 3 x 3 x 64
constexpr int kK4 = 1024, kN4 = 128;
constexpr int kK5 = 128, kN5 = kOut;
// flat parameter vector: W1 b1 W2 b2 W3 b3 W4 b4 W5 b5, W row-major (N x K)
constexpr long kOffW1 = 0, kOffB1 = kOffW1 + (long)kN1 * kK1;
constexpr long kOffW2 = kOffB1 + kN1, kOffB2 = kOffW2 + (long)kN2 * kK2;
constexpr long kOffW3 = kOffB2 + kN2, kOffB3 = kOffW3 + (long)kN3 * kK3;
constexpr long kOffW4 = kOffB3 + kN3, kOffB4 = kOffW4 + (long)kN4 * kK4;
constexpr long kOffW5 = kOffB4 + kN4, kOffB5 = kOffW5 + (long)kN5 * kK5;
constexpr long kNParams = kOffB5 + kN5;
// output scales: the network works on (F / 0.05, U / 16)
__constant__ float c_scale[kOut] = {0.05f, 0.05f, 0.05f, 0.05f, 16.f, 16.f};

// ---- operand loaders: element (row, col) of a GEMM operand, 0 outside ------------------
// Dense row-major: X[r * ld + c].
struct LdDense {
    const float *p;
    long ld;
    __device__ float operator()(long r, long c) const { return p[r * ld + c]; }
};
// Transposed dense: X[c * ld + r].
struct LdTrans {
    const float *p;
    long ld;
    __device__ float operator()(long r, long c) const { return p[c * ld + r]; }
};
// L1 patchify from the images: row = (patch, cy, cx), col = (ky, kx, channel); the patch
// of node q covers pixels x0 - 32 .. x0 + 31 (1-based node centre x0), clamped into the
// image.
struct LdPatchImg {
    const float *f, *g;
    const int *cx0, *cy0;  // 0-based pixel of the patch's top-left corner per patch
    int nx, ny;
    __device__ float operator()(long r, long c) const
    {
        const long p = r >> 8;
        const int cell = (int)(r & 255), cy = cell >> 4, cx = cell & 15;
        const int ky = (int)c >> 3, kx = ((int)c >> 1) & 3, ch = (int)c & 1;
        int x = cx0[p] + 4 * cx + kx, y = cy0[p] + 4 * cy + ky;
        x = min(max(x, 0), nx - 1);
        y = min(max(y, 0), ny - 1);
        return (ch ? g : f)[(long)y * nx + x];
    }
};
// L1 patchify from stored patch tensors (NHWC 64 x 64 x 2 per patch), rows of the batch
// mapped through idx.
struct LdPatchSet {
    const float *t;
    const int *idx;
    __device__ float operator()(long r, long c) const
    {
        const long p = idx[r >> 8];
        const int cell = (int)(r & 255), cy = cell >> 4, cx = cell & 15;
        const int ky = (int)c >> 3, kx = ((int)c >> 1) & 3, ch = (int)c & 1;
        return t[((p * kPatch + 4 * cy + ky) * kPatch + 4 * cx + kx) * 2 + ch];
    }
};
// 3 x 3 stride 2 pad 1 convolution input (NHWC, H x W x C per patch): row = (patch, oy,
// ox) over Ho x Wo = H/2 x W/2, col = (ky, kx, ci).
template <int H, int C>
struct LdConv {
    const float *x;
    __device__ float operator()(long r, long c) const
    {
        constexpr int Ho = H / 2, HW = Ho * Ho;
        const long p = r / HW;
        const int o = (int)(r % HW), oy = o / Ho, ox = o % Ho;
        const int t = (int)c / C, ci = (int)c % C, ky = t / 3, kx = t % 3;
        const int iy = 2 * oy + ky - 1, ix = 2 * ox + kx - 1;
        if (iy < 0 || iy >= H || ix < 0 || ix >= H) return 0.f;
        return x[((p * H + iy) * H + ix) * C + ci];
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

// C(i, j) = sum_r A(i, r) B(r, j) for i < M, j < N, r < K, with A(i, r) = la(i, r) and
// B(r, j) = lb(r, j); then + bias[j] and ReLU (MODE 1), + bias[j] (MODE 2), or plain.
// 256 threads, a 64 x 64 tile of C, each wave 32 x 32 (2 x 2 MFMA tiles
// of 16 x 16: A 16 x 16 with lane l holding (row l % 16, k 4 (l / 16) + e), B with lane l
// holding (k 4 (l / 16) + e, col l % 16), D with lane l, element e holding (row
// 4 (l / 16) + e, col l % 16)).  Operands are staged per k step as BF16 in LDS, A by rows
// and B by columns (so each lane's four k are one 8-byte read).
template <class LA, class LB, int MODE>
__global__ __launch_bounds__(256) void k_cnn_gemm(LA la, LB lb, long M, int N, long K,
                                                  const float *__restrict__ bias,
                                                  float *__restrict__ C, long ldc)
{
    __shared__ unsigned short As[kGT][kGK + kGPad];
    __shared__ unsigned short Bs[kGT][kGK + kGPad];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const long m0 = (long)blockIdx.x * kGT;
    const int n0 = blockIdx.y * kGT;
    const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
    cnn_f4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = cnn_f4{0.f, 0.f, 0.f, 0.f};
    // staging: thread t loads row t / 4, k (t % 4) * 4 .. + 3 of both tiles
    const int sr = t >> 2, sk = (t & 3) * 4;
    for (long k0 = 0; k0 < K; k0 += kGK) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const long k = k0 + sk + e;
            const long gm = m0 + sr;
            const int gn = n0 + sr;
            const float av = (gm < M && k < K) ? la(gm, k) : 0.f;
            const float bv = (gn < N && k < K) ? lb(k, gn) : 0.f;
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

template <int MODE, class LA, class LB>
hipError_t gemm(LA la, LB lb, long M, int N, long K, const float *bias, float *C, long ldc,
                hipStream_t st)
{
    if (M <= 0 || N <= 0) return hipSuccess;
    const dim3 grid((unsigned)((M + kGT - 1) / kGT), (unsigned)((N + kGT - 1) / kGT));
    hipLaunchKernelGGL((k_cnn_gemm<LA, LB, MODE>), grid, dim3(256), 0, st, la, lb, M, N, K, bias,
                       C, ldc);
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

// Gradient of a 3 x 3 stride 2 pad 1 input (NHWC, H x H x C per patch) from the gradient
// of its gathered operand dA (rows (patch, oy, ox), cols (ky, kx, ci)): each input element
// collects the (at most four) taps that read it.
template <int H, int C>
__global__ void k_col2im(const float *__restrict__ dA, float *__restrict__ dX, long npatch)
{
    constexpr int Ho = H / 2;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npatch * H * H * C) return;
    const int ci = (int)(i % C);
    const long pix = i / C;
    const int ix = (int)(pix % H), iy = (int)((pix / H) % H);
    const long p = pix / (H * H);
    float s = 0.f;
#pragma unroll
    for (int ky = 0; ky < 3; ++ky) {
        const int ty = iy + 1 - ky;
        if (ty < 0 || (ty & 1) || (ty >> 1) >= Ho) continue;
#pragma unroll
        for (int kx = 0; kx < 3; ++kx) {
            const int tx = ix + 1 - kx;
            if (tx < 0 || (tx & 1) || (tx >> 1) >= Ho) continue;
            const long row = (p * Ho + (ty >> 1)) * Ho + (tx >> 1);
            s += dA[row * (9 * C) + (ky * 3 + kx) * C + ci];
        }
    }
    dX[i] = s;
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

// patch tensors of nodes (NHWC 64 x 64 x 2 from f, g), clamped at the image border
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

__global__ void k_corner(const double *__restrict__ xy, int S, int *__restrict__ cx0,
                         int *__restrict__ cy0)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    // 1-based centre x0: the patch's first pixel is x0 - 32, 0-based x0 - 33
    cx0[i] = (int)floor(xy[2 * i]) - kPatch / 2 - 1;
    cy0[i] = (int)floor(xy[2 * i + 1]) - kPatch / 2 - 1;
}

__global__ void k_scale_tgt(const double *__restrict__ P, int n, float *__restrict__ t)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n * kOut) t[i] = (float)(P[i] / (double)c_scale[i % kOut]);
}

int nblk(long n) { return (int)((n + 255) / 256); }

}  // namespace

// ---- network passes -------------------------------------------------------------------
struct CnnWork {
    DBuf w, m, v, grad;              // parameters, Adam moments, gradient
    DBuf a1, a2, a3, a4, y;          // activations (post-ReLU), outputs
    DBuf d, dcol, dx;                // backward scratch
    DBuf set, tgt;                   // training set: patch tensors, scaled targets
    long n_set = 0, cap_set = 0;
    DBuf corner, idx, loss;
    long adam_t = 0;
    bool have_w = false;
    ~CnnWork()
    {
        for (DBuf *b : {&w, &m, &v, &grad, &a1, &a2, &a3, &a4, &y, &d, &dcol, &dx, &set, &tgt, &corner,
                        &idx, &loss})
            b->release();
    }
};

CnnWork *cnn_work_new() { return new (std::nothrow) CnnWork(); }
void cnn_work_free(CnnWork *c) { delete c; }
long cnn_nparams() { return kNParams; }

static hipError_t ensure_acts(CnnWork &c, long B)
{
    hipError_t e;
    if ((e = c.a1.ensure(sizeof(float) * B * kR1 * kN1)) || (e = c.a2.ensure(sizeof(float) * B * kR2 * kN2)) ||
        (e = c.a3.ensure(sizeof(float) * B * kR3 * kN3)) || (e = c.a4.ensure(sizeof(float) * B * kN4)) ||
        (e = c.y.ensure(sizeof(float) * B * kOut)))
        return e;
    return hipSuccess;
}

// forward of B patches whose L1 operand comes from loader l1
template <class L1>
static hipError_t forward(CnnWork &c, L1 l1, long B, hipStream_t st)
{
    hipError_t e;
    const float *w = c.w.as<float>();
    float *a1 = c.a1.as<float>(), *a2 = c.a2.as<float>(), *a3 = c.a3.as<float>(), *a4 = c.a4.as<float>();
    if ((e = gemm<1>(l1, LdTrans{w + kOffW1, kK1}, B * kR1, kN1, kK1, w + kOffB1, a1, kN1, st))) return e;
    if ((e = gemm<1>(LdConv<16, 32>{a1}, LdTrans{w + kOffW2, kK2}, B * kR2, kN2, kK2, w + kOffB2, a2, kN2, st)))
        return e;
    if ((e = gemm<1>(LdConv<8, 64>{a2}, LdTrans{w + kOffW3, kK3}, B * kR3, kN3, kK3, w + kOffB3, a3, kN3, st)))
        return e;
    if ((e = gemm<1>(LdDense{a3, kK4}, LdTrans{w + kOffW4, kK4}, B, kN4, kK4, w + kOffB4, a4, kN4, st))) return e;
    return gemm<2>(LdDense{a4, kK5}, LdTrans{w + kOffW5, kK5}, B, kN5, kK5, w + kOffB5, c.y.as<float>(), kOut, st);
}

// gradient of the mean squared error of a batch (rows idx of the training set) into
// c.grad; the forward pass must have run on the same batch
static hipError_t backward(CnnWork &c, long B, hipStream_t st)
{
    hipError_t e;
    const float *w = c.w.as<float>();
    float *g = c.grad.as<float>();
    float *a1 = c.a1.as<float>(), *a2 = c.a2.as<float>(), *a3 = c.a3.as<float>(), *a4 = c.a4.as<float>();
    float *d = c.d.as<float>(), *dcol = c.dcol.as<float>(), *dx = c.dx.as<float>();
    // L5: d = dy (B x 6)
    const float *dy = c.dx.as<float>();  // k_mse wrote dy into dx
    // dW5 = dy^T a4, db5, da4 = dy W5
    if ((e = gemm<0>(LdTrans{dy, kOut}, LdDense{a4, kK5}, kN5, kK5, B, nullptr, g + kOffW5, kK5, st))) return e;
    hipLaunchKernelGGL(k_colsum, dim3(kN5), dim3(256), 0, st, dy, B, kN5, g + kOffB5);
    if ((e = gemm<0>(LdDense{dy, kOut}, LdDense{w + kOffW5, kK5}, B, kK5, kN5, nullptr, d, kK5, st))) return e;
    // L4: relu, dW4 = d^T a3, db4, da3 = d W4
    hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(B * kN4)), dim3(256), 0, st, a4, d, B * kN4);
    if ((e = gemm<0>(LdTrans{d, kN4}, LdDense{a3, kK4}, kN4, kK4, B, nullptr, g + kOffW4, kK4, st))) return e;
    hipLaunchKernelGGL(k_colsum, dim3(kN4), dim3(256), 0, st, d, B, kN4, g + kOffB4);
    if ((e = gemm<0>(LdDense{d, kN4}, LdDense{w + kOffW4, kK4}, B, kK4, kN4, nullptr, dcol, kK4, st))) return e;
    // L3 (dcol = da3 as B*16 x 64): relu, dW3 = dcol^T col(a2), db3, d = dcol W3, col2im
    hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(B * kR3 * kN3)), dim3(256), 0, st, a3, dcol, B * kR3 * kN3);
    if ((e = gemm<0>(LdTrans{dcol, kN3}, LdConv<8, 64>{a2}, kN3, kK3, B * kR3, nullptr,
                     g + kOffW3, kK3, st)))
        return e;
    hipLaunchKernelGGL(k_colsum, dim3(kN3), dim3(256), 0, st, dcol, B * kR3, kN3, g + kOffB3);
    if ((e = gemm<0>(LdDense{dcol, kN3}, LdDense{w + kOffW3, kK3}, B * kR3, kK3, kN3, nullptr, d, kK3, st)))
        return e;
    hipLaunchKernelGGL((k_col2im<8, 64>), dim3(nblk(B * 8 * 8 * 64)), dim3(256), 0, st, d, dx, B);
    // L2: relu on a2, dW2 = d^T col(a1), db2, dcol2 = d W2, col2im
    hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(B * kR2 * kN2)), dim3(256), 0, st, a2, dx, B * kR2 * kN2);
    if ((e = gemm<0>(LdTrans{dx, kN2}, LdConv<16, 32>{a1}, kN2, kK2, B * kR2, nullptr,
                     g + kOffW2, kK2, st)))
        return e;
    hipLaunchKernelGGL(k_colsum, dim3(kN2), dim3(256), 0, st, dx, B * kR2, kN2, g + kOffB2);
    if ((e = gemm<0>(LdDense{dx, kN2}, LdDense{w + kOffW2, kK2}, B * kR2, kK2, kN2, nullptr, dcol, kK2, st)))
        return e;
    hipLaunchKernelGGL((k_col2im<16, 32>), dim3(nblk(B * 16 * 16 * 32)), dim3(256), 0, st, dcol, d, B);
    // L1: relu on a1, dW1 = d^T patches, db1
    hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(B * kR1 * kN1)), dim3(256), 0, st, a1, d, B * kR1 * kN1);
    LdPatchSet ps{c.set.as<float>(), c.idx.as<int>()};
    if ((e = gemm<0>(LdTrans{d, kN1}, ps, kN1, kK1, B * kR1, nullptr, g + kOffW1, kK1, st)))
        return e;
    hipLaunchKernelGGL(k_colsum, dim3(kN1), dim3(256), 0, st, d, B * kR1, kN1, g + kOffB1);
    return hipGetLastError();
}

static hipError_t ensure_back(CnnWork &c, long B)
{
    hipError_t e;
    const long big = std::max({B * kR1 * kN1, B * kR2 * kK2, B * kR3 * kK3, B * kK4});
    if ((e = c.d.ensure(sizeof(float) * big)) || (e = c.dcol.ensure(sizeof(float) * big)) ||
        (e = c.dx.ensure(sizeof(float) * big)) || (e = c.grad.ensure(sizeof(float) * kNParams)))
        return e;
    return hipSuccess;
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

// P (S x 6, device) for nodes xy (device, 2 S) from the images f, g, in batches
hipError_t cnn_forward(CnnWork &c, const float *f, const float *g, int nx, int ny,
                       const double *xy, int S, double *P, hipStream_t st)
{
    constexpr long kBatch = 8192;
    hipError_t e;
    if ((e = c.corner.ensure(sizeof(int) * 2 * (size_t)S))) return e;
    int *cx0 = c.corner.as<int>(), *cy0 = cx0 + S;
    hipLaunchKernelGGL(k_corner, dim3(nblk(S)), dim3(256), 0, st, xy, S, cx0, cy0);
    if ((e = ensure_acts(c, std::min<long>(kBatch, S)))) return e;
    for (long b0 = 0; b0 < S; b0 += kBatch) {
        const long B = std::min<long>(kBatch, S - b0);
        LdPatchImg l1{f, g, cx0 + b0, cy0 + b0, nx, ny};
        if ((e = forward(c, l1, B, st))) return e;
        hipLaunchKernelGGL(k_out_scale, dim3(nblk(B * kOut)), dim3(256), 0, st, c.y.as<float>(), (int)B,
                           P + kOut * b0);
    }
    return hipGetLastError();
}

// appends S patch tensors (from f, g at nodes xy, device) and their targets P (device,
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
    LdPatchSet l1{c.set.as<float>(), c.idx.as<int>()};
    if ((e = forward(c, l1, B, st))) return e;
    hipLaunchKernelGGL(k_mse, dim3(1), dim3(256), 0, st, c.y.as<float>(), c.tgt.as<float>(), c.idx.as<int>(), B,
                       c.dx.as<float>(), loss_dev);
    if ((e = backward(c, B, st))) return e;
    if (lr > 0.f) {
        ++c.adam_t;
        const float c1 = 1.f / (1.f - std::pow(0.9f, (float)c.adam_t));
        const float c2 = 1.f / (1.f - std::pow(0.999f, (float)c.adam_t));
        hipLaunchKernelGGL(k_adam, dim3(nblk(kNParams)), dim3(256), 0, st, c.w.as<float>(), c.grad.as<float>(),
                           c.m.as<float>(), c.v.as<float>(), kNParams, lr, c1, c2);
    }
    return hipGetLastError();
}

const float *cnn_grad(const CnnWork &c) { return c.grad.as<float>(); }
bool cnn_have_weights(const CnnWork &c) { return c.have_w; }
const float *cnn_outputs(const CnnWork &c) { return c.y.as<float>(); }

}  // namespace aldic
