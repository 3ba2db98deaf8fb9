// Image pre-pass: per-image ROI statistics, then normalisation fused with the
// 7-tap gradient of the reference image.
//   funNormalizeImg.m:22-40  (I - mean_ROI) / std_ROI, std with N-1
//   funImgGradient.m:53-78   correlation with [-1/60 3/20 -3/4 0 3/4 -3/20 1/60],
//                            valid for 4 <= x <= nx-3 and 4 <= y <= ny-3, times mask
// HBM traffic: the raw pair is read once per pass (stats, then prep) and the four
// FP32 planes are written once: N_px * (2 * 2 * b_in + 16) bytes.
#include "aldic_internal.h"
#include "common.h"

namespace aldic {

template <class T>
__device__ __forceinline__ double rawv(const void *p, size_t i)
{
    return (double)static_cast<const T *>(p)[i];
}

__device__ __forceinline__ double load_raw(const void *p, int dtype, size_t i)
{
    switch (dtype) {
    case ALDIC_U8: return rawv<uint8_t>(p, i);
    case ALDIC_U16: return rawv<uint16_t>(p, i);
    case ALDIC_F32: return rawv<float>(p, i);
    default: return rawv<double>(p, i);
    }
}

// Per-block partial sums over the ROI: part[2b] = sum, part[2b+1] = sum of squares
// (FP64; exact for integer input since every partial stays below 2^53).  Block b takes
// rows y1 + b, y1 + b + kStatParts, ...; the prep kernel reduces the partials.
__global__ __launch_bounds__(256) void k_stats(const void *raw, int dtype, int nx, int x1,
                                               int x2, int y1, int y2, double *part)
{
    double s = 0.0, s2 = 0.0;
    for (int y = y1 - 1 + blockIdx.x; y < y2; y += gridDim.x) {
        const size_t row = (size_t)y * nx;
        for (int x = x1 - 1 + threadIdx.x; x < x2; x += blockDim.x) {
            const double v = load_raw(raw, dtype, row + x);
            s += v;
            s2 += v * v;
        }
    }
    s = wave_sum(s);
    s2 = wave_sum(s2);
    __shared__ double red[2][4];
    const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (lane == 0) { red[0][wid] = s; red[1][wid] = s2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        part[2 * blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}

__device__ __forceinline__ void mean_std(double sum, double sum2, double n, double &mean,
                                         double &sd)
{
    mean = sum / n;
    double ss = sum2 - sum * mean;  // sum (x - mean)^2
    if (ss < 0) ss = 0;
    sd = sqrt(ss / (n - 1.0));
}

// One thread per pixel.  f, g normalised in FP64 then stored FP32.  The gradient is
// formed from raw differences (exact for integer input, so flat regions give exact
// zeros like the reference's FP64 filter) and scaled by 1/std.
__global__ __launch_bounds__(256) void k_prep(const void *rawf, const void *rawg, int dtype,
                                              int nx, int ny, const uint8_t *mask,
                                              const double *acc, double n, float *f, float *g,
                                              float *gx, float *gy)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;  // 0-based
    const int y = blockIdx.y;
    // reduce the per-block partials of both images (same order in every block)
    __shared__ double st4[4];
    if (threadIdx.x < 64) {
        double a0 = 0, a1 = 0, b0 = 0, b1 = 0;
        for (int k = threadIdx.x; k < kStatParts; k += 64) {
            a0 += acc[2 * k]; a1 += acc[2 * k + 1];
            b0 += acc[2 * kStatParts + 2 * k]; b1 += acc[2 * kStatParts + 2 * k + 1];
        }
        a0 = wave_sum(a0); a1 = wave_sum(a1); b0 = wave_sum(b0); b1 = wave_sum(b1);
        if (threadIdx.x == 0) { st4[0] = a0; st4[1] = a1; st4[2] = b0; st4[3] = b1; }
    }
    __syncthreads();
    if (x >= nx) return;
    double mf, sf, mg, sg;
    mean_std(st4[0], st4[1], n, mf, sf);
    mean_std(st4[2], st4[3], n, mg, sg);
    const size_t i = (size_t)y * nx + x;
    f[i] = (float)((load_raw(rawf, dtype, i) - mf) / sf);
    g[i] = (float)((load_raw(rawg, dtype, i) - mg) / sg);
    float ox = 0.f, oy = 0.f;
    if (x >= 3 && x <= nx - 4 && y >= 3 && y <= ny - 4) {
        const double dx = 0.75 * (load_raw(rawf, dtype, i + 1) - load_raw(rawf, dtype, i - 1))
                        - 0.15 * (load_raw(rawf, dtype, i + 2) - load_raw(rawf, dtype, i - 2))
                        + (1.0 / 60.0) * (load_raw(rawf, dtype, i + 3) - load_raw(rawf, dtype, i - 3));
        const size_t r = (size_t)nx;
        const double dy = 0.75 * (load_raw(rawf, dtype, i + r) - load_raw(rawf, dtype, i - r))
                        - 0.15 * (load_raw(rawf, dtype, i + 2 * r) - load_raw(rawf, dtype, i - 2 * r))
                        + (1.0 / 60.0) * (load_raw(rawf, dtype, i + 3 * r) - load_raw(rawf, dtype, i - 3 * r));
        const double m = mask ? (double)mask[i] : 1.0;
        ox = (float)(dx / sf * m);
        oy = (float)(dy / sf * m);
    }
    gx[i] = ox;
    gy[i] = oy;
}

// Gradient of an already normalised FP32 reference (no mask).
__global__ __launch_bounds__(256) void k_gradient(const float *f, int nx, int ny, float *gx,
                                                  float *gy)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= nx) return;
    const size_t i = (size_t)y * nx + x;
    float ox = 0.f, oy = 0.f;
    if (x >= 3 && x <= nx - 4 && y >= 3 && y <= ny - 4) {
        const size_t r = (size_t)nx;
        ox = 0.75f * (f[i + 1] - f[i - 1]) - 0.15f * (f[i + 2] - f[i - 2])
           + (1.0f / 60.0f) * (f[i + 3] - f[i - 3]);
        oy = 0.75f * (f[i + r] - f[i - r]) - 0.15f * (f[i + 2 * r] - f[i - 2 * r])
           + (1.0f / 60.0f) * (f[i + 3 * r] - f[i - 3 * r]);
    }
    gx[i] = ox;
    gy[i] = oy;
}

hipError_t launch_stats(const void *raw, int dtype, int nx, int ny, const int *roi, double *acc,
                        hipStream_t st)
{
    hipLaunchKernelGGL(k_stats, dim3(kStatParts), dim3(256), 0, st, raw, dtype, nx, roi[0],
                       roi[1], roi[2], roi[3], acc);
    return hipGetLastError();
}

hipError_t launch_prep(const void *rawf, const void *rawg, int dtype, int nx, int ny,
                         const uint8_t *mask, const double *acc, double n, float *f, float *g,
                         float *gx, float *gy, hipStream_t st)
{
    dim3 grid((nx + 255) / 256, ny);
    hipLaunchKernelGGL(k_prep, grid, dim3(256), 0, st, rawf, rawg, dtype, nx, ny, mask, acc, n, f,
                       g, gx, gy);
    return hipGetLastError();
}

hipError_t launch_gradient(const float *f, int nx, int ny, float *gx, float *gy, hipStream_t st)
{
    dim3 grid((nx + 255) / 256, ny);
    hipLaunchKernelGGL(k_gradient, grid, dim3(256), 0, st, f, nx, ny, gx, gy);
    return hipGetLastError();
}

}  // namespace aldic
