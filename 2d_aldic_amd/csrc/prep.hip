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

// Accumulator type of the ROI sums: exact integers for integer pixels.
template <class T> struct Acc { using type = double; };
template <> struct Acc<uint8_t> { using type = unsigned long long; };
template <> struct Acc<uint16_t> { using type = unsigned long long; };

// Per-block partial sums over the ROI of image blockIdx.y (raw0, then raw1):
// part[2b] = sum, part[2b+1] = sum of squares, image 1's partials after image 0's.
// Integer pixels accumulate in 64-bit integers (exact); block b takes rows y1 + b,
// y1 + b + gridDim.x, ...
template <class T>
__global__ __launch_bounds__(256) void k_stats(const T *raw0, const T *raw1, int nx, int x1,
                                               int x2, int y1, int y2, double *part)
{
    using A = typename Acc<T>::type;
    const T *raw = blockIdx.y ? raw1 : raw0;
    part += 2 * (size_t)gridDim.x * blockIdx.y;
    A s = 0, s2 = 0;
    for (int y = y1 - 1 + blockIdx.x; y < y2; y += gridDim.x) {
        const T *row = raw + (size_t)y * nx;
        for (int x = x1 - 1 + threadIdx.x; x < x2; x += blockDim.x) {
            const A v = (A)row[x];
            s += v;
            s2 += v * v;
        }
    }
    double ds = (double)s, ds2 = (double)s2;  // per-thread partials: exact below 2^53
    ds = wave_sum(ds);
    ds2 = wave_sum(ds2);
    __shared__ double red[2][4];
    const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (lane == 0) { red[0][wid] = ds; red[1][wid] = ds2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        part[2 * blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}

// Reduces both images' partials (fixed order) to ms = {mean_f, 1/std_f, mean_g, 1/std_g}.
__global__ __launch_bounds__(256) void k_stats_final(const double *part, int np, double n,
                                                     double *ms)
{
    __shared__ double red[4][4];
    double a[4] = {0, 0, 0, 0};
    for (int k = threadIdx.x; k < np; k += blockDim.x) {
        a[0] += part[2 * k]; a[1] += part[2 * k + 1];
        a[2] += part[2 * np + 2 * k]; a[3] += part[2 * np + 2 * k + 1];
    }
    const int wid = threadIdx.x / kWave;
    for (int q = 0; q < 4; ++q) {
        a[q] = wave_sum(a[q]);
        if ((threadIdx.x & 63) == 0) red[q][wid] = a[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int im = 0; im < 2; ++im) {
            const double sum = red[2 * im][0] + red[2 * im][1] + red[2 * im][2] + red[2 * im][3];
            const double sum2 =
                red[2 * im + 1][0] + red[2 * im + 1][1] + red[2 * im + 1][2] + red[2 * im + 1][3];
            const double mean = sum / n;
            double ss = sum2 - sum * mean;  // sum (x - mean)^2
            if (ss < 0) ss = 0;
            ms[2 * im] = mean;
            ms[2 * im + 1] = 1.0 / sqrt(ss / (n - 1.0));
        }
    }
}

// One thread per pixel.  f, g normalised ((I - mean) / std, FP64 then stored FP32).  The
// gradient is formed in FP32 from raw differences (exact for integer input, so flat
// regions give exact zeros like the reference's FP64 filter) and scaled by 1/std.
template <class T>
__global__ __launch_bounds__(256) void k_prep(const T *rf, const T *rg, int nx, int ny,
                                              const uint8_t *mask, const double *ms, float *f,
                                              float *g, float *gx, float *gy)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;  // 0-based
    const int y = blockIdx.y;
    if (x >= nx) return;
    const double mf = ms[0], isf = ms[1], mg = ms[2], isg = ms[3];
    const size_t i = (size_t)y * nx + x;
    f[i] = (float)(((double)rf[i] - mf) * isf);
    g[i] = (float)(((double)rg[i] - mg) * isg);
    float ox = 0.f, oy = 0.f;
    if (x >= 3 && x <= nx - 4 && y >= 3 && y <= ny - 4) {
        const T *p = rf + i;
        const size_t r = (size_t)nx;
        const float dx = 0.75f * ((float)p[1] - (float)p[-1]) - 0.15f * ((float)p[2] - (float)p[-2])
                       + (1.0f / 60.0f) * ((float)p[3] - (float)p[-3]);
        const float dy = 0.75f * ((float)p[r] - (float)p[-(long)r])
                       - 0.15f * ((float)p[2 * r] - (float)p[-2 * (long)r])
                       + (1.0f / 60.0f) * ((float)p[3 * r] - (float)p[-3 * (long)r]);
        const float m = mask ? (float)mask[i] : 1.0f;
        ox = (float)((double)dx * isf) * m;
        oy = (float)((double)dy * isf) * m;
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

template <class T>
static void stats_t(const void *rawf, const void *rawg, int nx, const int *roi, double *part,
                    hipStream_t st)
{
    hipLaunchKernelGGL(k_stats<T>, dim3(kStatParts, 2), dim3(256), 0, st, (const T *)rawf,
                       (const T *)rawg, nx, roi[0], roi[1], roi[2], roi[3], part);
}

hipError_t launch_stats(const void *rawf, const void *rawg, int dtype, int nx, int ny,
                        const int *roi, double *acc, hipStream_t st)
{
    (void)ny;
    switch (dtype) {
    case ALDIC_U8: stats_t<uint8_t>(rawf, rawg, nx, roi, acc, st); break;
    case ALDIC_U16: stats_t<uint16_t>(rawf, rawg, nx, roi, acc, st); break;
    case ALDIC_F32: stats_t<float>(rawf, rawg, nx, roi, acc, st); break;
    default: stats_t<double>(rawf, rawg, nx, roi, acc, st); break;
    }
    return hipGetLastError();
}

template <class T>
static void prep_t(const void *rf, const void *rg, int nx, int ny, const uint8_t *mask,
                   const double *ms, float *f, float *g, float *gx, float *gy, hipStream_t st)
{
    dim3 grid((nx + 255) / 256, ny);
    hipLaunchKernelGGL(k_prep<T>, grid, dim3(256), 0, st, (const T *)rf, (const T *)rg, nx, ny,
                       mask, ms, f, g, gx, gy);
}

// acc: 2*kStatParts partials of f then of g (from launch_stats); ms: 4 doubles out.
hipError_t launch_prep(const void *rawf, const void *rawg, int dtype, int nx, int ny,
                       const uint8_t *mask, const double *acc, double n_roi, double *ms,
                       float *f, float *g, float *gx, float *gy, hipStream_t st)
{
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(256), 0, st, acc, kStatParts, n_roi, ms);
    switch (dtype) {
    case ALDIC_U8: prep_t<uint8_t>(rawf, rawg, nx, ny, mask, ms, f, g, gx, gy, st); break;
    case ALDIC_U16: prep_t<uint16_t>(rawf, rawg, nx, ny, mask, ms, f, g, gx, gy, st); break;
    case ALDIC_F32: prep_t<float>(rawf, rawg, nx, ny, mask, ms, f, g, gx, gy, st); break;
    default: prep_t<double>(rawf, rawg, nx, ny, mask, ms, f, g, gx, gy, st); break;
    }
    return hipGetLastError();
}

hipError_t launch_gradient(const float *f, int nx, int ny, float *gx, float *gy, hipStream_t st)
{
    dim3 grid((nx + 255) / 256, ny);
    hipLaunchKernelGGL(k_gradient, grid, dim3(256), 0, st, f, nx, ny, gx, gy);
    return hipGetLastError();
}

}  // namespace aldic
