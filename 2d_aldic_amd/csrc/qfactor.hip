// Correlation quality factors of the integer search (funIntegerSearch.m:126-130,
// 169-177, 211-264): on each subset's full normxcorr2 map (every lag with any overlap,
// (tw - 1 + Dx) x (tw - 1 + Dy) values)
//   cc_min = cc - min(cc);  PCE = max(cc_min)^2 / mean(cc_min^2)
//   PPE    = 1 / entropy of the 30-bin histogram of cc_min (equal bins over
//            [0, max(cc_min)], the last one closed)
// a subset whose template is constant gets [Inf, Inf] (the reference's catch branch);
// then each column becomes (q - min q) / max(q - min q) over all subsets (:174-177).
//
// Only requested with the search's q-factor output (not on the frame path).  One
// workgroup per subset: template and window in LDS, the map in LDS as FP64.  NCC at a
// lag uses the sums over the overlap of template and window: with raw pixels centred by
// Z (t = f - Z, w = g - Z, exact integers in FP64) and d = Z - mean(g),
//   n num = n P - ST SW + d (n ST_in - n_in ST),
//   n var = (n SW2 - SW^2) + 2 d SW (n - n_in) + d^2 n_in (n - n_in),
// the normalised images' NCC (normxcorr2 pads the normalised image with zeros); for
// float (already normalised) images Z = d = 0, which is normxcorr2's formula itself.
#include <algorithm>

#include "aldic_internal.h"
#include "common.h"

namespace aldic {

namespace {

constexpr int kQfThreads = 256;

template <class T> struct QfVal { static constexpr float Z = 0.f; };
template <> struct QfVal<uint8_t> { static constexpr float Z = 128.f; };
template <> struct QfVal<uint16_t> { static constexpr float Z = 32896.f; };

__device__ __forceinline__ double block_reduce(double v, double *red, int op)  // 0 sum 1 min 2 max
{
    for (int m = 32; m >= 1; m >>= 1) {
        const double o = __shfl_xor(v, m, kWave);
        v = op == 0 ? v + o : (op == 1 ? fmin(v, o) : fmax(v, o));
    }
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double r = red[0];
    for (int k = 1; k < kQfThreads / 64; ++k)
        r = op == 0 ? r + red[k] : (op == 1 ? fmin(r, red[k]) : fmax(r, red[k]));
    return r;
}

template <class T>
__global__ __launch_bounds__(kQfThreads) void k_qfactors(const T *f, const T *g, int nx, int x1,
                                                         int y1, int Nx, int step, int ws, int sx,
                                                         int sy, const double *ms, double *qf)
{
    extern __shared__ __align__(16) uint8_t lds8[];
    __shared__ double red[kQfThreads / 64];
    __shared__ int hist[30];
    const int tw = ws + 1, Dx = tw + 2 * sx, Dy = tw + 2 * sy;
    const int Cx = tw - 1 + Dx, Cy = tw - 1 + Dy, nC = Cx * Cy;
    float *Tv = reinterpret_cast<float *>(lds8);
    float *Wv = Tv + tw * tw;
    double *C = reinterpret_cast<double *>(lds8 + ((sizeof(float) * (tw * tw + Dx * Dy) + 15) & ~15));
    const int q = blockIdx.x;
    const int ii = x1 + (q % Nx) * step, jj = y1 + (q / Nx) * step;  // 1-based corner
    const float Z = QfVal<T>::Z;
    const double d = ms ? (double)Z - ms[2] : 0.0;  // Z - mean(g) for raw pixels
    // template f(ii .. ii + w, jj .. jj + w), window g(ii - sx .. ii + w + sx, ...)
    for (int k = threadIdx.x; k < tw * tw; k += kQfThreads) {
        const int j = k / tw, i = k - j * tw;
        Tv[k] = (float)f[(size_t)(jj - 1 + j) * nx + (ii - 1 + i)] - Z;
    }
    for (int k = threadIdx.x; k < Dx * Dy; k += kQfThreads) {
        const int j = k / Dx, i = k - j * Dx;
        Wv[k] = (float)g[(size_t)(jj - 1 - sy + j) * nx + (ii - 1 - sx + i)] - Z;
    }
    if (threadIdx.x < 30) hist[threadIdx.x] = 0;
    __syncthreads();
    double st = 0.0, st2 = 0.0;
    for (int k = threadIdx.x; k < tw * tw; k += kQfThreads) {
        st += Tv[k];
        st2 += (double)Tv[k] * Tv[k];
    }
    st = block_reduce(st, red, 0);
    st2 = block_reduce(st2, red, 0);
    const double n = (double)tw * tw, stn = n * st2 - st * st;
    bool constant = true;
    for (int k = threadIdx.x; k < tw * tw; k += kQfThreads) constant = constant && Tv[k] == Tv[0];
    constant = __syncthreads_and(constant);
    if (constant) {  // normxcorr2 raises: [Inf, Inf] (funIntegerSearch.m:153-157)
        if (threadIdx.x == 0) { qf[2 * q] = INFINITY; qf[2 * q + 1] = INFINITY; }
        return;
    }
    // the full map: lag (ox, oy) = template corner in the window, -(tw-1) .. D-1
    double mn = INFINITY;
    for (int p = threadIdx.x; p < nC; p += kQfThreads) {
        const int oy = p / Cx - (tw - 1), ox = p % Cx - (tw - 1);
        const int j0 = max(0, -oy), j1 = min(tw, Dy - oy), i0 = max(0, -ox), i1 = min(tw, Dx - ox);
        double P = 0.0, sw = 0.0, sw2 = 0.0, stin = 0.0;
        for (int j = j0; j < j1; ++j) {
            const float *tr = Tv + j * tw, *wr = Wv + (oy + j) * Dx + ox;
            for (int i = i0; i < i1; ++i) {
                const double t = tr[i], w = wr[i];
                P = fma(t, w, P);
                sw += w;
                sw2 = fma(w, w, sw2);
                stin += t;
            }
        }
        const double nin = (double)(j1 - j0) * (i1 - i0), nout = n - nin;
        const double num = (n * P - st * sw) + d * (n * stin - nin * st);
        double var = (n * sw2 - sw * sw) + 2.0 * d * sw * nout + d * d * nin * nout;
        var = var > 0.0 ? var : 0.0;
        const double den = sqrt(stn * var);
        const double c = den > 0.0 ? num / den : 0.0;
        C[p] = c;
        mn = fmin(mn, c);
    }
    mn = block_reduce(mn, red, 1);
    double mx = -INFINITY, s2 = 0.0;
    for (int p = threadIdx.x; p < nC; p += kQfThreads) {
        const double c = C[p] - mn;
        mx = fmax(mx, c);
        s2 = fma(c, c, s2);
    }
    mx = block_reduce(mx, red, 2);
    s2 = block_reduce(s2, red, 0);
    for (int p = threadIdx.x; p < nC; p += kQfThreads) {
        int b = mx > 0.0 ? (int)floor((C[p] - mn) / mx * 30.0) : 0;
        b = b < 0 ? 0 : (b > 29 ? 29 : b);
        atomicAdd(&hist[b], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double e = 0.0;
        for (int b = 0; b < 30; ++b)
            if (hist[b]) {
                const double pb = (double)hist[b] / nC;
                e += pb * log(1.0 / pb);
            }
        qf[2 * q] = mx * mx / (s2 / nC);
        qf[2 * q + 1] = 1.0 / e;
    }
}

// (q - min q) / max(q - min q) per column over all S subsets (one workgroup per column)
__global__ __launch_bounds__(kQfThreads) void k_qf_normalize(double *qf, int S)
{
    __shared__ double red[kQfThreads / 64];
    const int col = blockIdx.x;
    double mn = INFINITY;
    for (int k = threadIdx.x; k < S; k += kQfThreads) mn = fmin(mn, qf[2 * k + col]);
    mn = block_reduce(mn, red, 1);
    double mx = -INFINITY;
    for (int k = threadIdx.x; k < S; k += kQfThreads) mx = fmax(mx, qf[2 * k + col] - mn);
    mx = block_reduce(mx, red, 2);
    for (int k = threadIdx.x; k < S; k += kQfThreads) qf[2 * k + col] = (qf[2 * k + col] - mn) / mx;
}

template <class T>
hipError_t launch_t(const SearchArgs &a, const T *f, const T *g, const double *ms, hipStream_t st)
{
    const int tw = a.ws + 1, Dx = tw + 2 * a.sx, Dy = tw + 2 * a.sy;
    const size_t lds = ((sizeof(float) * (tw * tw + Dx * Dy) + 15) & ~(size_t)15) +
                       sizeof(double) * (size_t)(tw - 1 + Dx) * (tw - 1 + Dy);
    if (lds > 150 * 1024) return hipErrorInvalidValue;
    auto kern = k_qfactors<T>;
    (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                              150 * 1024);
    hipLaunchKernelGGL(kern, dim3(a.Nx * a.Ny), dim3(kQfThreads), lds, st, f, g, a.nx, a.x1, a.y1,
                       a.Nx, a.step, a.ws, a.sx, a.sy, ms, a.qf);
    hipLaunchKernelGGL(k_qf_normalize, dim3(2), dim3(kQfThreads), 0, st, a.qf, a.Nx * a.Ny);
    return hipGetLastError();
}

}  // namespace

size_t qfactor_lds_bytes(int ws, int sx, int sy)
{
    const int tw = ws + 1, Dx = tw + 2 * sx, Dy = tw + 2 * sy;
    return ((sizeof(float) * (tw * tw + Dx * Dy) + 15) & ~(size_t)15) +
           sizeof(double) * (size_t)(tw - 1 + Dx) * (tw - 1 + Dy);
}

hipError_t launch_qfactors(const SearchArgs &a, hipStream_t st)
{
    if (a.Nx * a.Ny <= 0 || !a.qf) return hipSuccess;
    if (a.f16 && a.g16 && a.ms) return launch_t(a, a.f16, a.g16, a.ms, st);
    if (a.f8 && a.g8 && a.ms) return launch_t(a, a.f8, a.g8, a.ms, st);
    return launch_t(a, a.f, a.g, (const double *)nullptr, st);
}

}  // namespace aldic
