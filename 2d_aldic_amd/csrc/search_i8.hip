// Whole-field integer search for 8-bit images on the int8 matrix cores.
//
//   funIntegerSearch.m:82-200  subset loop (corner ii, jj; centre ii + w/2)
//   normxcorr2 (called :127)   NCC with zero-padded local sums
//   :134                       crop -> (2s+3)^2 lags, lag -s-1 .. s+1
//   findpeak.m:1-68            argmax (x-lag fastest) and 3x3 fit, 1e-3 rounding
//
// The reference correlates the normalised images (I - mean) / std.  NCC does not
// change under a positive affine map of either image, so every lag whose template
// footprint lies inside the search window is computed from the raw bytes, and the
// crop's edge lags (the footprint leaves the window by one row or column, where
// normxcorr2 pads the normalised image with zeros, i.e. the raw image with its
// mean) get a closed-form correction in the mean.  With t = f - 128, w = g - 128
// (int8, 0 on the padding) and sums over the footprint:
//   n num = n P - ST SW + d (n ST_in - n_in ST),          P = sum t w
//   n var = (n SW2 - SW^2) + 2 d SW (n - n_in) + d^2 n_in (n - n_in)
//   n st  = n ST2 - ST^2,         NCC = n num / sqrt(n st * n var),   d = 128 - mean(g)
// (ST, ST2: template sums; ST_in, n_in over the part inside the window; SW, SW2 box
// sums of w).  P, ST, ST2, SW, SW2 are exact integers (int8 MFMA, int32 sums), so the
// only rounding is the final FP64 arithmetic: the map equals the FP64 reference's up
// to a few ulps, and the constant-template test n st == 0 is exact.
//
// P for all lags of a subset is one 32 x 32 GEMM per wave: rows = x-lag lx, columns =
// y-lag ly, and for template row j the reduction runs over window columns c:
//   P(lx, ly) = sum_j sum_c T(j, c - lx) W(ly + j, c).
// The template row sits in LDS with zero padding, so A(lx, (j, c)) = T(j, c - lx) is
// a contiguous byte run starting at lane lx's own offset; a lane reads the aligned
// dwords around it and funnel-shifts them (v_alignbyte).  B((j, c), ly) = W(ly + j, c)
// is an aligned 16-byte LDS read.  The MFMA's K slots are filled with the same window
// column in A and B, which is all the product needs.  Box sums: a lane per window
// column keeps running vertical sums, a DPP wave scan gives horizontal prefix sums.
//
// One wave per subset; each wave owns its LDS region and loops over subsets (taken in
// contiguous runs per XCD), so the zero padding is written once.
#include "aldic_internal.h"
#include "common.h"
#include "search_peak.h"

namespace aldic {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kI8Waves = 2;  // waves (subsets in flight) per workgroup
constexpr int kPL = 32;      // template row left padding (> largest x-lag index)

struct I8Geom {
    int tw, sx, sy, nrx, nry, Dx, Dy, DH, VW;
    int nk;          // MFMAs per template row (32 window columns each)
    int TS, WS;      // template / window row strides in bytes
    int tgrp, wgrp;  // dwords per template row / window row rewritten per subset
    int off_T, off_W, off_S, off_N, per_wave;  // byte offsets in a wave's LDS region
};

int r16(int v) { return (v + 15) & ~15; }

bool i8_geom(int ws, int sx, int sy, I8Geom &G)
{
    G.tw = ws + 1;
    G.sx = sx;
    G.sy = sy;
    G.nrx = 2 * sx + 3;
    G.nry = 2 * sy + 3;
    G.Dx = G.tw + 2 * sx;
    G.Dy = G.tw + 2 * sy;
    G.DH = G.Dy + 2;
    G.VW = G.Dx + 2;  // = nrx + tw - 1
    if (G.nrx > 32 || G.nry > 32 || G.VW > 64 || G.tw < 2) return false;
    G.nk = (G.VW + 31) / 32;
    // lane run: bytes [kPL - lx + 16 nk h, + 16 nk), read as 4 nk + 1 aligned dwords
    G.TS = kPL + 32 * G.nk + 8;
    G.WS = 32 * G.nk + 16;  // odd multiple of 16: rows of neighbouring lanes spread banks
    G.tgrp = (G.tw + 3) / 4;
    G.wgrp = (G.VW + 3) / 4;
    G.off_T = 0;
    G.off_W = r16(G.off_T + G.tw * G.TS);
    G.off_S = r16(G.off_W + G.DH * G.WS);
    G.off_N = r16(G.off_S + 8 * G.nry * G.nrx);
    G.per_wave = r16(G.off_N + 8 * G.nry * G.nrx);
    return kI8Waves * G.per_wave <= 64 * 1024;
}

// inclusive prefix sum over the 64 lanes (DPP row shifts, then row broadcasts)
__device__ __forceinline__ int wave_scan(int x)
{
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}

__device__ __forceinline__ int wave_sum_i(int v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

__device__ __forceinline__ unsigned pack4(const uint8_t *p, int k0, int k1)
{
    // bytes p[k0 .. k1) shifted to int8 (b ^ 0x80), zero outside
    unsigned v = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (k >= k0 && k < k1) v |= (unsigned)(p[k] ^ 0x80u) << (8 * k);
    return v;
}

__global__ __launch_bounds__(64 * kI8Waves) void k_search_i8(SearchArgs a, I8Geom G)
{
    extern __shared__ __align__(16) uint8_t lds8[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint8_t *base = lds8 + wv * G.per_wave;
    uint8_t *Tl = base + G.off_T;
    uint8_t *Wl = base + G.off_W;
    int *S1 = reinterpret_cast<int *>(base + G.off_S);
    int *S2 = S1 + G.nry * G.nrx;
    double *Nc = reinterpret_cast<double *>(base + G.off_N);
    for (int k = lane; k < G.per_wave / 4; k += 64) reinterpret_cast<unsigned *>(base)[k] = 0u;
    wave_sync_lds();

    const int tw = G.tw, nrx = G.nrx, nry = G.nry, nx = a.nx;
    const long long n = (long long)tw * tw;
    const double d = 128.0 - a.ms[2];
    // MFMA lane roles
    const int lq = lane & 31, h = lane >> 5;
    const int o = kPL - lq + 16 * G.nk * h, b0 = o & ~3, s3 = o & 3;
    const int rowB = min(lq, nry - 1);

    const int S = a.Nx * a.Ny;
    const int nb = gridDim.x, b = blockIdx.x;
    int q0, q1, qstep;
    if (nb >= 8) {
        // subsets in contiguous runs per XCD (workgroups are dealt to the 8 XCDs in turn)
        const int xcd = b & 7, nbx = (nb - xcd + 7) >> 3, per = S >> 3, rm = S & 7;
        const int s0 = xcd * per + min(xcd, rm);
        q0 = s0 + (b >> 3) * kI8Waves + wv;
        q1 = s0 + per + (xcd < rm ? 1 : 0);
        qstep = nbx * kI8Waves;
    } else {
        q0 = b * kI8Waves + wv;
        q1 = S;
        qstep = nb * kI8Waves;
    }
    for (int q = q0; q < q1; q += qstep) {
        const int kx = q % a.Nx, ky = q / a.Nx;
        const int ii = a.x1 + kx * a.step, jj = a.y1 + ky * a.step;  // 1-based corner
        // template rows -> T(j, i) at Tl[j TS + kPL + i], and its integer statistics
        // (|sums| <= 128^2 * 65^2 < 2^31)
        int st = 0, st2 = 0, r0 = 0, rl = 0, c0 = 0, cl = 0, k00 = 0, k0l = 0, kl0 = 0, kll = 0;
        for (int it = lane; it < tw * G.tgrp; it += 64) {
            const int j = it / G.tgrp, g4 = it - j * G.tgrp;
            const uint8_t *src = a.f8 + (size_t)(jj - 1 + j) * nx + (ii - 1) + 4 * g4;
            const int k1 = min(4, tw - 4 * g4);
            const unsigned v = pack4(src, 0, k1);
            *reinterpret_cast<unsigned *>(Tl + j * G.TS + kPL + 4 * g4) = v;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (k >= k1) break;
                const int t = (int)(int8_t)(v >> (8 * k));
                const int i = 4 * g4 + k;
                st += t;
                st2 += t * t;
                if (j == 0) r0 += t;
                if (j == tw - 1) rl += t;
                if (i == 0) c0 += t;
                if (i == tw - 1) cl += t;
                if (j == 0 && i == 0) k00 = t;
                if (j == 0 && i == tw - 1) k0l = t;
                if (j == tw - 1 && i == 0) kl0 = t;
                if (j == tw - 1 && i == tw - 1) kll = t;
            }
        }
        // window rows 1 .. Dy, columns 1 .. Dx (row 0, row Dy + 1, column 0 and columns
        // past Dx stay zero): Wl[r WS + c] = g(ii - sx - 2 + c, jj - sy - 2 + r) - 128
        for (int it = lane; it < G.Dy * G.wgrp; it += 64) {
            const int r = 1 + it / G.wgrp, g4 = it - (r - 1) * G.wgrp;
            const int c = 4 * g4;
            const uint8_t *src =
                a.g8 + (size_t)(jj - G.sy - 2 + r) * nx + (ii - G.sx - 2) + c;
            const int k0 = c == 0 ? 1 : 0, k1 = min(4, G.Dx + 1 - c);
            *reinterpret_cast<unsigned *>(Wl + r * G.WS + c) = pack4(src, k0, k1);
        }
        st = wave_sum_i(st);
        st2 = wave_sum_i(st2);
        r0 = wave_sum_i(r0);
        rl = wave_sum_i(rl);
        c0 = wave_sum_i(c0);
        cl = wave_sum_i(cl);
        k00 = wave_sum_i(k00);
        k0l = wave_sum_i(k0l);
        kl0 = wave_sum_i(kl0);
        kll = wave_sum_i(kll);
        wave_sync_lds();

        const double xc = (ii + ii + a.ws) / 2.0, yc = (jj + jj + a.ws) / 2.0;
        const long long stn = n * st2 - (long long)st * st;
        if (stn == 0) {  // constant template: normxcorr2 raises (funIntegerSearch.m:153-157)
            if (lane == 0) {
                a.x[q] = xc; a.y[q] = yc;
                a.u[q] = a.v[q] = a.ccmax[q] = NAN;
                a.status[q] = -1;
                if (a.qf) { a.qf[2 * q] = INFINITY; a.qf[2 * q + 1] = INFINITY; }
                if (a.U0) { a.U0[2 * q] = NAN; a.U0[2 * q + 1] = NAN; }
            }
            continue;
        }

        // box sums of w and w^2 per lag: vertical running sums (a lane per window
        // column), horizontal sums from a wave prefix scan
        {
            const bool colok = lane < G.VW;
            int v1 = 0, v2 = 0;
            if (colok)
                for (int r = 0; r < tw; ++r) {
                    const int w = (int8_t)Wl[r * G.WS + lane];
                    v1 += w;
                    v2 += w * w;
                }
            for (int ly = 0; ly < nry; ++ly) {
                if (ly > 0 && colok) {
                    const int wi = (int8_t)Wl[(ly + tw - 1) * G.WS + lane];
                    const int wo = (int8_t)Wl[(ly - 1) * G.WS + lane];
                    v1 += wi - wo;
                    v2 += wi * wi - wo * wo;
                }
                const int p1 = wave_scan(v1), p2 = wave_scan(v2);
                const int hi = min(lane + tw - 1, 63), lo = max(lane - 1, 0);
                const int a1 = __shfl(p1, hi, 64), a2 = __shfl(p2, hi, 64);
                const int z1 = __shfl(p1, lo, 64), z2 = __shfl(p2, lo, 64);
                if (lane < nrx) {
                    S1[ly * nrx + lane] = a1 - (lane > 0 ? z1 : 0);
                    S2[ly * nrx + lane] = a2 - (lane > 0 ? z2 : 0);
                }
            }
        }

        // P = sum_j A_j B_j on the matrix cores
        v16i acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = 0;
        for (int j = 0; j < tw; ++j) {
            const unsigned *tq = reinterpret_cast<const unsigned *>(Tl + j * G.TS + b0);
            const uint8_t *wr = Wl + (rowB + j) * G.WS + 16 * G.nk * h;
            if (G.nk == 1) {
                unsigned dw[5];
#pragma unroll
                for (int k = 0; k < 5; ++k) dw[k] = tq[k];
                v4i A, B = *reinterpret_cast<const v4i *>(wr);
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    A[k] = (int)__builtin_amdgcn_alignbyte(dw[k + 1], dw[k], s3);
                acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, B, acc, 0, 0, 0);
            } else {
                unsigned dw[9];
#pragma unroll
                for (int k = 0; k < 9; ++k) dw[k] = tq[k];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    v4i A, B = *reinterpret_cast<const v4i *>(wr + 16 * t);
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        A[k] = (int)__builtin_amdgcn_alignbyte(dw[4 * t + k + 1], dw[4 * t + k], s3);
                    acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, B, acc, 0, 0, 0);
                }
            }
        }
        wave_sync_lds();

        // P of the lags this lane holds (acc[r] is (lx, ly) = (8 (r / 4) + 4 h + r % 4, lq)),
        // exact as doubles, into the NCC map
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int lx = 8 * (r >> 2) + 4 * h + (r & 3);
            if (lx < nrx && lq < nry) Nc[lq * nrx + lx] = (double)acc[r];
        }
        wave_sync_lds();

        // NCC per lag in place, and the argmax (ties -> smallest column-major index,
        // x-lag fastest)
        const int nl = nrx * nry;
        double best = -INFINITY;
        int bi = 0x7fffffff;
        for (int p = lane; p < nl; p += 64) {
            const int ly = p / nrx, lx = p - ly * nrx;
            const long long P = (long long)Nc[p];
            const long long sw = S1[p], sw2 = S2[p];
            const bool top = ly == 0, bot = ly == nry - 1, lft = lx == 0, rgt = lx == nrx - 1;
            const int ro = (top || bot) ? 1 : 0, co = (lft || rgt) ? 1 : 0;
            const int rs = top ? r0 : (bot ? rl : 0);
            const int cs = lft ? c0 : (rgt ? cl : 0);
            const int kc = (ro && co) ? (top ? (lft ? k00 : k0l) : (lft ? kl0 : kll)) : 0;
            const long long stin = st - rs - cs + kc;
            const long long nin = (long long)(tw - ro) * (tw - co);
            const double num =
                (double)(n * P - (long long)st * sw) + d * (double)(n * stin - nin * st);
            double var = (double)(n * sw2 - sw * sw) + 2.0 * d * (double)sw * (double)(n - nin) +
                         d * d * (double)nin * (double)(n - nin);
            var = var > 0.0 ? var : 0.0;
            const double den = sqrt((double)stn * var);
            const double c = den > 0.0 ? num / den : 0.0;
            Nc[p] = c;
            if (c > best) { best = c; bi = p; }
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            const double ob = __shfl_xor(best, m, 64);
            const int oi = __shfl_xor(bi, m, 64);
            if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        wave_sync_lds();
        if (lane == 0) {
            double xp, yp, mf;
            findpeak_fit(Nc, nrx, nry, bi, best, xp, yp, mf);
            const double u = yp - (G.sx + 2), v = xp - (G.sy + 2);
            a.x[q] = xc;
            a.y[q] = yc;
            a.u[q] = u;
            a.v[q] = v;
            a.ccmax[q] = mf;
            a.status[q] = 0;
            if (a.qf) { a.qf[2 * q] = NAN; a.qf[2 * q + 1] = NAN; }
            if (a.U0) { a.U0[2 * q] = u; a.U0[2 * q + 1] = v; }
        }
        wave_sync_lds();
    }
}

}  // namespace

bool search_i8_supported(int ws, int sx, int sy)
{
    I8Geom G;
    return i8_geom(ws, sx, sy, G);
}

hipError_t launch_search_i8(const SearchArgs &a, hipStream_t st)
{
    const int S = a.Nx * a.Ny;
    if (S <= 0) return hipSuccess;
    I8Geom G;
    if (!i8_geom(a.ws, a.sx, a.sy, G) || !a.f8 || !a.g8 || !a.ms) return hipErrorInvalidValue;
    const size_t lds = (size_t)kI8Waves * G.per_wave;
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)k_search_i8,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024);
        attr_set = true;
    }
    // resident workgroups: LDS-limited, and enough to give every wave a few subsets
    const int per_cu = std::max(1, std::min(8, (int)((160 * 1024) / lds)));
    const int nwg = std::min((S + kI8Waves - 1) / kI8Waves, 256 * per_cu);
    hipLaunchKernelGGL(k_search_i8, dim3(nwg), dim3(64 * kI8Waves), lds, st, a, G);
    return hipGetLastError();
}

}  // namespace aldic
