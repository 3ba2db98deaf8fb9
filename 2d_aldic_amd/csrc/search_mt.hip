// Whole-field integer search on the int8 matrix cores for 16-bit images and for search
// radii whose lag map exceeds one 32 x 32 tile (8-bit images with small radii use
// search_i8.hip; float images use search.hip).
//
//   funIntegerSearch.m:82-200  subset loop;  normxcorr2 (called :127);  :134 crop to
//   (2s+3)^2 lags;  findpeak.m:1-68  argmax (x-lag fastest) and 3x3 fit.
//
// Exactness follows search_i8.hip: NCC is computed from the raw pixels centred by Z
// (t = f - Z, w = g - Z) with exact integer sums and the closed-form mean correction on
// the crop's edge lags:
//   n num = n P - ST SW + d (n ST_in - n_in ST),   d = Z - mean(g)
//   n var = (n SW2 - SW^2) + 2 d SW (n - n_in) + d^2 n_in (n - n_in)
//   NCC   = n num / sqrt((n ST2 - ST^2) n var)
// * 16-bit pixels are two int8 digits, v - 32896 = 256 (hi - 128) + (lo - 128), so
//   P = 65536 P_hh + 256 (P_hl + P_lh) + P_ll from four int8 GEMMs (int32 exact).
// * The lag map is covered by 32 x 32 tiles: tile (ox, oy) is the GEMM
//   P(ox + i, oy + k) = sum_j sum_c T(j, c - i) W(oy + k + j, ox + c), whose A operand
//   (the zero-padded template row read at lane i's offset) is the same for every tile.
// * The window box sums SW, SW2 come from planes computed once per frame over the whole
//   image (sliding tw x tw sums, exact int32 / int64); an edge lag's footprint leaves the
//   window by one row or column, whose strip sum (row- and column-strip planes) is
//   subtracted, the shared corner added back.
// * The argmax compares num |num| / den2 by cross-multiplication; afterwards the tile
//   holding the peak's 3 x 3 neighbourhood is recomputed for the exact FP64 fit.
// One wave per workgroup; each wave loops over subsets taken in contiguous runs per XCD.
#include <algorithm>

#include "aldic_internal.h"
#include "common.h"
#include "search_box.h"
#include "search_peak.h"

namespace aldic {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kMtPL = 32;  // template row left padding (> largest lane lag index)

struct MTGeom {
    int tw, sx, sy, nrx, nry, Dx, Dy, DH, VW, ntx, nty, nk;
    int TS, WS;          // template / window plane row strides in bytes
    int tplane, wplane;  // bytes per digit plane
    int tgrp, wgrp;      // destination dwords (4 pixels) per template / window row
    int off_T, off_W, off_P, off_N, total;
};

int r16(int v) { return (v + 15) & ~15; }

__device__ __forceinline__ long long wave_sum_ll(long long v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

bool mt_geom(int ws, int sx, int sy, int dig, MTGeom &G)
{
    G.tw = ws + 1;
    G.sx = sx;
    G.sy = sy;
    G.nrx = 2 * sx + 3;
    G.nry = 2 * sy + 3;
    G.Dx = G.tw + 2 * sx;
    G.Dy = G.tw + 2 * sy;
    G.DH = G.Dy + 2;
    G.VW = G.Dx + 2;
    G.ntx = (G.nrx + 31) / 32;
    G.nty = (G.nry + 31) / 32;
    G.nk = (31 + G.tw + 31) / 32;  // 32-column MFMA steps per template row of one tile
    if (G.tw < 2 || G.nk > 3) return false;
    G.TS = kMtPL + 32 * G.nk + 8;
    // window columns read: a tile origin (a multiple of 32, or of 16 <= nrx - 2 for the
    // peak's neighbourhood) plus 32 nk
    const int wcols = std::max(32 * (G.ntx - 1), 16 * ((G.nrx - 2) / 16)) + 32 * G.nk;
    G.tgrp = (G.tw + 3) / 4;
    G.wgrp = (G.VW + 3) / 4;
    G.WS = r16(std::max(wcols, 4 * G.wgrp));
    if ((G.WS / 16) % 2 == 0) G.WS += 16;  // odd multiple of 16: neighbouring rows spread banks
    G.tplane = G.tw * G.TS;
    G.wplane = G.DH * G.WS;
    G.off_T = 0;
    G.off_W = r16(dig * G.tplane);
    G.off_P = r16(G.off_W + dig * G.wplane);
    G.off_N = G.off_P;  // (no P tile in LDS: lags are evaluated in the MFMA layout)
    G.total = r16(G.off_N + 16 * 8);
    return G.total <= 64 * 1024;
}

// ---- box planes of the (centred) window image --------------------------------------

// Z centres the pixels (the digit planes' offset); bits (nullable) selects zsmall instead
// when every pixel of both images is below 4096 (the 12-bit search's raw planes).
template <class T>
__global__ __launch_bounds__(256) void k_box_rows(const T *g, int nx, int ny, int tw, int Z,
                                                  const unsigned *bits, int zsmall, int *rs1,
                                                  long long *rs2, int pitch)
{
    __shared__ int v[256 + 128];
    if (bits && *bits < 4096u) Z = zsmall;
    const int ye = (int)blockIdx.y - 1;
    const int xb = (int)blockIdx.x * 256 - 1;
    for (int k = threadIdx.x; k < 256 + tw - 1; k += 256) {
        const int xe = xb + k;
        v[k] = (ye >= 0 && ye < ny && xe >= 0 && xe < nx) ? (int)g[(size_t)ye * nx + xe] - Z : 0;
    }
    __syncthreads();
    const int xe = xb + (int)threadIdx.x;
    if (xe > nx) return;
    int s1 = 0;
    long long s2 = 0;
    for (int i = 0; i < tw; ++i) {
        const int w = v[threadIdx.x + i];
        s1 += w;
        s2 += (long long)w * w;
    }
    const size_t o = (size_t)(ye + 1) * pitch + (xe + 1);
    rs1[o] = s1;
    rs2[o] = s2;
}

template <class T>
__global__ __launch_bounds__(256) void k_box_cols(const T *g, int nx, int ny, int tw, int Z,
                                                  const unsigned *bits, int zsmall,
                                                  const int *rs1, const long long *rs2, int *cs1,
                                                  double *cs2, int *bs1, double *bs2, int pitch)
{
    if (bits && *bits < 4096u) Z = zsmall;
    const int xe = (int)(blockIdx.x * 256 + threadIdx.x) - 1;
    if (xe > nx) return;
    const int y0 = (int)blockIdx.y * 64 - 1, y1 = min(y0 + 64, ny + 1);
    const bool xin = xe >= 0 && xe < nx;
    int c1 = 0, b1 = 0;
    long long c2 = 0, b2 = 0;
    for (int j = 0; j < tw; ++j) {
        const int ye = y0 + j;
        const int w = (xin && ye >= 0 && ye < ny) ? (int)g[(size_t)ye * nx + xe] - Z : 0;
        c1 += w;
        c2 += (long long)w * w;
        if (ye <= ny) {
            const size_t o = (size_t)(ye + 1) * pitch + xe + 1;
            b1 += rs1[o];
            b2 += rs2[o];
        }
    }
    for (int ye = y0; ye < y1; ++ye) {
        const size_t o = (size_t)(ye + 1) * pitch + xe + 1;
        cs1[o] = c1;
        cs2[o] = (double)c2;
        bs1[o] = b1;
        bs2[o] = (double)b2;
        // slide down one row: drop ye, take ye + tw
        const int yi = ye + tw;
        const int wo = (xin && ye >= 0 && ye < ny) ? (int)g[(size_t)ye * nx + xe] - Z : 0;
        const int wi = (xin && yi >= 0 && yi < ny) ? (int)g[(size_t)yi * nx + xe] - Z : 0;
        c1 += wi - wo;
        c2 += (long long)wi * wi - (long long)wo * wo;
        b1 -= rs1[o];
        b2 -= rs2[o];
        if (yi <= ny) {
            const size_t oi = (size_t)(yi + 1) * pitch + xe + 1;
            b1 += rs1[oi];
            b2 += rs2[oi];
        }
    }
}

// OR of every pixel of both 16-bit images into *bits (cleared by k_bits_clear first): below
// 4096 means 12-bit data.
__global__ void k_bits_clear(unsigned *bits) { *bits = 0u; }

__global__ __launch_bounds__(256) void k_bits_or(const uint16_t *f, const uint16_t *g, size_t n,
                                                 unsigned *bits)
{
    unsigned acc = 0;
    const size_t n8 = n / 8, stride = (size_t)gridDim.x * blockDim.x;
    const uint4 *f4 = reinterpret_cast<const uint4 *>(f), *g4 = reinterpret_cast<const uint4 *>(g);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n8; i += stride) {
        const uint4 a = f4[i], b = g4[i];
        acc |= a.x | a.y | a.z | a.w | b.x | b.y | b.z | b.w;
    }
    acc = (acc & 0xffffu) | (acc >> 16);  // each dword held two pixels
    if (blockIdx.x == 0)
        for (size_t i = 8 * n8 + threadIdx.x; i < n; i += blockDim.x) acc |= (unsigned)f[i] | g[i];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) acc |= __shfl_xor(acc, m, kWave);
    if ((threadIdx.x & 63) == 0 && acc) atomicOr(bits, acc);
}

// ---- the search kernel --------------------------------------------------------------

// Pixel values of a row segment as centred digit planes: 4 pixels from pixel index pix
// (pixels with k outside [k0, k1) become 0).  DIG 1: one byte plane (v - 128); DIG 2:
// high and low byte planes (hi - 128, lo - 128).
template <int DIG, class T>
__device__ __forceinline__ void load4(const T *img, long long pix, int k0, int k1, unsigned &p0,
                                      unsigned &p1)
{
    if (DIG == 1) {
        const long long sa = pix, lo = sa & ~3LL;
        const unsigned d0 = *reinterpret_cast<const unsigned *>(
            reinterpret_cast<const uint8_t *>(img) + (lo < 0 ? 0 : lo));
        const unsigned d1 =
            *reinterpret_cast<const unsigned *>(reinterpret_cast<const uint8_t *>(img) + lo + 4);
        const unsigned m = (k1 >= 4 ? 0xffffffffu : ((1u << (8 * k1)) - 1u)) &
                           (k0 >= 1 ? 0xffffff00u : 0xffffffffu);
        p0 = (__builtin_amdgcn_alignbyte(d1, d0, (unsigned)(sa & 3)) ^ 0x80808080u) & m;
        p1 = 0;
    } else {
        const long long sa = 2 * pix, lo = sa & ~3LL;
        const uint8_t *b = reinterpret_cast<const uint8_t *>(img);
        const unsigned d0 = *reinterpret_cast<const unsigned *>(b + (lo < 0 ? 0 : lo));
        const unsigned d1 = *reinterpret_cast<const unsigned *>(b + lo + 4);
        const unsigned d2 = *reinterpret_cast<const unsigned *>(b + lo + 8);
        const unsigned s = (unsigned)(sa & 3);
        unsigned w0 = __builtin_amdgcn_alignbyte(d1, d0, s);  // pixels 0 | 1
        unsigned w1 = __builtin_amdgcn_alignbyte(d2, d1, s);  // pixels 2 | 3
        // pixels outside [k0, k1) -> 0x8080, whose digits centre to 0
        w0 = ((k0 <= 0 && k1 >= 1) ? (w0 & 0xffffu) : 0x8080u) |
             ((k0 <= 1 && k1 >= 2) ? (w0 & 0xffff0000u) : 0x80800000u);
        w1 = ((k0 <= 2 && k1 >= 3) ? (w1 & 0xffffu) : 0x8080u) |
             ((k0 <= 3 && k1 >= 4) ? (w1 & 0xffff0000u) : 0x80800000u);
        p0 = __builtin_amdgcn_perm(w1, w0, 0x07050301u) ^ 0x80808080u;  // high bytes
        p1 = __builtin_amdgcn_perm(w1, w0, 0x06040200u) ^ 0x80808080u;  // low bytes
    }
}

// Centred value of pixel k (0..3) of a digit-plane dword pair.
template <int DIG>
__device__ __forceinline__ int pixval(unsigned p0, unsigned p1, int k)
{
    const int hi = (int)(int8_t)(p0 >> (8 * k));
    return DIG == 1 ? hi : 256 * hi + (int)(int8_t)(p1 >> (8 * k));
}

// P of one 32 x 32 lag tile at (ox, oy) (ox a multiple of 16).  acc0 (and for 16-bit
// digits acc1 = hl + lh, acc2 = ll) in the MFMA layout: register r of lane l holds lag
// (ox + 8 (r / 4) + 4 (l / 32) + r % 4, oy + l % 32).
template <int DIG, int NK>
__device__ __forceinline__ void tile_gemm(const uint8_t *Tl, const uint8_t *Wl, const MTGeom &G,
                                          int ox, int oy, int lq, int h, int b0, int s3,
                                          v16i &acc0, v16i &acc1, v16i &acc2)
{
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        acc0[r] = 0;
        acc1[r] = 0;
        acc2[r] = 0;
    }
    const int rowB = min(oy + lq, G.nry - 1);
    const uint8_t *tq0 = Tl + b0;
    const uint8_t *wr0 = Wl + rowB * G.WS + ox + 16 * NK * h;
    for (int j = 0; j < G.tw; ++j) {
        v4i A[DIG][NK], B[DIG][NK];
#pragma unroll
        for (int pl = 0; pl < DIG; ++pl) {
            const unsigned *tq = reinterpret_cast<const unsigned *>(tq0 + pl * G.tplane + j * G.TS);
            const uint8_t *wr = wr0 + pl * G.wplane + j * G.WS;
            unsigned dw[4 * NK + 1];
#pragma unroll
            for (int k = 0; k < 4 * NK + 1; ++k) dw[k] = tq[k];
#pragma unroll
            for (int t = 0; t < NK; ++t) {
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    A[pl][t][k] = (int)__builtin_amdgcn_alignbyte(dw[4 * t + k + 1], dw[4 * t + k], s3);
                B[pl][t] = *reinterpret_cast<const v4i *>(wr + 16 * t);
            }
        }
#pragma unroll
        for (int t = 0; t < NK; ++t) {
            if (DIG == 1) {
                acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[0][t], B[0][t], acc0, 0, 0, 0);
            } else {
                acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[0][t], B[0][t], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[0][t], B[DIG - 1][t], acc1, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[DIG - 1][t], B[0][t], acc1, 0, 0, 0);
                acc2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[DIG - 1][t], B[DIG - 1][t], acc2, 0, 0, 0);
            }
        }
    }
}

// Division-free argmax over one tile's lags straight from the accumulators: lane l's
// register r is lag (ox + 8 (r / 4) + 4 (l / 32) + r % 4, oy + l % 32).  The lane's 16
// box sums are loaded before any is used so their latencies overlap.
template <int DIG, bool WIDE, class T>
__device__ __forceinline__ void tile_argmax(const MTStats &s, const MTGeom &G, const BoxPlanes &bp,
                                            const T *g, int nx, int ny, int Z, int ox, int oy,
                                            int lq, int h, const v16i &acc0, const v16i &acc1,
                                            const v16i &acc2, double &bk, double &bd, int &bi)
{
    const int ly = oy + lq;
    const int lyc = ly < G.nry ? ly : G.nry - 1;
    const size_t rowo = (size_t)(s.y0 + lyc + 1) * bp.pitch + (s.x0 + 1);
    int sw1[16];
    double sw2[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int lx = min(ox + 8 * (r >> 2) + 4 * h + (r & 3), G.nrx - 1);
        sw1[r] = bp.bs1[rowo + lx];
        sw2[r] = bp.bs2[rowo + lx];
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int lx = ox + 8 * (r >> 2) + 4 * h + (r & 3);
        if (lx >= G.nrx || ly >= G.nry) continue;
        const double P = DIG == 1 ? (double)acc0[r]
                                  : fma(65536.0, (double)acc0[r],
                                        fma(256.0, (double)acc1[r], (double)acc2[r]));
        double num, den2;
        mt_terms_sw<WIDE>(s, G, bp, g, nx, ny, Z, lx, ly, P, (double)sw1[r], sw2[r], num, den2);
        double key = num * fabs(num);
        if (!(den2 > 0.0)) { key = 0.0; den2 = 1.0; }
        const int p = ly * G.nrx + lx;
        const double l = key * bd, rr = bk * den2;
        if (l > rr || (l == rr && p < bi)) { bk = key; bd = den2; bi = p; }
    }
}

// skip (nullable): the 12-bit search handles the frame when *skip < 4096.
template <int DIG, int NK, bool WIDE, class T>
__global__ __launch_bounds__(64, 2) void k_search_mt(SearchArgs a, MTGeom G, BoxPlanes bp,
                                                  const T *fimg, const T *gimg,
                                                  const unsigned *skip)
{
    if (skip && *skip < 4096u) return;
    extern __shared__ __align__(16) uint8_t lds8[];
    const int lane = threadIdx.x;
    uint8_t *Tl = lds8 + G.off_T;
    uint8_t *Wl = lds8 + G.off_W;
    double *N9 = reinterpret_cast<double *>(lds8 + G.off_N);
    for (int k = lane; k < G.total / 4; k += 64) reinterpret_cast<unsigned *>(lds8)[k] = 0u;
    wave_sync_lds();

    const int tw = G.tw, nx = a.nx, ny = a.ny;
    const int Z = DIG == 1 ? 128 : 32896;
    MTStats s;
    s.n = (long long)tw * tw;
    s.d = (double)Z - a.ms[2];
    const int lq = lane & 31, h = lane >> 5;
    const int o = kMtPL - lq + 16 * NK * h, b0 = o & ~3, s3 = o & 3;

    const int S = a.Nx * a.Ny;
    const int nb = gridDim.x, b = blockIdx.x;
    int q0, q1, qstep;
    if (nb >= 8) {
        // subsets in contiguous runs per XCD (workgroups are dealt to the 8 XCDs in turn)
        const int xcd = b & 7, nbx = (nb - xcd + 7) >> 3, per = S >> 3, rm = S & 7;
        const int s0 = xcd * per + min(xcd, rm);
        q0 = s0 + (b >> 3);
        q1 = s0 + per + (xcd < rm ? 1 : 0);
        qstep = nbx;
    } else {
        q0 = b;
        q1 = S;
        qstep = nb;
    }
    for (int q = q0; q < q1; q += qstep) {
        const int kx = q % a.Nx, ky = q / a.Nx;
        const int ii = a.x1 + kx * a.step, jj = a.y1 + ky * a.step;  // 1-based corner
        s.x0 = ii - G.sx - 2;
        s.y0 = jj - G.sy - 2;
        // template T(j, i) at plane[j TS + kPL + i] and its integer statistics
        long long st = 0, st2 = 0;
        int r0 = 0, rl = 0, c0 = 0, cl = 0, k00 = 0, k0l = 0, kl0 = 0, kll = 0;
#pragma unroll 2
        for (int it = lane; it < tw * G.tgrp; it += 64) {
            const int j = it / G.tgrp, g4 = it - j * G.tgrp;
            unsigned p0, p1;
            load4<DIG>(fimg, (long long)(jj - 1 + j) * nx + (ii - 1) + 4 * g4, 0,
                       min(4, tw - 4 * g4), p0, p1);
            uint8_t *dst = Tl + j * G.TS + kMtPL + 4 * g4;
            *reinterpret_cast<unsigned *>(dst) = p0;
            if (DIG == 2) *reinterpret_cast<unsigned *>(dst + G.tplane) = p1;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int v = pixval<DIG>(p0, p1, k), i = 4 * g4 + k;
                st += v;
                st2 += (long long)v * v;
                if (j == 0) r0 += v;
                if (j == tw - 1) rl += v;
                if (i == 0) c0 += v;
                if (i == tw - 1) cl += v;
                if (j == 0 && i == 0) k00 = v;
                if (j == 0 && i == tw - 1) k0l = v;
                if (j == tw - 1 && i == 0) kl0 = v;
                if (j == tw - 1 && i == tw - 1) kll = v;
            }
        }
        // window rows 1 .. Dy, columns 1 .. Dx (the ring and the columns past Dx stay 0):
        // window (r, c) = image (x0 + c, y0 + r)
#pragma unroll 4
        for (int it = lane; it < G.Dy * G.wgrp; it += 64) {
            const int r = 1 + it / G.wgrp, g4 = it - (r - 1) * G.wgrp;
            unsigned p0, p1;
            load4<DIG>(gimg, (long long)(s.y0 + r) * nx + s.x0 + 4 * g4, g4 == 0 ? 1 : 0,
                       min(4, G.Dx + 1 - 4 * g4), p0, p1);
            uint8_t *dst = Wl + r * G.WS + 4 * g4;
            *reinterpret_cast<unsigned *>(dst) = p0;
            if (DIG == 2) *reinterpret_cast<unsigned *>(dst + G.wplane) = p1;
        }
        s.st = wave_sum_ll(st);
        s.st2 = wave_sum_ll(st2);
        s.r0 = (int)wave_sum_ll(r0);
        s.rl = (int)wave_sum_ll(rl);
        s.c0 = (int)wave_sum_ll(c0);
        s.cl = (int)wave_sum_ll(cl);
        s.k00 = (int)wave_sum_ll(k00);
        s.k0l = (int)wave_sum_ll(k0l);
        s.kl0 = (int)wave_sum_ll(kl0);
        s.kll = (int)wave_sum_ll(kll);
        s.stn = s.n * s.st2 - s.st * s.st;
        s.nd = (double)s.n;
        s.std = (double)s.st;
        s.stnd = (double)s.stn;
        wave_sync_lds();

        const double xc = (ii + ii + a.ws) / 2.0, yc = (jj + jj + a.ws) / 2.0;
        if (s.stn == 0) {  // constant template: normxcorr2 raises (funIntegerSearch.m:153-157)
            if (lane == 0) {
                a.x[q] = xc; a.y[q] = yc;
                a.u[q] = a.v[q] = a.ccmax[q] = NAN;
                a.status[q] = -1;
                if (a.qf) { a.qf[2 * q] = INFINITY; a.qf[2 * q + 1] = INFINITY; }
                if (a.U0) { a.U0[2 * q] = NAN; a.U0[2 * q + 1] = NAN; }
            }
            continue;
        }

        // every tile: P, then the division-free argmax (ties -> smallest column-major
        // index, x-lag fastest)
        double bk = -INFINITY, bd = 1.0;
        int bi = 0x7fffffff;
        v16i acc0, acc1, acc2;
        for (int ty = 0; ty < G.nty; ++ty)
            for (int tx = 0; tx < G.ntx; ++tx) {
                const int ox = 32 * tx, oy = 32 * ty;
                tile_gemm<DIG, NK>(Tl, Wl, G, ox, oy, lq, h, b0, s3, acc0, acc1, acc2);
                tile_argmax<DIG, WIDE>(s, G, bp, gimg, nx, ny, Z, ox, oy, lq, h, acc0, acc1, acc2, bk,
                                 bd, bi);
            }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            const double ok = __shfl_xor(bk, m, 64), od = __shfl_xor(bd, m, 64);
            const int oi = __shfl_xor(bi, m, 64);
            const double l = ok * bd, rr = bk * od;
            if (l > rr || (l == rr && oi < bi)) { bk = ok; bd = od; bi = oi; }
        }
        // the tile around the peak again, for the exact NCC of its 3 x 3 neighbourhood
        const int bx = bi % G.nrx, by = bi / G.nrx;
        const int ox = (max(bx - 1, 0) / 16) * 16, oy = max(by - 1, 0);
        if (lane < 9) N9[lane] = 0.0;  // neighbours off the map stay 0
        tile_gemm<DIG, NK>(Tl, Wl, G, ox, oy, lq, h, b0, s3, acc0, acc1, acc2);
        wave_sync_lds();
        {
            // u9[(dc + 1) 3 + (dr + 1)]: dc = y-lag offset, dr = x-lag offset; the lanes
            // holding the peak's neighbourhood write it
            const int ly = oy + lq;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int lx = ox + 8 * (r >> 2) + 4 * h + (r & 3);
                const int dr = lx - bx, dc = ly - by;
                if (dr < -1 || dr > 1 || dc < -1 || dc > 1 || lx >= G.nrx || ly >= G.nry) continue;
                const double P = DIG == 1 ? (double)acc0[r]
                                          : fma(65536.0, (double)acc0[r],
                                                fma(256.0, (double)acc1[r], (double)acc2[r]));
                double num, den2;
                mt_terms<WIDE>(s, G, bp, gimg, nx, ny, Z, lx, ly, P, num, den2);
                const double den = sqrt(den2);
                N9[(dc + 1) * 3 + (dr + 1)] = den > 0.0 ? num / den : 0.0;
            }
        }
        wave_sync_lds();
        const double best = N9[4];
        if (lane == 0) {
            const int r = bx + 1, c = by + 1;  // 1-based row = x-lag, column = y-lag
            double xp = c, yp = r, mf = best;
            if (!(c == 1 || c == G.nry || r == 1 || r == G.nrx)) findpeak_fit9(N9, r, c, best, xp, yp, mf);
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

// Work buffer layout: rs2, cs2, bs2 (8-byte), rs1, cs1, bs1 (4-byte), 256 spare bytes (a
// lane may read a few elements past the last row), then the bits word.
size_t planes_bytes(size_t np) { return np * (3 * sizeof(int) + 3 * sizeof(long long)) + 256; }

template <class T>
hipError_t box_planes(const T *g, int nx, int ny, int tw, int Z, const unsigned *bits, int zsmall,
                      DBuf &buf, BoxPlanes &bp, hipStream_t st)
{
    const int pitch = nx + 2;
    const size_t np = (size_t)pitch * (ny + 2);
    hipError_t e;
    if ((e = buf.ensure(planes_bytes(np) + 256)) != hipSuccess) return e;
    long long *l = buf.as<long long>();
    long long *rs2 = l;
    double *cs2 = reinterpret_cast<double *>(l + np), *bs2 = reinterpret_cast<double *>(l + 2 * np);
    int *i = reinterpret_cast<int *>(l + 3 * np);
    int *rs1 = i, *cs1 = i + np, *bs1 = i + 2 * np;
    hipLaunchKernelGGL(k_box_rows<T>, dim3((pitch + 255) / 256, ny + 2), dim3(256), 0, st, g, nx,
                       ny, tw, Z, bits, zsmall, rs1, rs2, pitch);
    hipLaunchKernelGGL(k_box_cols<T>, dim3((pitch + 255) / 256, (ny + 2 + 63) / 64), dim3(256), 0,
                       st, g, nx, ny, tw, Z, bits, zsmall, (const int *)rs1,
                       (const long long *)rs2, cs1, cs2, bs1, bs2, pitch);
    bp.bs1 = bs1;
    bp.rs1 = rs1;
    bp.cs1 = cs1;
    bp.bs2 = bs2;
    bp.rs2 = rs2;
    bp.cs2 = cs2;
    bp.pitch = pitch;
    return hipGetLastError();
}

template <int DIG, int NK, bool WIDE, class T>
hipError_t launch_mt(const SearchArgs &a, const MTGeom &G, const BoxPlanes &bp, const T *f,
                     const T *g, int S, const unsigned *skip, hipStream_t st)
{
    auto kern = k_search_mt<DIG, NK, WIDE, T>;
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  64 * 1024);
        attr_set = true;
    }
    const int per_cu = std::max(1, std::min(16, (160 * 1024) / G.total));
    const int nwg = std::min(S, 256 * per_cu);
    hipLaunchKernelGGL(kern, dim3(nwg), dim3(64), (size_t)G.total, st, a, G, bp, f, g, skip);
    return hipGetLastError();
}

template <int DIG, class T>
hipError_t launch_dig(const SearchArgs &a, const MTGeom &G, const BoxPlanes &bp, const T *f,
                      const T *g, int S, const unsigned *skip, hipStream_t st)
{
    // 16-bit templates from 54 x 54 up: n P and n SW2 may pass 2^53, formed in int64
    const double n = (double)G.tw * G.tw;
    const bool wide = DIG == 2 && n * n * 32896.0 * 32896.0 >= 9007199254740992.0;
    switch (G.nk) {
    case 1: return launch_mt<DIG, 1, false>(a, G, bp, f, g, S, skip, st);
    case 2: return launch_mt<DIG, 2, false>(a, G, bp, f, g, S, skip, st);
    default:
        return wide ? launch_mt<DIG, 3, true>(a, G, bp, f, g, S, skip, st)
                    : launch_mt<DIG, 3, false>(a, G, bp, f, g, S, skip, st);
    }
}

}  // namespace

bool search_mt_supported(int ws, int sx, int sy, int bits)
{
    MTGeom G;
    return (bits == 8 || bits == 16) && mt_geom(ws, sx, sy, bits / 8, G);
}

hipError_t launch_search_mt(const SearchArgs &a, hipStream_t st)
{
    const int S = a.Nx * a.Ny;
    if (S <= 0) return hipSuccess;
    const int dig = a.g16 ? 2 : 1;
    MTGeom G;
    if (!mt_geom(a.ws, a.sx, a.sy, dig, G) || !a.ms || !a.planes) return hipErrorInvalidValue;
    BoxPlanes bp;
    hipError_t e;
    if (dig == 2) {
        // 12-bit data (every pixel of both images below 4096, decided on the device) runs
        // the three-GEMM search on raw digits (box planes uncentred); otherwise the
        // four-GEMM search on centred 16-bit digits
        const bool k3 = search_k3_supported(a.ws, a.sx, a.sy, a.step);
        const size_t np = (size_t)(a.nx + 2) * (a.ny + 2);
        if ((e = a.planes->ensure(planes_bytes(np) + 256)) != hipSuccess) return e;
        unsigned *bits = reinterpret_cast<unsigned *>(a.planes->as<uint8_t>() + planes_bytes(np));
        hipLaunchKernelGGL(k_bits_clear, dim3(1), dim3(1), 0, st, bits);
        hipLaunchKernelGGL(k_bits_or, dim3(1024), dim3(256), 0, st, a.f16, a.g16,
                           (size_t)a.nx * a.ny, bits);
        if ((e = box_planes(a.g16, a.nx, a.ny, G.tw, 32896, bits, k3 ? 0 : 32896, *a.planes, bp,
                            st)) != hipSuccess)
            return e;
        if (a.ev_kernel && (e = hipEventRecord(a.ev_kernel, st)) != hipSuccess) return e;
        if (k3 && (e = launch_search_k3(a, bp, bits, st)) != hipSuccess) return e;
        return launch_dig<2>(a, G, bp, a.f16, a.g16, S, k3 ? bits : nullptr, st);
    }
    if (!a.f8 || !a.g8) return hipErrorInvalidValue;
    if ((e = box_planes(a.g8, a.nx, a.ny, G.tw, 128, nullptr, 128, *a.planes, bp, st)) !=
        hipSuccess)
        return e;
    if (a.ev_kernel && (e = hipEventRecord(a.ev_kernel, st)) != hipSuccess) return e;
    return launch_dig<1>(a, G, bp, a.f8, a.g8, S, nullptr, st);
}

}  // namespace aldic
