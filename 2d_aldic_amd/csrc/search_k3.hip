// Whole-field integer search for images whose pixels are all below 4096 (12-bit data
// staged as 16-bit), on the int8 matrix cores with three GEMMs per lag tile.
//
//   funIntegerSearch.m:82-200  subset loop;  normxcorr2 (called :127);  :134 crop to
//   (2s+3)^2 lags;  findpeak.m:1-68  argmax (x-lag fastest) and 3x3 fit.
//
// * A pixel v < 4096 is two base-64 digits, v = 64 h + l with h, l in [0, 63], so
//   P = 4096 P_hh + 64 (P_hl + P_lh) + P_ll, and with s = h + l in [0, 126] (still an
//   int8) the middle term is P_ss - P_hh - P_ll (Karatsuba): three int8 GEMMs, exact
//   in int32.  The digits are raw (uncentred), so the window box-sum planes are too.
// * A lag tile is the 32 x 32 x 32 GEMM of search_mt.hip: the template row j read at
//   lane i's offset (A, the same for every tile) times 32 window rows (B).  Two
//   neighbouring x tiles run in one pass and share their B operands (K step t of tile
//   ox reads the window columns of step t + 1 of tile ox - 32).
// * A workgroup is four waves, one subset each, on four subsets stacked in y: their
//   windows overlap and are staged once (shared rows m step + r).  The zero ring of
//   subset m's window rows 0 and DH - 1 is another subset's data in the shared copy, so
//   the two lag rows that reach it read a zero row instead.
// * Interior lags (whose footprint stays inside the window) take the argmax straight
//   from the accumulators: n num = n P - ST SW and n var = n SW2 - SW^2 are exact
//   integers below 2^53; the edge lags' P go to LDS and take the closed-form padding
//   correction of search_box.h afterwards.
// * The peak's 3 x 3 neighbourhood is recomputed exactly with byte dot products.
#include <algorithm>

#include "aldic_internal.h"
#include "common.h"
#include "search_box.h"
#include "search_peak.h"

namespace aldic {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef int v4i_u __attribute__((ext_vector_type(4), aligned(4)));
typedef double v4d_u __attribute__((ext_vector_type(4), aligned(8)));

constexpr int kPL = 32;  // template plane: zero bytes before row 0 (> largest lane shift)
constexpr int kNT = 2;   // x tiles per pass

struct K3Geom {
    int tw, sx, sy, nrx, nry, Dx, Dy, DH, VW, ntx, nty, nk;
    int step, MW, SH;   // subset offset in y, subsets (waves) per workgroup, shared rows
    int WS, wplane;     // window row stride, bytes per window digit plane (SH + 1 rows)
    int TS, tplane;     // template row stride, bytes per template digit plane
    int tgrp, wgrp;     // 4-pixel groups per template / window row
    int NE;             // edge lags per subset
    int off_T, per_wave, off_E, off_X, total;
};

int r16(int v) { return (v + 15) & ~15; }

bool k3_geom(int ws, int sx, int sy, int step, K3Geom &G)
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
    G.nk = (62 + G.tw) / 32;  // K = 32 nk >= 31 + tw template columns per lane offset
    if (G.tw < 2 || G.nk > 3 || step < 1 || sx < 0 || sy < 0) return false;
    G.step = step;
    // template rows of TS bytes: row j at kPL + TS j; the TS - tw bytes after a row (or
    // the kPL before row 0) are zero and serve as the next row's left padding
    G.TS = (std::max(32 * G.nk, G.tw + 31) + 3) & ~3;
    G.tplane = r16(kPL + G.TS * G.tw + 16);
    G.tgrp = (G.tw + 3) / 4;
    G.wgrp = (G.VW + 3) / 4;
    // B reads reach column 32 (ntx + nk - 1); the peak's dot products Dx + 8
    G.WS = r16(std::max({32 * (G.ntx + G.nk - 1), 4 * G.wgrp, G.Dx + 12}));
    if ((G.WS / 16) % 2 == 0) G.WS += 16;  // odd multiple of 16: neighbouring rows spread banks
    G.NE = 2 * G.nrx + 2 * (G.nry - 2);
    G.off_E = 2 * G.tplane;
    G.off_X = G.off_E + r16(8 * G.NE);
    G.per_wave = r16(G.off_X + 9 * 8 + 64 * 3 * 4);
    for (int mw = 4; mw >= 1; mw >>= 1) {
        G.MW = mw;
        G.SH = G.DH + (mw - 1) * step;
        G.wplane = (G.SH + 1) * G.WS;
        G.off_T = r16(3 * G.wplane);  // window digit planes h, l and s = h + l
        G.total = G.off_T + mw * G.per_wave;
        if (G.total <= 80 * 1024) return true;  // two workgroups per CU
    }
    return false;
}

// Four 16-bit pixels from pixel index pix: the three dwords around them and the byte
// shift (raw4_load), then (raw4_digits) the pixels outside [k0, k1) -> 0 and base-64
// digit bytes h4 = v >> 6, l4 = v & 63 (pixels below 4096).  Split so that a thread can
// have several rows' loads in flight before it converts any.
struct Raw4 {
    unsigned d0, d1, d2, s;
};

__device__ __forceinline__ Raw4 raw4_load(const uint16_t *img, long long pix)
{
    const long long sa = 2 * pix, lo = sa & ~3LL;
    const uint8_t *b = reinterpret_cast<const uint8_t *>(img);
    Raw4 r;
    r.d0 = *reinterpret_cast<const unsigned *>(b + (lo < 0 ? 0 : lo));
    r.d1 = *reinterpret_cast<const unsigned *>(b + lo + 4);
    r.d2 = *reinterpret_cast<const unsigned *>(b + lo + 8);
    r.s = (unsigned)(sa & 3);
    return r;
}

__device__ __forceinline__ void raw4_digits(const Raw4 &r, int k0, int k1, unsigned &h4,
                                            unsigned &l4)
{
    unsigned w0 = __builtin_amdgcn_alignbyte(r.d1, r.d0, r.s);  // pixels 0 | 1
    unsigned w1 = __builtin_amdgcn_alignbyte(r.d2, r.d1, r.s);  // pixels 2 | 3
    w0 &= ((k0 <= 0 && k1 >= 1) ? 0xffffu : 0u) | ((k0 <= 1 && k1 >= 2) ? 0xffff0000u : 0u);
    w1 &= ((k0 <= 2 && k1 >= 3) ? 0xffffu : 0u) | ((k0 <= 3 && k1 >= 4) ? 0xffff0000u : 0u);
    const unsigned l0 = w0 & 0x003f003fu, l1 = w1 & 0x003f003fu;
    const unsigned h0 = (w0 >> 6) & 0x003f003fu, h1 = (w1 >> 6) & 0x003f003fu;
    l4 = __builtin_amdgcn_perm(l1, l0, 0x06040200u);
    h4 = __builtin_amdgcn_perm(h1, h0, 0x06040200u);
}

constexpr int kLB = 1;  // rows whose loads a thread has in flight at once

__device__ __forceinline__ unsigned udot4(unsigned a, unsigned b, unsigned c)
{
    return __builtin_amdgcn_udot4(a, b, c, false);
}

__device__ __forceinline__ int wave_sum_i(int v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

// A operand of K step t for lane (i = lq, half h): bytes T(j, 32 t + 16 h + b - i),
// b = 0..15, from a template digit plane (row pointer tr = plane + kPL + TS j, lane
// byte offset o = 16 h - lq, split as o = b0 + s3 with b0 % 4 == 0 relative to tr).
template <int NK>
__device__ __forceinline__ void a_operand(const uint8_t *tr, int s3, v4i (&A)[NK])
{
#pragma unroll
    for (int t = 0; t < NK; ++t) {
        const unsigned *q = reinterpret_cast<const unsigned *>(tr + 32 * t);
        unsigned dw[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) dw[k] = q[k];
#pragma unroll
        for (int k = 0; k < 4; ++k) A[t][k] = (int)__builtin_amdgcn_alignbyte(dw[k + 1], dw[k], s3);
    }
}

// One template row j of a pass: NT x tiles (ox, ox + 32), NK K steps, three digit GEMMs.
// b: this lane's window row in the h plane (column ox + 16 h); the l and s planes follow
// at wp and 2 wp bytes.
template <int NK, int NT>
__device__ __forceinline__ void pass_row(const uint8_t *th, const uint8_t *tl, int s3,
                                         const uint8_t *b, int wp, v16i (&acc)[3][NT])
{
    v4i Ah[NK], Al[NK], As[NK];
    a_operand<NK>(th, s3, Ah);
    a_operand<NK>(tl, s3, Al);
#pragma unroll
    for (int t = 0; t < NK; ++t) As[t] = Ah[t] + Al[t];  // digit sums <= 126: no carries
    constexpr int NC = NT + NK - 1;
    v4i Bh[NC], Bl[NC], Bs[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        Bh[c] = *reinterpret_cast<const v4i *>(b + 32 * c);
        Bl[c] = *reinterpret_cast<const v4i *>(b + wp + 32 * c);
        Bs[c] = *reinterpret_cast<const v4i *>(b + 2 * wp + 32 * c);
    }
#pragma unroll
    for (int tt = 0; tt < NT; ++tt)
#pragma unroll
        for (int t = 0; t < NK; ++t) {
            acc[0][tt] = __builtin_amdgcn_mfma_i32_32x32x32_i8(Ah[t], Bh[tt + t], acc[0][tt], 0, 0, 0);
            acc[1][tt] = __builtin_amdgcn_mfma_i32_32x32x32_i8(Al[t], Bl[tt + t], acc[1][tt], 0, 0, 0);
            acc[2][tt] = __builtin_amdgcn_mfma_i32_32x32x32_i8(As[t], Bs[tt + t], acc[2][tt], 0, 0, 0);
        }
}

struct Best {
    double k, d;  // NCC ordering key num |num| and its divisor (var > 0)
    int i;        // lag index ly nrx + lx
};

__device__ __forceinline__ void best_take(Best &b, double key, double den, int p)
{
    const double l = key * b.d, rr = b.k * den;
    if (l > rr || (l == rr && p < b.i)) { b.k = key; b.d = den; b.i = p; }
}

// Edge-lag slot of lag (lx, ly) in the per-subset edge list, or -1 for an interior lag.
__device__ __forceinline__ int edge_slot(const K3Geom &G, int lx, int ly)
{
    if (ly == 0) return lx;
    if (ly == G.nry - 1) return G.nrx + lx;
    if (lx == 0) return 2 * G.nrx + ly - 1;
    if (lx == G.nrx - 1) return 2 * G.nrx + G.nry - 2 + ly - 1;
    return -1;
}

// P of accumulator slot r of tile tt: 4096 P_hh + 64 (P_ss - P_hh - P_ll) + P_ll.
template <int NT>
__device__ __forceinline__ double slot_p(const v16i (&acc)[3][NT], int tt, int r)
{
    const int q = 64 * acc[2][tt][r] - 63 * acc[1][tt][r];  // |q| < 2^31
    return fma(4032.0, (double)acc[0][tt][r], (double)q);
}

// The argmax over one pass's interior lags, straight from the accumulators.  Lane
// (lq, h) holds lag (ox + 32 tt + 8 rg + 4 h + e, ly) in register 4 rg + e of tile tt;
// its slots come in increasing lag index, so a tie keeps the earlier lag.  The window
// box sums are fetched one 4-lag group ahead.  EDGE: the pass holds lag row 0 or nry - 1,
// whose lanes store every P to the edge list.
template <int NT, bool EDGE>
__device__ __forceinline__ void pass_argmax(const K3Geom &G, const v16i (&acc)[3][NT], int ox,
                                            int ly, int h, bool rowin, size_t rowo,
                                            const MTStats &s, const BoxPlanes &bp, double *E,
                                            Best &b)
{
    const bool rowedge = ly == 0 || ly == G.nry - 1;
    const int ebase = ly == 0 ? 0 : G.nrx;
    const int lo = 1 - 4 * h, hi = G.nrx - 2 - 4 * h;  // valid lxb + e for this lane
    constexpr int NG = 4 * NT;
    v4i_u w1 = *reinterpret_cast<const v4i_u *>(bp.bs1 + rowo + ox + 4 * h);
    v4d_u w2 = *reinterpret_cast<const v4d_u *>(bp.bs2 + rowo + ox + 4 * h);
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const int tt = g >> 2, rg = g & 3;
        const int lxb = ox + 32 * tt + 8 * rg;
        if (lxb >= G.nrx) break;
        v4i_u n1 = w1;
        v4d_u n2 = w2;
        if (g + 1 < NG && lxb + 8 < G.nrx) {
            n1 = *reinterpret_cast<const v4i_u *>(bp.bs1 + rowo + lxb + 8 + 4 * h);
            n2 = *reinterpret_cast<const v4d_u *>(bp.bs2 + rowo + lxb + 8 + 4 * h);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int r = 4 * rg + e;
            const double P = slot_p<NT>(acc, tt, r);
            if (EDGE && rowedge && lxb + e + 4 * h < G.nrx) E[ebase + lxb + e + 4 * h] = P;
            const double sw = (double)w1[e];
            const double num = fma(s.nd, P, -(s.std * sw));
            const double var = fma(s.nd, w2[e], -(sw * sw));
            const double key = num * fabs(num), den = fmax(var, 0.5);
            const double l = key * b.d, rr = b.k * den;
            if (rowin && lxb + e >= lo && lxb + e <= hi && l > rr) {
                b.k = key;
                b.d = den;
                b.i = ly * G.nrx + lxb + e + 4 * h;
            }
        }
        w1 = n1;
        w2 = n2;
        __builtin_amdgcn_sched_barrier(0);  // one group of box-sum loads in flight at a time
    }
}

template <int NK, int NT>
__device__ __forceinline__ void run_pass(const K3Geom &G, const uint8_t *Th, const uint8_t *Tl,
                                         const uint8_t *Wh, int mrow, int ox, int oy, int lq,
                                         int h, const MTStats &s, const BoxPlanes &bp, double *E,
                                         Best &b)
{
    v16i acc[3][NT];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int tt = 0; tt < NT; ++tt)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[p][tt][r] = 0;
    const int ly = oy + lq, lyc = min(ly, G.nry - 1);
    const int o = kPL + 16 * h - lq, b0 = o & ~3, s3 = o & 3;
    const uint8_t *th = Th + b0, *tl = Tl + b0;
    const int wp = G.wplane;
    const uint8_t *bw = Wh + (mrow + lyc) * G.WS + ox + 16 * h;
    const uint8_t *zw = Wh + G.SH * G.WS + ox + 16 * h;
    const int tw = G.tw;
    // j = 0 and j = tw - 1 reach the window's zero ring for lag rows 0 and nry - 1
    pass_row<NK, NT>(th, tl, s3, lyc == 0 ? zw : bw, wp, acc);
    for (int j = 1; j < tw - 1; ++j)
        pass_row<NK, NT>(th + j * G.TS, tl + j * G.TS, s3, bw + j * G.WS, wp, acc);
    {
        const int j = tw - 1;
        pass_row<NK, NT>(th + j * G.TS, tl + j * G.TS, s3,
                         lyc == G.nry - 1 ? zw : bw + j * G.WS, wp, acc);
    }
    const bool rowin = ly >= 1 && ly <= G.nry - 2;
    const size_t rowo = (size_t)(s.y0 + lyc + 1) * bp.pitch + (s.x0 + 1);
    if (oy == 0 || oy + 31 >= G.nry - 1)
        pass_argmax<NT, true>(G, acc, ox, ly, h, rowin, rowo, s, bp, E, b);
    else
        pass_argmax<NT, false>(G, acc, ox, ly, h, rowin, rowo, s, bp, E, b);
    // the interior rows' edge columns: lag 0 (register 0 of tile 0, h = 0) and lag
    // nrx - 1 (register r1 of tile t1, h = h1)
    if (ox == 0 && rowin && h == 0) E[2 * G.nrx + ly - 1] = slot_p<NT>(acc, 0, 0);
    const int t1 = ((G.nrx - 1) >> 5) - (ox >> 5);
    if (t1 >= 0 && t1 < NT) {
        const int rem = (G.nrx - 1) & 31, r1 = 4 * (rem >> 3) + (rem & 3), h1 = (rem >> 2) & 1;
#pragma unroll
        for (int tt = 0; tt < NT; ++tt)
#pragma unroll
            for (int r = 0; r < 16; ++r)
                if (tt == t1 && r == r1 && rowin && h == h1)
                    E[2 * G.nrx + G.nry - 2 + ly - 1] = slot_p<NT>(acc, tt, r);
    }
}

template <int NK>
__global__ __launch_bounds__(256, 2) void k_search_k3(SearchArgs a, K3Geom G, BoxPlanes bp,
                                                      const uint16_t *fimg, const uint16_t *gimg,
                                                      const unsigned *bits)
{
    if (*bits >= 4096u) return;  // a pixel above 12 bits: the 16-bit search runs
    extern __shared__ __align__(16) uint8_t lds8[];
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    for (int k = tid; k < G.total / 4; k += blockDim.x) reinterpret_cast<unsigned *>(lds8)[k] = 0u;
    __syncthreads();
    uint8_t *Wh = lds8, *Wl = lds8 + G.wplane, *Ws = lds8 + 2 * G.wplane;
    uint8_t *Tb = lds8 + G.off_T + wv * G.per_wave;
    uint8_t *Th = Tb, *Tl = Tb + G.tplane;
    double *E = reinterpret_cast<double *>(Tb + G.off_E);
    double *N9 = reinterpret_cast<double *>(Tb + G.off_X);
    int *part = reinterpret_cast<int *>(Tb + G.off_X + 9 * 8);

    const int tw = G.tw, nx = a.nx, ny = a.ny;
    const int lq = lane & 31, h = lane >> 5;
    // load layouts without per-item division: thread -> (row, 4-pixel group)
    const int wg4 = tid % G.wgrp, wr0 = tid / G.wgrp, wrs = (int)blockDim.x / G.wgrp;
    const bool wact = wr0 < wrs;
    const int wk0 = wg4 == 0 ? 1 : 0, wk1 = min(4, G.Dx + 1 - 4 * wg4);
    const int tg4 = lane % G.tgrp, tj0 = lane / G.tgrp, tjs = 64 / G.tgrp;
    const bool tact = tj0 < tjs;
    const int tk1 = min(4, tw - 4 * tg4);
    MTStats s;
    s.n = (long long)tw * tw;
    s.nd = (double)s.n;
    s.d = -a.ms[2];  // padding (the normalised image's 0) is mean_g in raw units

    const int NyG = (a.Ny + G.MW - 1) / G.MW, NG = a.Nx * NyG;
    const int nb = gridDim.x, blk = blockIdx.x;
    int q0, q1, qstep;
    if (nb >= 8) {
        // groups in contiguous runs per XCD (workgroups are dealt to the 8 XCDs in turn)
        const int xcd = blk & 7, nbx = (nb - xcd + 7) >> 3, per = NG >> 3, rm = NG & 7;
        const int s0 = xcd * per + min(xcd, rm);
        q0 = s0 + (blk >> 3);
        q1 = s0 + per + (xcd < rm ? 1 : 0);
        qstep = nbx;
    } else {
        q0 = blk;
        q1 = NG;
        qstep = nb;
    }
    for (int grp = q0; grp < q1; grp += qstep) {
        const int kx = grp % a.Nx, ky0 = (grp / a.Nx) * G.MW;
        const int mv = min(G.MW, a.Ny - ky0);
        const int ii = a.x1 + kx * a.step;        // 1-based template corner
        const int jj0 = a.y1 + ky0 * a.step;
        s.x0 = ii - G.sx - 2;
        const int y00 = jj0 - G.sy - 2;
        // shared window rows 1 .. SHv - 2 (shared row R = image row y00 + R), columns
        // 1 .. Dx; the ring columns and the zero row stay 0
        const int SHv = G.DH + (mv - 1) * a.step;
        if (wact) {
            const long long dpix = (long long)wrs * nx;
            long long pix = (long long)(y00 + 1 + wr0) * nx + s.x0 + 4 * wg4;
            for (int R = 1 + wr0; R < SHv - 1; R += kLB * wrs, pix += kLB * dpix) {
                Raw4 raw[kLB];
#pragma unroll
                for (int k = 0; k < kLB; ++k)
                    if (R + k * wrs < SHv - 1) raw[k] = raw4_load(gimg, pix + k * dpix);
#pragma unroll
                for (int k = 0; k < kLB; ++k)
                    if (R + k * wrs < SHv - 1) {
                        unsigned h4, l4;
                        raw4_digits(raw[k], wk0, wk1, h4, l4);
                        const int o = (R + k * wrs) * G.WS + 4 * wg4;
                        *reinterpret_cast<unsigned *>(Wh + o) = h4;
                        *reinterpret_cast<unsigned *>(Wl + o) = l4;
                        *reinterpret_cast<unsigned *>(Ws + o) = h4 + l4;  // digit sums: no carries
                    }
            }
        }
        const bool active = wv < mv;
        const int jj = jj0 + wv * a.step;
        s.y0 = jj - G.sy - 2;
        int sh = 0, sl = 0, shh = 0, shl = 0, sll = 0, r0 = 0, rl = 0;
        if (active && tact) {
            const long long dpix = (long long)tjs * nx;
            long long pix = (long long)(jj - 1 + tj0) * nx + (ii - 1) + 4 * tg4;
            for (int j = tj0; j < tw; j += kLB * tjs, pix += kLB * dpix) {
                Raw4 raw[kLB];
#pragma unroll
                for (int k = 0; k < kLB; ++k)
                    if (j + k * tjs < tw) raw[k] = raw4_load(fimg, pix + k * dpix);
#pragma unroll
                for (int k = 0; k < kLB; ++k) {
                    const int jr = j + k * tjs;
                    if (jr < tw) {
                        unsigned h4, l4;
                        raw4_digits(raw[k], 0, tk1, h4, l4);
                        *reinterpret_cast<unsigned *>(Th + kPL + jr * G.TS + 4 * tg4) = h4;
                        *reinterpret_cast<unsigned *>(Tl + kPL + jr * G.TS + 4 * tg4) = l4;
                        const int gh = (int)udot4(h4, 0x01010101u, 0u);
                        const int gl = (int)udot4(l4, 0x01010101u, 0u);
                        sh += gh;
                        sl += gl;
                        shh = (int)udot4(h4, h4, (unsigned)shh);
                        shl = (int)udot4(h4, l4, (unsigned)shl);
                        sll = (int)udot4(l4, l4, (unsigned)sll);
                        if (jr == 0) r0 += 64 * gh + gl;
                        if (jr == tw - 1) rl += 64 * gh + gl;
                    }
                }
            }
        }
        __syncthreads();  // window and templates staged
        if (active) {
            int c0 = 0, cl = 0;
            for (int j = lane; j < tw; j += 64) {
                const int o0 = kPL + j * G.TS, o1 = o0 + tw - 1;
                c0 += 64 * Th[o0] + Tl[o0];
                cl += 64 * Th[o1] + Tl[o1];
            }
            sh = wave_sum_i(sh);
            sl = wave_sum_i(sl);
            shh = wave_sum_i(shh);
            shl = wave_sum_i(shl);
            sll = wave_sum_i(sll);
            s.r0 = wave_sum_i(r0);
            s.rl = wave_sum_i(rl);
            s.c0 = wave_sum_i(c0);
            s.cl = wave_sum_i(cl);
            const int oL = kPL + (tw - 1) * G.TS;
            s.k00 = 64 * Th[kPL] + Tl[kPL];
            s.k0l = 64 * Th[kPL + tw - 1] + Tl[kPL + tw - 1];
            s.kl0 = 64 * Th[oL] + Tl[oL];
            s.kll = 64 * Th[oL + tw - 1] + Tl[oL + tw - 1];
            s.st = 64LL * sh + sl;
            s.st2 = 4096LL * shh + 128LL * shl + sll;
            s.stn = s.n * s.st2 - s.st * s.st;
            s.std = (double)s.st;
            s.stnd = (double)s.stn;
            const int q = kx + (ky0 + wv) * a.Nx;
            const double xc = (ii + ii + a.ws) / 2.0, yc = (jj + jj + a.ws) / 2.0;
            if (s.stn == 0) {  // constant template: normxcorr2 raises (funIntegerSearch.m:153-157)
                if (lane == 0) {
                    a.x[q] = xc;
                    a.y[q] = yc;
                    a.u[q] = a.v[q] = a.ccmax[q] = NAN;
                    a.status[q] = -1;
                    if (a.qf) { a.qf[2 * q] = INFINITY; a.qf[2 * q + 1] = INFINITY; }
                    if (a.U0) { a.U0[2 * q] = NAN; a.U0[2 * q + 1] = NAN; }
                }
            } else {
                const int mrow = wv * a.step;
                Best b{-INFINITY, 1.0, 0x7fffffff};
                for (int ty = 0; ty < G.nty; ++ty) {
                    int tx = 0;
                    for (; tx + kNT <= G.ntx; tx += kNT)
                        run_pass<NK, kNT>(G, Th, Tl, Wh, mrow, 32 * tx, 32 * ty, lq, h, s, bp, E,
                                          b);
                    if (tx < G.ntx)
                        run_pass<NK, 1>(G, Th, Tl, Wh, mrow, 32 * tx, 32 * ty, lq, h, s, bp, E, b);
                }
                wave_sync_lds();  // edge P in E
                // edge lags: the closed-form correction for the zero padding
                for (int e = lane; e < G.NE; e += 64) {
                    int lx, ly;
                    if (e < G.nrx) { lx = e; ly = 0; }
                    else if (e < 2 * G.nrx) { lx = e - G.nrx; ly = G.nry - 1; }
                    else if (e < 2 * G.nrx + G.nry - 2) { lx = 0; ly = e - 2 * G.nrx + 1; }
                    else { lx = G.nrx - 1; ly = e - 2 * G.nrx - (G.nry - 2) + 1; }
                    double num, den2;
                    mt_terms<false>(s, G, bp, gimg, nx, ny, 0, lx, ly, E[e], num, den2);
                    double key = num * fabs(num), den = den2 / s.stnd;
                    if (!(den > 0.0)) { key = 0.0; den = 1.0; }
                    best_take(b, key, den, ly * G.nrx + lx);
                }
#pragma unroll
                for (int m = 32; m >= 1; m >>= 1) {
                    const double ok = __shfl_xor(b.k, m, 64), od = __shfl_xor(b.d, m, 64);
                    const int oi = __shfl_xor(b.i, m, 64);
                    best_take(b, ok, od, oi);
                }
                const int bx = b.i % G.nrx, by = b.i / G.nrx;
                const bool inner = bx >= 1 && bx <= G.nrx - 2 && by >= 1 && by <= G.nry - 2;
                if (inner) {
                    // P of the 3 x 3 neighbourhood: lane n + 9 jj (jj < 7) sums template
                    // rows jj, jj + 7, ... of lag n = (dc + 1) 3 + (dr + 1)
                    int ph = 0, pl = 0, ps = 0;
                    const int n = lane % 9, j0 = lane / 9;
                    if (lane < 63) {
                        const int lx = bx + n % 3 - 1, ly = by + n / 3 - 1;
                        for (int j = j0; j < tw; j += 7) {
                            if (ly + j == 0 || ly + j == G.DH - 1) continue;  // zero ring row
                            const int wb = (mrow + ly + j) * G.WS + lx, wa = wb & ~3, ws3 = wb & 3;
                            const unsigned *wh = reinterpret_cast<const unsigned *>(Wh + wa);
                            const unsigned *wl = reinterpret_cast<const unsigned *>(Wl + wa);
                            const unsigned *wsp = reinterpret_cast<const unsigned *>(Ws + wa);
                            const unsigned *tH = reinterpret_cast<const unsigned *>(Th + kPL + j * G.TS);
                            const unsigned *tL = reinterpret_cast<const unsigned *>(Tl + kPL + j * G.TS);
                            unsigned hp = wh[0], lp = wl[0], sp = wsp[0];
                            for (int g4 = 0; g4 < G.tgrp; ++g4) {
                                const unsigned hn = wh[g4 + 1], ln = wl[g4 + 1], sn = wsp[g4 + 1];
                                const unsigned t_h = tH[g4], t_l = tL[g4];
                                ph = (int)udot4(t_h, __builtin_amdgcn_alignbyte(hn, hp, ws3),
                                                (unsigned)ph);
                                pl = (int)udot4(t_l, __builtin_amdgcn_alignbyte(ln, lp, ws3),
                                                (unsigned)pl);
                                ps = (int)udot4(t_h + t_l, __builtin_amdgcn_alignbyte(sn, sp, ws3),
                                                (unsigned)ps);
                                hp = hn;
                                lp = ln;
                                sp = sn;
                            }
                        }
                    }
                    part[3 * lane] = ph;
                    part[3 * lane + 1] = pl;
                    part[3 * lane + 2] = ps;
                    wave_sync_lds();
                    if (lane < 9) {
                        long long H = 0, L = 0, SS = 0;
                        for (int k = 0; k < 7; ++k) {
                            H += part[3 * (lane + 9 * k)];
                            L += part[3 * (lane + 9 * k) + 1];
                            SS += part[3 * (lane + 9 * k) + 2];
                        }
                        const double P = (double)(4096 * H + 64 * (SS - H - L) + L);
                        const int lx = bx + lane % 3 - 1, ly = by + lane / 3 - 1;
                        double num, den2;
                        mt_terms<false>(s, G, bp, gimg, nx, ny, 0, lx, ly, P, num, den2);
                        const double den = sqrt(den2);
                        N9[lane] = den > 0.0 ? num / den : 0.0;
                    }
                } else if (lane == 0) {
                    // a peak on the crop's edge: no fit, its own NCC only (P from E)
                    double num, den2;
                    mt_terms<false>(s, G, bp, gimg, nx, ny, 0, bx, by, E[edge_slot(G, bx, by)],
                                    num, den2);
                    const double den = sqrt(den2);
                    N9[4] = den > 0.0 ? num / den : 0.0;
                }
                wave_sync_lds();
                if (lane == 0) {
                    const double best = N9[4];
                    const int r = bx + 1, c = by + 1;  // 1-based row = x-lag, column = y-lag
                    double xp = c, yp = r, mf = best;
                    if (inner) findpeak_fit9(N9, r, c, best, xp, yp, mf);
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
            }
        }
        __syncthreads();  // before the next group overwrites the window and templates
    }
}

template <int NK>
hipError_t launch_k3(const SearchArgs &a, const K3Geom &G, const BoxPlanes &bp,
                     const unsigned *bits, hipStream_t st)
{
    auto kern = k_search_k3<NK>;
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  80 * 1024);
        attr_set = true;
    }
    const int NG = a.Nx * ((a.Ny + G.MW - 1) / G.MW);
    const int per_cu = std::max(1, std::min(2, (160 * 1024) / G.total));
    const int nwg = std::min(NG, 256 * per_cu);
    hipLaunchKernelGGL(kern, dim3(nwg), dim3(64 * G.MW), (size_t)G.total, st, a, G, bp, a.f16,
                       a.g16, bits);
    return hipGetLastError();
}

}  // namespace

bool search_k3_supported(int ws, int sx, int sy, int step)
{
    K3Geom G;
    return k3_geom(ws, sx, sy, step, G);
}

hipError_t launch_search_k3(const SearchArgs &a, const BoxPlanes &bp, const unsigned *bits,
                            hipStream_t st)
{
    K3Geom G;
    if (!k3_geom(a.ws, a.sx, a.sy, a.step, G) || !a.f16 || !a.g16 || !a.ms)
        return hipErrorNotSupported;
    if (a.Nx * a.Ny <= 0) return hipSuccess;
    switch (G.nk) {
    case 1: return launch_k3<1>(a, G, bp, bits, st);
    case 2: return launch_k3<2>(a, G, bp, bits, st);
    default: return launch_k3<3>(a, G, bp, bits, st);
    }
}

}  // namespace aldic
