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
    // (the N region first holds the vertical box sums, 2 nry VW ints)
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
    G.per_wave = r16(G.off_N + std::max(8 * G.nry * G.nrx, 8 * G.nry * G.VW));
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
    return __builtin_amdgcn_readlane(wave_scan(v), 63);
}

// Template statistics of one subset for the NCC of a lag (see the header comment).
struct TplStats {
    int tw, nrx, nry, st, r0, rl, c0, cl, k00, k0l, kl0, kll;
    double dn, dst, dstn, d;
};

// num and den2 = n st * n var of lag p (NCC = num / sqrt(den2)); P, sw, sw2 its sums.
// Every term is an exact integer below 2^53 in FP64 except the mean correction's
// one rounding.
__device__ __forceinline__ void ncc_terms(const TplStats &t, int p, double P, int isw, int isw2,
                                          double &num, double &den2)
{
    const int ly = p / t.nrx, lx = p - ly * t.nrx;
    const double sw = (double)isw, sw2 = (double)isw2;
    const bool top = ly == 0, bot = ly == t.nry - 1, lft = lx == 0, rgt = lx == t.nrx - 1;
    const int ro = (top || bot) ? 1 : 0, co = (lft || rgt) ? 1 : 0;
    const int rs = top ? t.r0 : (bot ? t.rl : 0);
    const int cs = lft ? t.c0 : (rgt ? t.cl : 0);
    const int kc = (ro && co) ? (top ? (lft ? t.k00 : t.k0l) : (lft ? t.kl0 : t.kll)) : 0;
    const double stin = (double)(t.st - rs - cs + kc);
    const double nin = (double)((t.tw - ro) * (t.tw - co)), nout = t.dn - nin;
    num = (t.dn * P - t.dst * sw) + t.d * (t.dn * stin - nin * t.dst);
    const double var = (t.dn * sw2 - sw * sw) + 2.0 * t.d * sw * nout + t.d * t.d * nin * nout;
    den2 = t.dstn * (var > 0.0 ? var : 0.0);
}

// The rows of one subset's template and window, fetched into registers ahead of use
// (the next subset's loads are in flight while the current one is computed).  Item
// k of a lane is dword (row, g) = divmod(lane + 64 k, dwords per row) of the template
// (KT items) or window (KW items); each is two aligned dwords around the row's bytes.
template <int KT, int KW> struct Prefetch {
    unsigned t0[KT], t1[KT], w0[KW], w1[KW];
    int ii, jj;  // 1-based corner of the subset they belong to
};

struct Walk {
    int r, g;
    __device__ __forceinline__ void next(int dr, int dg, int grp)
    {
        g += dg;
        r += dr;
        if (g >= grp) { g -= grp; ++r; }
    }
};

template <int KT, int KW>
__device__ __forceinline__ void issue_rows(const SearchArgs &a, const I8Geom &G, int q, int lane,
                                           Walk wt, int djt, int dgt, Walk ww, int djw, int dgw,
                                           Prefetch<KT, KW> &pf)
{
    const int kx = q % a.Nx, ky = q / a.Nx;
    pf.ii = a.x1 + kx * a.step;
    pf.jj = a.y1 + ky * a.step;
    const int nx = a.nx;
    const int tb = (pf.jj - 1) * nx + (pf.ii - 1);
    const int nt = G.tw * G.tgrp;
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        if (lane + 64 * k < nt) {
            const int lo = (tb + wt.r * nx + 4 * wt.g) & ~3;
            pf.t0[k] = *reinterpret_cast<const unsigned *>(a.f8 + lo);
            pf.t1[k] = *reinterpret_cast<const unsigned *>(a.f8 + lo + 4);
        }
        wt.next(djt, dgt, G.tgrp);
    }
    // window row r (1 .. Dy) starts at image byte wb + r nx; its column 0 (the zero
    // ring) may lie one byte before the image, so the offset can be -1
    const int wb = (pf.jj - G.sy - 2) * nx + (pf.ii - G.sx - 2);
    const int nw = G.Dy * G.wgrp;
#pragma unroll
    for (int k = 0; k < KW; ++k) {
        if (lane + 64 * k < nw) {
            const int lo = (wb + (1 + ww.r) * nx + 4 * ww.g) & ~3;
            pf.w0[k] = *reinterpret_cast<const unsigned *>(a.g8 + (lo < 0 ? 0 : lo));
            pf.w1[k] = *reinterpret_cast<const unsigned *>(a.g8 + lo + 4);
        }
        ww.next(djw, dgw, G.wgrp);
    }
}

template <int KT, int KW>
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
    // staging items it = lane + 64 k: (row, dword) walked without divisions
    Walk wt0, ww0;
    wt0.r = lane / G.tgrp;
    wt0.g = lane - wt0.r * G.tgrp;
    const int djt = 64 / G.tgrp, dgt = 64 - djt * G.tgrp;
    ww0.r = lane / G.wgrp;
    ww0.g = lane - ww0.r * G.wgrp;
    const int djw = 64 / G.wgrp, dgw = 64 - djw * G.wgrp;

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
    Prefetch<KT, KW> pf;
    if (q0 < q1) issue_rows(a, G, q0, lane, wt0, djt, dgt, ww0, djw, dgw, pf);
    for (int q = q0; q < q1; q += qstep) {
        const int ii = pf.ii, jj = pf.jj;  // 1-based corner
        // template rows -> T(j, i) at Tl[j TS + kPL + i] (b ^ 0x80 per byte) and its
        // integer statistics (|sums| <= 128^2 * 65^2 < 2^31)
        int st = 0, st2 = 0, r0 = 0, rl = 0, c0 = 0, cl = 0, k00 = 0, k0l = 0, kl0 = 0, kll = 0;
        {
            Walk w = wt0;
            const int tb = (jj - 1) * nx + (ii - 1);
            const int nt = tw * G.tgrp;
#pragma unroll
            for (int k = 0; k < KT; ++k) {
                if (lane + 64 * k < nt) {
                    const int j = w.r, g4 = w.g;
                    const int off = tb + j * nx + 4 * g4;
                    const int k1 = min(4, tw - 4 * g4);
                    const unsigned m = k1 >= 4 ? 0xffffffffu : ((1u << (8 * k1)) - 1u);
                    const unsigned v = (__builtin_amdgcn_alignbyte(pf.t1[k], pf.t0[k], (unsigned)(off & 3)) ^
                                        0x80808080u) & m;
                    *reinterpret_cast<unsigned *>(Tl + j * G.TS + kPL + 4 * g4) = v;
                    const int s1 = __builtin_amdgcn_sdot4((int)v, 0x01010101, 0, false);
                    st += s1;
                    st2 = __builtin_amdgcn_sdot4((int)v, (int)v, st2, false);
                    if (j == 0) r0 += s1;
                    if (j == tw - 1) rl += s1;
                    const int first = (int)(int8_t)(v & 0xffu);
                    const int last = (int)(int8_t)(v >> (8 * ((tw - 1) & 3)));
                    const bool gl = g4 == ((tw - 1) >> 2);
                    if (g4 == 0) c0 += first;
                    if (gl) cl += last;
                    if (j == 0 && g4 == 0) k00 = first;
                    if (j == 0 && gl) k0l = last;
                    if (j == tw - 1 && g4 == 0) kl0 = first;
                    if (j == tw - 1 && gl) kll = last;
                }
                w.next(djt, dgt, G.tgrp);
            }
        }
        // window rows 1 .. Dy, columns 1 .. Dx (row 0, row Dy + 1, column 0 and columns
        // past Dx stay zero): Wl[r WS + c] = g(ii - sx - 2 + c, jj - sy - 2 + r) - 128
        {
            Walk w = ww0;
            const int wb = (jj - G.sy - 2) * nx + (ii - G.sx - 2);
            const int nw = G.Dy * G.wgrp;
#pragma unroll
            for (int k = 0; k < KW; ++k) {
                if (lane + 64 * k < nw) {
                    const int r = 1 + w.r, g4 = w.g;
                    const int off = wb + r * nx + 4 * g4;  // >= -1
                    const int k0 = g4 == 0 ? 1 : 0, k1 = min(4, G.Dx + 1 - 4 * g4);
                    const unsigned m = (k1 >= 4 ? 0xffffffffu : ((1u << (8 * k1)) - 1u)) &
                                       (k0 ? 0xffffff00u : 0xffffffffu);
                    *reinterpret_cast<unsigned *>(Wl + r * G.WS + 4 * g4) =
                        (__builtin_amdgcn_alignbyte(pf.w1[k], pf.w0[k], (unsigned)(off & 3)) ^
                         0x80808080u) & m;
                }
                w.next(djw, dgw, G.wgrp);
            }
        }
        wave_sync_lds();
        // the next subset's rows load while this one is computed
        if (q + qstep < q1) issue_rows(a, G, q + qstep, lane, wt0, djt, dgt, ww0, djw, dgw, pf);
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

        // box sums of w and w^2 per lag: vertical running sums over tw rows (a lane per
        // window column) into V, then horizontal running sums over tw columns (a lane
        // per lag row and quantity) into S
        {
            int *V1 = reinterpret_cast<int *>(Nc), *V2 = V1 + nry * G.VW;
            if (lane < G.VW) {
                int v1 = 0, v2 = 0;
                for (int r = 0; r < tw; ++r) {
                    const int w = (int8_t)Wl[r * G.WS + lane];
                    v1 += w;
                    v2 += w * w;
                }
                V1[lane] = v1;
                V2[lane] = v2;
                for (int ly = 1; ly < nry; ++ly) {
                    const int wi = (int8_t)Wl[(ly + tw - 1) * G.WS + lane];
                    const int wo = (int8_t)Wl[(ly - 1) * G.WS + lane];
                    v1 += wi - wo;
                    v2 += wi * wi - wo * wo;
                    V1[ly * G.VW + lane] = v1;
                    V2[ly * G.VW + lane] = v2;
                }
            }
            wave_sync_lds();
            if (lane < 2 * nry) {
                const int qn = lane >= nry ? 1 : 0, ly = lane - qn * nry;
                const int *vr = (qn ? V2 : V1) + ly * G.VW;
                int *so = (qn ? S2 : S1) + ly * nrx;
                int hs = 0;
                for (int i = 0; i < tw; ++i) hs += vr[i];
                so[0] = hs;
                for (int lx = 1; lx < nrx; ++lx) {
                    hs += vr[lx - 1 + tw] - vr[lx - 1];
                    so[lx] = hs;
                }
            }
        }

        // P = sum_j A_j B_j on the matrix cores; row j + 1's operands load before row
        // j's products issue
        v16i acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = 0;
        if (G.nk == 1) {
            const unsigned *tq = reinterpret_cast<const unsigned *>(Tl + b0);
            const uint8_t *wr = Wl + rowB * G.WS + 16 * h;
            unsigned dw[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) dw[k] = tq[k];
            v4i B = *reinterpret_cast<const v4i *>(wr);
            for (int j = 0; j < tw; ++j) {
                v4i A;
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    A[k] = (int)__builtin_amdgcn_alignbyte(dw[k + 1], dw[k], s3);
                const v4i Bc = B;
                if (j + 1 < tw) {
                    tq += G.TS / 4;
                    wr += G.WS;
#pragma unroll
                    for (int k = 0; k < 5; ++k) dw[k] = tq[k];
                    B = *reinterpret_cast<const v4i *>(wr);
                }
                acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, Bc, acc, 0, 0, 0);
            }
        } else {
            const unsigned *tq = reinterpret_cast<const unsigned *>(Tl + b0);
            const uint8_t *wr = Wl + rowB * G.WS + 32 * h;
            unsigned dw[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) dw[k] = tq[k];
            v4i B0 = *reinterpret_cast<const v4i *>(wr), B1 = *reinterpret_cast<const v4i *>(wr + 16);
            for (int j = 0; j < tw; ++j) {
                v4i A0, A1;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    A0[k] = (int)__builtin_amdgcn_alignbyte(dw[k + 1], dw[k], s3);
                    A1[k] = (int)__builtin_amdgcn_alignbyte(dw[k + 5], dw[k + 4], s3);
                }
                const v4i C0 = B0, C1 = B1;
                if (j + 1 < tw) {
                    tq += G.TS / 4;
                    wr += G.WS;
#pragma unroll
                    for (int k = 0; k < 9; ++k) dw[k] = tq[k];
                    B0 = *reinterpret_cast<const v4i *>(wr);
                    B1 = *reinterpret_cast<const v4i *>(wr + 16);
                }
                acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(A0, C0, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(A1, C1, acc, 0, 0, 0);
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

        // argmax of NCC without dividing: NCC = num / sqrt(den2) has the order of
        // key / den2 with key = num |num|, compared as key_a den2_b > key_b den2_a
        // (ties -> smallest column-major index, x-lag fastest)
        TplStats ts;
        ts.tw = tw; ts.nrx = nrx; ts.nry = nry;
        ts.st = st; ts.r0 = r0; ts.rl = rl; ts.c0 = c0; ts.cl = cl;
        ts.k00 = k00; ts.k0l = k0l; ts.kl0 = kl0; ts.kll = kll;
        ts.dn = (double)n; ts.dst = (double)st; ts.dstn = (double)stn; ts.d = d;
        const int nl = nrx * nry;
        double bk = -INFINITY, bd = 1.0;
        int bi = 0x7fffffff;
        for (int p = lane; p < nl; p += 64) {
            double num, den2;
            ncc_terms(ts, p, Nc[p], S1[p], S2[p], num, den2);
            double key = num * fabs(num);
            if (!(den2 > 0.0)) { key = 0.0; den2 = 1.0; }
            if (key * bd > bk * den2) { bk = key; bd = den2; bi = p; }
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            const double ok = __shfl_xor(bk, m, 64), od = __shfl_xor(bd, m, 64);
            const int oi = __shfl_xor(bi, m, 64);
            const double l = ok * bd, r = bk * od;
            if (l > r || (l == r && oi < bi)) { bk = ok; bd = od; bi = oi; }
        }
        // exact NCC at the peak and its 3 x 3 neighbourhood (lane k: offset (k%3-1, k/3-1))
        wave_sync_lds();
        double best = 0.0;
        {
            const int rx = bi % nrx, cy = bi / nrx;
            const int dxk = lane % 3 - 1, dyk = lane / 3 - 1;
            const int px = rx + dxk, py = cy + dyk;
            double c = 0.0;
            const bool mine = lane < 9 && px >= 0 && px < nrx && py >= 0 && py < nry;
            if (mine) {
                double num, den2;
                const int pp = py * nrx + px;
                ncc_terms(ts, pp, Nc[pp], S1[pp], S2[pp], num, den2);
                const double den = sqrt(den2);
                c = den > 0.0 ? num / den : 0.0;
            }
            best = __shfl(c, 4, 64);
            wave_sync_lds();
            if (mine) Nc[py * nrx + px] = c;
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

template <int KT, int KW>
hipError_t launch_kt(const SearchArgs &a, const I8Geom &G, int S, hipStream_t st)
{
    const size_t lds = (size_t)kI8Waves * G.per_wave;
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)k_search_i8<KT, KW>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024);
        attr_set = true;
    }
    // resident workgroups: LDS-limited, and enough to give every wave a few subsets
    const int per_cu = std::max(1, std::min(8, (int)((160 * 1024) / lds)));
    const int nwg = std::min((S + kI8Waves - 1) / kI8Waves, 256 * per_cu);
    auto kern = k_search_i8<KT, KW>;
    hipLaunchKernelGGL(kern, dim3(nwg), dim3(64 * kI8Waves), lds, st, a, G);
    return hipGetLastError();
}

hipError_t launch_search_i8(const SearchArgs &a, hipStream_t st)
{
    const int S = a.Nx * a.Ny;
    if (S <= 0) return hipSuccess;
    I8Geom G;
    if (!i8_geom(a.ws, a.sx, a.sy, G) || !a.f8 || !a.g8 || !a.ms) return hipErrorInvalidValue;
    // prefetch registers per lane: template and window dwords over 64 lanes
    const int kt = (G.tw * G.tgrp + 63) / 64, kw = (G.Dy * G.wgrp + 63) / 64;
    if (kt <= 3 && kw <= 6) return launch_kt<3, 6>(a, G, S, st);
    if (kt <= 5 && kw <= 10) return launch_kt<5, 10>(a, G, S, st);
    if (kt <= 8 && kw <= 16) return launch_kt<8, 16>(a, G, S, st);
    return launch_kt<16, 24>(a, G, S, st);
}

}  // namespace aldic
