// Window box-sum planes and the exact NCC terms of one lag, shared by the 16-bit and the
// 12-bit matrix-core searches (search_mt.hip, search_k3.hip).
//   normxcorr2 (called at funIntegerSearch.m:127) on the window zero-padded by one
//   pixel: the crop of :134 keeps the lags whose footprint leaves the window by at most
//   one row or column (the edge lags).
#pragma once
#include "aldic_internal.h"

namespace aldic {

struct BoxPlanes {
    const int *bs1, *rs1, *cs1;
    const long long *rs2;
    const double *bs2, *cs2;  // exact integers (< 2^53), kept as FP64 for the lag arithmetic
    int pitch;  // nx + 2: element (ye, xe) at (ye + 1) pitch + xe + 1, ye in [-1, ny], xe in [-1, nx]
};

struct MTStats {
    long long st, st2, stn, n;
    double nd, std, stnd;  // n, st, stn as FP64 (exact)
    int r0, rl, c0, cl, k00, k0l, kl0, kll;  // |.| <= 2^16 * 66
    double d;
    int x0, y0;  // 0-based image position of window element (0, 0) (the zero ring's corner)
};

// num and den2 (NCC = num / sqrt(den2)) of lag (lx, ly); P its correlation sum, sw / sw2
// the window box sums of the lag's full footprint (box-sum planes at (ly, lx)).
// GEO: the search's geometry (fields tw, nrx, nry, Dx, Dy).
template <bool WIDE, class T, class GEO>
__device__ __forceinline__ void mt_terms_sw(const MTStats &s, const GEO &G, const BoxPlanes &bp,
                                            const T *g, int nx, int ny, int Z, int lx, int ly,
                                            double P, double sw, double sw2, double &num,
                                            double &den2)
{
    // P, SW, SW2 and the statistics are integers below 2^53; so are n P, ST SW, n SW2 and
    // SW^2 unless WIDE (16-bit pixels and n^2 32896^2 >= 2^53, templates above 53 x 53),
    // so FP64 products of them are exact; WIDE forms those in int64
    const bool top = ly == 0, bot = ly == G.nry - 1, lft = lx == 0, rgt = lx == G.nrx - 1;
    const int ye = s.y0 + ly, xe = s.x0 + lx;
    const int yr = top ? s.y0 : s.y0 + G.Dy + 1, xr = lft ? s.x0 : s.x0 + G.Dx + 1;
    if (top || bot) {
        const size_t orow = (size_t)(yr + 1) * bp.pitch + (xe + 1);
        sw -= (double)bp.rs1[orow];
        sw2 -= (double)bp.rs2[orow];
    }
    if (lft || rgt) {
        const size_t ocol = (size_t)(ye + 1) * bp.pitch + (xr + 1);
        sw -= (double)bp.cs1[ocol];
        sw2 -= bp.cs2[ocol];
    }
    const int ro = (top || bot) ? 1 : 0, co = (lft || rgt) ? 1 : 0;
    if (ro && co) {  // the corner pixel was taken off twice
        const double cv =
            (yr >= 0 && yr < ny && xr >= 0 && xr < nx) ? (double)((int)g[(size_t)yr * nx + xr] - Z) : 0.0;
        sw += cv;
        sw2 += cv * cv;
    }
    const long long rs = top ? s.r0 : (bot ? s.rl : 0);
    const long long cs = lft ? s.c0 : (rgt ? s.cl : 0);
    const long long kc = (ro && co) ? (top ? (lft ? s.k00 : s.k0l) : (lft ? s.kl0 : s.kll)) : 0;
    const long long stin = s.st - rs - cs + kc;
    const long long nin = (long long)(G.tw - ro) * (G.tw - co), nout = s.n - nin;
    double t1, t2;
    if constexpr (WIDE) {
        const long long Pi = (long long)P, swi = (long long)sw, sw2i = (long long)sw2;
        t1 = (double)(s.n * Pi - s.st * swi);
        t2 = (double)(s.n * sw2i - swi * swi);
    } else {
        const double a1 = s.nd * P, a2 = s.std * sw, b1 = s.nd * sw2, b2 = sw * sw;
        t1 = a1 - a2;
        t2 = b1 - b2;
    }
    num = t1 + s.d * (double)(s.n * stin - nin * s.st);
    const double var = t2 + 2.0 * s.d * sw * (double)nout + s.d * s.d * (double)nin * (double)nout;
    den2 = s.stnd * (var > 0.0 ? var : 0.0);
}

template <bool WIDE, class T, class GEO>
__device__ __forceinline__ void mt_terms(const MTStats &s, const GEO &G, const BoxPlanes &bp,
                                         const T *g, int nx, int ny, int Z, int lx, int ly,
                                         double P, double &num, double &den2)
{
    const size_t o = (size_t)(s.y0 + ly + 1) * bp.pitch + (s.x0 + lx + 1);
    mt_terms_sw<WIDE>(s, G, bp, g, nx, ny, Z, lx, ly, P, (double)bp.bs1[o], bp.bs2[o], num, den2);
}

// Launches the 12-bit search (search_k3.hip) when its geometry fits: returns
// hipErrorNotSupported (nothing launched) otherwise.  bits is the device word holding
// the OR of every pixel of both images; the kernel returns at once when it is 4096 or
// more, leaving the work to the 16-bit search.
bool search_k3_supported(int ws, int sx, int sy, int step);
hipError_t launch_search_k3(const SearchArgs &a, const BoxPlanes &bp, const unsigned *bits,
                            hipStream_t st);

}  // namespace aldic
