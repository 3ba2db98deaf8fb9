// findpeak.m:30-66 on the cropped NCC map of one subset (shared by the search kernels).
#pragma once

namespace aldic {

// The fit of findpeak.m:37-66 on a 3x3 neighbourhood u[(dc + 1) * 3 + (dr + 1)]
// (dc = column offset = y-lag, dr = row offset = x-lag) around the 1-based peak (r, c).
__device__ inline void findpeak_fit9(const double *u9, int r, int c, double best, double &xp,
                                     double &yp, double &mf)
{
    xp = c;
    yp = r;
    mf = best;
    double s0 = 0, sxu = 0, syu = 0, sxy = 0, sxx = 0, syy = 0;
    for (int dc = -1; dc <= 1; ++dc)
        for (int dr = -1; dr <= 1; ++dr) {
            const double u = u9[(dc + 1) * 3 + (dr + 1)];
            s0 += u; sxu += dc * u; syu += dr * u; sxy += dc * dr * u;
            sxx += dc * dc * u; syy += dr * dr * u;
        }
    const double A1 = (5.0 * s0 - 3.0 * sxx - 3.0 * syy) / 9.0;
    const double A2 = sxu / 6.0, A3 = syu / 6.0, A4 = sxy / 4.0;
    const double A5 = sxx / 2.0 - s0 / 3.0, A6 = syy / 2.0 - s0 / 3.0;
    const double den = A4 * A4 - 4.0 * A5 * A6;
    double xo = (-A3 * A4 + 2.0 * A6 * A2) / den;
    double yo = -1.0 / den * (A4 * A2 - 2.0 * A5 * A3);
    if (!(fabs(xo) > 1.0 || fabs(yo) > 1.0) && !isnan(xo) && !isnan(yo)) {
        xo = xo < 0 ? -floor(-1000.0 * xo + 0.5) / 1000.0 : floor(1000.0 * xo + 0.5) / 1000.0;
        yo = yo < 0 ? -floor(-1000.0 * yo + 0.5) / 1000.0 : floor(1000.0 * yo + 0.5) / 1000.0;
        xp = c + xo;
        yp = r + yo;
        mf = A1 + A2 * xo + A3 * yo + A4 * xo * yo + A5 * xo * xo + A6 * yo * yo;
    }
}

// ncc[(c - 1) * nrx + (r - 1)]: row r = x-lag, column c = y-lag (both 1-based); bi is the
// index of the first maximum in that (column-major) order and best its value.  Returns
// the 1-based peak (xp = column, yp = row) and the fitted maximum mf: off the map's edge
// the least-squares quadratic of the 3x3 neighbourhood moves the peak by (xo, yo) when
// both stay within one pixel, rounded to 1e-3 as findpeak.m:57-58 does.
template <class T>
__device__ inline void findpeak_fit(const T *ncc, int nrx, int nry, int bi, double best,
                                    double &xp, double &yp, double &mf)
{
    const int r = bi % nrx + 1, c = bi / nrx + 1;
    xp = c;
    yp = r;
    mf = best;
    if (c == 1 || c == nry || r == 1 || r == nrx) return;
    double u9[9];
    for (int dc = -1; dc <= 1; ++dc)
        for (int dr = -1; dr <= 1; ++dr)
            u9[(dc + 1) * 3 + (dr + 1)] = ncc[(c - 1 + dc) * nrx + (r - 1 + dr)];
    findpeak_fit9(u9, r, c, best, xp, yp, mf);
}

}  // namespace aldic
