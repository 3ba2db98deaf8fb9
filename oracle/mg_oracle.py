"""CPU restatement of funIntegerSearchMg (func/funIntegerSearchMg.m:3-285), the multigrid
integer search of IntegerSearch.m:120 (InitFFTSearchMethod 0), for the tests only.

Arrays follow MATLAB's orientation: an image is indexed [x, y] (dim 1 = x, the
reference's transposed image, ReadImage.m:79), ranges are 1-based and inclusive.

  * normxcorr2 is the oracle's definition (aldic_oracle.c ncc_full: sum (T - mean T) A over
    the footprint, A zero-padded, over sqrt(sum (T - mean T)^2 * max(sum A^2 - (sum A)^2 /
    n, 0)), 0 where that is 0), here for rectangular templates: the level-1 map in full
    (numerator by FFT, window sums by prefix sums), the later levels only on the 13 x 14
    lags their findpeak reads (:172-176), directly.
  * findpeak is the oracle's (findpeak.m:1-68).
  * The final fit is oracle.gridfit (gridfit.m with 'regularizer','springs'), after
    gridfit.m:332's removal of NaN points.
  * The q-factors of :168-170 are not restated (only the outlier test with qDICOrNot 1
    reads them, and that test is not part of this path); cc.qfactors is returned as NaN.
  * MATLAB's try/catch (:163-182): a region whose g window (+-5 px) leaves the image, has
    a NaN bound or a constant template gives NaN, NaN and Phi 0.
"""
import math

import numpy as np

import oracle


def _ceil(v):
    return math.ceil(v) if np.isfinite(v) else float("nan")


def ncc_full(T, A):
    """normxcorr2(T, A) in full, (m + M - 1) x (n + N - 1); None for a constant template."""
    T = np.asarray(T, np.float64)
    A = np.asarray(A, np.float64)
    m, n = T.shape
    M, N = A.shape
    if np.all(T == T.flat[0]):
        return None
    Tp = T - T.mean()
    st = float(np.sum(Tp * Tp))
    R, Cc = m + M - 1, n + N - 1
    P = 1 << max(R - 1, 1).bit_length()
    Q = 1 << max(Cc - 1, 1).bit_length()
    ft = np.fft.fft2(Tp, (P, Q))
    fa = np.fft.fft2(A, (P, Q))
    corr = np.fft.ifft2(np.conj(ft) * fa).real
    # lag k = r - (m - 1): corr[k mod P]
    rr = (np.arange(R) - (m - 1)) % P
    cc = (np.arange(Cc) - (n - 1)) % Q
    num = corr[np.ix_(rr, cc)]
    # window sums of A and A^2 over the footprint at every lag (A zero-padded)
    def wsum(X):
        Z = np.zeros((M + 2 * m, N + 2 * n))
        Z[m:m + M, n:n + N] = X
        S = np.zeros((Z.shape[0] + 1, Z.shape[1] + 1))
        S[1:, 1:] = Z.cumsum(0).cumsum(1)
        # footprint of output (r, c): Z rows r + 1 .. r + m, cols c + 1 .. c + n
        r0 = np.arange(R)[:, None] + 1
        c0 = np.arange(Cc)[None, :] + 1
        return S[r0 + m, c0 + n] - S[r0, c0 + n] - S[r0 + m, c0] + S[r0, c0]
    sa = wsum(A)
    sa2 = wsum(A * A)
    va = np.maximum(sa2 - sa * sa / (m * n), 0.0)
    den = np.sqrt(st * va)
    out = np.zeros_like(num)
    nz = den > 0
    out[nz] = num[nz] / den[nz]
    return out


def ncc_crop(T, A, ox, oy):
    """normxcorr2(T, A) at template offsets ox (x, 0-based position of T(1,1) in A) x oy;
    A zero-padded.  None for a constant template."""
    T = np.asarray(T, np.float64)
    A = np.asarray(A, np.float64)
    m, n = T.shape
    M, N = A.shape
    if np.all(T == T.flat[0]):
        return None
    Tp = T - T.mean()
    st = float(np.sum(Tp * Tp))
    out = np.zeros((len(ox), len(oy)))
    for i, a in enumerate(ox):
        for j, b in enumerate(oy):
            x0, x1 = max(a, 0), min(a + m, M)
            y0, y1 = max(b, 0), min(b + n, N)
            if x1 <= x0 or y1 <= y0:
                continue
            W = A[x0:x1, y0:y1]
            Tw = Tp[x0 - a:x1 - a, y0 - b:y1 - b]
            num = float(np.sum(Tw * W))
            sa, sa2 = float(W.sum()), float(np.sum(W * W))
            va = max(sa2 - sa * sa / (m * n), 0.0)
            den = math.sqrt(st * va)
            out[i, j] = num / den if den > 0 else 0.0
    return out


def _clamp_region(r, nx, ny, gap):
    """Border checks of :59-66 / :202-229 on one region [xf1 xf2 yf1 yf2 xg1 xg2 yg1 yg2]
    and the odd-width fix of :68-69 / :231-239."""
    xf1, xf2, yf1, yf2, xg1, xg2, yg1, yg2 = r
    if xg1 < gap:
        t = xg1; xg1 = gap; xf1 = xf1 + gap - t
    if yg1 < gap:
        t = yg1; yg1 = gap; yf1 = yf1 + gap - t
    if xg2 > nx - gap:
        t = xg2; xg2 = nx - gap; xf2 = xf2 + (nx - gap) - t
    if yg2 > ny - gap:
        t = yg2; yg2 = ny - gap; yf2 = yf2 + (ny - gap) - t
    if np.isfinite(xf2 - xf1) and (xf2 - xf1) % 2 == 1:
        xf2 -= 1; xg2 -= 1
    if np.isfinite(yf2 - yf1) and (yf2 - yf1) % 2 == 1:
        yf2 -= 1; yg2 -= 1
    return [xf1, xf2, yf1, yf2, xg1, xg2, yg1, yg2]


def _region_peak(f, g, r, nx, ny):
    """:163-176 for one region: (du, dv, Phi) relative to the region's g offset, NaN on
    the reference's catch."""
    xf1, xf2, yf1, yf2, xg1, xg2, yg1, yg2 = r
    vals = np.array(r, np.float64)
    if not np.all(np.isfinite(vals)) or np.any(vals != np.round(vals)):
        return float("nan"), float("nan"), 0.0
    xf1, xf2, yf1, yf2, xg1, xg2, yg1, yg2 = [int(v) for v in r]
    if (xf1 < 1 or yf1 < 1 or xf2 > nx or yf2 > ny or xf2 < xf1 or yf2 < yf1 or
            xg1 - 5 < 1 or yg1 - 5 < 1 or xg2 + 5 > nx or yg2 + 5 > ny):
        return float("nan"), float("nan"), 0.0
    C = f[xf1 - 1:xf2, yf1 - 1:yf2]
    D = g[xg1 - 6:xg2 + 5, yg1 - 6:yg2 + 5]
    w1, w2 = C.shape[0] - 1, C.shape[1] - 1
    # crop rows w1 .. end-w1+1 and columns w2 .. end-w2+2 of the full map: offsets
    # -1 .. 11 (x) and -1 .. 12 (y); zero displacement at crop index 7
    crop = ncc_crop(C, D, np.arange(-1, 12), np.arange(-1, 13))
    if crop is None:
        return float("nan"), float("nan"), 0.0
    xp, yp, mf = oracle.findpeak(crop)
    return yp - 7.0, xp - 7.0, mf


def search_mg(f, g, gridx, gridy, winsize, winstepsize):
    """funIntegerSearchMg(f, g, gridx, gridy, winsize, winstepsize, winstepsize) on images
    f, g given as (ny, nx) arrays (row-major [y][x]).  Returns dict(x, y, u, v, ccmax)
    (each (ny_grid, nx_grid) like gridfit's output), the final-level region centres and
    values (tempx, tempy, tempu, tempv, tempPhi) and the level count."""
    f = np.asarray(f, np.float64).T      # [x, y]
    g = np.asarray(g, np.float64).T
    nx, ny = f.shape
    gridx = list(gridx)
    gridy = list(gridy)
    gap = 1 + int(np.floor(0.75 * winsize + 0.5))   # MATLAB round
    if gridx[0] < gap: gridx[0] = gap
    if gridy[0] < gap: gridy[0] = gap
    if gridx[1] > nx - gap: gridx[1] = nx - gap
    if gridy[1] > ny - gap: gridy[1] = ny - gap
    # level 1
    C = f[gridx[0] - 1:gridx[1], gridy[0] - 1:gridy[1]]
    D = g[gridx[0] - 1:gridx[1], gridy[0] - 1:gridy[1]]
    full = ncc_full(C, D)
    if full is None:
        raise ValueError("funIntegerSearchMg: constant reference ROI (normxcorr2 raises)")
    v1, u1, _ = oracle.findpeak(full)
    m, n = C.shape
    zx, zy = math.ceil((full.shape[0] + 1) / 2), math.ceil((full.shape[1] + 1) / 2)
    u = u1 - zx
    v = v1 - zy
    r0 = _clamp_region([gridx[0], gridx[1], gridy[0], gridy[1],
                        gridx[0] + _ceil(u), gridx[1] + _ceil(u),
                        gridy[0] + _ceil(v), gridy[1] + _ceil(v)], nx, ny, gap)
    regions = [r0]
    uc, vc = [_ceil(u)], [_ceil(v)]
    wx, wy = r0[1] - r0[0], r0[3] - r0[2]
    levels = 1
    while True:
        levels += 1
        ratio = wx / wy
        new, un, vn = [], [], []
        for r, uu, vv in zip(regions, uc, vc):
            xf1, xf2, yf1, yf2, xg1, xg2, yg1, yg2 = r
            if ratio > 2:
                if xf2 - xf1 > winsize:
                    mf, mg = 0.5 * (xf1 + xf2), 0.5 * (xg1 + xg2)
                    kids = [[xf1, mf, yf1, yf2, xg1, mg, yg1, yg2], [mf, xf2, yf1, yf2, mg, xg2, yg1, yg2]]
                else:
                    kids = [list(r), list(r)]
            elif ratio < 0.5:
                if yf2 - yf1 > winsize:
                    mf, mg = 0.5 * (yf1 + yf2), 0.5 * (yg1 + yg2)
                    kids = [[xf1, xf2, yf1, mf, xg1, xg2, yg1, mg], [xf1, xf2, mf, yf2, xg1, xg2, mg, yg2]]
                else:
                    kids = [list(r), list(r)]
            else:
                if xf2 - xf1 > winsize and yf2 - yf1 > winsize:
                    mfx, mgx = 0.5 * (xf1 + xf2), 0.5 * (xg1 + xg2)
                    mfy, mgy = 0.5 * (yf1 + yf2), 0.5 * (yg1 + yg2)
                    kids = [[xf1, mfx, yf1, mfy, xg1, mgx, yg1, mgy],
                            [mfx, xf2, yf1, mfy, mgx, xg2, yg1, mgy],
                            [xf1, mfx, mfy, yf2, xg1, mgx, mgy, yg2],
                            [mfx, xf2, mfy, yf2, mgx, xg2, mgy, yg2]]
                else:
                    kids = [list(r)] * 4
            for k in kids:
                new.append(k)
                un.append(uu)
                vn.append(vv)
        if ratio > 2:
            wxn, wyn = wx / 2, wy
        elif ratio < 0.5:
            wxn, wyn = wx, wy / 2
        else:
            wxn, wyn = wx / 2, wy / 2
        phi = []
        for i, r in enumerate(new):
            du, dv, mf = _region_peak(f, g, r, nx, ny)
            un[i] = un[i] + du
            vn[i] = vn[i] + dv
            phi.append(mf)
        if wxn < winsize and wyn < winsize:
            break
        regions, uc, vc = [], [], []
        for r, uu, vv in zip(new, un, vn):
            cu, cv = _ceil(uu), _ceil(vv)
            rr = _clamp_region([r[0], r[1], r[2], r[3], r[0] + cu, r[1] + cu, r[2] + cv, r[3] + cv],
                               nx, ny, gap)
            regions.append(rr)
            uc.append(cu)
            vc.append(cv)
        wx, wy = wxn, wyn
    tempx = np.array([0.5 * (r[0] + r[1]) for r in new])
    tempy = np.array([0.5 * (r[2] + r[3]) for r in new])
    tempu, tempv, tempphi = np.array(un), np.array(vn), np.array(phi)
    # :262-270 the grid, away from the borders the final displacements need
    i1, i2 = int(np.argmin(tempx)), int(np.argmax(tempx))
    j1, j2 = int(np.argmin(tempy)), int(np.argmax(tempy))
    gx1 = _ceil(1 + 1.25 * winsize + tempu[i1])
    gx2 = _ceil(1 + 1.25 * winsize + tempu[i2])
    gy1 = _ceil(1 + 1.25 * winsize + tempv[j1])
    gy2 = _ceil(1 + 1.25 * winsize + tempv[j2])
    if gridx[0] < gx1: gridx[0] = gx1
    if gridy[0] < gy1: gridy[0] = gy1
    if gridx[1] > nx - gx2 + 1: gridx[1] = nx - gx2 + 1
    if gridy[1] > ny - gy2 + 1: gridy[1] = ny - gy2 + 1
    xlist = np.arange(gridx[0], gridx[1] + 1e-9, winstepsize)
    ylist = np.arange(gridy[0], gridy[1] + 1e-9, winstepsize)
    sel = ((tempx > xlist.min()) & (tempx < xlist.max()) & (tempy > ylist.min()) &
           (tempy < ylist.max()))
    out = {}
    for name, z in (("u", tempu), ("v", tempv), ("ccmax", tempphi)):
        k = sel & np.isfinite(z)
        out[name] = oracle.gridfit(tempx[k], tempy[k], z[k], xlist, ylist)
    X, Y = np.meshgrid(xlist, ylist)
    out.update(x=X, y=Y, tempx=tempx, tempy=tempy, tempu=tempu, tempv=tempv, tempphi=tempphi,
               levels=levels, sel=sel, xlist=xlist, ylist=ylist)
    return out
