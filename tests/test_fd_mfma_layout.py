"""The FP64 matrix-core GEMM of the structured Subpb2 solve (fem.hip, k_gemm_mfma, used
for grids with more than 128 nodes per row) emulated lane by lane on the CPU: every lane's
clamped fragment addresses and the zero factor past K as the kernel computes them, the
v_mfma_f64_16x16x4f64 operand and result layout (A 16 x 4: lane l holds (l % 16, l / 16);
B 4 x 16: lane l holds (l / 16, l % 16); D 16 x 16: lane l, element v holds
(l / 16 + 4 v, l % 16)), the chunks of 16 k steps summed per tile, and the guarded
stores, against numpy's product on the
strided, batched shapes of the four passes (row counts and K not multiples of 4 or 16).
A wrong lane mapping, a missing zero past K or a transposed stride fails here (rows
and columns past M, N are read clamped and never stored).

The accumulator layout was measured on the device: stored as row 4 (l / 16) + v, a
single product comes back with rows 1 and 4 (2 and 8, 3 and 12, ...) of every 16-row
tile exchanged.  That map is an involution, so the fused two-product pass of the small
grids, which applies it to both products, was right all along, while every single
product was not; test_row_map_cancels_in_two_products restates this."""
import numpy as np
import pytest


def _mfma(a, b, acc):
    """One v_mfma_f64_16x16x4f64 on 64-lane register images: a[l], b[l], acc[l][v]."""
    A = np.zeros((16, 4))
    B = np.zeros((4, 16))
    for l in range(64):
        A[l % 16, l // 16] = a[l]
        B[l // 16, l % 16] = b[l]
    D = A @ B
    out = acc.copy()
    for l in range(64):
        for v in range(4):
            out[l, v] += D[l // 16 + 4 * v, l % 16]
    return out


def _gemm_mfma(TM, M, N, K, A, sam, sak, B, sbk, sbn, C, scm, scn, scale=None):
    """k_gemm_mfma<TM, SCALE> for one batch entry, wave by wave: chunks of 16 k steps,
    one MFMA chain per output tile from zero, chunk results summed."""
    MT = TM // 16
    tilesN, tilesM = (N + 31) // 32, (M + TM - 1) // TM
    ks = (K + 3) // 4
    for tile in range(tilesM * tilesN):
        m0, n0 = (tile // tilesN) * TM, (tile % tilesN) * 32
        total = [[np.zeros((64, 4)) for _ in range(2)] for _ in range(MT)]
        for s0 in range(0, ks, 16):
            a = np.zeros((16, MT, 64))
            b = np.zeros((16, 2, 64))
            for u in range(16):
                for l in range(64):
                    li, lk = l & 15, l >> 4
                    k = 4 * (s0 + u) + lk
                    kc = min(k, K - 1)
                    km = 1.0 if k < K else 0.0
                    for i in range(MT):
                        a[u, i, l] = A[min(m0 + 16 * i + li, M - 1) * sam + kc * sak] * km
                    for j in range(2):
                        b[u, j, l] = B[kc * sbk + min(n0 + 16 * j + li, N - 1) * sbn]
            for i in range(MT):
                for j in range(2):
                    part = np.zeros((64, 4))
                    for u in range(16):
                        part = _mfma(a[u, i], b[u, j], part)
                    total[i][j] += part
        for l in range(64):
            li, lk = l & 15, l >> 4
            for i in range(MT):
                for j in range(2):
                    c = n0 + 16 * j + li
                    if c >= N:
                        continue
                    for v in range(4):
                        r = m0 + 16 * i + lk + 4 * v
                        if r >= M:
                            continue
                        o = total[i][j][l, v]
                        if scale is not None:
                            o /= scale(r, c)
                        C[r * scm + c * scn] = o


@pytest.mark.parametrize("TM,gy", [(16, 37), (32, 37), (16, 70)])
def test_first_and_last_pass_shapes(TM, gy):
    # T1 = Vy^T B and X = Vy T1 on a gy x 2gx nodal array (gx = 21; gy = 70 takes two
    # chunks of 16 k steps)
    rng = np.random.default_rng(1)
    gx = 21
    Vy = rng.normal(size=(gy, gy))
    b = rng.normal(size=(gy, 2 * gx))
    T1 = np.full(gy * 2 * gx, np.nan)
    # A(m, k) = Vy[k][m]: sam 1, sak gy;  B(k, n) = b[k][n]: sbk 2gx, sbn 1
    _gemm_mfma(TM, gy, 2 * gx, gy, Vy.ravel(), 1, gy, b.ravel(), 2 * gx, 1, T1, 2 * gx, 1)
    assert np.allclose(T1.reshape(gy, 2 * gx), Vy.T @ b, rtol=1e-13, atol=1e-13)
    X = np.full(gy * 2 * gx, np.nan)
    _gemm_mfma(TM, gy, 2 * gx, gy, Vy.ravel(), gy, 1, T1, 2 * gx, 1, X, 2 * gx, 1)
    assert np.allclose(X.reshape(gy, 2 * gx), Vy @ (Vy.T @ b), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("TM", [16, 32])
def test_middle_passes_per_component_with_scale(TM):
    # T2 = (T1 Vx) ./ D per component: A(m, k) = T1[m][2k + comp], B(k, n) = Vx[k][n]
    rng = np.random.default_rng(2)
    gy, gx = 19, 35
    Vx = rng.normal(size=(gx, gx))
    T1 = rng.normal(size=(gy, 2 * gx))
    ly, lx = rng.uniform(0, 2, gy), rng.uniform(0, 2, gx)
    mu, cK = 1e-3, 0.7
    T2 = np.full(gy * 2 * gx, np.nan)
    for comp in range(2):                       # the batch offset is 1 on A and C
        _gemm_mfma(TM, gy, gx, gx, T1.ravel()[comp:], 2 * gx, 2, Vx.ravel(), gx, 1,
                   T2[comp:], 2 * gx, 2, scale=lambda r, c: mu + cK * (ly[r] + lx[c]))
    D = mu + cK * (ly[:, None] + lx[None, :])
    for comp in range(2):
        ref = (T1[:, comp::2] @ Vx) / D
        assert np.allclose(T2.reshape(gy, 2 * gx)[:, comp::2], ref, rtol=1e-12, atol=1e-13)
    # T1' = T2 Vx^T: B(k, n) = Vx[n][k], sbk 1, sbn gx
    T3 = np.full(gy * 2 * gx, np.nan)
    for comp in range(2):
        _gemm_mfma(TM, gy, gx, gx, T2[comp:], 2 * gx, 2, Vx.ravel(), 1, gx, T3[comp:], 2 * gx, 2)
    for comp in range(2):
        ref = T2.reshape(gy, 2 * gx)[:, comp::2] @ Vx.T
        assert np.allclose(T3.reshape(gy, 2 * gx)[:, comp::2], ref, rtol=1e-12, atol=1e-12)


def _strip_product(L, B, rowmap):
    """One 16-row strip of L B by emulated MFMAs (L 16 x K, B K x n, n a multiple of 16),
    element v of lane l stored to row rowmap(l // 16, v)."""
    K, n = B.shape
    out = np.zeros((16, n))
    for c0 in range(0, n, 16):
        acc = np.zeros((64, 4))
        for st in range(0, K, 4):
            a = np.array([L[l % 16, st + l // 16] for l in range(64)])
            b = np.array([B[st + l // 16, c0 + l % 16] for l in range(64)])
            acc = _mfma(a, b, acc)
        for l in range(64):
            for v in range(4):
                out[rowmap(l // 16, v), c0 + l % 16] = acc[l, v]
    return out


def test_row_map_cancels_in_two_products():
    rng = np.random.default_rng(3)
    L, B, R = rng.normal(size=(16, 8)), rng.normal(size=(8, 32)), rng.normal(size=(32, 16))
    right = lambda lk, v: lk + 4 * v
    wrong = lambda lk, v: 4 * lk + v
    assert np.allclose(_strip_product(L, B, right), L @ B, rtol=1e-13, atol=1e-13)
    P = _strip_product(L, B, wrong)
    assert not np.allclose(P, L @ B)
    assert np.allclose(P[[0, 4, 8, 12, 1, 5, 9, 13, 2, 6, 10, 14, 3, 7, 11, 15]], L @ B)
    # the second product takes the strip as its A rows and stores with the same map
    assert np.allclose(_strip_product(P, R, wrong), L @ B @ R, rtol=1e-12, atol=1e-12)
