"""CPU restatement of the learned initial guess's network (2d_aldic_amd/csrc/cnn.hip), for
the tests only: forward pass, mean squared error and backward pass in numpy.

The network (no counterpart in the reference, whose initial guess is the FFT search of
IntegerSearch.m) is a siamese correlation network on 4 x 4 pixel cells:

  E1  4 x 4 stride 4 patchify of one image, 1 -> 16 channels, bias, ReLU
  E2  3 x 3 stride 1 'valid' convolution, 16 -> 16, bias (linear)
  N   each cell's 16 features divided by (their Euclidean norm + 1e-3)
  C   cost volumes: quadrant q of the 6 x 6 central cells p of the f patch (cells 5..10 of
      16), displacement d = (dx, dy) in -5..5 cells: Q[q, d] = 1/9 sum_{p in q} sum_ch
      nf[p] ng[p + d]; the global volume G = mean of the four
  S   soft-argmax of each of the five volumes: w = softmax(beta V), m = sum_d w (4 dx, 4 dy)
  L   y = Wl m + bl (10 -> 6), the scaled outputs (u_x, v_x, u_y, v_y) / 0.05, (u, v) / 4

Input per node: a 72 x 72 pair (f, g) whose central 64 x 64 is the node's patch (first pixel
x0 - 32 for the 1-based node x0); the 4-pixel ring is the context of E2 at the patch's
edge cells, so the features of a patch equal those of the same cells in a feature map of
the whole image (the device's inference path).  Quadrants are numbered (qy, qx) ->
2 qy + qx; displacements d index (dy + 5) * 11 + (dx + 5).

With dev=True it reproduces the device's arithmetic: E1 and E2 run on BF16 operands
(round to nearest even) with float64 sums here (FP32 on the device, far below one BF16
step), activations stored in FP32; the rest is FP32 on the device and float64 here.
With dev=False everything is float64, for finite-difference checks.

Layout: patches (B, 72, 72, 2) NHWC; flat parameters W1 b1 W2 b2 beta Wl bl with W
row-major (out x in), E2 inputs ordered (ky, kx, ci).
"""
import numpy as np

C = 16
R = 5
ND = (2 * R + 1) ** 2
SHAPES = [(C, 16), (C, 9 * C)]                 # (N, K) of the two GEMM layers
SCALE = np.array([0.05, 0.05, 0.05, 0.05, 4.0, 4.0])
EPS = 1e-3
BETA0 = 20.0
PATCH = 72
N1 = PATCH // 4                                # 18 x 18 E1 cells
N2 = N1 - 2                                    # 16 x 16 E2 cells
P0 = 5                                         # first central cell (of 16)
DX = np.tile(np.arange(-R, R + 1), 2 * R + 1).astype(np.float64)
DY = np.repeat(np.arange(-R, R + 1), 2 * R + 1).astype(np.float64)


def nparams():
    return C * 16 + C + C * 9 * C + C + 1 + 6 * 10 + 6


def unpack(w):
    w = np.asarray(w, np.float64).ravel()
    o = 0

    def take(n):
        nonlocal o
        r = w[o:o + n]
        o += n
        return r
    W1 = take(C * 16).reshape(C, 16)
    b1 = take(C)
    W2 = take(C * 9 * C).reshape(C, 9 * C)
    b2 = take(C)
    beta = float(take(1)[0])
    Wl = take(60).reshape(6, 10)
    bl = take(6)
    return W1, b1, W2, b2, beta, Wl, bl


def pack(W1, b1, W2, b2, beta, Wl, bl):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, [beta], np.ravel(Wl), bl])


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


def patchify(img):
    """(M, 72, 72) -> (M * 324, 16): rows (cy, cx), cols (ky, kx)."""
    M = img.shape[0]
    t = img.reshape(M, N1, 4, N1, 4).transpose(0, 1, 3, 2, 4)
    return t.reshape(M * N1 * N1, 16)


def conv_cols(x):
    """3 x 3 'valid' operand of x (M, 18, 18, C): rows (oy, ox) over 16 x 16, cols (ky, kx, ci)."""
    M, H, _, Ci = x.shape
    Ho = H - 2
    cols = np.empty((M, Ho, Ho, 3, 3, Ci))
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, ky, kx] = x[:, ky:ky + Ho, kx:kx + Ho]
    return cols.reshape(M * Ho * Ho, 9 * Ci)


def conv_cols_adjoint(d, M, H, Ci):
    """Adjoint of conv_cols: (M * (H-2)^2, 9 Ci) -> (M, H, H, Ci)."""
    Ho = H - 2
    d = d.reshape(M, Ho, Ho, 3, 3, Ci)
    x = np.zeros((M, H, H, Ci))
    for ky in range(3):
        for kx in range(3):
            x[:, ky:ky + Ho, kx:kx + Ho] += d[:, :, :, ky, kx]
    return x


def encode(W1, b1, W2, b2, imgs, dev):
    """imgs (M, 72, 72) -> E1 operand, E1 output h1 (M, 18, 18, C), E2 operand and output
    e (M, 16, 16, C), normalised n."""
    rb = bf16 if dev else (lambda a: np.asarray(a, np.float64))
    st = _f32 if dev else (lambda a: a)
    M = imgs.shape[0]
    A1 = patchify(np.asarray(imgs, np.float64))
    h1 = st(np.maximum(rb(A1) @ rb(W1).T + b1, 0.0)).reshape(M, N1, N1, C)
    A2 = conv_cols(h1)
    e = st(rb(A2) @ rb(W2).T + b2).reshape(M, N2, N2, C)
    r = np.sqrt(np.sum(e * e, axis=-1, keepdims=True))
    n = st(e / (r + EPS))
    return A1, h1, A2, e, r, n


def correlate(nf, ng):
    """Quadrant cost volumes Q (B, 4, ND) from normalised features (B, 16, 16, C)."""
    B = nf.shape[0]
    Q = np.zeros((B, 4, ND))
    fc = nf[:, P0:P0 + 6, P0:P0 + 6]
    for k in range(ND):
        dx, dy = int(DX[k]), int(DY[k])
        gs = ng[:, P0 + dy:P0 + dy + 6, P0 + dx:P0 + dx + 6]
        c = np.sum(fc * gs, axis=-1)                       # (B, 6, 6)
        for q in range(4):
            qy, qx = divmod(q, 2)
            Q[:, q, k] = c[:, 3 * qy:3 * qy + 3, 3 * qx:3 * qx + 3].sum(axis=(1, 2)) / 9.0
    return Q


def correlate_adjoint(dQ, nf, ng):
    """Gradients of sum(dQ * Q) with respect to nf and ng."""
    dnf = np.zeros_like(nf)
    dng = np.zeros_like(ng)
    fc = nf[:, P0:P0 + 6, P0:P0 + 6]
    # per central cell, the gradient of its quadrant's volume
    dqc = np.zeros((nf.shape[0], 6, 6, ND))
    for q in range(4):
        qy, qx = divmod(q, 2)
        dqc[:, 3 * qy:3 * qy + 3, 3 * qx:3 * qx + 3] = dQ[:, q][:, None, None, :] / 9.0
    for k in range(ND):
        dx, dy = int(DX[k]), int(DY[k])
        gs = ng[:, P0 + dy:P0 + dy + 6, P0 + dx:P0 + dx + 6]
        w = dqc[:, :, :, k][..., None]
        dnf[:, P0:P0 + 6, P0:P0 + 6] += w * gs
        dng[:, P0 + dy:P0 + dy + 6, P0 + dx:P0 + dx + 6] += w * fc
    return dnf, dng


def volumes(Q):
    """The five volumes (B, 5, ND): global mean, then the four quadrants."""
    return np.concatenate([Q.mean(axis=1, keepdims=True), Q], axis=1)


def softargmax(V, beta):
    """m (B, 5, 2) in pixels and the weights w (B, 5, ND)."""
    z = beta * V
    z = z - z.max(axis=-1, keepdims=True)
    w = np.exp(z)
    w /= w.sum(axis=-1, keepdims=True)
    m = np.stack([4.0 * (w * DX).sum(-1), 4.0 * (w * DY).sum(-1)], axis=-1)
    return m, w


def forward(w, patches, dev=True):
    """Scaled outputs y (B, 6) and the cached intermediates."""
    W1, b1, W2, b2, beta, Wl, bl = unpack(w)
    p = np.asarray(patches, np.float64)
    B = p.shape[0]
    imgs = np.concatenate([p[..., 0], p[..., 1]], axis=0)      # f patches, then g patches
    A1, h1, A2, e, r, n = encode(W1, b1, W2, b2, imgs, dev)
    nf, ng = n[:B], n[B:]
    Q = correlate(nf, ng)
    V = volumes(Q)
    m, sw = softargmax(V, beta)
    feats = m.reshape(B, 10)
    y = feats @ Wl.T + bl
    if dev:
        y = _f32(y)
    return y, dict(A1=A1, h1=h1, A2=A2, e=e, r=r, n=n, Q=Q, V=V, w=sw, m=m, feats=feats)


def init_weights(seed=0):
    """Random encoder (He-normal), beta = 20 and a head that starts as the geometric
    estimate: the global soft-argmax for (u, v), differences of the quadrant ones over
    their 12 px spacing for the gradients (FP32)."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(scale=np.sqrt(2.0 / 16), size=(C, 16))
    W2 = rng.normal(scale=np.sqrt(1.0 / (9 * C)), size=(C, 9 * C))
    Wl = np.zeros((6, 10))
    Wl[4, 0] = 1.0 / SCALE[4]
    Wl[5, 1] = 1.0 / SCALE[5]
    for comp, row_x, row_y in ((0, 0, 2), (1, 1, 3)):
        for q in range(4):
            qy, qx = divmod(q, 2)
            Wl[row_x, 2 + 2 * q + comp] = (1.0 if qx else -1.0) / 24.0 / SCALE[row_x]
            Wl[row_y, 2 + 2 * q + comp] = (1.0 if qy else -1.0) / 24.0 / SCALE[row_y]
    return pack(W1, np.zeros(C), W2, np.zeros(C), BETA0, Wl, np.zeros(6)).astype(np.float32)


def loss_and_grad(w, patches, targets, dev=True):
    """Mean squared error over the B x 6 scaled outputs and its gradient (flat, like w)."""
    W1, b1, W2, b2, beta, Wl, bl = unpack(w)
    rb = bf16 if dev else (lambda a: np.asarray(a, np.float64))
    st = _f32 if dev else (lambda a: a)
    B = np.asarray(patches).shape[0]
    y, c = forward(w, patches, dev)
    err = y - targets
    loss = float(np.mean(err ** 2))
    dy = err * (2.0 / (B * 6))
    dWl = dy.T @ c["feats"]
    dbl = dy.sum(axis=0)
    dm = (dy @ Wl).reshape(B, 5, 2)
    # soft-argmax: m = sum w a, w = softmax(beta V)
    a = np.stack([4.0 * DX, 4.0 * DY], axis=-1)                # (ND, 2)
    s = (a[None, None] * dm[:, :, None, :]).sum(-1)             # (B, 5, ND): a_d . dm
    sbar = (c["w"] * s).sum(-1, keepdims=True)
    dz = c["w"] * (s - sbar)                                    # d loss / d (beta V)
    dV = beta * dz
    dbeta = float(np.sum(dz * c["V"]))
    dQ = dV[:, 1:] + dV[:, :1] / 4.0
    nf, ng = c["n"][:B], c["n"][B:]
    dnf, dng = correlate_adjoint(dQ, nf, ng)
    dn = st(np.concatenate([dnf, dng], axis=0))
    e, r = c["e"], c["r"]
    de = dn / (r + EPS) - e * np.sum(e * dn, axis=-1, keepdims=True) / (
        np.maximum(r, 1e-30) * (r + EPS) ** 2)
    de = st(de).reshape(-1, C)
    dW2 = rb(de).T @ rb(c["A2"])
    db2 = de.sum(axis=0)
    dA2 = st(rb(de) @ rb(W2))
    M = 2 * B
    dh1 = st(conv_cols_adjoint(dA2, M, N1, C)).reshape(-1, C)
    dh1 = np.where(c["h1"].reshape(-1, C) > 0, dh1, 0.0)
    dW1 = rb(dh1).T @ rb(c["A1"])
    db1 = dh1.sum(axis=0)
    flat = pack(dW1, db1, dW2, db2, dbeta, dWl, dbl)
    return loss, flat, y
