"""CPU restatement of the learned initial guess's network (2d_aldic_amd/csrc/cnn.hip), for
the tests only: forward pass, mean squared error and backward pass in numpy.

With dev=True it reproduces the device's arithmetic: activations held in FP32, every GEMM
operand rounded to BF16 (round to nearest even) before the products, products and sums
in float64 (the device sums FP32 products in FP32, a difference far below one BF16 step
of the next operand).  With dev=False everything is float64, for finite-difference
checks of the backward pass.  This network has no counterpart in the reference; it is
pinned by finite differences (backward) and by its own definition (layer shapes below).

Layout: patches (B, 64, 64, 2) NHWC; parameters flat: W1 b1 .. W5 b5 with W row-major
(out x in), conv inputs ordered (ky, kx, ci).
  L1 4x4 stride 4, 2 -> 32, ReLU; L2 3x3 s2 p1, 32 -> 64, ReLU; L3 3x3 s2 p1, 64 -> 64,
  ReLU; L4 dense 1024 -> 128, ReLU; L5 dense 128 -> 6.
"""
import numpy as np

SHAPES = [(32, 32), (64, 288), (64, 576), (128, 1024), (6, 128)]   # (N, K) per layer
SCALE = np.array([0.05, 0.05, 0.05, 0.05, 16.0, 16.0])


def nparams():
    return sum(n * k + n for n, k in SHAPES)


def unpack(w):
    w = np.asarray(w, np.float64).ravel()
    out, o = [], 0
    for n, k in SHAPES:
        W = w[o:o + n * k].reshape(n, k)
        o += n * k
        b = w[o:o + n]
        o += n
        out.append((W, b))
    return out


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


def patchify(p):
    """(B, 64, 64, 2) -> (B*256, 32): rows (cy, cx), cols (ky, kx, c)."""
    B = p.shape[0]
    t = p.reshape(B, 16, 4, 16, 4, 2).transpose(0, 1, 3, 2, 4, 5)
    return t.reshape(B * 256, 32)


def conv_cols(x):
    """3x3 stride 2 pad 1 operand of x (B, H, H, C): rows (oy, ox), cols (ky, kx, ci)."""
    B, H, _, C = x.shape
    Ho = H // 2
    xp = np.zeros((B, H + 2, H + 2, C))
    xp[:, 1:H + 1, 1:H + 1] = x
    cols = np.empty((B, Ho, Ho, 3, 3, C))
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, ky, kx] = xp[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Ho:2]
    return cols.reshape(B * Ho * Ho, 9 * C)


def conv_cols_adjoint(d, B, H, C):
    """Adjoint of conv_cols: (B*Ho*Ho, 9C) -> (B, H, H, C)."""
    Ho = H // 2
    d = d.reshape(B, Ho, Ho, 3, 3, C)
    xp = np.zeros((B, H + 2, H + 2, C))
    for ky in range(3):
        for kx in range(3):
            xp[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Ho:2] += d[:, :, :, ky, kx]
    return xp[:, 1:H + 1, 1:H + 1]


def forward(w, patches, dev=True):
    """Network outputs y (B, 6) (scaled: F / 0.05, U / 16) and the cached operands."""
    L = unpack(w)
    rb = bf16 if dev else (lambda a: np.asarray(a, np.float64))
    st = _f32 if dev else (lambda a: a)
    B = patches.shape[0]
    cache = []
    A = patchify(np.asarray(patches, np.float64))
    h = None
    for li, (W, b) in enumerate(L):
        if li == 1:
            A = conv_cols(h.reshape(B, 16, 16, 32))
        elif li == 2:
            A = conv_cols(h.reshape(B, 8, 8, 64))
        elif li == 3:
            A = h.reshape(B, 1024)
        elif li == 4:
            A = h
        z = rb(A) @ rb(W).T + b
        h = st(np.maximum(z, 0.0)) if li < 4 else st(z)
        cache.append((A, h))
    return h, cache


def loss_and_grad(w, patches, targets, dev=True):
    """Mean squared error over B x 6 scaled outputs and its gradient (flat, like w)."""
    L = unpack(w)
    rb = bf16 if dev else (lambda a: np.asarray(a, np.float64))
    st = _f32 if dev else (lambda a: a)
    B = patches.shape[0]
    y, cache = forward(w, patches, dev)
    e = y - targets
    loss = float(np.mean(e ** 2))
    d = st(e * (2.0 / (B * 6)))
    grads = [None] * 5
    for li in range(4, -1, -1):
        W, b = L[li]
        A, h = cache[li]
        if li < 4:
            d = np.where(h > 0, d, 0.0)
        dW = rb(d).T @ rb(A)
        db = d.sum(axis=0)
        grads[li] = (dW, db)
        if li == 0:
            break
        dA = st(rb(d) @ rb(W))
        if li == 4:
            d = dA
        elif li == 3:
            d = dA.reshape(B * 16, 64)
        elif li == 2:
            d = st(conv_cols_adjoint(dA, B, 8, 64)).reshape(B * 64, 64)
        elif li == 1:
            d = st(conv_cols_adjoint(dA, B, 16, 32)).reshape(B * 256, 32)
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1]]) for g in grads])
    return loss, flat, y


def init_weights(seed=0):
    """He-normal weights, zero biases (FP32)."""
    rng = np.random.default_rng(seed)
    parts = []
    for n, k in SHAPES:
        parts.append(rng.normal(scale=np.sqrt(2.0 / k), size=n * k))
        parts.append(np.zeros(n))
    return np.concatenate(parts).astype(np.float32)
