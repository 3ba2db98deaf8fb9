"""aldic_prep_noalloc (the allocation-free pre-pass on caller-owned device buffers) against
aldic_set_images, the entry point whose kernels it shares: bit-identical FP32 planes, on
two calls in a row that reuse the same device buffers, the second with a smaller image
(so the library's fixed work array and the outputs hold the first call's contents).
Device memory comes from the HIP runtime the library itself links (no torch)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    h = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def _ok(rc, what):
    assert rc == 0, f"{what} returned {rc}"


def test_prep_noalloc_matches_set_images(pkg):
    ctx = pkg.Context(0)
    L = ctx.L
    hip = _hip()
    cases = [  # (nx, ny, dtype, with mask, roi): the second input is smaller
        (257, 193, np.uint8, True, [5, 250, 3, 190]),
        (130, 97, np.uint16, False, None),
        (64, 40, np.float64, True, [1, 64, 1, 40]),
    ]
    nmax = cases[0][0] * cases[0][1]
    bufs = {}
    for name, nbytes in (("f", 8 * nmax), ("g", 8 * nmax), ("m", nmax), ("fo", 4 * nmax),
                         ("go", 4 * nmax), ("gxo", 4 * nmax), ("gyo", 4 * nmax)):
        p = C.c_void_p()
        _ok(hip.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        bufs[name] = p
    try:
        rng = np.random.default_rng(7)
        for nx, ny, dt, with_mask, roi in cases:
            if dt == np.float64:
                f = rng.normal(100.0, 20.0, (ny, nx))
                g = rng.normal(100.0, 20.0, (ny, nx))
            else:
                hi = 255 if dt == np.uint8 else 4095
                f = rng.integers(0, hi + 1, (ny, nx)).astype(dt)
                g = rng.integers(0, hi + 1, (ny, nx)).astype(dt)
            mask = (rng.random((ny, nx)) > 0.2).astype(np.uint8) if with_mask else None
            ctx.set_images(f, g, mask=mask, roi=roi)
            ref = ctx.get_images()

            _ok(hip.hipMemcpy(bufs["f"], f.ctypes.data, f.nbytes, H2D), "hipMemcpy f")
            _ok(hip.hipMemcpy(bufs["g"], g.ctypes.data, g.nbytes, H2D), "hipMemcpy g")
            if mask is not None:
                _ok(hip.hipMemcpy(bufs["m"], mask.ctypes.data, mask.nbytes, H2D), "hipMemcpy m")
            for k in ("fo", "go", "gxo", "gyo"):  # NaN pattern: every output must be written
                _ok(hip.hipMemset(bufs[k], 0xFF, 4 * nmax), "hipMemset")
            r = None if roi is None else np.asarray(roi, np.int32)
            rc = L.aldic_prep_noalloc(ctx.h, bufs["f"], bufs["g"], nx, ny,
                                      {np.uint8: 0, np.uint16: 1, np.float64: 3}[dt],
                                      bufs["m"] if mask is not None else None,
                                      None if r is None else r.ctypes.data,
                                      bufs["fo"], bufs["go"], bufs["gxo"], bufs["gyo"], None)
            _ok(rc, "aldic_prep_noalloc")
            _ok(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
            for k, want in zip(("fo", "go", "gxo", "gyo"), ref):
                got = np.empty((ny, nx), np.float32)
                _ok(hip.hipMemcpy(got.ctypes.data, bufs[k], got.nbytes, D2H), "hipMemcpy out")
                assert np.array_equal(got.astype(np.float64), want), \
                    f"{k} differs for {nx}x{ny} {np.dtype(dt).name}"
        # sizes beyond the limit are rejected before anything is launched
        rc = L.aldic_prep_noalloc(ctx.h, bufs["f"], bufs["g"], 1 << 22, 8, 0, None, None,
                                  bufs["fo"], bufs["go"], bufs["gxo"], bufs["gyo"], None)
        assert rc == -1
        rc = L.aldic_prep_noalloc(ctx.h, bufs["f"], bufs["g"], 4, 4, 0, None, None,
                                  bufs["fo"], bufs["go"], bufs["gxo"], bufs["gyo"], None)
        assert rc == -1
    finally:
        _ok(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
        for p in bufs.values():
            hip.hipFree(p)
        ctx.close()
