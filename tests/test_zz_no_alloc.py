"""The no-allocation entry point wm_pc_apply_noalloc against wm_pc_apply,
through the C ABI: the same parameters on the same images give the same bits,
on two calls in a row (the second with smaller images) through the one
static scratch array.  Device buffers come from the HIP runtime the library
itself uses."""
import ctypes

import numpy as np
import pytest

import wumingpic2d_amd as wm
from wmhelpers import make_ctx, state

pytestmark = pytest.mark.gpu


class _Dev:
    """hipMalloc'd buffer of the test (the entry point allocates nothing)."""

    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(nbytes)) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0

    def get(self, shape):
        a = np.empty(shape)
        assert self.hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(a.nbytes), 2) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def _hip():
    wm.load_library()
    hip = ctypes.CDLL("libamdhip64.so.7")       # already loaded by libwmpic
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def _pc_ctx(nx, ny, theta):
    p = state(nx, ny, 1, 3)[0]
    ctx = make_ctx(p)
    ctx.pc_set_params(theta)
    return ctx


def test_noalloc_apply_matches_apply_on_two_calls():
    hip = _hip()
    rng = np.random.default_rng(5)
    ctx64 = make_ctx(state(64, 48, 1, 3)[0])
    ctx64.pc_init(11)
    theta = ctx64.pc_params()
    theta = theta + np.float32(0.05) * rng.standard_normal(theta.shape).astype(np.float32)
    ctx64.pc_set_params(theta)
    ctx32 = _pc_ctx(32, 24, theta)
    big = rng.standard_normal((3, 48, 64))
    small = rng.standard_normal((3, 24, 32)) * 3.0
    rb, zb = _Dev(hip, big.nbytes), _Dev(hip, big.nbytes)
    try:
        for ctx, img, (ny, nx) in ((ctx64, big, (48, 64)), (ctx64, small, (24, 32))):
            ref = (ctx64 if nx == 64 else ctx32).pc_apply(img)
            r = _Dev(hip, img.nbytes)
            z = _Dev(hip, img.nbytes)
            r.put(img)
            ctx.pc_apply_noalloc(r.p.value, z.p.value, nx, ny)
            out = z.get(img.shape)
            r.free()
            z.free()
            assert np.array_equal(out, ref), np.abs(out - ref).max()
            assert np.abs(out).max() > 0
        # too large for the scratch array, or not multiples of 8: refused
        for nx, ny in ((1024, 512), (60, 48)):
            with pytest.raises(wm.WmError) as e:
                ctx64.pc_apply_noalloc(rb.p.value, zb.p.value, nx, ny)
            assert e.value.code == wm.WM_ERR_ARG
    finally:
        rb.free()
        zb.free()
        ctx64.close()
        ctx32.close()
