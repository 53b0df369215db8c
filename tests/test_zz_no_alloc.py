"""The allocation-free G x (gra_G_vec_mul_noalloc) called through the C ABI
with device pointers, against gra_G_vec_mul on the same G: the same kernels
and launch shape on the same stored values, so the products must agree bit
for bit -- for FP64, FP32 and BF16 storage, and for two calls in a row with
the second on a smaller G (the shared scratch keeps the first call's
contents: the staged x's zero padding and the partial sums are rewritten by
every call).  A G beyond the scratch's fixed size is refused."""
import ctypes

import numpy as np
import pytest

from graada3d_amd import mesh as M
from graada3d_amd.engine import Context, Fwd, GraError

pytestmark = pytest.mark.gpu

_hip = ctypes.CDLL("libamdhip64.so")
_hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
_hip.hipFree.argtypes = [ctypes.c_void_p]
_hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
H2D, D2H = 1, 2


def _noalloc(f, x):
    """G x by the allocation-free entry point (device buffers made here, by
    the test, around the call)."""
    dx, dy = ctypes.c_void_p(), ctypes.c_void_p()
    assert _hip.hipMalloc(ctypes.byref(dx), 8 * f.Nm) == 0
    assert _hip.hipMalloc(ctypes.byref(dy), 8 * f.Nd) == 0
    try:
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert _hip.hipMemcpy(dx, x.ctypes.data_as(ctypes.c_void_p), 8 * f.Nm, H2D) == 0
        L = f.ctx.L
        f.ctx.check(L.gra_G_vec_mul_noalloc(f.ctx.h, dx, dy), "G_vec_mul_noalloc")
        f.ctx.check(L.gra_sync(f.ctx.h), "sync")
        y = np.empty(f.Nd)
        assert _hip.hipMemcpy(y.ctypes.data_as(ctypes.c_void_p), dy, 8 * f.Nd, D2H) == 0
        return y
    finally:
        _hip.hipFree(dx)
        _hip.hipFree(dy)


@pytest.fixture(scope="module")
def own_ctx():
    c = Context(0)
    yield c
    c.close()


def test_noalloc_matches_the_allocating_product(own_ctx):
    rng = np.random.default_rng(3)
    big = Fwd(M.regular_mesh((0.0, 1000.0), 20, (0.0, 900.0), 18, (0.0, 450.0), 9),
              M.grid_points(40, -50.0, 1050.0, 40, -50.0, 950.0, -5.0), 0x3FE, ctx=own_ctx)
    small = Fwd(M.regular_mesh((0.0, 300.0), 6, (0.0, 250.0), 5, (0.0, 200.0), 4),
                M.grid_points(7, -25.0, 325.0, 6, 0.0, 250.0, -10.0), 0x12, ctx=own_ctx)
    for store in ("f64", "f32", "bf16"):
        x = rng.standard_normal(big.Nm)
        big.compute_G(store)
        y_ref = big.G_vec_mul(x)
        y = _noalloc(big, x)
        assert np.array_equal(y, y_ref), store
        # a second call in a row, on a smaller G
        xs = rng.standard_normal(small.Nm)
        small.compute_G(store)
        ys_ref = small.G_vec_mul(xs)
        ys = _noalloc(small, xs)
        assert np.array_equal(ys, ys_ref), store


def test_noalloc_refuses_a_G_beyond_its_scratch(own_ctx):
    # 4.3 M rows of 64 columns: ld + Nd exceeds the 4,194,304 doubles
    pr = M.regular_mesh((0.0, 400.0), 4, (0.0, 400.0), 4, (0.0, 200.0), 4)
    pts = M.grid_points(2100, -10.0, 410.0, 2050, -10.0, 410.0, -5.0)
    f = Fwd(pr, pts, 0x2, ctx=own_ctx)
    f.compute_G("bf16")
    with pytest.raises(GraError, match="allocation-free"):
        _noalloc(f, np.ones(f.Nm))
