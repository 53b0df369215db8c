"""The engine's per-(point, prism) math, compiled for the host (test build of
graada3d_amd/csrc/prism_math.hpp with g++ and the same libm as the oracle), is
bit-identical to the oracle on random, grid-aligned, on-face and degenerate
geometry, for every component subset.  This pins the axis-aligned
specialisation of the reference's vector algebra; on the GPU only the device
atan/log differ."""
import ctypes
import os

import numpy as np
import pytest

import oracle

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hostcheck", "libhostcheck.so")


@pytest.fixture(scope="module")
def hc():
    if not os.path.exists(LIB):
        from graada3d_amd.build import build_hostcheck
        build_hostcheck()
    L = ctypes.CDLL(LIB)
    dp = ctypes.POINTER(ctypes.c_double)
    L.hc_prism_field.argtypes = [dp, dp, ctypes.c_double, ctypes.c_uint32, dp]
    return L


def call(L, pt, pr, rho, flag):
    dp = ctypes.POINTER(ctypes.c_double)
    out = np.zeros(10)
    pt = np.ascontiguousarray(pt, float)
    pr = np.ascontiguousarray(pr, float)
    st = L.hc_prism_field(pt.ctypes.data_as(dp), pr.ctypes.data_as(dp), rho, flag,
                          out.ctypes.data_as(dp))
    return out, st


FLAGS = [0x3FE, 0x12, 0x2, 0x10, 0x80, 0x100, 0x20, 0x40, 0x200, 0x4, 0x8, 0x180, 0x240, 0x72]


def test_bit_exact_vs_oracle(hc):
    rng = np.random.default_rng(1)
    checked = degenerate = 0
    for trial in range(6000):
        pr = np.sort(rng.uniform(-500, 1500, (3, 2)), axis=1).ravel()
        if trial % 4 == 0:
            pr = np.round(pr / 50) * 50
            if np.any(pr[1::2] <= pr[0::2]):
                continue
        pt = rng.uniform(-1000, 2500, 3)
        if trial % 3 == 0:
            pt[0] = pr[rng.integers(0, 2)]
        if trial % 5 == 0:
            pt[1] = pr[2 + rng.integers(0, 2)]
        if trial % 7 == 0:
            pt[2] = pr[4 + rng.integers(0, 2)]
        if trial % 11 == 0:
            pt = np.round(pt / 50) * 50
        flag = FLAGS[trial % len(FLAGS)]
        rho = 1.0 if trial % 2 else rng.uniform(-900, 900)
        a, s1 = oracle.prism_field(pt, pr, rho, flag)
        b, s2 = call(hc, pt, pr, rho, flag)
        assert s1 == s2, (pt, pr, flag)
        checked += 1
        if s1:
            degenerate += 1
            continue
        assert np.array_equal(a, b, equal_nan=True), (pt, pr, hex(flag), a - b)
    assert checked > 5000 and degenerate > 20


def test_reference_grid_geometry_bit_exact(hc):
    """Points of the reference's golden grids sit on cell face planes and on
    edge extensions: every branch of the reference is taken."""
    from graada3d_amd import mesh as M
    prisms = M.regular_mesh((0.0, 2000.0), 40, (0.0, 2000.0), 40, (0.0, 1000.0), 20)
    pts = M.grid_points(41, 0.0, 2000.0, 41, 0.0, 2000.0, -0.1)
    rng = np.random.default_rng(2)
    for _ in range(1500):
        pt = pts[rng.integers(len(pts))]
        pr = prisms[rng.integers(len(prisms))]
        a, s1 = oracle.prism_field(pt, pr, 1.0, 0x3FE)
        b, s2 = call(hc, pt, pr, 1.0, 0x3FE)
        assert s1 == s2 == 0
        assert np.array_equal(a, b)
