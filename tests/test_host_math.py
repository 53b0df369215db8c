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


PAIRS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hostcheck",
                     "libhostcheck_pairs.so")


@pytest.fixture(scope="module")
def hcp():
    if not os.path.exists(PAIRS):
        from graada3d_amd.build import build_hostcheck
        build_hostcheck()
    L = ctypes.CDLL(PAIRS)
    dp = ctypes.POINTER(ctypes.c_double)
    L.hc_prism_field.argtypes = [dp, dp, ctypes.c_double, ctypes.c_uint32, dp]
    return L


def _gauss_field(pt, pr, n=40):
    """gz, gx, gy (mGal, unit density) of a prism by n^3-point Gauss-Legendre
    quadrature of the point-mass kernel: accurate to ~1e-15 far from the
    prism, where the closed forms lose digits to cancellation."""
    x, w = np.polynomial.legendre.leggauss(n)
    ax = [0.5 * (pr[2 * k + 1] - pr[2 * k]) * x + 0.5 * (pr[2 * k + 1] + pr[2 * k])
          for k in range(3)]
    aw = [0.5 * (pr[2 * k + 1] - pr[2 * k]) * w for k in range(3)]
    d = [ax[0][:, None, None] - pt[0], ax[1][None, :, None] - pt[1], ax[2][None, None, :] - pt[2]]
    r3 = (d[0] ** 2 + d[1] ** 2 + d[2] ** 2) ** 1.5
    W = aw[0][:, None, None] * aw[1][None, :, None] * aw[2][None, None, :]
    c = 6.673e-11 * 1e5
    return c * np.sum(W * d[2] / r3), c * np.sum(W * d[0] / r3), c * np.sum(W * d[1] / r3)


def test_pairwise_atan_form_vs_oracle(hcp):
    """The GPU's formulation (each atan(x) - atan(y) as one atan, the table
    log, the polynomial atan) built for the host against the reference's
    order: equal to the last bits on ordinary geometry (median ~1e-15 of the
    evaluation's largest component), at most ~3e-10 on far prisms, where
    both forms lose digits to cancellation (next test)."""
    rng = np.random.default_rng(1)
    dev = []
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
        a, s1 = oracle.prism_field(pt, pr, 1.0, flag)
        b, s2 = call(hcp, pt, pr, 1.0, flag)
        assert s1 == s2
        if s1:
            continue
        assert np.all(np.isfinite(b)) == np.all(np.isfinite(a))
        # relative to the evaluation's largest component, plus 1e-15
        # absolute: symmetric cases give exact zeros in the reference's order
        # and ~1e-17 here
        err, scale = np.abs(a - b).max(), np.abs(a).max()
        assert err <= 1e-9 * scale + 1e-15, (pt, pr, hex(flag), a, b)
        if scale > 1e-9:
            dev.append(err / scale)
    assert np.median(dev) < 1e-14


@pytest.mark.parametrize("pt,pr,bit", [
    ([2313.39794426, 410.01708091, 2291.57087313],
     [-447.90722066, -434.92850845, 882.50485342, 1380.14728176, 340.03089171, 343.19178741], 2),
    ([990.82114741, -886.42430457, -808.78812371],
     [528.32040258, 990.82114741, 894.12889471, 1173.68758559, 1027.0179202, 1027.27737156], 1),
    ([5000.0, 4000.0, -100.0], [0.0, 50.0, 0.0, 50.0, 300.0, 350.0], 1),
])
def test_far_field_both_forms_at_the_cancellation_floor(hc, hcp, pt, pr, bit):
    """Far from a small prism the face/edge closed form cancels large terms
    (the reference's order errs by 1e-10 at 2 km from a 13 m slab and by 1e-8
    at 6 km from a 50 m cube): against an independent quadrature of the same
    field, the pairwise form is no further off than the reference's order
    (within a factor 3, or 3e-10)."""
    pt, pr = np.array(pt), np.array(pr)
    flag = 1 << bit
    q = dict(zip((1, 2, 3), _gauss_field(pt, pr)))[bit]
    a, _ = call(hc, pt, pr, 1.0, flag)
    b, _ = call(hcp, pt, pr, 1.0, flag)
    assert abs(b[bit] - q) <= max(3e-10 * abs(q), 3.0 * abs(a[bit] - q))


def test_table_log_accuracy(hc):
    """fastlog.hpp's table_log (the GPU's log in the edge terms and the
    node-sum forward) against 50-digit logarithms: within 1.5 ulp where
    |ln v| >= 0.5 (1.002 ulp measured) and within 1.2e-16 absolutely on
    [0.5, 2]."""
    from decimal import Decimal, getcontext
    getcontext().prec = 50
    hc.hc_table_log.argtypes = [ctypes.c_double]
    hc.hc_table_log.restype = ctypes.c_double
    rng = np.random.default_rng(5)
    vals = np.concatenate([np.exp(rng.uniform(np.log(1e-9), np.log(1e9), 4000)),
                           rng.uniform(0.5, 2.0, 2000), 1.0 + rng.uniform(-1e-7, 1e-7, 200),
                           [1.0, 2.0, 0.5, 1.0078125, 1.9999999999]])
    for v in vals:
        exact = Decimal(float(v)).ln()
        got = hc.hc_table_log(float(v))
        err = abs(Decimal(got) - exact)
        if abs(exact) >= Decimal("0.5"):
            assert err <= Decimal(1.5 * np.spacing(abs(float(exact)))), v
        else:
            assert err <= Decimal("1.2e-16"), v


def test_atan_ratio_accuracy_and_special_values(hc):
    """fastatan.hpp's atan_ratio(n, d) = atan(n / d) (the GPU's atan in the
    assembly and the node-sum forward): within 2 ulp of a 50-digit atan of
    the exact quotient, and the same special values as atan(n / d) (signed
    zeros, infinities, NaN)."""
    from decimal import Decimal, getcontext
    getcontext().prec = 50
    hc.hc_atan_ratio.argtypes = [ctypes.c_double, ctypes.c_double]
    hc.hc_atan_ratio.restype = ctypes.c_double

    def datan(x):
        x, k = Decimal(x), 0
        while abs(x) > Decimal("0.05"):
            x, k = x / (1 + (1 + x * x).sqrt()), k + 1
        s, term, n, x2 = Decimal(0), x, 0, x * x
        while abs(term) > Decimal(10) ** -45:
            s += term / (2 * n + 1) if n % 2 == 0 else -term / (2 * n + 1)
            term, n = term * x2, n + 1
        return s * (2 ** k)

    rng = np.random.default_rng(9)
    for _ in range(3000):
        n = float(rng.choice([-1, 1]) * 10 ** rng.uniform(-8, 8))
        d = float(rng.choice([-1, 1]) * 10 ** rng.uniform(-8, 8))
        ex = datan(Decimal(n) / Decimal(d))
        got = hc.hc_atan_ratio(n, d)
        assert abs(Decimal(got) - ex) <= Decimal(2 * np.spacing(abs(float(ex)))), (n, d)
    inf, nan = np.inf, np.nan
    with np.errstate(divide="ignore", invalid="ignore"):
        for n, d in [(1.0, 0.0), (1.0, -0.0), (-2.0, 0.0), (0.0, 3.0), (-0.0, 3.0), (0.0, -3.0),
                     (inf, 2.0), (-inf, -2.0), (2.0, inf), (0.0, 0.0), (inf, inf), (nan, 1.0),
                     (1.0, nan), (5.0, 5.0), (-5.0, 5.0)]:
            want = np.arctan(np.float64(n) / np.float64(d))
            got = hc.hc_atan_ratio(n, d)
            if np.isnan(want):
                assert np.isnan(got), (n, d)
            else:
                assert abs(got - want) <= 2 * np.spacing(abs(want)) and \
                    np.signbit(got) == np.signbit(want), (n, d, got, want)
