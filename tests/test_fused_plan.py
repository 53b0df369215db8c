"""The fused product's launch planner (graada3d_amd/csrc/fused_plan.hpp),
built for the host: the grid never exceeds the co-resident limit (occupancy
per CU x CUs), every row and column is covered by exactly one (group, tile),
XCD-aligned grids give each XCD whole groups, and impossible shapes are
refused (the solver then runs the two-pass iteration)."""
import ctypes
import os

import numpy as np
import pytest

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hostcheck", "libhostcheck.so")

# (TW, RB) of the kernel per storage type: 64 KB blocks; 4,096-column tiles
# (FP64 / FP32), 8,192 (BF16)
KERNELS = {"f64": (4096, 2), "f32": (4096, 4), "bf16": (8192, 4)}
MAX_POLL = 4


@pytest.fixture(scope="module")
def hc():
    if not os.path.exists(LIB):
        from graada3d_amd.build import build_hostcheck
        build_hostcheck()
    lib = ctypes.CDLL(LIB)
    lib.hc_fused_plan.restype = ctypes.c_int
    lib.hc_fused_plan.argtypes = [ctypes.c_int64, ctypes.c_int64] + [ctypes.c_int] * 8 + \
        [ctypes.POINTER(ctypes.c_int64)]
    return lib


def plan(hc, nd, ld, store, cus=256, n_xcd=8, per_cu=1, target=0, min_blocks=0):
    tw, rb = KERNELS[store]
    out = (ctypes.c_int64 * 6)()
    ok = hc.hc_fused_plan(nd, ld, tw, rb, cus, n_xcd, per_cu, target, MAX_POLL, min_blocks, out)
    if not ok:
        return None
    return dict(zip(("NT", "S", "per_xcd", "rows", "nrb", "grid"), list(out))) | {"RB": rb, "TW": tw}


def check_cover(p, nd, ld, cus, n_xcd, per_cu):
    assert p["grid"] == p["NT"] * p["S"] <= per_cu * cus
    assert p["NT"] * p["TW"] >= ld > (p["NT"] - 1) * p["TW"]          # tiles cover the columns
    assert p["rows"] % p["RB"] == 0 and p["nrb"] * p["RB"] == p["rows"]
    assert p["S"] * p["rows"] >= nd                                  # groups cover the rows
    assert (p["S"] - 1) * p["rows"] < nd or p["per_xcd"] > 0         # no empty group unless aligned
    assert p["NT"] * p["RB"] <= 64 * MAX_POLL
    if p["per_xcd"]:
        # workgroup b -> lin = (b % n_xcd) * per_xcd + b // n_xcd: a bijection
        # onto 0..grid-1 in which each XCD holds whole groups
        g = p["grid"]
        assert g % n_xcd == 0 and p["per_xcd"] * n_xcd == g and p["per_xcd"] % p["NT"] == 0
        b = np.arange(g)
        lin = (b % n_xcd) * p["per_xcd"] + b // n_xcd
        assert sorted(lin) == list(range(g))
        group = lin // p["NT"]
        for x in range(n_xcd):
            groups = set(group[b % n_xcd == x])
            assert all(set(group[group == q].tolist()) for q in groups)
            assert not groups & set(group[b % n_xcd != x])


@pytest.mark.parametrize("store", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("nd,ld", [(90000, 131072), (5000, 32000), (2000, 256), (7, 4096),
                                   (123457, 65536), (64, 8192)])
def test_plan_covers_and_fits(hc, store, nd, ld):
    for cus, n_xcd, per_cu in ((256, 8, 1), (256, 8, 2), (304, 8, 1), (80, 1, 1)):
        p = plan(hc, nd, ld, store, cus, n_xcd, per_cu)
        if p is None:
            tw, rb = KERNELS[store]
            nt = -(-ld // tw)
            assert nt * rb > 64 * MAX_POLL or nt > per_cu * cus
            continue
        check_cover(p, nd, ld, cus, n_xcd, per_cu)


def test_plan_config3_bf16_groups_stay_in_an_xcd(hc):
    # 16 tiles of 8,192 columns: two row groups per XCD, 32 workgroups each
    p = plan(hc, 90000, 131072, "bf16")
    assert (p["NT"], p["S"], p["per_xcd"], p["grid"]) == (16, 16, 32, 256)
    assert p["per_xcd"] % p["NT"] == 0
    assert p["rows"] == 5628 and p["nrb"] == 1407 and p["rows"] * p["S"] >= 90000


def test_plan_refuses_what_cannot_be_resident(hc):
    # no occupancy: the kernel does not fit a CU
    assert plan(hc, 5000, 32000, "bf16", per_cu=0) is None
    # more tiles than co-resident workgroups
    assert plan(hc, 5000, 8192 * 40, "bf16", cus=32, n_xcd=1) is None
    # a block's partials exceed the polling wave (NT * RB > 256)
    assert plan(hc, 90000, 8192 * 65, "bf16") is None
    assert plan(hc, 90000, 4096 * 65, "f32") is None
    # a requested grid is capped at the co-resident limit
    p = plan(hc, 90000, 131072, "f64", target=100000)
    assert p["grid"] <= 256


def test_plan_minimum_row_blocks(hc):
    """With a minimum of row blocks per group the planner refuses short
    groups (the iteration then runs two-pass) and keeps long ones."""
    assert plan(hc, 5000, 32000, "bf16")["nrb"] == 20
    assert plan(hc, 5000, 32000, "bf16", min_blocks=24) is None
    assert plan(hc, 5000, 32000, "f64", min_blocks=24)["nrb"] == 79
    assert plan(hc, 90000, 131072, "bf16", min_blocks=24)["nrb"] == 1407
