"""Multi-right-hand-side products on the matrix cores (gra_G_mat_mul,
gra_GT_mat_mul) against numpy products of the stored G.

Bars: FP64-stored G (FP64 MFMA): ||y - y_ref|| <= 1e-13 ||y_ref|| per
column.  FP32 / BF16-stored G (FP32 MFMA, the vectors rounded to FP32, FP32
partial sums of at most 32 products added into FP64): elementwise
|y - y_ref| <= 4e-6 (|G| |x|)_i, the rounding bound of such sums, against the
product of the stored G with the FP32-rounded vectors in FP64.  Sizes cover
row / column counts that are not multiples of the 16-wide tiles, right-hand
side counts 1, 5, 16, 17 and 40, and G large enough to be split over the
summation range.  Results are bit-for-bit repeatable."""
import numpy as np
import pytest

from graada3d_amd import Fwd, GraError
from graada3d_amd import mesh as M
from helpers import rel_norm

pytestmark = pytest.mark.gpu


def _fwd(ctx, dims, npts, flag, store):
    nx, ny, nz = dims
    prisms = M.regular_mesh((0.0, 40.0 * nx), nx, (0.0, 40.0 * ny), ny, (0.0, 30.0 * nz), nz)
    pts = M.grid_points(npts, -20.0, 40.0 * nx + 20, npts + 1, -20.0, 40.0 * ny + 20, -5.0)
    f = Fwd(prisms, pts, flag, ctx=ctx)
    f.compute_G(store)
    return f, f.get_G()


def _check(got, G, X, store):
    if store == "f64":
        ref = G @ X
        for j in range(X.shape[1]):
            assert rel_norm(got[:, j], ref[:, j]) < 1e-13, j
    else:
        X32 = X.astype(np.float32).astype(np.float64)
        ref = G @ X32
        bound = 4e-6 * (np.abs(G) @ np.abs(X32))
        assert np.all(np.abs(got - ref) <= bound)


@pytest.mark.parametrize("store", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("k", [1, 5, 16, 17, 40])
def test_small_g_vs_numpy(ctx, store, k):
    print(f"mat_mul small store {store} k {k}", flush=True)
    f, G = _fwd(ctx, (12, 10, 6), 6, 0x12, store)       # 720 cells, 84 rows
    rng = np.random.default_rng(k)
    X = rng.standard_normal((f.Nm, k))
    R = rng.standard_normal((f.Nd, k))
    Y, Z = f.G_mat_mul(X), f.GT_mat_mul(R)
    assert Y.shape == (f.Nd, k) and Z.shape == (f.Nm, k)
    _check(Y, G, X, store)
    _check(Z, G.T, R, store)
    assert np.array_equal(Y, f.G_mat_mul(X)) and np.array_equal(Z, f.GT_mat_mul(R))


@pytest.mark.parametrize("store", ["f64", "f32", "bf16"])
def test_split_summation_vs_numpy(ctx, store):
    """G of 3,200 x 13,500: both products split over their summation range."""
    print(f"mat_mul split store {store}", flush=True)
    f, G = _fwd(ctx, (30, 30, 15), 39, 0x12, store)
    rng = np.random.default_rng(5)
    X = rng.standard_normal((f.Nm, 24))
    R = rng.standard_normal((f.Nd, 24))
    _check(f.G_mat_mul(X), G, X, store)
    _check(f.GT_mat_mul(R), G.T, R, store)


@pytest.mark.parametrize("store", ["f64", "f32", "bf16"])
def test_small_survey_g_vs_numpy(ctx, store):
    """The 5,000 x 32,000 G of the small inversion (gz + Tzz on the 40 x 40 x
    20 mesh) with 16 right-hand sides.  Its 5,000 rows once split into spans
    of 1,667 and 1,666 rows, which broke the earlier G^T R kernel: lanes that
    left its row loop one turn early still fed their stale registers to the
    wave's last MFMAs (4e-2 relative error)."""
    from graada3d_amd.workloads import forward_problem_small
    print(f"mat_mul small survey store {store}", flush=True)
    fp = forward_problem_small()
    f = Fwd(fp.prisms, fp.points, fp.flag, ctx=ctx)
    f.compute_G(store)
    G = f.get_G()
    assert G.shape == (5000, 32000)
    rng = np.random.default_rng(11)
    X = rng.standard_normal((f.Nm, 16))
    R = rng.standard_normal((f.Nd, 16))
    _check(f.G_mat_mul(X), G, X, store)
    _check(f.GT_mat_mul(R), G.T, R, store)


@pytest.mark.parametrize("store", ["f64", "f32", "bf16"])
def test_tiny_g_one_partial_chunk(ctx, store):
    """32 cells (one 64-column row pitch, less than one G X chunk for FP32 /
    BF16, so only the masked tail chunk runs) and 24 rows (one partial
    64-row workgroup)."""
    print(f"mat_mul tiny store {store}", flush=True)
    f, G = _fwd(ctx, (4, 4, 2), 3, 0x12, store)
    assert G.shape == (24, 32)
    rng = np.random.default_rng(3)
    X = rng.standard_normal((f.Nm, 3))
    R = rng.standard_normal((f.Nd, 3))
    _check(f.G_mat_mul(X), G, X, store)
    _check(f.GT_mat_mul(R), G.T, R, store)


def test_columns_equal_single_vector_products(ctx):
    f, G = _fwd(ctx, (12, 10, 6), 6, 0x3FE, "f64")
    rng = np.random.default_rng(9)
    X = rng.standard_normal((f.Nm, 7))
    R = rng.standard_normal((f.Nd, 7))
    Y, Z = f.G_mat_mul(X), f.GT_mat_mul(R)
    for j in range(7):
        assert rel_norm(Y[:, j], f.G_vec_mul(X[:, j])) < 1e-13
        assert rel_norm(Z[:, j], f.GT_vec_mul(R[:, j])) < 1e-13


def test_rejections(ctx):
    f, _ = _fwd(ctx, (4, 4, 2), 3, 0x2, "f64")
    with pytest.raises(GraError):
        f.G_mat_mul(np.zeros((f.Nm, 0)))
    with pytest.raises(GraError):
        f.GT_mat_mul(np.zeros((f.Nd, 1025)))
    with pytest.raises(ValueError):
        f.G_mat_mul(np.zeros((f.Nm + 1, 2)))
