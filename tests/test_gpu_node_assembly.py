"""G of a tensor-product mesh by corner sums (gra_assemble_rectilinear,
Fwd.compute_G(method="nodes")) against the oracle's restatement of the
reference's per-prism face/edge formulation.

The two formulations are the same integral evaluated differently.  Per
component block, every entry agrees with the oracle to TOL = 1e-12 of the
block's largest value, except where the reference's formulation itself is
off: there (points over a cell's face plane, far along the plane) its values
miss the integral by up to 1e-8 of the block scale, even in sign.  Each such
entry is checked against the prism integral evaluated at 60 digits
(helpers.corner_exact): the corner sum must be within 1e-12 of the block
scale of it (its own rounding reaches ~1e-13 for the g components, whose
corner terms y atan(...) grow with the distance) and nearer to it than the
oracle.  Meshes cover uniform and
non-uniform node spacing, tiles cut by the mesh edge, points over node
planes and node lines (the x = 0 / x = y = 0 limits) and points off the
mesh's footprint.  Storage in FP32 / BF16 rounds the FP64 corner sums
exactly as gra_assemble rounds its values."""
import numpy as np
import pytest

import oracle
from graada3d_amd import Fwd, GraError
from graada3d_amd import mesh as M
from helpers import corner_exact, rel_norm

pytestmark = pytest.mark.gpu
TOL = 1e-12
EXACT_TOL = 1e-12
MAX_EXACT = 400          # disagreeing entries checked at 60 digits per block


def _blocks(G, nc, n_obs):
    return [G[k * n_obs:(k + 1) * n_obs] for k in range(nc)]


def _check(prisms, pts, flag, ctx):
    f = Fwd(prisms, pts, flag, ctx=ctx)
    f.compute_G(method="nodes")
    G = f.get_G()
    Go, st = oracle.assemble(prisms, pts, flag)
    assert st == 0
    bits = [b for b in range(1, 10) if flag & (1 << b)]
    rng = np.random.default_rng(0)
    for bit, a, b in zip(bits, _blocks(G, len(bits), len(pts)), _blocks(Go, len(bits), len(pts))):
        scale = np.abs(b).max()
        off = np.argwhere(np.abs(a - b) > TOL * scale)
        assert len(off) <= 0.01 * a.size, (bit, len(off))
        if len(off) > MAX_EXACT:
            off = off[rng.choice(len(off), MAX_EXACT, replace=False)]
        for ip, j in off:
            ex = corner_exact(prisms[j], pts[ip], bit)
            assert abs(a[ip, j] - ex) <= EXACT_TOL * scale, (bit, ip, j, a[ip, j], ex)
            assert abs(a[ip, j] - ex) < abs(b[ip, j] - ex), (bit, ip, j)
        assert rel_norm(a, b) < 1e-7, bit
    return f, G


def _nonuniform(rng, nx, ny, nz):
    xs = np.sort(rng.uniform(0, 1000, nx + 1))
    ys = np.sort(rng.uniform(0, 800, ny + 1))
    zs = np.sort(rng.uniform(0, 500, nz + 1))
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    prisms = np.stack([xs[i], xs[i + 1], ys[j], ys[j + 1], zs[k], zs[k + 1]], axis=1)
    return xs, ys, zs, prisms


@pytest.mark.parametrize("flag", [0x3FE, 0x12, 0x2, 0x10, 0x240, 0x80 | 0x200 | 0x4])
def test_nonuniform_mesh_vs_face_edge_oracle(ctx, flag):
    rng = np.random.default_rng(41)
    xs, ys, zs, prisms = _nonuniform(rng, 23, 17, 6)    # tiles cut by both mesh edges
    pts = np.column_stack([rng.uniform(-300, 1300, 24), rng.uniform(-300, 1100, 24),
                           rng.uniform(-150, -0.5, 24)])
    pts[0, :2] = [xs[3], ys[2]]          # above a node line
    pts[1, 0] = xs[0]                    # over a node plane
    pts[2, 1] = ys[17]
    pts[3, :2] = [xs[23], ys[0]]         # above a corner node line
    print(f"node assembly 23x17x6 flag {flag:#x}", flush=True)
    _check(prisms, pts, flag, ctx)


@pytest.mark.parametrize("flag", [0x12, 0x3FE])
def test_reference_grid_vs_face_edge_oracle(ctx, flag):
    """40x40x20 mesh of the reference examples, points on the golden grid at
    z = -0.1 m (over face planes and edge extensions)."""
    prisms = M.regular_mesh((0.0, 2000.0), 40, (0.0, 2000.0), 40, (0.0, 1000.0), 20)
    pts = M.grid_points(41, 0.0, 2000.0, 41, 0.0, 2000.0, -0.1)[::97]
    print(f"node assembly 40x40x20 flag {flag:#x}", flush=True)
    _check(prisms, pts, flag, ctx)


def test_stores_round_the_fp64_corner_sums(ctx):
    prisms = M.regular_mesh((0.0, 500.0), 10, (0.0, 500.0), 10, (0.0, 250.0), 5)
    pts = M.grid_points(9, 0.0, 500.0, 9, 0.0, 500.0, -3.0)
    f = Fwd(prisms, pts, 0x12, ctx=ctx)
    f.compute_G("f64", method="nodes")
    G64 = f.get_G()
    f.compute_G("f32", method="nodes")
    assert np.array_equal(f.get_G(), G64.astype(np.float32).astype(np.float64))
    f.compute_G("bf16", method="nodes")
    Gb = f.get_G()
    u = G64.astype(np.float32).view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16     # round-to-nearest-even
    assert np.array_equal(Gb, rounded.astype(np.uint32).view(np.float32).astype(np.float64))


def test_products_and_forward_on_the_node_G(ctx):
    """The node-assembled G drives the products like any G, and G rho equals
    the node-sum forward of the same model."""
    from graada3d_amd import forward_rectilinear
    rng = np.random.default_rng(42)
    xs, ys, zs, prisms = _nonuniform(rng, 12, 9, 7)
    pts = np.column_stack([rng.uniform(0, 1000, 30), rng.uniform(0, 800, 30),
                           rng.uniform(-80, -1, 30)])
    f = Fwd(prisms, pts, 0x3FE, ctx=ctx)
    f.compute_G(method="nodes")
    G = f.get_G()
    x = rng.standard_normal(f.Nm)
    w = rng.standard_normal(f.Nd)
    assert rel_norm(f.G_vec_mul(x), G @ x) < 1e-14
    assert rel_norm(f.GT_vec_mul(w), G.T @ w) < 1e-14
    rho = rng.uniform(-500, 500, f.Nm)
    fw = forward_rectilinear(xs, ys, zs, rho, pts, 0x3FE, ctx=ctx)
    d = (G @ rho).reshape(9, -1).T
    for k in range(9):
        assert np.abs(d[:, k] - fw[:, k]).max() <= 1e-10 * np.abs(fw[:, k]).max(), k


def test_rejections(ctx):
    prisms = M.regular_mesh((0.0, 100.0), 4, (0.0, 100.0), 3, (0.0, 50.0), 2)
    bad = prisms.copy()
    bad[[0, 1]] = bad[[1, 0]]                       # not in regular_mesh order
    with pytest.raises(ValueError):
        Fwd(bad, np.array([[10.0, 10.0, -1.0]]), 0x2, ctx=ctx).compute_G(method="nodes")
    with pytest.raises(GraError):                    # a point inside the mesh's depth range
        Fwd(prisms, np.array([[10.0, 10.0, 5.0]]), 0x2, ctx=ctx).compute_G(method="nodes")
    with pytest.raises(ValueError):
        Fwd(prisms, np.array([[10.0, 10.0, -1.0]]), 0x2, ctx=ctx).compute_G(method="edges")
    # the face/edge path still accepts the point inside the depth range
    f = Fwd(prisms, np.array([[10.0, 10.0, 5.0]]), 0x2, ctx=ctx)
    f.compute_G()
    assert np.isfinite(f.get_G()).all()


@pytest.mark.parametrize("name", ["model1", "model2", "model3"])
def test_node_paths_vs_reference_golden(ctx, name):
    """The reference's own forward outputs (its committed golden data: 40x40x20
    and 60x60x35 regular meshes, points at z = -0.1 m over every node line)
    from the corner-sum G times the density and from the node-sum forward:
    per component, max error <= 1e-12 of the largest value and relative
    2-norm error <= 1e-12 (measured: <= 4.5e-13)."""
    from graada3d_amd import forward_rectilinear
    from helpers import load, max_rel_to_scale, model_from_fixture
    fx = load(f"ref_forward_{name}.npz")
    prisms, rho = model_from_fixture(fx)
    axes = M.rectilinear_axes(prisms)
    assert axes is not None
    pts, flag, gold = fx["points"], int(fx["flag"]), fx["values"]
    print(f"node paths vs golden {name}", flush=True)
    f = Fwd(prisms, pts, flag, ctx=ctx)
    f.compute_G(method="nodes")
    nc = bin(flag).count("1")
    d = f.G_vec_mul(rho).reshape(nc, -1).T
    fw = forward_rectilinear(*axes, rho, pts, flag, ctx=ctx)
    for k in range(nc):
        for got in (d[:, k], fw[:, k]):
            assert max_rel_to_scale(got, gold[:, k]) <= 1e-12, k
            assert rel_norm(got, gold[:, k]) <= 1e-12, k
