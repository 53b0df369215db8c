"""G of any prism mesh by per-prism corner sums (gra_assemble_corners,
Fwd.compute_G(method="corners")).

Pins: on tensor-product meshes every entry equals the shared-node corner-sum
G (gra_assemble_rectilinear, itself pinned against the reference's golden
forward outputs) bit for bit; on refined octree meshes every entry agrees
with the oracle's restatement of the reference's face/edge formulation to
1e-12 of the component block's largest value, except where the reference's
formulation is off (points over a cell's face plane far along it): there the
entry must be within 1e-12 of the block scale of the prism integral
evaluated at 60 digits and nearer to it than the oracle, and -- independent
of the corner formula and its limit convention -- within the range of the
oracle's face/edge values at the point moved off the planes diagonally by
+-delta in x and y (delta = 1e-2 and 1e-3 m): the field is continuous above
the body, so those values bracket its value at the point to O(delta^2).
FP32 / BF16 storage rounds the FP64 sums."""
import numpy as np
import pytest

import oracle
from graada3d_amd import Fwd, GraError
from graada3d_amd import mesh as M
from graada3d_amd.octree import OctreeMesh
from helpers import corner_exact, rel_norm

pytestmark = pytest.mark.gpu
TOL = 1e-12
EXACT_TOL = 1e-12
MAX_EXACT = 300


def _octree(seed, rounds=2):
    rng = np.random.default_rng(seed)
    mesh = OctreeMesh((0.0, 800.0), 8, (0.0, 640.0), 8, (0.0, 400.0), 5)
    for _ in range(rounds):
        n = mesh.n_cells
        mesh.refine(np.sort(rng.choice(n, n // 6, replace=False)))
    return mesh.prisms()


def _within_moved_point_values(prism, p, bit, value, scale):
    """value lies within the spread of the face/edge oracle's values at p
    moved by (+-delta, +-delta) in x and y, off every face plane through p."""
    for d in (1e-2, 1e-3):
        q = np.array([[p[0] + sx * d, p[1] + sy * d, p[2]] for sx in (1, -1) for sy in (1, -1)])
        v, st = oracle.assemble(prism[None], q, 1 << bit)
        assert st == 0
        v = v[:, 0]
        mid = v.mean()
        assert abs(value - mid) <= np.abs(v - mid).max() + 1e-11 * scale, (bit, d, value, v)


def _check(prisms, pts, flag, ctx):
    f = Fwd(prisms, pts, flag, ctx=ctx)
    f.compute_G(method="corners")
    G = f.get_G()
    Go, st = oracle.assemble(prisms, pts, flag)
    assert st == 0
    bits = [b for b in range(1, 10) if flag & (1 << b)]
    rng = np.random.default_rng(0)
    n = len(pts)
    for k, bit in enumerate(bits):
        a, b = G[k * n:(k + 1) * n], Go[k * n:(k + 1) * n]
        scale = np.abs(b).max()
        off = np.argwhere(np.abs(a - b) > TOL * scale)
        assert len(off) <= 0.01 * a.size, (bit, len(off))
        if len(off) > MAX_EXACT:
            off = off[rng.choice(len(off), MAX_EXACT, replace=False)]
        for ip, j in off:
            ex = corner_exact(prisms[j], pts[ip], bit)
            assert abs(a[ip, j] - ex) <= EXACT_TOL * scale, (bit, ip, j, a[ip, j], ex)
            assert abs(a[ip, j] - ex) < abs(b[ip, j] - ex), (bit, ip, j)
            _within_moved_point_values(prisms[j], pts[ip], bit, a[ip, j], scale)
        assert rel_norm(a, b) < 1e-7, bit
    return f, G


@pytest.mark.parametrize("flag", [0x12, 0x3FE, 0x2, 0x80 | 0x200 | 0x4])
def test_octree_vs_face_edge_oracle(ctx, flag):
    prisms = _octree(5)
    rng = np.random.default_rng(6)
    pts = np.column_stack([rng.uniform(-200, 1000, 30), rng.uniform(-200, 840, 30),
                           rng.uniform(-120, -0.5, 30)])
    pts[0, :2] = prisms[7, [0, 2]]           # above a corner of a cell
    pts[1, 0] = prisms[40, 1]                # over a cell's face plane
    print(f"corner assembly octree {len(prisms)} cells flag {flag:#x}", flush=True)
    _check(prisms, pts, flag, ctx)


@pytest.mark.parametrize("flag", [0x12, 0x3FE, 0x10])
def test_tensor_mesh_equals_shared_node_G(ctx, flag):
    rng = np.random.default_rng(9)
    xs = np.sort(rng.uniform(0, 1000, 21))
    ys = np.sort(rng.uniform(0, 800, 14))
    zs = np.sort(rng.uniform(0, 500, 8))
    i, j, k = np.meshgrid(np.arange(20), np.arange(13), np.arange(7), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    prisms = np.stack([xs[i], xs[i + 1], ys[j], ys[j + 1], zs[k], zs[k + 1]], axis=1)
    pts = np.column_stack([rng.uniform(-300, 1300, 40), rng.uniform(-300, 1100, 40),
                           rng.uniform(-150, -0.5, 40)])
    pts[0, :2] = [xs[3], ys[2]]
    pts[1, 0] = xs[0]
    print(f"corner vs node assembly flag {flag:#x}", flush=True)
    f = Fwd(prisms, pts, flag, ctx=ctx)
    f.compute_G(method="nodes")
    Gn = f.get_G()
    f.compute_G(method="corners")
    np.testing.assert_array_equal(f.get_G(), Gn)


def test_stores_and_products(ctx):
    prisms = _octree(7, rounds=1)
    rng = np.random.default_rng(8)
    pts = np.column_stack([rng.uniform(0, 800, 25), rng.uniform(0, 640, 25),
                           rng.uniform(-60, -1, 25)])
    f = Fwd(prisms, pts, 0x3FE, ctx=ctx)
    f.compute_G("f64", method="corners")
    G64 = f.get_G()
    x = rng.standard_normal(f.Nm)
    assert rel_norm(f.G_vec_mul(x), G64 @ x) < 1e-14
    f.compute_G("f32", method="corners")
    assert np.array_equal(f.get_G(), G64.astype(np.float32).astype(np.float64))
    f.compute_G("bf16", method="corners")
    u = G64.astype(np.float32).view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    assert np.array_equal(f.get_G(), rounded.astype(np.uint32).view(np.float32).astype(np.float64))


def test_rejections(ctx):
    prisms = M.regular_mesh((0.0, 100.0), 4, (0.0, 100.0), 3, (50.0, 150.0), 2)
    with pytest.raises(GraError):            # below the top of the shallowest prism
        Fwd(prisms, np.array([[10.0, 10.0, 50.0]]), 0x2, ctx=ctx).compute_G(method="corners")
    with pytest.raises(GraError):
        Fwd(prisms, np.array([[10.0, 10.0, -1.0]]), 0x1, ctx=ctx).compute_G(method="corners")
    f = Fwd(prisms, np.array([[10.0, 10.0, 49.0]]), 0x2, ctx=ctx)
    f.compute_G(method="corners")
    assert np.isfinite(f.get_G()).all()


def test_batched_corners_equal_per_instance(ctx):
    """gra_assemble_batch_corners fills each context's G exactly as
    gra_assemble_corners does for that instance (uneven octree instances,
    different component sets, every storage type); the batch rejects an item
    with a point below a prism's top."""
    from graada3d_amd import Context
    from graada3d_amd.engine import assemble_batch
    from graada3d_amd.workloads import batch_instances
    inst = batch_instances(n=5, seed=77)
    for b in inst:
        b.points = b.points[:: max(1, b.points.shape[0] // 60)]
    for store in ("f64", "f32", "bf16"):
        print(f"batched corners {store}", flush=True)
        fw = [Fwd(b.prisms, b.points, b.flag, ctx=Context(0)) for b in inst]
        try:
            assemble_batch(fw, store, method="corners")
            for b, f in zip(inst, fw):
                single = Fwd(b.prisms, b.points, b.flag, ctx=ctx)
                single.compute_G(store, method="corners")
                assert np.array_equal(f.get_G(), single.get_G()), (store, b.seed)
        finally:
            for f in fw:
                f.ctx.close()
    bad = inst[1].points.copy()
    bad[0, 2] = inst[1].prisms[:, 4].min() + 1.0
    fw = [Fwd(inst[0].prisms, inst[0].points, inst[0].flag, ctx=Context(0)),
          Fwd(inst[1].prisms, bad, inst[1].flag, ctx=Context(0))]
    try:
        with pytest.raises(GraError, match="item 1"):
            assemble_batch(fw, "f64", method="corners")
    finally:
        for f in fw:
            f.ctx.close()
