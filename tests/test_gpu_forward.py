"""GPU prism kernels (through the C ABI) against the oracle and the
reference's golden vectors.

Bar: the GPU evaluates the reference's FP64 face/edge formulation with the
same operation order (bit-exact on the host, test_host_math.py); only the
device atan/log differ from glibc's by an ulp, amplified by the face
differences' cancellation.  Tolerances: per component block
||x - x_ref||/||x_ref|| <= 1e-12 and max|x - x_ref|/max|x_ref| <= 1e-12 against
the oracle; against the 16-digit golden files 1e-12 as well.  The required
bound for the plain FP64 path is 1e-10.
"""
import numpy as np
import pytest

import oracle
from graada3d_amd import GraError, Fwd, forward
from graada3d_amd import mesh as M
from helpers import load, max_rel_to_scale, model_from_fixture, rel_norm

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _blocks(G, nc, n_obs):
    return [G[k * n_obs:(k + 1) * n_obs] for k in range(nc)]


def test_assemble_small_G_all_components_vs_reference(ctx):
    fx = load("ref_G_small.npz")
    f = Fwd(fx["prisms"], fx["points"], int(fx["flag"]), ctx=ctx)
    f.compute_G("f64")
    G = f.get_G()
    n_obs = fx["points"].shape[0]
    for k, (a, b) in enumerate(zip(_blocks(G, 9, n_obs), _blocks(fx["G"], 9, n_obs))):
        assert rel_norm(a, b) < TOL, k
        assert max_rel_to_scale(a, b) < TOL, k
    # column-major copy is the same matrix
    assert np.array_equal(f.get_G(col_major=True), G)


def test_assemble_reference_pairs(ctx):
    fx = load("ref_pairs.npz")
    vals = fx["values"]
    # one (point, prism) per call batch: points as rows, one prism at a time is
    # slow, so evaluate every point against every prism and take the diagonal
    n = 256
    f = Fwd(fx["prisms"][:n], fx["points"][:n], 0x3FE, ctx=ctx)
    f.compute_G()
    G = f.get_G()
    for k, bit in enumerate(range(1, 10)):
        got = np.array([G[k * n + q, q] for q in range(n)])
        ref = vals[:n, bit]
        assert max_rel_to_scale(got, ref) < 1e-11, bit
        assert rel_norm(got, ref) < 1e-11, bit


@pytest.mark.parametrize("flag", [0x12, 0x3FE, 0x2, 0x10, 0x240])
def test_assemble_vs_oracle_reference_grid(ctx, flag):
    """40x40x20 mesh of the reference examples, points on the golden grid
    (on face planes and edge extensions), all component sets that have their
    own kernel instantiation plus a generic one."""
    prisms = M.regular_mesh((0.0, 2000.0), 40, (0.0, 2000.0), 40, (0.0, 1000.0), 20)
    pts = M.grid_points(41, 0.0, 2000.0, 41, 0.0, 2000.0, -0.1)[::97]
    f = Fwd(prisms, pts, flag, ctx=ctx)
    f.compute_G()
    G = f.get_G()
    Go, st = oracle.assemble(prisms, pts, flag)
    assert st == 0
    nc = bin(flag).count("1")
    for k, (a, b) in enumerate(zip(_blocks(G, nc, len(pts)), _blocks(Go, nc, len(pts)))):
        assert rel_norm(a, b) < TOL, k
        assert max_rel_to_scale(a, b) < TOL, k


def test_store_f32_bf16_are_rounded_f64(ctx):
    prisms = M.regular_mesh((0.0, 500.0), 10, (0.0, 500.0), 10, (0.0, 250.0), 5)
    pts = M.grid_points(9, 0.0, 500.0, 9, 0.0, 500.0, -3.0)
    f = Fwd(prisms, pts, 0x12, ctx=ctx)
    f.compute_G("f64")
    G64 = f.get_G()
    f.compute_G("f32")
    G32 = f.get_G()
    assert np.array_equal(G32, G64.astype(np.float32).astype(np.float64))
    f.compute_G("bf16")
    Gb = f.get_G()
    f32 = G64.astype(np.float32)
    u = f32.view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16     # round-to-nearest-even
    ref = rounded.astype(np.uint32).view(np.float32).astype(np.float64)
    assert np.array_equal(Gb, ref)


@pytest.mark.parametrize("name", ["model1", "model3", "model2"])
def test_forward_vs_reference_golden(ctx, name):
    fx = load(f"ref_forward_{name}.npz")
    prisms, rho = model_from_fixture(fx)
    flag = int(fx["flag"])
    print(f"[gpu] forward {name}: {prisms.shape[0]} cells x {fx['points'].shape[0]} points",
          flush=True)
    out = forward(prisms, rho, fx["points"], flag, ctx=ctx)   # all cells, as compute does
    bits = sorted(fx["bits"].tolist())
    out = out[:, [bits.index(b) for b in fx["bits"]]]
    gold = fx["values"]
    for k in range(gold.shape[1]):
        assert rel_norm(out[:, k], gold[:, k]) < TOL, k
        assert max_rel_to_scale(out[:, k], gold[:, k]) < TOL, k


def test_forward_equals_G_times_rho(ctx):
    prisms = M.regular_mesh((0.0, 1000.0), 20, (0.0, 800.0), 16, (0.0, 500.0), 10)
    rng = np.random.default_rng(4)
    rho = rng.uniform(-500, 500, prisms.shape[0])
    pts = M.grid_points(12, -100.0, 1100.0, 10, -100.0, 900.0, -20.0)
    out = forward(prisms, rho, pts, 0x3FE, ctx=ctx)
    f = Fwd(prisms, pts, 0x3FE, ctx=ctx)
    f.compute_G()
    d = f.G_vec_mul(rho).reshape(9, -1).T
    assert rel_norm(out, d) < 1e-12


def test_degenerate_and_bad_arguments_raise(ctx):
    pr = np.array([[0.0, 50.0, 0.0, 50.0, 0.0, 50.0]])
    with pytest.raises(GraError) as e:
        # on the vertical edge x=0,y=0: degenerate for x-face integrals (Txz), not for Tzz
        forward(pr, np.ones(1), np.array([[0.0, 0.0, 20.0]]), 0x20, ctx=ctx)
    assert e.value.status == 5
    with pytest.raises(GraError) as e:
        forward(np.array([[0.0, -1.0, 0.0, 1.0, 0.0, 1.0]]), np.ones(1),
                np.array([[0.0, 0.0, -1.0]]), 0x2, ctx=ctx)
    assert e.value.status == 1
    with pytest.raises(GraError) as e:
        forward(pr, np.ones(1), np.array([[0.0, 0.0, -1.0]]), 0x1, ctx=ctx)
    assert e.value.status == 1
    # the context stays usable after errors
    v = forward(pr, np.ones(1), np.array([[25.0, 25.0, -10.0]]), 0x2, ctx=ctx)
    o, _ = oracle.forward(pr, np.ones(1), np.array([[25.0, 25.0, -10.0]]), 0x2)
    assert abs(v[0, 0] - o[0, 0]) < 1e-13 * abs(o[0, 0])


# ---- node-sum forward on rectilinear meshes --------------------------------

def _rectilinear(rng, nx, ny, nz, uniform=False):
    if uniform:
        xs, ys, zs = (np.linspace(0, 4000, nx + 1), np.linspace(0, 3000, ny + 1),
                      np.linspace(0, 2000, nz + 1))
    else:
        xs = np.sort(rng.uniform(0, 1000, nx + 1))
        ys = np.sort(rng.uniform(0, 800, ny + 1))
        zs = np.sort(rng.uniform(0, 500, nz + 1))
    prisms = np.array([[xs[i], xs[i + 1], ys[j], ys[j + 1], zs[k], zs[k + 1]]
                       for i in range(nx) for j in range(ny) for k in range(nz)])
    return xs, ys, zs, prisms


@pytest.mark.parametrize("flag", [0x3FE, 0x12, 0x2, 0x10, 0x80 | 0x200 | 0x4, 0x40 | 0x8])
def test_rectilinear_forward_matches_face_edge_oracle(ctx, flag):
    """The node-sum form against the reference's per-prism face/edge values
    (the oracle): per component, max error <= 1e-9 of the block's largest
    value, on a non-uniform mesh with points over node planes and node
    lines (the x = y = 0 limit) and points off the mesh's footprint."""
    from graada3d_amd import forward_rectilinear
    rng = np.random.default_rng(31)
    xs, ys, zs, prisms = _rectilinear(rng, 7, 6, 5)
    rho = rng.uniform(-500, 500, prisms.shape[0])
    pts = np.column_stack([rng.uniform(-300, 1300, 40), rng.uniform(-300, 1100, 40),
                           rng.uniform(-150, -0.5, 40)])
    pts[0, :2] = [xs[3], ys[2]]          # above a node line
    pts[1, 0] = xs[0]                    # over a node plane
    pts[2, 1] = ys[6]
    pts[3, :2] = [xs[7], ys[0]]          # above a corner node line
    got = forward_rectilinear(xs, ys, zs, rho, pts, flag, ctx=ctx)
    ref, st = oracle.forward(prisms, rho, pts, flag)
    assert st == 0
    for k in range(got.shape[1]):
        assert np.abs(got[:, k] - ref[:, k]).max() <= 1e-9 * np.abs(ref[:, k]).max(), k


def test_rectilinear_forward_survey_grid_and_rejections(ctx):
    from graada3d_amd import GraError, forward_rectilinear
    rng = np.random.default_rng(32)
    xs, ys, zs, prisms = _rectilinear(rng, 16, 12, 8, uniform=True)
    rho = rng.uniform(-500, 500, prisms.shape[0])
    pts = M.grid_points(13, 0.0, 4000.0, 10, 0.0, 3000.0, -20.0)   # on node planes
    got = forward_rectilinear(xs, ys, zs, rho, pts, 0x3FE, ctx=ctx)
    ref, _ = oracle.forward(prisms, rho, pts, 0x3FE)
    for k in range(9):
        assert rel_norm(got[:, k], ref[:, k]) < 1e-10, k
    # uniform density: the node-sum form reduces to the single outer prism
    one = forward_rectilinear(xs, ys, zs, np.full(prisms.shape[0], 300.0), pts[:5], 0x12,
                              ctx=ctx)
    big, _ = oracle.forward(np.array([[0, 4000, 0, 3000, 0, 2000.0]]), np.array([300.0]),
                            pts[:5], 0x12)
    assert np.abs(one - big).max() <= 1e-11 * np.abs(big).max()
    with pytest.raises(GraError):      # a point inside the mesh's depth range
        forward_rectilinear(xs, ys, zs, rho, np.array([[10.0, 10.0, 5.0]]), 0x2, ctx=ctx)
    with pytest.raises(GraError):      # non-increasing nodes
        forward_rectilinear(xs[::-1], ys, zs, rho, pts[:2], 0x2, ctx=ctx)


def test_batched_assembly_equals_per_instance(ctx):
    """gra_assemble_batch fills each context's G exactly as gra_assemble does
    (same per-entry arithmetic), for uneven octree instances with different
    component sets, in every storage type."""
    from graada3d_amd import Context
    from graada3d_amd.engine import assemble_batch
    from graada3d_amd.workloads import batch_instances
    inst = batch_instances(n=5, seed=77)
    for b in inst:                                    # keep the test small
        b.points = b.points[:: max(1, b.points.shape[0] // 60)]
    for store in ("f64", "f32", "bf16"):
        fw = [Fwd(b.prisms, b.points, b.flag, ctx=Context(0)) for b in inst]
        try:
            assemble_batch(fw, store)
            for b, f in zip(inst, fw):
                got = f.get_G()
                single = Fwd(b.prisms, b.points, b.flag, ctx=ctx)
                single.compute_G(store)
                assert np.array_equal(got, single.get_G()), (store, b.seed)
                if store == "f64":
                    ref, st = oracle.assemble(b.prisms, b.points, b.flag)
                    assert st == 0 and np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
        finally:
            for f in fw:
                f.ctx.close()
