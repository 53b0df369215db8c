"""expand_G on the GPU (AdaptiveInversion::expand_G, AdaptiveInversion.cpp:
23-106): the expanded G equals a fresh assembly over the refined leaf list by
the same formulation bit for bit (same kernels, columns moved as stored
bits; the children of a corner-sum G are corner sums), and each parent's
column equals the sum of its children's (superposition; the face/edge form's
rounding, <= 1e-9 of the column's largest entry)."""
import numpy as np
import pytest

from graada3d_amd import Fwd
from graada3d_amd import mesh as M
from graada3d_amd.octree import OctreeMesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("store,flag,method", [("f64", 0x12, "faces"), ("f64", 0x3FE, "faces"),
                                               ("f32", 0x2 | 0x80, "faces"), ("bf16", 0x10, "faces"),
                                               ("f64", 0x3FE, "corners"), ("f32", 0x12, "corners")])
def test_expand_equals_fresh_assembly_and_superposition(ctx, store, flag, method):
    rng = np.random.default_rng(9)
    o = OctreeMesh((0, 800), 8, (0, 800), 8, (0, 400), 4)
    pts = M.grid_points(9, -50.0, 850.0, 8, -50.0, 850.0, -7.5)
    f = Fwd(o.prisms(), pts, flag, ctx=ctx)
    f.compute_G(store, method)
    for rnd in range(3):
        G_old = f.get_G()
        targets = rng.choice(o.n_cells, size=o.n_cells // 8, replace=False)
        split, kids = o.refine(targets)
        nm = f.expand_G(split, kids)
        assert nm == o.n_cells and np.array_equal(f.prisms, o.prisms())
        G_new = f.get_G()
        fresh = Fwd(o.prisms(), pts, flag, ctx=None)
        try:
            fresh.compute_G(store, method)
            assert np.array_equal(G_new, fresh.get_G()), rnd
        finally:
            fresh.ctx.close()
        # unsplit columns moved unchanged; children sum to their parent
        for s, col in enumerate(split):
            start = col + 7 * s
            child_sum = G_new[:, start:start + 8].sum(axis=1)
            if store == "f64":
                scale = np.abs(G_old[:, col]).max()
                assert np.abs(child_sum - G_old[:, col]).max() <= 1e-9 * scale, (rnd, col)
        keep = np.setdiff1d(np.arange(G_old.shape[1]), split)
        newpos = keep + 7 * np.searchsorted(split, keep)
        assert np.array_equal(G_new[:, newpos], G_old[:, keep])
    # the grown G still works in the products
    x = rng.standard_normal(f.Nm)
    assert np.allclose(f.G_vec_mul(x), f.get_G() @ x, rtol=1e-12, atol=0)


def test_expand_of_a_corner_sum_G_needs_points_above_the_children(ctx):
    from graada3d_amd import GraError
    o = OctreeMesh((0, 400), 4, (0, 400), 4, (0, 200), 2)
    f = Fwd(o.prisms(), M.grid_points(4, 0.0, 400.0, 4, 0.0, 400.0, -5.0), 0x2, ctx=ctx)
    f.compute_G("f64", "corners")
    kids = np.tile(o.prisms()[:1], (8, 1))
    kids[:, 4] = -10.0                      # a child reaching above the points
    with pytest.raises(GraError):
        f.expand_G([0], kids)


def test_expand_rejects_bad_split_lists(ctx):
    from graada3d_amd import GraError
    o = OctreeMesh((0, 400), 4, (0, 400), 4, (0, 200), 2)
    f = Fwd(o.prisms(), M.grid_points(4, 0.0, 400.0, 4, 0.0, 400.0, -5.0), 0x2, ctx=ctx)
    f.compute_G()
    kids = np.tile(o.prisms()[:1], (16, 1))
    with pytest.raises(GraError):
        f.expand_G([3, 3], kids)            # not strictly increasing
    with pytest.raises(GraError):
        f.expand_G([5, 40], kids)           # beyond Nm
