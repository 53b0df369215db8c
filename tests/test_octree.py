"""Octree refinement (host): the reference's leaf order and 2:1 face balance
(Mesh.cpp:1462-1976), pinned by construction invariants."""
import numpy as np

from graada3d_amd import mesh as M
from graada3d_amd.octree import OctreeMesh


def _volume(p):
    return float(np.sum((p[:, 1] - p[:, 0]) * (p[:, 3] - p[:, 2]) * (p[:, 5] - p[:, 4])))


def test_level0_is_the_regular_mesh():
    o = OctreeMesh((0, 400), 4, (0, 300), 3, (0, 200), 2)
    assert np.array_equal(o.prisms(), M.regular_mesh((0, 400), 4, (0, 300), 3, (0, 200), 2))


def test_split_replaces_parent_in_place_with_children_in_xyz_order():
    o = OctreeMesh((0, 400), 4, (0, 300), 3, (0, 200), 2)
    before = o.prisms()
    split, kids = o.refine([5])
    assert list(split) == [5] and kids.shape == (8, 6)
    after = o.prisms()
    assert after.shape[0] == before.shape[0] + 7
    assert np.array_equal(after[:5], before[:5]) and np.array_equal(after[13:], before[6:])
    p = before[5]
    xm, ym, zm = 0.5 * (p[1] + p[0]), 0.5 * (p[3] + p[2]), 0.5 * (p[5] + p[4])
    for c, child in enumerate(after[5:13]):
        i, j, k = c >> 2, (c >> 1) & 1, c & 1
        assert child[0] == (p[0], xm)[i] and child[1] == (xm, p[1])[i]
        assert child[2] == (p[2], ym)[j] and child[3] == (ym, p[3])[j]
        assert child[4] == (p[4], zm)[k] and child[5] == (zm, p[5])[k]
    assert np.array_equal(kids, after[5:13])


def test_balance_splits_coarser_face_neighbours_first():
    o = OctreeMesh((0, 400), 4, (0, 400), 4, (0, 400), 4)
    o.refine([0])                               # cell (0,0,0) -> 8 children
    # child 7 of cell 0 (upper x,y,z octant) touches level-0 cells (1,0,0),
    # (0,1,0), (0,0,1); splitting it forces those three to split
    n0 = o.n_cells
    split, kids = o.refine([7])
    assert o.is_balanced()
    assert len(split) == 4 and kids.shape == (32, 6)
    assert 7 in split and o.n_cells == n0 + 28
    # cells (0,0,1), (0,1,0), (1,0,0) sit at old indices 8, 11, 23 (level-0 id + 7)
    assert sorted(split) == [7, 8, 11, 23]


def test_random_refinement_keeps_volume_balance_and_disjointness():
    rng = np.random.default_rng(4)
    o = OctreeMesh((0, 800), 4, (0, 800), 4, (0, 400), 2)
    vol = 800.0 * 800.0 * 400.0
    for _ in range(4):
        targets = rng.choice(o.n_cells, size=max(1, o.n_cells // 10), replace=False)
        prev = o.prisms()
        split, kids = o.refine(targets)
        assert np.all(np.diff(split) > 0)
        p = o.prisms()
        assert p.shape[0] == prev.shape[0] + 7 * split.size
        assert abs(_volume(p) - vol) < 1e-6 * vol
        assert o.is_balanced()
        # children of each split cell tile their parent exactly
        for s, col in enumerate(split):
            assert abs(_volume(kids[8 * s:8 * s + 8]) - _volume(prev[col:col + 1])) < 1e-9 * vol
    # no two leaves overlap: cell centres are distinct and inside exactly one leaf
    c = 0.5 * (p[:, 0::2] + p[:, 1::2])
    inside = ((c[:, None, 0] > p[None, :, 0]) & (c[:, None, 0] < p[None, :, 1]) &
              (c[:, None, 1] > p[None, :, 2]) & (c[:, None, 1] < p[None, :, 3]) &
              (c[:, None, 2] > p[None, :, 4]) & (c[:, None, 2] < p[None, :, 5]))
    assert np.array_equal(inside.sum(axis=1), np.ones(p.shape[0]))


def test_indicator_matches_brute_force_face_overlaps():
    """indicator_calculator: sum over face-adjacent leaf pairs of
    (m_n - m_c)^2 * overlap area, found here by brute force geometry."""
    rng = np.random.default_rng(6)
    o = OctreeMesh((0, 800), 4, (0, 800), 4, (0, 400), 2)
    for _ in range(3):
        o.refine(rng.choice(o.n_cells, size=o.n_cells // 6, replace=False))
    p = o.prisms()
    m = rng.standard_normal(p.shape[0])
    got = o.indicator(m)
    ref = np.zeros(p.shape[0])
    for a in range(3):
        o1, o2 = [b for b in range(3) if b != a]
        lo, hi = p[:, 2 * a], p[:, 2 * a + 1]
        touch = np.isclose(hi[:, None], lo[None, :], rtol=0, atol=1e-9)      # c below n along a
        ox = np.minimum(p[:, 2 * o1 + 1][:, None], p[:, 2 * o1 + 1][None, :]) - \
            np.maximum(p[:, 2 * o1][:, None], p[:, 2 * o1][None, :])
        oy = np.minimum(p[:, 2 * o2 + 1][:, None], p[:, 2 * o2 + 1][None, :]) - \
            np.maximum(p[:, 2 * o2][:, None], p[:, 2 * o2][None, :])
        A = np.where(touch & (ox > 0) & (oy > 0), np.clip(ox, 0, None) * np.clip(oy, 0, None), 0.0)
        d2 = (m[:, None] - m[None, :]) ** 2
        ref += (A * d2).sum(axis=1) + (A * d2).sum(axis=0)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-9 * ref.max())


def test_difference_matrices_level0_equal_regular_operators():
    from graada3d_amd import operators as O
    o = OctreeMesh((0, 400), 5, (0, 300), 4, (0, 200), 3)
    for a, b in zip(o.difference_matrices(), O.difference_matrices(o.prisms(), 5, 4, 3)):
        assert np.array_equal(a.toarray(), b.toarray())


def test_difference_matrices_on_refined_octree():
    """Every row differentiates a linear field exactly (neighbour centres lie
    0.5*(d0+d1) away along the axis; 4 finer neighbours average to the point
    across the face), annihilates constants, and matches a brute-force
    geometric construction of the reference's stencil rules."""
    rng = np.random.default_rng(8)
    o = OctreeMesh((0, 800), 4, (0, 800), 4, (0, 400), 2)
    for _ in range(3):
        o.refine(rng.choice(o.n_cells, size=o.n_cells // 5, replace=False))
    p = o.prisms()
    c = 0.5 * (p[:, 0::2] + p[:, 1::2])
    d = p[:, 1::2] - p[:, 0::2]
    D = o.difference_matrices()
    n = p.shape[0]
    for a in range(3):
        assert np.allclose(D[a] @ c[:, a], 1.0, rtol=0, atol=1e-12)
        assert np.allclose(D[a] @ np.ones(n), 0.0, atol=1e-15)
        # brute force: + face neighbours (touching planes, positive overlap)
        o1, o2 = [b for b in range(3) if b != a]
        ref = np.zeros((n, n))
        for i in range(n):
            def across(side):
                plane = p[i, 2 * a + 1] if side else p[i, 2 * a]
                other = p[:, 2 * a] if side else p[:, 2 * a + 1]
                ov1 = np.minimum(p[i, 2 * o1 + 1], p[:, 2 * o1 + 1]) - np.maximum(p[i, 2 * o1], p[:, 2 * o1])
                ov2 = np.minimum(p[i, 2 * o2 + 1], p[:, 2 * o2 + 1]) - np.maximum(p[i, 2 * o2], p[:, 2 * o2])
                return np.flatnonzero(np.isclose(other, plane, atol=1e-9, rtol=0) & (ov1 > 0) & (ov2 > 0))
            nb, sgn = across(1), 1.0
            if nb.size == 0:
                nb, sgn = across(0), -1.0
            if nb.size == 0:
                continue
            finer = d[nb[0], a] < d[i, a]
            h = 1.0 / (0.5 * (d[i, a] + d[nb[0], a]))
            ref[i, i] = -sgn * h
            ref[i, nb] = sgn * h * (0.25 if finer else 1.0)
        assert np.array_equal(D[a].toarray(), ref), a


def test_batch_instances_and_their_operators():
    from graada3d_amd.workloads import batch_instances, batch_regularisation, smooth_random_field
    inst = batch_instances(n=3, seed=5)
    assert [b.cross_gradient for b in inst] == [False, True, False]
    for b in inst:
        n = b.prisms.shape[0]
        assert 2000 * 0.5 < n < 60000 * 1.5 and b.mesh.is_balanced()
        assert b.points.shape[0] in [m * m for m in range(20, 101)]
        assert b.flag & 1 == 0 and 0 < b.flag <= 0x3FE and b.lp in (0, 2)
        W, T = batch_regularisation(b)
        assert all(w.shape == (n, n) for w in W)
        assert (T is not None) == b.cross_gradient
    f = smooth_random_field(np.random.default_rng(0).uniform(0, 4000, (5000, 3)), 300.0, 1)
    assert abs(f.mean()) < 1e-12 and abs(f.std() - 1.0) < 1e-12


def test_rectilinear_axes_recovers_tensor_meshes():
    """mesh.rectilinear_axes: the node coordinates of a regular_mesh-ordered
    tensor-product mesh, None for anything else (bitwise checks)."""
    from graada3d_amd import mesh as M
    for dims in [(40, 40, 20), (1, 4, 3), (3, 1, 1), (1, 1, 1), (5, 3, 1)]:
        p = M.regular_mesh((0.0, 200.0), dims[0], (-50.0, 10.0), dims[1], (0.0, 30.0), dims[2])
        axes = M.rectilinear_axes(p)
        assert axes is not None and [a.size - 1 for a in axes] == list(dims)
        xs, ys, zs = axes
        assert xs[0] == 0.0 and xs[-1] == 200.0 and ys[0] == -50.0 and zs[-1] == 30.0
    p = M.regular_mesh((0.0, 200.0), 4, (0.0, 10.0), 3, (0.0, 30.0), 2)
    q = p.copy()
    q[5, 1] = np.nextafter(q[5, 1], 1e9)          # one face off by an ulp
    assert M.rectilinear_axes(q) is None
    assert M.rectilinear_axes(p[::-1]) is None
    assert M.rectilinear_axes(p[:-1]) is None
    assert M.rectilinear_axes(np.zeros((0, 6))) is None
