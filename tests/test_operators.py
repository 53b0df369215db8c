"""Regularisation operators on regular meshes (InversionBase.cpp:189-528),
pinned by closed forms: first differences are exact on linear fields,
cross-gradients of a field with itself vanish, W = a V D Z entrywise."""
import numpy as np
import scipy.sparse as sp

from graada3d_amd import mesh as M
from graada3d_amd import operators as O


def test_differences_exact_on_linear_fields():
    prisms = M.regular_mesh((0.0, 300.0), 6, (0.0, 200.0), 4, (0.0, 150.0), 3)
    Dx, Dy, Dz = O.difference_matrices(prisms, 6, 4, 3)
    c = M.cell_centres(prisms)
    f = 2.0 * c[:, 0] - 3.0 * c[:, 1] + 0.5 * c[:, 2] + 7.0
    assert np.allclose(Dx @ f, 2.0, atol=1e-12)
    assert np.allclose(Dy @ f, -3.0, atol=1e-12)
    assert np.allclose(Dz @ f, 0.5, atol=1e-12)
    for D in (Dx, Dy, Dz):
        assert np.all(np.diff(D.indptr) == 2)


def test_boundary_rows_use_backward_difference():
    prisms = M.regular_mesh((0.0, 100.0), 2, (0.0, 100.0), 2, (0.0, 100.0), 2)
    Dx, _, _ = O.difference_matrices(prisms, 2, 2, 2)
    d = Dx.toarray()
    # cell id 4 (i=1) is at the +x boundary: +1/h on itself, -1/h on id 0
    assert d[4, 4] == 1.0 / 50.0 and d[4, 0] == -1.0 / 50.0
    assert d[0, 0] == -1.0 / 50.0 and d[0, 4] == 1.0 / 50.0


def test_single_layer_has_empty_z_rows():
    prisms = M.regular_mesh((0.0, 100.0), 2, (0.0, 100.0), 2, (0.0, 50.0), 1)
    _, _, Dz = O.difference_matrices(prisms, 2, 2, 1)
    assert Dz.nnz == 0


def test_weighting_operator_entries():
    prisms = M.regular_mesh((0.0, 300.0), 3, (0.0, 300.0), 3, (0.0, 200.0), 2)
    Z = O.depth_weights_offset(prisms, 50.0, 3.0)
    Ws, Wx, Wy, Wz = O.weighting_operators(prisms, 3, 3, 2, Z, a_s=1e-3, a_x=2.0)
    V = np.sqrt(M.cell_volumes(prisms))
    Dx, _, _ = O.difference_matrices(prisms, 3, 3, 2)
    ref = 2.0 * sp.diags(V) @ Dx @ sp.diags(Z)
    assert np.allclose(Wx.toarray(), ref.toarray(), rtol=1e-15, atol=0)
    assert np.allclose(Ws.diagonal(), 1e-3 * V * Z, rtol=1e-15)
    zc = 0.5 * (prisms[:, 4] + prisms[:, 5])
    assert np.allclose(Z, (zc + 50.0) ** -1.5, rtol=1e-15)


def test_depth_weighting_reference_form():
    prisms = M.regular_mesh((0.0, 100.0), 2, (0.0, 100.0), 2, (0.0, 100.0), 2)
    pts = M.grid_points(3, 0.0, 100.0, 3, 0.0, 100.0, -1.0)
    Z = O.depth_weights_reference(prisms, pts, 2.0)
    zc = 0.5 * (prisms[:, 4] + prisms[:, 5])
    assert np.allclose(Z, 1.0 / np.abs(-1.0 - zc), rtol=1e-15)


def test_cross_gradient_of_field_with_itself_vanishes():
    prisms = M.regular_mesh((0.0, 200.0), 4, (0.0, 200.0), 4, (0.0, 100.0), 3)
    rng = np.random.default_rng(0)
    m0 = rng.standard_normal(prisms.shape[0])
    Tx, Ty, Tz = O.cross_gradient_operators(prisms, 4, 4, 3, m0)
    for T in (Tx, Ty, Tz):
        assert np.abs(T @ m0).max() < 1e-12 * np.abs(m0).max()
        assert T.nnz > 0


def test_cross_gradient_mask_box_with_tolerant_faces():
    from graada3d_amd import mesh as M
    from graada3d_amd import operators as O
    pr = M.regular_mesh((0.0, 100.0), 10, (0.0, 100.0), 10, (0.0, 50.0), 5)
    s = O.cross_gradient_mask(pr, (25.0, 55.0), (5.0, 95.0), (0.0, 50.0))
    c = O.cell_centres(pr)
    # centres at 5, 15, ..., 95; x in [25, 55] -> 25, 35, 45, 55 (faces included)
    want = (np.isin(c[:, 0], [25.0, 35.0, 45.0, 55.0]) & (c[:, 1] >= 5.0) & (c[:, 1] <= 95.0))
    assert np.array_equal(s, want.astype(float))
    assert np.array_equal(O.cross_gradient_mask(pr, (25.0 + 1e-11, 55.0), (0, 100), (0, 50)),
                          O.cross_gradient_mask(pr, (25.0, 55.0), (0, 100), (0, 50)))
