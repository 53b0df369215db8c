"""Regularisation operators as CSR (regular meshes here; octrees supply their
own difference matrices through the ``*_from`` builders).

Host-side builders of the inputs that ``solve_cg`` consumes, following
``InversionBase`` on an unrefined mesh:

* ``difference_matrices``  - D_x1, D_y1, D_z1 first differences
  (InversionBase::set_difference_matrix, src/Inversion/InversionBase.cpp:189-434):
  forward difference to the +neighbour over 0.5*(d0+d1); at the + boundary a
  backward difference to the -neighbour; a z row is empty when the cell has
  no z neighbour at all (nz == 1).
* ``V`` = diag(sqrt(volume))   (set_Vmatrix, :452-466)
* depth weighting Z            (set_depth_weighting, :497-528)
* ``W* = a * S * V * D * Z`` with S = I   (GaussNewtonInversion.cpp:711-718,
  set_S :95-144)
* cross-gradient ``T_z = V (Dys D_x1 - Dxs D_y1)`` etc. (set_Tmatrix :468-495)
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .mesh import cell_centres, cell_volumes


def _sizes(prisms):
    return np.abs(prisms[:, 1::2] - prisms[:, 0::2])   # (Nm, 3): dx, dy, dz


def difference_matrices(prisms: np.ndarray, nx: int, ny: int, nz: int):
    """(D_x1, D_y1, D_z1) for a regular nx*ny*nz mesh in leaf order
    id = i*ny*nz + j*nz + k."""
    n = nx * ny * nz
    if prisms.shape[0] != n:
        raise ValueError("prism count does not match nx*ny*nz")
    if nx < 2 or ny < 2:
        raise ValueError("the reference's x/y difference stencils need at least 2 cells")
    d = _sizes(prisms)
    ids = np.arange(n)
    i, rem = np.divmod(ids, ny * nz)
    j, k = np.divmod(rem, nz)
    out = []
    for ax, (idx, cnt, stride) in enumerate([(i, nx, ny * nz), (j, ny, nz), (k, nz, 1)]):
        rows, cols, vals = [], [], []
        fwd = idx < cnt - 1
        nb = np.where(fwd, ids + stride, ids - stride)
        has = fwd | (idx > 0)
        r = ids[has]
        nbr = nb[has]
        h = 1.0 / (0.5 * (d[r, ax] + d[nbr, ax]))
        sgn = np.where(fwd[has], 1.0, -1.0)
        rows = np.concatenate([r, r])
        cols = np.concatenate([r, nbr])
        vals = np.concatenate([-sgn * h, sgn * h])
        out.append(sp.csr_matrix((vals, (rows, cols)), shape=(n, n)))
    return tuple(out)


def depth_weights_reference(prisms: np.ndarray, points: np.ndarray, beta: float) -> np.ndarray:
    """Z_ii = 1/|mean(z_obs) - z_c|^(beta/2)  (set_depth_weighting flag 0)."""
    zp = 0.0
    for z in points[:, 2]:
        zp += z
    zp = zp / points.shape[0]
    zc = 0.5 * (prisms[:, 4] + prisms[:, 5])
    return 1.0 / np.power(np.abs(zp - zc), beta * 0.5)


def depth_weights_sensitivity(col_norms: np.ndarray, Nd: int) -> np.ndarray:
    """Z_ii = sqrt(||G(:, i)||_2 / Nd)  (set_depth_weighting flag 1,
    InversionBase.cpp:519-526); col_norms from Fwd.col_norms()."""
    return np.sqrt(np.asarray(col_norms, float) / Nd)


def depth_weights_offset(prisms: np.ndarray, z0: float, beta: float) -> np.ndarray:
    """Z_ii = (z_c + z0)^(-beta/2), the explicit depth weighting of the
    benchmark inversions."""
    zc = 0.5 * (prisms[:, 4] + prisms[:, 5])
    return 1.0 / np.power(zc + z0, beta * 0.5)


def weighting_operators(prisms, nx, ny, nz, Z, a_s=1e-3, a_x=1.0, a_y=1.0, a_z=1.0):
    """(Ws, Wx, Wy, Wz) with W = a * V * D * Z, entries ((a*V_i)*D_ij)*Z_j."""
    return weighting_operators_from(prisms, difference_matrices(prisms, nx, ny, nz), Z, a_s, a_x,
                                    a_y, a_z)


def weighting_operators_from(prisms, D, Z, a_s=1e-3, a_x=1.0, a_y=1.0, a_z=1.0):
    """weighting_operators for any mesh, given its (D_x1, D_y1, D_z1) -- e.g.
    an octree's (octree.OctreeMesh.difference_matrices)."""
    Dx, Dy, Dz = D
    n = prisms.shape[0]
    V = np.sqrt(cell_volumes(prisms))
    Ds = sp.identity(n, format="csr")

    def w(a, D):
        D = sp.csr_matrix(D)
        D.sort_indices()
        rows = np.repeat(np.arange(n), np.diff(D.indptr))
        vals = ((a * V[rows]) * D.data) * Z[D.indices]
        return sp.csr_matrix((vals, D.indices.copy(), D.indptr.copy()), shape=(n, n))

    return w(a_s, Ds), w(a_x, Dx), w(a_y, Dy), w(a_z, Dz)


def cross_gradient_operators(prisms, nx, ny, nz, m0_s):
    """(T_x, T_y, T_z) of InversionBase::set_Tmatrix for the reference model
    m0_s."""
    return cross_gradient_operators_from(prisms, difference_matrices(prisms, nx, ny, nz), m0_s)


def cross_gradient_operators_from(prisms, D, m0_s):
    """cross_gradient_operators for any mesh, given its (D_x1, D_y1, D_z1)."""
    Dx, Dy, Dz = D
    V = sp.diags(np.sqrt(cell_volumes(prisms)))
    Dxs, Dys, Dzs = (sp.diags(D @ m0_s) for D in (Dx, Dy, Dz))
    Tz = (V @ (Dys @ Dx - Dxs @ Dy)).tocsr()
    Tx = (V @ (Dzs @ Dy - Dys @ Dz)).tocsr()
    Ty = (V @ (Dxs @ Dz - Dzs @ Dx)).tocsr()
    for T in (Tx, Ty, Tz):
        T.sum_duplicates()
        T.sort_indices()
    return Tx, Ty, Tz


def cross_gradient_mask(prisms, x_lim, y_lim, z_lim) -> np.ndarray:
    """S_crg of InversionBase::update_S_crg (InversionBase.cpp:710-728): 1 for
    the cells whose centre lies in the constraint box (its faces included to
    within 1e-10), else 0."""
    c = cell_centres(prisms)
    s = np.ones(prisms.shape[0])
    for a, (lo, hi) in enumerate((x_lim, y_lim, z_lim)):
        v = c[:, a]
        s *= (((v < hi) | (np.abs(v - hi) < 1e-10)) & ((v > lo) | (np.abs(v - lo) < 1e-10)))
    return s


__all__ = ["cross_gradient_mask", "difference_matrices", "depth_weights_reference", "depth_weights_offset",
           "depth_weights_sensitivity", "weighting_operators", "weighting_operators_from",
           "cross_gradient_operators", "cross_gradient_operators_from", "cell_centres"]
