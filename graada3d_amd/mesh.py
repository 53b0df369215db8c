"""Regular prism meshes and observation grids (host-side input builders).

These produce the two arrays the engine consumes, in the reference's order:

* prisms: float64 (Nm, 6) rows ``x0, x1, y0, y1, z0, z1`` in leaf-cell order
  of a level-0 regular mesh, x outermost, z innermost
  (``Mesh::generate_regular_mesh``, src/Forward_modelling/Mesh.cpp:796-846).
* points: float64 (N_obs, 3) rows ``x, y, z`` with x varying fastest
  (``Observation::generate``, src/Forward_modelling/Observation.cpp:48-77).

z points down; observation points above the ground sit at negative z.
"""
from __future__ import annotations

import numpy as np


def linspaced(n: int, low: float, high: float) -> np.ndarray:
    """Grid coordinates as Eigen's ``VectorXd::LinSpaced(n, low, high)`` forms
    them (``low + i*step``, last point exactly ``high``, mirrored from the high
    end when ``|high| < |low|``), as used by Mesh.cpp:507-509."""
    if n == 1:
        return np.array([high], dtype=np.float64)
    step = (high - low) / float(n - 1)
    i = np.arange(n, dtype=np.float64)
    if abs(high) < abs(low):
        out = high - (float(n - 1) - i) * step
        out[0] = low
    else:
        out = low + i * step
        out[-1] = high
    return out


def regular_mesh(x_lim, nx, y_lim, ny, z_lim, nz) -> np.ndarray:
    """Prisms of a regular nx*ny*nz mesh; cell id = i*ny*nz + j*nz + k."""
    xs = linspaced(nx + 1, *x_lim)
    ys = linspaced(ny + 1, *y_lim)
    zs = linspaced(nz + 1, *z_lim)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    return np.stack([xs[i], xs[i + 1], ys[j], ys[j + 1], zs[k], zs[k + 1]], axis=1)


def padded_nodes(lim, n: int, n_pad: int, factor: float, both_sides: bool = True) -> np.ndarray:
    """Node coordinates of one axis of Mesh::generate_regular_mesh_with_padding
    (Mesh.cpp:541-577): n equal cells over lim (Eigen LinSpaced), then n_pad
    cells on each side (x, y) or below (z: both_sides=False) whose sizes grow
    by `factor` per cell from the inner cell size, in the reference's order
    of additions."""
    inner = linspaced(n + 1, *lim)
    lead = n_pad if both_sides else 0
    out = np.empty(n + 1 + lead + n_pad)
    out[lead:lead + n + 1] = inner
    d = inner[1] - inner[0]
    for i in range(n_pad):
        d *= factor
        out[lead + n + 1 + i] = out[lead + n + i] + d
        if both_sides:
            out[lead - 1 - i] = out[lead - i] - d
    return out


def rectilinear_axes(prisms: np.ndarray):
    """(xs, ys, zs) if the prisms are exactly the tensor-product cells of those
    node coordinates in the order of ``regular_mesh`` (cell id = i*ny*nz +
    j*nz + k), else None.  Bitwise comparison: shared faces must carry
    identical coordinates."""
    p = np.ascontiguousarray(prisms, dtype=np.float64).reshape(-1, 6)
    n = p.shape[0]
    if n == 0:
        return None
    # z levels: the leading run of cells in the first (i, j) column
    same_col = (p[:, 0] == p[0, 0]) & (p[:, 2] == p[0, 2])
    nz = int(np.argmin(same_col)) if not same_col.all() else n
    # y columns: the leading run of cells in the first x slab
    same_slab = p[:, 0] == p[0, 0]
    nyz = int(np.argmin(same_slab)) if not same_slab.all() else n
    if nz == 0 or nyz % nz or n % nyz:
        return None
    ny, nx = nyz // nz, n // nyz
    zs = np.append(p[:nz, 4], p[nz - 1, 5])
    ys = np.append(p[0:nyz:nz, 2], p[nyz - nz, 3])
    xs = np.append(p[0:n:nyz, 0], p[n - nyz, 1])
    if not (np.all(np.diff(xs) > 0) and np.all(np.diff(ys) > 0) and np.all(np.diff(zs) > 0)):
        return None
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    ref = np.stack([xs[i], xs[i + 1], ys[j], ys[j + 1], zs[k], zs[k + 1]], axis=1)
    return (xs, ys, zs) if np.array_equal(ref, p) else None


def set_region(prisms: np.ndarray, values: np.ndarray, box, value: float) -> None:
    """Assign ``value`` to every cell fully inside ``box`` = (x0,x1,y0,y1,z0,z1),
    with the 1e-10 boundary tolerance of ``Mesh::set_parameter_in_a_region``
    (Mesh.cpp:1079-1106).  Later boxes overwrite earlier ones."""
    x0, x1, y0, y1, z0, z1 = box
    c = prisms

    def ge(a, b):
        return (a > b) | (np.abs(a - b) < 1e-10)

    def le(a, b):
        return (a < b) | (np.abs(a - b) < 1e-10)

    inside = (ge(c[:, 4], z0) & le(c[:, 5], z1) & ge(c[:, 0], x0) & le(c[:, 1], x1)
              & ge(c[:, 2], y0) & le(c[:, 3], y1))
    values[inside] = value


def model_from_boxes(prisms: np.ndarray, boxes) -> np.ndarray:
    """Density vector from a list of (x0,x1,y0,y1,z0,z1,rho) anomaly boxes, as
    the reference's makeModel tool builds it (src/makeModel.cpp:138-199)."""
    rho = np.zeros(prisms.shape[0])
    for b in boxes:
        set_region(prisms, rho, b[:6], b[6])
    return rho


def grid_points(nx, x0, x1, ny, y0, y1, z) -> np.ndarray:
    """Observation::generate: nx*ny points, x fastest, spacing (end-start)/(n-1)."""
    xs = x0 + np.arange(nx) * ((x1 - x0) / (nx - 1) if nx > 1 else 0.0)
    ys = y0 + np.arange(ny) * ((y1 - y0) / (ny - 1) if ny > 1 else 0.0)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    return np.stack([X.ravel(), Y.ravel(), np.full(X.size, float(z))], axis=1)


def cell_centres(prisms: np.ndarray) -> np.ndarray:
    return np.stack([0.5 * (prisms[:, 1] + prisms[:, 0]), 0.5 * (prisms[:, 3] + prisms[:, 2]),
                     0.5 * (prisms[:, 5] + prisms[:, 4])], axis=1)


def cell_volumes(prisms: np.ndarray) -> np.ndarray:
    d = np.abs(prisms[:, 1::2] - prisms[:, 0::2])
    return d[:, 0] * d[:, 1] * d[:, 2]
