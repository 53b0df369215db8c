"""Gridded model files and their multilinear interpolation: the reference
(petrophysical) model and the cross-gradient structural model of a GraAda3D
inversion.

File format (InversionBase::create_ref_model_from_data /
create_crg_model_from_data, InversionBase.cpp:1342-1740 / 913-1340): blank
separated numbers, four per record -- three coordinates in `data_order` (a
permutation of "xyz", e.g. "zyx" = z y x val) and the value -- for every node
of an nx x ny x nz grid.  `fast_dimension` (a permutation of "xyz") names the
loop order of the records: its first axis changes fastest, its last slowest.
The grid coordinates of each axis are read from the records, as the reference
takes them: the fastest axis from the first run of records, the middle axis
from the last record of each fast run in the first slow block, the slowest
axis from the last record of each slow block.

Interpolation (Linterp's InterpMultilinear<3, double>, 3rd_party_lib/Linterp/
linterp.h:174-182, 185-218, 310-366): per axis the cell is the last grid node
<= the coordinate (upper_bound - 1); a coordinate below the first node or at
or above the last takes that end node's value (constant extension); inside,
the weight is (x - g[c]) / (g[c+1] - g[c]) clamped to [0, 1], and the eight
node values are combined axis by axis, z innermost, then y, then x, each step
as lo + w (hi - lo).
"""
from __future__ import annotations

import numpy as np

AXES = "xyz"


def _perm(s: str, what: str) -> list:
    if sorted(s) != sorted(AXES) or len(s) != 3:
        raise ValueError(f"{what} must be a permutation of 'xyz' (xyz yxz zxy zyx xzy yzx), "
                         f"got {s!r}")
    return [AXES.index(ch) for ch in s]


def read_gridded_model(path: str, n_xyz, data_order: str, fast_dimension: str):
    """(gx, gy, gz, f) of a gridded model file; f[i, j, k] is the value at
    (gx[i], gy[j], gz[k])."""
    col = _perm(data_order, "data_order")         # column c holds axis col[c]
    loop = _perm(fast_dimension, "fast_dimension")  # loop[0] fastest
    n = [int(v) for v in n_xyz]
    if min(n) < 1:
        raise ValueError(f"grid sizes must be positive, got {n}")
    total = n[0] * n[1] * n[2]
    with open(path) as fh:
        tok = np.array(fh.read().split(), dtype=float)
    if tok.size < 4 * total:
        raise ValueError(f"{path}: {tok.size // 4} records, the {n[0]}x{n[1]}x{n[2]} grid "
                         f"needs {total}")
    rec = tok[:4 * total].reshape(total, 4)
    coord = np.empty((total, 3))
    for c in range(3):
        coord[:, col[c]] = rec[:, c]
    fa, ma, sa = loop                            # fast, middle, slow axis
    nf, nm, ns = n[fa], n[ma], n[sa]
    # record r = s * nm * nf + m * nf + f
    blk = coord.reshape(ns, nm, nf, 3)
    grids = [None, None, None]
    grids[fa] = blk[0, 0, :, fa].copy()
    grids[ma] = blk[0, :, nf - 1, ma].copy()
    grids[sa] = blk[:, nm - 1, nf - 1, sa].copy()
    vals = rec[:, 3].reshape(ns, nm, nf)
    # axes of vals are (slow, middle, fast); reorder to (x, y, z)
    f = np.transpose(vals, [[sa, ma, fa].index(a) for a in range(3)])
    return grids[0], grids[1], grids[2], np.ascontiguousarray(f)


class MultilinearInterpolator:
    """Multilinear interpolation on a tensor grid with constant extension
    outside it."""

    def __init__(self, gx, gy, gz, f):
        self.g = [np.asarray(v, float) for v in (gx, gy, gz)]
        self.f = np.asarray(f, float)
        if self.f.shape != tuple(len(v) for v in self.g):
            raise ValueError(f"values {self.f.shape} do not match the grid "
                             f"{tuple(len(v) for v in self.g)}")

    def _axis(self, d, x):
        g = self.g[d]
        n = len(g)
        c = np.searchsorted(g, x, side="right") - 1      # upper_bound - 1
        below, above = x < g[0], x >= g[-1]
        inner = ~(below | above)
        ci = np.clip(c, 0, max(n - 2, 0))
        w = np.zeros_like(x)
        if n > 1:
            w[inner] = (x[inner] - g[ci[inner]]) / (g[ci[inner] + 1] - g[ci[inner]])
        w = np.clip(w, 0.0, 1.0)
        w[below] = 1.0
        w[above] = 0.0
        lo = np.where(below, 0, np.where(above, n - 1, ci))
        hi = np.where(below, 0, np.where(above, n - 1, ci + 1))
        return lo, hi, w

    def __call__(self, x, y, z):
        x, y, z = (np.atleast_1d(np.asarray(v, float)) for v in (x, y, z))
        (xl, xh, wx), (yl, yh, wy), (zl, zh, wz) = (self._axis(d, v) for d, v in
                                                    enumerate((x, y, z)))
        f = self.f

        def along_z(i, j):
            a, b = f[i, j, zl], f[i, j, zh]
            return a + wz * (b - a)

        def along_y(i):
            a, b = along_z(i, yl), along_z(i, yh)
            return a + wy * (b - a)

        a, b = along_y(xl), along_y(xh)
        return a + wx * (b - a)

    def at_centres(self, prisms):
        """Values at the centres of (x0, x1, y0, y1, z0, z1) cells."""
        p = np.asarray(prisms, float)
        return self(0.5 * (p[:, 0] + p[:, 1]), 0.5 * (p[:, 2] + p[:, 3]),
                    0.5 * (p[:, 4] + p[:, 5]))


def model_from_config(tokens, base_dir: str = "."):
    """The interpolator of a configuration line "file data_order n1 n2 n3
    fast_dimension", the three sizes given in data_order (GraAda3D.cpp:596-660)."""
    import os
    if len(tokens) < 6:
        raise ValueError(f"model line needs 'file data_order n1 n2 n3 fast_dimension', "
                         f"got {' '.join(tokens)!r}")
    path, order = tokens[0], tokens[1]
    col = _perm(order, "data_order")
    n = [0, 0, 0]
    for c in range(3):
        n[col[c]] = int(tokens[2 + c])
    if not os.path.isabs(path):
        path = os.path.normpath(os.path.join(base_dir, path))
    return MultilinearInterpolator(*read_gridded_model(path, n, order, tokens[5]))
