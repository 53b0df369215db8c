"""Octree meshes with the reference's leaf (= G column) order and 2:1 balance.

* Level 0 is the regular mesh of Mesh::generate_regular_mesh
  (Mesh.cpp:796-846): cells in x-outer, y, z-inner order.
* Mesh::refinement (Mesh.cpp:1462-1976) splits a leaf at its centre
  (RectPrism::get_center, 0.5*(a+b)) into 8 children c_child[4i+2j+k] over the
  (x, y, z) halves (Mesh.cpp:1579-1598), which replace the parent in place in
  the leaf list (Mesh.cpp:1934-1941).  Before a cell is split, every face
  neighbour coarser than it is split first (the flag1/flag2/flag3 face tests,
  Mesh.cpp:1474-1556), so face neighbours never differ by more than one
  level.

A second, "reordered" list of the leaves (Mesh::ro_leaf_cells) gives the
column order of the wavelet-compressed G: on the regular mesh each (x, y)
column of cells is walked down or up in z, the direction alternating from
column to column in (x-outer, y-inner) order (Mesh.cpp:855-893); a split
cell is replaced there by its children in the path 0, 2, 6, 4, 5, 7, 3, 1
when its column ran down, and in the reverse path otherwise; children keep
their parent's direction (Mesh.cpp:1605, :1946-1960).  `reordered()` maps a
reordered position to the leaf index (Mesh::get_reordered_id).

`refine(targets)` splits the given leaves (indices in the current leaf list)
and the coarser neighbours their balance needs, and returns what
`Fwd.expand_G` (AdaptiveInversion::expand_G) consumes: the ascending old
indices of every split cell and their children's prisms in child order.
Because children replace parents in place, the new leaf order does not depend
on the order of the splits.
"""
from __future__ import annotations

import numpy as np

from .mesh import linspaced


# child paths of the reordered list (child index 4i + 2j + k)
_PATH_DOWN = (0, 2, 6, 4, 5, 7, 3, 1)
_PATH_UP = (1, 3, 7, 5, 4, 6, 2, 0)


class _Leaf:
    __slots__ = ("level", "key", "bounds", "orig", "children", "down")

    def __init__(self, level, key, bounds, orig=None, down=True):
        self.level, self.key, self.bounds, self.orig = level, key, bounds, orig
        self.children = None
        self.down = down


class OctreeMesh:
    def __init__(self, x_lim, nx, y_lim, ny, z_lim, nz, nodes=None):
        """Level-0 cells: nx x ny x nz equal cells over the limits, or, with
        `nodes` = (xs, ys, zs), the tensor product of those node coordinates
        (Mesh::generate_regular_mesh over given grid points, Mesh.cpp:828-846,
        as the padded mesh builds it)."""
        if nodes is not None:
            xs, ys, zs = (np.asarray(a, dtype=np.float64) for a in nodes)
            nx, ny, nz = xs.size - 1, ys.size - 1, zs.size - 1
        else:
            xs, ys, zs = (linspaced(nx + 1, *x_lim), linspaced(ny + 1, *y_lim),
                          linspaced(nz + 1, *z_lim))
        self.base = (nx, ny, nz)
        self.leaves = []
        for i in range(nx):
            for j in range(ny):
                for k in range(nz):
                    self.leaves.append(_Leaf(0, (i, j, k),
                                             (xs[i], xs[i + 1], ys[j], ys[j + 1], zs[k], zs[k + 1])))
        self._index = {(0,) + lf.key: lf for lf in self.leaves}
        self.ro = []
        down = True
        for i in range(nx):
            for j in range(ny):
                col = self.leaves[(i * ny + j) * nz:(i * ny + j + 1) * nz]
                for lf in col:
                    lf.down = down
                self.ro.extend(col if down else col[::-1])
                down = not down

    @property
    def n_cells(self) -> int:
        return len(self.leaves)

    def prisms(self) -> np.ndarray:
        return np.array([lf.bounds for lf in self.leaves], dtype=float).reshape(-1, 6)

    def levels(self) -> np.ndarray:
        return np.array([lf.level for lf in self.leaves], dtype=np.int64)

    # -- neighbours ---------------------------------------------------------
    def _coarser_neighbour(self, lf: _Leaf, axis: int, side: int):
        L = lf.level
        pos = list(lf.key)
        pos[axis] += 1 if side else -1
        if pos[axis] < 0 or pos[axis] >= self.base[axis] << L:
            return None                                   # domain boundary
        for lv in range(L - 1, -1, -1):
            sh = L - lv
            key = (lv, pos[0] >> sh, pos[1] >> sh, pos[2] >> sh)
            nb = self._index.get(key)
            if nb is not None:
                return nb
        return None

    # -- splitting ----------------------------------------------------------
    def _split(self, lf: _Leaf):
        if lf.children is not None:
            return
        for axis in range(3):
            for side in (0, 1):
                nb = self._coarser_neighbour(lf, axis, side)
                if nb is not None:
                    self._split(nb)
        x0, x1, y0, y1, z0, z1 = lf.bounds
        xm, ym, zm = 0.5 * (x1 + x0), 0.5 * (y1 + y0), 0.5 * (z1 + z0)
        xb, yb, zb = (x0, xm, x1), (y0, ym, y1), (z0, zm, z1)
        L = lf.level + 1
        kids = []
        for i in range(2):
            for j in range(2):
                for k in range(2):
                    key = (2 * lf.key[0] + i, 2 * lf.key[1] + j, 2 * lf.key[2] + k)
                    c = _Leaf(L, key, (xb[i], xb[i + 1], yb[j], yb[j + 1], zb[k], zb[k + 1]),
                              down=lf.down)
                    kids.append(c)
                    self._index[(L,) + key] = c
        del self._index[(lf.level,) + lf.key]
        lf.children = kids

    def refine(self, targets):
        """Split the leaves at `targets` (current indices) plus the coarser
        face neighbours 2:1 balance requires.  Returns (split_cols, child_prisms):
        ascending old leaf indices of all split cells and an (8*n, 6) array of
        their children in child order."""
        for idx, lf in enumerate(self.leaves):
            lf.orig = idx
        for t in sorted(set(int(t) for t in targets)):
            self._split(self.leaves[t])
        split, kids, new = [], [], []
        for lf in self.leaves:
            if lf.children is None:
                new.append(lf)
                continue
            if any(c.children is not None for c in lf.children):
                raise AssertionError("a child was split within the same refinement")
            split.append(lf.orig)
            kids.extend(c.bounds for c in lf.children)
            new.extend(lf.children)
        self.leaves = new
        ro = []
        for lf in self.ro:
            if lf.children is None:
                ro.append(lf)
            else:
                ro.extend(lf.children[c] for c in (_PATH_DOWN if lf.down else _PATH_UP))
        self.ro = ro
        return np.array(split, dtype=np.int64), np.array(kids, dtype=float).reshape(-1, 6)

    def reordered(self) -> np.ndarray:
        """Leaf index of each position of the reordered list."""
        idx = {id(lf): i for i, lf in enumerate(self.leaves)}
        return np.array([idx[id(lf)] for lf in self.ro], dtype=np.int64)

    def _finer_across(self, lf: _Leaf, axis: int, side: int):
        """The 4 children (level L+1) adjacent to lf's face when the cell
        across it is split, else None."""
        L = lf.level
        pos = list(lf.key)
        pos[axis] += 1 if side else -1
        if pos[axis] < 0 or pos[axis] >= self.base[axis] << L:
            return None
        if (L,) + tuple(pos) in self._index:
            return None
        other = [a for a in range(3) if a != axis]
        out = []
        for u in range(2):
            for v in range(2):
                key = [0, 0, 0]
                key[axis] = 2 * pos[axis] + (0 if side else 1)
                key[other[0]] = 2 * pos[other[0]] + u
                key[other[1]] = 2 * pos[other[1]] + v
                c = self._index.get((L + 1,) + tuple(key))
                if c is None:
                    return None
                out.append(c)
        return out

    def indicator(self, m) -> np.ndarray:
        """AdaptiveInversion::indicator_calculator (AdaptiveInversion.cpp:108-204):
        for each leaf, the sum over its face neighbours of (m_nb - m_c)^2 times
        the shared face area (a split neighbour contributes each of its 4 child
        faces); faces on the mesh boundary contribute nothing."""
        m = np.asarray(m, float)
        idx = {id(lf): i for i, lf in enumerate(self.leaves)}
        out = np.zeros(len(self.leaves))

        def area(b, axis):
            d = (b[1] - b[0], b[3] - b[2], b[5] - b[4])
            return d[(axis + 1) % 3] * d[(axis + 2) % 3]

        for i, lf in enumerate(self.leaves):
            mc = m[i]
            acc = 0.0
            for side in (0, 1):
                for axis in range(3):
                    fine = self._finer_across(lf, axis, side)
                    if fine is not None:
                        for nb in fine:
                            d = m[idx[id(nb)]] - mc
                            acc += d * d * area(nb.bounds, axis)
                        continue
                    pos = list(lf.key)
                    pos[axis] += 1 if side else -1
                    if pos[axis] < 0 or pos[axis] >= self.base[axis] << lf.level:
                        continue
                    nb = self._index.get((lf.level,) + tuple(pos))
                    if nb is None:
                        nb = self._coarser_neighbour(lf, axis, side)
                    d = m[idx[id(nb)]] - mc
                    acc += d * d * area(lf.bounds, axis)
            out[i] = acc
        return out

    def difference_matrices(self):
        """(D_x1, D_y1, D_z1) of InversionBase::set_difference_matrix
        (InversionBase.cpp:189-434) on this octree, CSR in leaf order.  Row c,
        axis a: a forward difference to the + neighbour over
        h = 0.5*(d_c + d_n) (a same-size or coarser neighbour; for a split +
        face the mean of its 4 finer neighbours, weight 0.25 each); at the
        + boundary the backward difference to the - neighbour(s) with the same
        rule; a z row is empty when the cell has no z neighbour at all."""
        import scipy.sparse as sp
        n = len(self.leaves)
        idx = {id(lf): i for i, lf in enumerate(self.leaves)}
        out = []
        for axis in range(3):
            rows, cols, vals = [], [], []
            for i, lf in enumerate(self.leaves):
                d0 = lf.bounds[2 * axis + 1] - lf.bounds[2 * axis]

                def neighbours(side):
                    fine = self._finer_across(lf, axis, side)
                    if fine is not None:
                        return fine
                    pos = list(lf.key)
                    pos[axis] += 1 if side else -1
                    if pos[axis] < 0 or pos[axis] >= self.base[axis] << lf.level:
                        return None
                    nb = self._index.get((lf.level,) + tuple(pos))
                    return [nb if nb is not None else self._coarser_neighbour(lf, axis, side)]

                nbs, sgn = neighbours(1), 1.0
                if nbs is None:
                    nbs, sgn = neighbours(0), -1.0
                    if nbs is None:
                        if axis == 2:
                            continue
                        raise ValueError("x/y differences need two cells along the axis")
                d1 = nbs[0].bounds[2 * axis + 1] - nbs[0].bounds[2 * axis]
                h = 1.0 / (0.5 * (d0 + d1))
                rows.append(i)
                cols.append(i)
                vals.append(-sgn * h)
                w = sgn * h * 0.25 if len(nbs) == 4 else sgn * h
                for nb in nbs:
                    rows.append(i)
                    cols.append(idx[id(nb)])
                    vals.append(w)
            D = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
            D.sort_indices()
            out.append(D)
        return tuple(out)

    def is_balanced(self) -> bool:
        """Every face neighbour of every leaf is at most one level coarser."""
        for lf in self.leaves:
            for axis in range(3):
                for side in (0, 1):
                    nb = self._coarser_neighbour(lf, axis, side)
                    if nb is not None and nb.level < lf.level - 1:
                        return False
        return True
