"""Synthetic benchmark workloads (seeded, reproducible).

* ``forward_problem_small``: the 40x40x20 regular mesh (50 m cells) over
  x,y in [0,2000] m, z in [0,1000] m with a +1000 kg/m^3 200 m cube centred at
  (1000,1000,400) m and a -500 kg/m^3 400x200x100 m slab centred at
  (1400,600,600) m; a 50x50 observation grid at z = -1 m; components gz and
  Tzz; Gaussian noise with sigma = 1% of max|d| per component.
* ``inversion_small``: the L2 Gauss-Newton inversion of that data with a
  dense 5,000 x 32,000 FP64 G, Ws/Wx/Wy/Wz first-difference regularisation,
  depth weighting (z + 50)^-1.5, lambda_k = 100 * 0.5^k for 10 outer
  iterations of 50 CG iterations each.
* ``tensor_problem_large``: the 64x64x32 mesh (62.5 m cells) over a
  4000x4000x2000 m volume, 100x100 points at z = -50 m, all 9 components,
  12 seeded random boxes (sides log-uniform in [100, 800] m, contrasts
  uniform in [-800, 800] kg/m^3).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import mesh as M
from .operators import depth_weights_offset, weighting_operators

GZ, TZZ = 1 << 1, 1 << 4
ALL9 = 0x3FE


@dataclass
class ForwardProblem:
    prisms: np.ndarray
    rho: np.ndarray
    points: np.ndarray
    flag: int
    shape: tuple
    seed: int
    sigma: np.ndarray = field(default=None)   # per component


def forward_problem_small(seed: int = 20240601) -> ForwardProblem:
    prisms = M.regular_mesh((0.0, 2000.0), 40, (0.0, 2000.0), 40, (0.0, 1000.0), 20)
    rho = M.model_from_boxes(prisms, [
        (900.0, 1100.0, 900.0, 1100.0, 300.0, 500.0, 1000.0),
        (1200.0, 1600.0, 500.0, 700.0, 550.0, 650.0, -500.0),
    ])
    pts = M.grid_points(50, 0.0, 2000.0, 50, 0.0, 2000.0, -1.0)
    return ForwardProblem(prisms, rho, pts, GZ | TZZ, (40, 40, 20), seed)


def add_noise(d: np.ndarray, n_obs: int, n_comp: int, seed: int, frac: float = 0.01):
    """d (component-blocked, length n_comp*n_obs) + N(0, (frac*max|d_k|)^2)."""
    rng = np.random.default_rng(seed)
    d = d.reshape(n_comp, n_obs)
    sigma = frac * np.abs(d).max(axis=1)
    noisy = d + sigma[:, None] * rng.standard_normal(d.shape)
    return noisy.ravel(), sigma


@dataclass
class InversionProblem:
    fwd: ForwardProblem
    dobs: np.ndarray
    Wd: np.ndarray
    sigma: np.ndarray
    Ws: object
    Wx: object
    Wy: object
    Wz: object
    m_min: np.ndarray
    m_max: np.ndarray
    mref: np.ndarray
    m_ini: np.ndarray
    lambdas: list
    cg_maxit: int
    cg_tol: float


def inversion_small(d_clean: np.ndarray, seed: int = 20240601) -> InversionProblem:
    """Builds the inversion inputs from the noise-free data d_clean
    (component-blocked gz then Tzz, as G's rows)."""
    fp = forward_problem_small(seed)
    n_obs = fp.points.shape[0]
    dobs, sigma = add_noise(d_clean, n_obs, 2, seed)
    Wd = np.repeat(1.0 / sigma, n_obs)
    Z = depth_weights_offset(fp.prisms, 50.0, 3.0)
    Ws, Wx, Wy, Wz = weighting_operators(fp.prisms, 40, 40, 20, Z, a_s=1e-3)
    nm = fp.prisms.shape[0]
    return InversionProblem(
        fp, dobs, Wd, sigma, Ws, Wx, Wy, Wz,
        m_min=np.full(nm, -1500.0), m_max=np.full(nm, 1500.0), mref=np.zeros(nm),
        m_ini=np.zeros(nm), lambdas=[100.0 * 0.5 ** k for k in range(10)], cg_maxit=50,
        cg_tol=0.0)


def tensor_problem_large(seed: int = 20240602) -> ForwardProblem:
    prisms = M.regular_mesh((0.0, 4000.0), 64, (0.0, 4000.0), 64, (0.0, 2000.0), 32)
    rng = np.random.default_rng(seed)
    boxes = []
    for _ in range(12):
        sides = np.exp(rng.uniform(np.log(100.0), np.log(800.0), 3))
        sides[2] = min(sides[2], 2000.0)
        lo = rng.uniform([0, 0, 0], [4000 - sides[0], 4000 - sides[1], 2000 - sides[2]])
        boxes.append((lo[0], lo[0] + sides[0], lo[1], lo[1] + sides[1], lo[2], lo[2] + sides[2],
                      rng.uniform(-800.0, 800.0)))
    rho = M.model_from_boxes(prisms, boxes)
    pts = M.grid_points(100, 0.0, 4000.0, 100, 0.0, 4000.0, -50.0)
    return ForwardProblem(prisms, rho, pts, ALL9, (64, 64, 32), seed)


@dataclass
class BatchInstance:
    mesh: object            # octree.OctreeMesh
    prisms: np.ndarray
    rho: np.ndarray
    points: np.ndarray
    flag: int
    lp: int                 # 0 (L0) or 2 (L2)
    cross_gradient: bool
    seed: int


def _crosses_boundary(prisms: np.ndarray, box) -> np.ndarray:
    """Cells that overlap the box without lying inside it (they straddle a
    box face)."""
    x0, x1, y0, y1, z0, z1 = box
    overlap = ((prisms[:, 0] < x1) & (prisms[:, 1] > x0) & (prisms[:, 2] < y1) &
               (prisms[:, 3] > y0) & (prisms[:, 4] < z1) & (prisms[:, 5] > z0))
    inside = ((prisms[:, 0] >= x0) & (prisms[:, 1] <= x1) & (prisms[:, 2] >= y0) &
              (prisms[:, 3] <= y1) & (prisms[:, 4] >= z0) & (prisms[:, 5] <= z1))
    return overlap & ~inside


def batch_instances(n: int = 64, seed: int = 20240604, extent=(4000.0, 4000.0, 2000.0)):
    """The uneven batch: per instance, a 2:1 octree (cell count log-uniform in
    [2k, 60k]: a regular base of about 40 % of the target, refined up to two
    levels at the faces of two seeded boxes), a square point grid of a
    log-uniform count in [400, 10000] at a height uniform in [0.5, 100] m, a
    random non-empty component subset, L0 or L2 regularisation and, for half
    of the instances, a cross-gradient term."""
    from .octree import OctreeMesh
    rng = np.random.default_rng(seed)
    X, Y, Z = extent
    out = []
    for b in range(n):
        target = float(np.exp(rng.uniform(np.log(2e3), np.log(6e4))))
        nz = max(2, int(round((0.4 * target / 4.0) ** (1.0 / 3.0))))
        o = OctreeMesh((0.0, X), 2 * nz, (0.0, Y), 2 * nz, (0.0, Z), nz)
        boxes = []
        for _ in range(2):
            s = np.exp(rng.uniform(np.log(300.0), np.log(1500.0), 3))
            s[2] = min(s[2], 0.8 * Z)
            lo = rng.uniform([0, 0, 0], [X - s[0], Y - s[1], Z - s[2]])
            boxes.append((lo[0], lo[0] + s[0], lo[1], lo[1] + s[1], lo[2], lo[2] + s[2],
                          float(rng.uniform(-500.0, 500.0))))
        for level in range(2):
            if o.n_cells >= 0.8 * target:
                break
            p = o.prisms()
            hit = np.zeros(p.shape[0], bool)
            for bx in boxes:
                hit |= _crosses_boundary(p, bx[:6])
            hit &= o.levels() == level
            idx = np.flatnonzero(hit)
            room = int((1.3 * target - o.n_cells) // 7)
            if idx.size == 0 or room <= 0:
                break
            if idx.size > room:
                idx = rng.choice(idx, room, replace=False)
            o.refine(idx)
        prisms = o.prisms()
        npts = float(np.exp(rng.uniform(np.log(400.0), np.log(1e4))))
        m = max(20, min(100, int(round(np.sqrt(npts)))))
        pts = M.grid_points(m, 0.0, X, m, 0.0, Y, -float(rng.uniform(0.5, 100.0)))
        ncomp = int(rng.integers(1, 10))
        bits = rng.choice(np.arange(1, 10), ncomp, replace=False)
        flag = int(sum(1 << int(q) for q in bits))
        out.append(BatchInstance(o, prisms, M.model_from_boxes(prisms, boxes), pts, flag,
                                 int(rng.choice([0, 2])), bool(b % 2), seed * 1000 + b))
    return out


def smooth_random_field(centres: np.ndarray, length: float, seed: int, modes: int = 32):
    """A smooth Gaussian-like random field (a sum of random Fourier modes with
    wavelengths around 2*pi*length), unit variance, at the given points."""
    rng = np.random.default_rng(seed)
    k = rng.standard_normal((modes, 3)) / length
    phase = rng.uniform(0.0, 2.0 * np.pi, modes)
    f = np.cos(centres @ k.T + phase).sum(axis=1)
    return (f - f.mean()) / (f.std() + 1e-300)


def batch_regularisation(b: BatchInstance, z0: float = 50.0, beta: float = 3.0):
    """Operators of one batch instance on its octree: (W, T or None), with the
    depth weighting (z + z0)^(-beta/2) and, for the cross-gradient instances,
    T built from a smooth random reference model m0_s (length scale 300 m)."""
    from .operators import (cross_gradient_operators_from, depth_weights_offset,
                            weighting_operators_from)
    D = b.mesh.difference_matrices()
    W = weighting_operators_from(b.prisms, D, depth_weights_offset(b.prisms, z0, beta))
    T = None
    if b.cross_gradient:
        m0s = smooth_random_field(M.cell_centres(b.prisms), 300.0, b.seed)
        T = cross_gradient_operators_from(b.prisms, D, m0s)
    return W, T
