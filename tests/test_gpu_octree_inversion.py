"""The resident Gauss-Newton loop with the cross-gradient term, and an L0
(IRLS) inversion on a refined octree with the octree operators, each step
against the CPU oracle (solve_cg, barrier update, misfit, IRLS weights) at
the FP64 bar of 1e-10."""
import numpy as np
import pytest

import oracle
from graada3d_amd import GaussNewtonInversion
from graada3d_amd import mesh as M
from graada3d_amd import operators as O
from graada3d_amd.octree import OctreeMesh
from helpers import rel_norm

pytestmark = pytest.mark.gpu


def _run(inv, G, W, T, dobs, Wd, steps, lp, eps2, crg=None, its=12):
    nm = inv.Nm
    lo, hi = np.full(nm, -1000.0), np.full(nm, 1000.0)
    kw = {}
    if crg is not None:
        kw = dict(cross_gradient=True, a_crg=crg["a"], S_crg=crg["S"])
    inv.gn_setup(dobs, Wd, lo, hi, np.zeros(nm), np.zeros(nm), lp=lp, eps2=eps2, **kw)
    m = np.zeros(nm)
    WL = {}
    for k in range(steps):
        lam = 20.0 * 0.5 ** k
        mis, err, got_its = inv.gn_step(lam, 0.0, its)
        P = oracle.barrier_P(m, lo, hi)
        c = None if crg is None else dict(a=crg["a"], S=crg["S"], Tx=T[0], Ty=T[1], Tz=T[2])
        dxo, erro, itso = oracle.solve_cg(G, 0.0, its, lam, m, np.zeros(nm), Wd, dobs, P, *W,
                                          WL=WL, crg=c, quirks=True)
        m = oracle.barrier_update(m, dxo, lo, hi, m.copy())
        miso = oracle.misfit(G, m, Wd, dobs)
        if lp != 2:
            WL = {key: oracle.irls_weights(Wq, m, lp, eps2)
                  for key, Wq in zip(("s", "x", "y", "z"), W)}
        assert got_its == itso == its
        assert abs(mis - miso) < 1e-10 * miso, k
        assert rel_norm(inv.gn_model(), m) < 1e-10, k


def test_resident_loop_with_cross_gradient(ctx):
    nx, ny, nz = 10, 9, 5
    prisms = M.regular_mesh((0.0, 500.0), nx, (0.0, 450.0), ny, (0.0, 250.0), nz)
    pts = M.grid_points(8, -20.0, 520.0, 8, -20.0, 470.0, -4.0)
    inv = GaussNewtonInversion(prisms, pts, 0x12, ctx=ctx)
    inv.compute_G()
    G = inv.get_G()
    rng = np.random.default_rng(21)
    W = O.weighting_operators(prisms, nx, ny, nz, O.depth_weights_offset(prisms, 50.0, 3.0))
    T = O.cross_gradient_operators(prisms, nx, ny, nz, rng.standard_normal(prisms.shape[0]))
    inv.set_regularization(*W, *T)
    S = O.cross_gradient_mask(prisms, (100.0, 400.0), (0.0, 450.0), (0.0, 200.0))
    rho = M.model_from_boxes(prisms, [(150, 300, 150, 300, 50, 150, 300.0)])
    dobs = G @ rho
    Wd = np.full(inv.Nd, 1.0 / np.abs(dobs).max())
    _run(inv, G, W, T, dobs, Wd, 3, 2, 0.0, crg=dict(a=0.6, S=S))


def test_l0_inversion_on_refined_octree(ctx):
    rng = np.random.default_rng(22)
    o = OctreeMesh((0.0, 800.0), 8, (0.0, 800.0), 8, (0.0, 400.0), 4)
    for _ in range(2):
        o.refine(rng.choice(o.n_cells, size=o.n_cells // 6, replace=False))
    prisms = o.prisms()
    pts = M.grid_points(9, -40.0, 840.0, 9, -40.0, 840.0, -6.0)
    inv = GaussNewtonInversion(prisms, pts, 0x2 | 0x10 | 0x40, ctx=ctx)
    inv.compute_G()
    G = inv.get_G()
    D = o.difference_matrices()
    W = O.weighting_operators_from(prisms, D, O.depth_weights_offset(prisms, 50.0, 3.0))
    inv.set_regularization(*W)
    rho = M.model_from_boxes(prisms, [(200, 500, 300, 600, 100, 250, -250.0)])
    dobs = G @ rho
    Wd = np.full(inv.Nd, 1.0 / np.abs(dobs).max())
    _run(inv, G, W, None, dobs, Wd, 3, 0, 1e-2)
