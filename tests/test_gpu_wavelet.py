"""The wavelet-compressed G on the GPU (gra_assemble_wavelet and the products,
CG and Gauss-Newton loop that use it) against the oracle.

Bars: n_dy and the number of kept coefficients identical; G x and G^T r
within 1e-12 (relative to the largest entry; the transforms round as the
oracle's, the sparse sums run in another order); at threshold 0 the
products equal the dense G's to 1e-12; solve_cg dx within 1e-10 of the
oracle's CGLS on the same operator (its dense equivalent) with identical
iteration counts; the resident Gauss-Newton loop's misfits within 1e-10.
Sizes cover n_dy inside the LDS window, above it (levels through HBM),
an exact power of two and a single cell."""
import numpy as np
import pytest

import oracle
from graada3d_amd import GaussNewtonInversion
from graada3d_amd import mesh as M
from graada3d_amd import operators as O
from graada3d_amd.octree import OctreeMesh
from helpers import rel_norm

pytestmark = pytest.mark.gpu


def _problem(nx, ny, nz, npts, seed=0):
    lim = ((0.0, 40.0 * nx), (0.0, 40.0 * ny), (0.0, 30.0 * nz))
    prisms = M.regular_mesh(lim[0], nx, lim[1], ny, lim[2], nz)
    rng = np.random.default_rng(seed)
    pts = np.column_stack([rng.uniform(-20, lim[0][1] + 20, npts),
                           rng.uniform(-20, lim[1][1] + 20, npts), rng.uniform(-60, -5, npts)])
    order = OctreeMesh(lim[0], nx, lim[1], ny, lim[2], nz).reordered()
    return prisms, pts, order


def _maxrel(a, b):
    return np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("dims,npts,flag,rel", [
    ((5, 4, 3), 7, 0x2, 0.0),                 # n_dy 64, every coefficient kept
    ((8, 6, 5), 9, 0x12, 1e-3),               # gz + Tzz
    ((4, 4, 4), 6, 0x2, 0.01),                # Nm = n_dy = 64
    ((1, 1, 1), 3, 0x2, 0.0),                 # one cell: n_dy = 1
    ((25, 20, 12), 11, 0x2, 2e-4),            # n_dy 8192
    ((40, 25, 20), 6, 0x2, 5e-4),             # n_dy 32768: levels above the LDS window
])
def test_wavelet_matches_oracle(ctx, dims, npts, flag, rel):
    prisms, pts, order = _problem(*dims, npts)
    print(f"wavelet {dims} pts {npts} flag {flag:#x} rel {rel}", flush=True)
    f = GaussNewtonInversion(prisms, pts, flag, ctx=ctx)
    n_dy, nnz = f.compute_G_wavelet(rel, order)
    w = oracle.WaveletG(prisms, pts, flag, order, rel)
    assert (n_dy, nnz) == (w.n_dy, w.nnz)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(f.Nm)
    r = rng.standard_normal(f.Nd)
    y, z = f.G_vec_mul(x), f.GT_vec_mul(r)
    assert _maxrel(y, w.G_vec_mul(x)) < 1e-12
    assert _maxrel(z, w.GT_vec_mul(r)) < 1e-12
    assert np.array_equal(y, f.G_vec_mul(x)) and np.array_equal(z, f.GT_vec_mul(r))
    if rel == 0.0:
        G, _ = oracle.assemble(prisms, pts, flag)
        assert _maxrel(y, G @ x) < 1e-12 and _maxrel(z, G.T @ r) < 1e-12


def test_wavelet_then_dense_again(ctx):
    """gra_assemble returns the context to the dense G; dense-only calls
    refuse the compressed one."""
    prisms, pts, order = _problem(6, 5, 4, 8)
    f = GaussNewtonInversion(prisms, pts, 0x2, ctx=ctx)
    f.compute_G_wavelet(1e-3, order)
    with pytest.raises(Exception, match="wavelet"):
        f.col_norms()
    with pytest.raises(Exception, match="wavelet"):
        f.get_G()
    f.compute_G()
    G = f.get_G()
    Go, _ = oracle.assemble(prisms, pts, 0x2)
    assert _maxrel(G, Go) < 1e-10
    x = np.random.default_rng(3).standard_normal(f.Nm)
    assert _maxrel(f.G_vec_mul(x), G @ x) < 1e-14


def _inversion(ctx, rel):
    nx, ny, nz = 12, 10, 6
    prisms, pts, order = _problem(nx, ny, nz, 40, seed=4)
    inv = GaussNewtonInversion(prisms, pts, 0x12, ctx=ctx)
    inv.compute_G_wavelet(rel, order)
    Z = O.depth_weights_offset(prisms, 50.0, 3.0)
    W = O.weighting_operators(prisms, nx, ny, nz, Z, a_s=1e-3)
    inv.set_regularization(*W)
    w = oracle.WaveletG(prisms, pts, 0x12, order, rel)
    Gw = w.dense()
    rho = M.model_from_boxes(prisms, [(120, 300, 100, 260, 40, 120, 400.0)])
    rng = np.random.default_rng(5)
    d = Gw @ rho
    d = d + 0.01 * np.abs(d).max() * rng.standard_normal(d.size)
    return inv, Gw, W, d


@pytest.mark.parametrize("rel", [0.0, 1e-3])
def test_wavelet_solve_cg_vs_oracle(ctx, rel):
    inv, Gw, W, d = _inversion(ctx, rel)
    nm = inv.Nm
    rng = np.random.default_rng(6)
    mk = rng.uniform(-50, 50, nm)
    Wd = np.full(inv.Nd, 1.0 / (0.01 * np.abs(d).max()))
    P = oracle.barrier_P(mk, np.full(nm, -1000.0), np.full(nm, 1000.0))
    WL = {"s": rng.uniform(0.5, 2, nm), "y": rng.uniform(0.5, 2, nm)}
    dx, err, its = inv.solve_cg(1e-30, 25, 2.0, mk, np.zeros(nm), Wd, d, P, WL_s=WL["s"],
                                WL_y=WL["y"])
    dxo, erro, itso = oracle.solve_cg(Gw, 1e-30, 25, 2.0, mk, np.zeros(nm), Wd, d, P, *W, WL=WL)
    assert its == itso == 25
    assert rel_norm(dx, dxo) < 1e-10
    assert abs(inv.misfit(mk, Wd, d) - oracle.misfit(Gw, mk, Wd, d)) < 1e-12 * oracle.misfit(
        Gw, mk, Wd, d)


def test_wavelet_resident_loop_vs_oracle(ctx):
    """8 CG iterations per step: with 15 this problem's iterates move by
    ~1e-9 under a 1e-16 relative perturbation of G (rounding amplified by the
    CG recurrences), with 8 by ~1e-13, which leaves the 1e-10 bar room."""
    inv, Gw, W, d = _inversion(ctx, 1e-3)
    nm = inv.Nm
    lo, hi = np.full(nm, -1000.0), np.full(nm, 1000.0)
    Wd = np.full(inv.Nd, 1.0 / (0.01 * np.abs(d).max()))
    inv.gn_setup(d, Wd, lo, hi, np.zeros(nm), np.zeros(nm))
    m = np.zeros(nm)
    for k in range(3):
        lam = 50.0 * 0.5 ** k
        mis, err, its = inv.gn_step(lam, 0.0, 8)
        P = oracle.barrier_P(m, lo, hi)
        dxo, _, itso = oracle.solve_cg(Gw, 0.0, 8, lam, m, np.zeros(nm), Wd, d, P, *W)
        m = oracle.barrier_update(m, dxo, lo, hi, m.copy())
        miso = oracle.misfit(Gw, m, Wd, d)
        assert its == itso == 8
        assert abs(mis - miso) < 1e-10 * miso, k
        assert rel_norm(inv.gn_model(), m) < 1e-10, k
