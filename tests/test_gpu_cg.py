"""GPU G products and the device-resident CGLS (through the C ABI) against
the CPU oracle.

Bars (FP64 throughout; the reductions run in a different but fixed order):
G*x and G^T*x <= 1e-14 relative; solve_cg dx <= 1e-10 relative (the plain
FP64 path's stated bound) with identical iteration counts; the resident
Gauss-Newton loop's misfits and model <= 1e-10 relative; every GPU result
bit-for-bit repeatable.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from graada3d_amd import GaussNewtonInversion
from graada3d_amd import mesh as M
from graada3d_amd import operators as O
from helpers import rel_norm

pytestmark = pytest.mark.gpu


def small_inversion(ctx, nx=12, ny=10, nz=6, npt=9, flag=0x12, crg=False, seed=0):
    prisms = M.regular_mesh((0.0, 600.0), nx, (0.0, 500.0), ny, (0.0, 300.0), nz)
    pts = M.grid_points(npt, -50.0, 650.0, npt, -50.0, 550.0, -5.0)
    inv = GaussNewtonInversion(prisms, pts, flag, ctx=ctx)
    inv.compute_G()
    Z = O.depth_weights_offset(prisms, 50.0, 3.0)
    Ws, Wx, Wy, Wz = O.weighting_operators(prisms, nx, ny, nz, Z, a_s=1e-3)
    rng = np.random.default_rng(seed)
    T = None
    if crg:
        m0s = rng.standard_normal(prisms.shape[0])
        T = O.cross_gradient_operators(prisms, nx, ny, nz, m0s)
        inv.set_regularization(Ws, Wx, Wy, Wz, *T)
        inv.use_cross_gradient_constraint = True
        inv.a_crg = 0.8
        inv.S_crg = rng.uniform(0.5, 1.5, prisms.shape[0])
    else:
        inv.set_regularization(Ws, Wx, Wy, Wz)
    rho = M.model_from_boxes(prisms, [(200, 350, 150, 300, 60, 160, 400.0)])
    G = inv.get_G()
    d = G @ rho
    d = d + 0.01 * np.abs(d).max() * rng.standard_normal(d.size)
    return inv, G, dict(Ws=Ws, Wx=Wx, Wy=Wy, Wz=Wz, T=T, dobs=d, rng=rng)


def test_gemv_vs_oracle(ctx):
    inv, G, _ = small_inversion(ctx, nx=20, ny=18, nz=9, npt=15)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(inv.Nm)
    y = rng.standard_normal(inv.Nd)
    assert rel_norm(inv.G_vec_mul(x), oracle.gemv(G, x)) < 1e-14
    assert rel_norm(inv.GT_vec_mul(y), oracle.gemv(G, y, trans=True)) < 1e-14
    a = inv.G_vec_mul(x)
    assert np.array_equal(a, inv.G_vec_mul(x))           # repeatable


@pytest.mark.parametrize("store", ["f64", "f32", "bf16"])
def test_col_norms_vs_numpy(ctx, store):
    """||G(:, j)||_2 of the stored values (set_depth_weighting flag 1)."""
    inv, G, _ = small_inversion(ctx, nx=20, ny=18, nz=9, npt=15)
    inv.compute_G(store)
    try:
        Gs = inv.get_G()
        n = inv.col_norms()
    finally:
        inv.compute_G("f64")
    ref = np.linalg.norm(Gs, axis=0)
    assert np.abs(n / ref - 1.0).max() < 1e-14
    if store == "f64":
        assert np.array_equal(Gs, G)


def _cg_inputs(inv, aux, wl=False):
    nm, nd = inv.Nm, inv.Nd
    rng = aux["rng"]
    mk = rng.uniform(-50, 50, nm)
    mref = np.zeros(nm)
    Wd = 1.0 / (0.01 * np.abs(aux["dobs"]).max() + 0 * aux["dobs"])
    P = oracle.barrier_P(mk, np.full(nm, -1000.0), np.full(nm, 1000.0))
    WL = {}
    if wl:
        for k in ("s", "x", "y", "z", "tx", "ty", "tz"):
            WL[k] = rng.uniform(0.5, 2.0, nm)
    return mk, mref, Wd, P, WL


@pytest.mark.parametrize("quirks,wl,crg", [(True, False, False), (True, True, False),
                                           (False, True, False), (True, True, True),
                                           (False, True, True)])
def test_solve_cg_vs_oracle(ctx, quirks, wl, crg):
    inv, G, aux = small_inversion(ctx, crg=crg)
    mk, mref, Wd, P, WL = _cg_inputs(inv, aux, wl)
    lam = 3.0
    kw = {f"WL_{k}": v for k, v in WL.items()}
    dx, err, its = inv.solve_cg(1e-30, 30, lam, mk, mref, Wd, aux["dobs"], P, quirks=quirks, **kw)
    crg_d = None
    if crg:
        Tx, Ty, Tz = aux["T"]
        crg_d = dict(a=inv.a_crg, S=inv.S_crg, Tx=Tx, Ty=Ty, Tz=Tz)
    dxo, erro, itso = oracle.solve_cg(G, 1e-30, 30, lam, mk, mref, Wd, aux["dobs"], P, aux["Ws"],
                                      aux["Wx"], aux["Wy"], aux["Wz"], WL=WL, crg=crg_d,
                                      quirks=quirks)
    assert its == itso == 30
    assert rel_norm(dx, dxo) < 1e-10
    assert abs(err - erro) < 1e-8 * abs(erro)
    dx2, err2, its2 = inv.solve_cg(1e-30, 30, lam, mk, mref, Wd, aux["dobs"], P, quirks=quirks,
                                   **kw)
    assert np.array_equal(dx, dx2) and err == err2


def test_solve_cg_tolerance_stop(ctx):
    inv, G, aux = small_inversion(ctx)
    mk, mref, Wd, P, _ = _cg_inputs(inv, aux)
    dx, err, its = inv.solve_cg(1e-5, 500, 1.0, mk, mref, Wd, aux["dobs"], P)
    dxo, erro, itso = oracle.solve_cg(G, 1e-5, 500, 1.0, mk, mref, Wd, aux["dobs"], P, aux["Ws"],
                                      aux["Wx"], aux["Wy"], aux["Wz"])
    assert its == itso < 500 and err < 1e-5
    assert rel_norm(dx, dxo) < 1e-9


def test_gauss_newton_resident_loop_vs_oracle(ctx):
    inv, G, aux = small_inversion(ctx)
    nm = inv.Nm
    lo, hi = np.full(nm, -1000.0), np.full(nm, 1000.0)
    Wd = np.full(inv.Nd, 1.0 / (0.01 * np.abs(aux["dobs"]).max()))
    inv.gn_setup(aux["dobs"], Wd, lo, hi, np.zeros(nm), np.zeros(nm))
    m = np.zeros(nm)
    for k in range(4):
        lam = 100.0 * 0.5 ** k
        mis, err, its = inv.gn_step(lam, 0.0, 20)
        P = oracle.barrier_P(m, lo, hi)
        dxo, erro, itso = oracle.solve_cg(G, 0.0, 20, lam, m, np.zeros(nm), Wd, aux["dobs"], P,
                                          aux["Ws"], aux["Wx"], aux["Wy"], aux["Wz"])
        m = oracle.barrier_update(m, dxo, lo, hi, m.copy())
        miso = oracle.misfit(G, m, Wd, aux["dobs"])
        assert its == itso == 20
        assert abs(mis - miso) < 1e-10 * miso, k
        assert rel_norm(inv.gn_model(), m) < 1e-10, k


def test_gauss_newton_l1_irls_vs_oracle(ctx):
    """Resident loop with IRLS (p = 1): the weights of step k come from the
    model accepted at step k-1 and enter solve_cg with the reference's
    operator asymmetries (t_s without WL_s)."""
    inv, G, aux = small_inversion(ctx)
    nm = inv.Nm
    lo, hi = np.full(nm, -1000.0), np.full(nm, 1000.0)
    Wd = np.full(inv.Nd, 1.0 / (0.01 * np.abs(aux["dobs"]).max()))
    inv.gn_setup(aux["dobs"], Wd, lo, hi, np.zeros(nm), np.zeros(nm), lp=1, eps2=1e-3)
    m = np.zeros(nm)
    WL = {}
    for k in range(4):
        lam = 50.0 * 0.5 ** k
        mis, err, its = inv.gn_step(lam, 0.0, 15)
        P = oracle.barrier_P(m, lo, hi)
        dxo, erro, itso = oracle.solve_cg(G, 0.0, 15, lam, m, np.zeros(nm), Wd, aux["dobs"], P,
                                          aux["Ws"], aux["Wx"], aux["Wy"], aux["Wz"], WL=WL,
                                          quirks=True)
        m = oracle.barrier_update(m, dxo, lo, hi, m.copy())
        miso = oracle.misfit(G, m, Wd, aux["dobs"])
        WL = {"s": oracle.irls_weights(aux["Ws"], m, 1, 1e-3),
              "x": oracle.irls_weights(aux["Wx"], m, 1, 1e-3),
              "y": oracle.irls_weights(aux["Wy"], m, 1, 1e-3),
              "z": oracle.irls_weights(aux["Wz"], m, 1, 1e-3)}
        assert its == itso == 15
        assert abs(mis - miso) < 1e-10 * miso, k
        assert rel_norm(inv.gn_model(), m) < 1e-10, k


def test_launch_shapes_do_not_change_results_beyond_rounding(ctx):
    inv, G, _ = small_inversion(ctx, nx=20, ny=18, nz=9, npt=15)
    rng = np.random.default_rng(2)
    x, y = rng.standard_normal(inv.Nm), rng.standard_normal(inv.Nd)
    a, b = inv.G_vec_mul(x), inv.GT_vec_mul(y)
    for kw in (dict(n_block=0, n_rows=8, n_unroll=1, n_splits=3, nontemporal=0),
               dict(n_block=0, n_rows=2, n_unroll=2, n_splits=1, t_chunks=1, t_unroll=4),
               dict(t_chunks=37, t_unroll=8, nontemporal=1),
               dict(n_block=1, n_rows=4, n_unroll=1, n_splits=1),
               dict(n_block=1, n_rows=2, n_unroll=2, n_splits=3, nontemporal=0),
               dict(n_block=1, n_rows=8, n_unroll=2, n_splits=2, nontemporal=1)):
        inv.ctx.tune(**kw)
        assert rel_norm(inv.G_vec_mul(x), a) < 1e-14
        assert rel_norm(inv.GT_vec_mul(y), b) < 1e-14
    inv.ctx.tune(n_rows=0, n_unroll=0, n_splits=0, t_chunks=0, t_unroll=0, nontemporal=-1, n_block=-1)


def test_autotune_keeps_results(ctx):
    inv, G, _ = small_inversion(ctx, nx=20, ny=18, nz=9, npt=15)
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(inv.Nm), rng.standard_normal(inv.Nd)
    try:
        ms_n, ms_t = inv.ctx.autotune(reps=2)
        t = inv.ctx.tuning()
        assert ms_n > 0 and ms_t > 0
        assert t["n_rows"] in (2, 4, 8) and t["n_unroll"] in (1, 2) and 1 <= t["n_splits"] <= 16
        assert t["n_block"] in (0, 1)
        assert 1 <= t["t_chunks"] <= 32 and t["t_unroll"] in (4, 8) and 0 <= t["nontemporal"] <= 3
        assert rel_norm(inv.G_vec_mul(x), oracle.gemv(G, x)) < 1e-14
        assert rel_norm(inv.GT_vec_mul(y), oracle.gemv(G, y, trans=True)) < 1e-14
    finally:
        inv.ctx.tune(n_rows=0, n_unroll=0, n_splits=0, t_chunks=0, t_unroll=0, nontemporal=-1, n_block=-1)


def test_helpers_vs_oracle(ctx):
    inv, G, aux = small_inversion(ctx)
    nm = inv.Nm
    rng = np.random.default_rng(3)
    m = rng.uniform(-900, 900, nm)
    lo, hi = np.full(nm, -1000.0), np.full(nm, 1000.0)
    assert np.array_equal(inv.barrier_P(m, lo, hi), oracle.barrier_P(m, lo, hi))
    dx = rng.standard_normal(nm)
    dx[5] = 1000.0      # exp overflow -> sequential fallback
    dx[0] = 2000.0
    mt0 = rng.uniform(-500, 500, nm)
    a = inv.barrier_update(m, dx, lo, hi, mt0)
    b = oracle.barrier_update(m, dx, lo, hi, mt0)
    assert rel_norm(a, b) < 1e-15 and abs(a[5] - b[5]) < 1e-12 and abs(a[0] - b[0]) < 1e-12
    Wd = np.full(inv.Nd, 2.0)
    assert abs(inv.misfit(m, Wd, aux["dobs"]) - oracle.misfit(G, m, Wd, aux["dobs"])) < 1e-13 * \
        oracle.misfit(G, m, Wd, aux["dobs"])
    for p in (0, 1):
        assert rel_norm(inv.irls_weights(aux["Wx"], m, p, 1e-3),
                        oracle.irls_weights(aux["Wx"], m, p, 1e-3)) < 1e-14


def test_full_size_products_properties(ctx):
    """At the 5,000 x 32,000 benchmark size: sampled rows of G against the
    oracle, and the adjoint identity (G x).y == x.(G^T y)."""
    from graada3d_amd.workloads import forward_problem_small
    fp = forward_problem_small()
    print("[gpu] full-size G 5000 x 32000", flush=True)
    inv = GaussNewtonInversion(fp.prisms, fp.points, fp.flag, ctx=ctx)
    inv.compute_G()
    rng = np.random.default_rng(5)
    x = rng.standard_normal(inv.Nm)
    y = rng.standard_normal(inv.Nd)
    Gx = inv.G_vec_mul(x)
    GTy = inv.GT_vec_mul(y)
    assert abs(Gx @ y - x @ GTy) < 1e-12 * np.linalg.norm(Gx) * np.linalg.norm(y)
    rows = rng.choice(fp.points.shape[0], 12, replace=False)
    Go, _ = oracle.assemble(fp.prisms, fp.points[rows], fp.flag)
    ref = Go.reshape(2, 12, -1)
    for k in range(2):
        for q, r in enumerate(rows):
            e = np.zeros(inv.Nd)
            e[k * fp.points.shape[0] + r] = 1.0
            row = inv.GT_vec_mul(e)
            assert rel_norm(row, ref[k, q]) < 1e-12
