"""The single-read CG iteration (one fused pass over G per iteration) against
the CPU oracle and against the two-pass iteration.

The fused iteration updates the gradient by the recurrence r <- r + alpha A^T t
instead of recomputing r = A^T s, so it agrees with the reference's iteration
to rounding, not bit for bit; CG amplifies that rounding with the iteration
count and the conditioning (on the cross-gradient problem below, a 1e-16
relative perturbation of G moves dx by 1e-13 after 20 iterations and by
3e-10 after 30).  Bars after 20 FP64 iterations: dx <= 1e-10 relative to the
oracle run with the same recurrence (the bar of the two-pass path against the
oracle's reference iteration) and <= 1e-8 relative to the oracle's reference
iteration; the same iteration counts, bit-for-bit repeatable results, and the
fused product actually launched once per iteration.
"""
import numpy as np
import pytest

import oracle
from graada3d_amd.engine import TIME_FUSED, TIME_GEMV_N
from helpers import rel_norm
from test_gpu_cg import _cg_inputs, small_inversion

pytestmark = pytest.mark.gpu

FUSED_TOL = 1e-10       # against the oracle's recurrence
REF_TOL = 1e-8          # against the oracle's reference iteration


def _solve(inv, aux, quirks, WL, fused, its=20, lam=3.0, args=None):
    mk, mref, Wd, P = args if args is not None else _cg_inputs(inv, aux)[:4]
    kw = {f"WL_{k}": v for k, v in WL.items()}
    inv.ctx.timing(True)
    inv.ctx.timing_reset()
    out = inv.solve_cg(1e-30, its, lam, mk, mref, Wd, aux["dobs"], P, quirks=quirks,
                       fused=fused, **kw)
    n_fused = inv.ctx.timing_get(TIME_FUSED)[1]
    n_gemv = inv.ctx.timing_get(TIME_GEMV_N)[1]
    inv.ctx.timing(False)
    return out, (mk, mref, Wd, P), n_fused, n_gemv


def _oracle(inv, G, aux, args, quirks, WL, crg, its=20, lam=3.0, recurrence=True):
    mk, mref, Wd, P = args
    crg_d = None
    if crg:
        Tx, Ty, Tz = aux["T"]
        crg_d = dict(a=inv.a_crg, S=inv.S_crg, Tx=Tx, Ty=Ty, Tz=Tz)
    return oracle.solve_cg(G, 1e-30, its, lam, mk, mref, Wd, aux["dobs"], P, aux["Ws"],
                           aux["Wx"], aux["Wy"], aux["Wz"], WL=WL, crg=crg_d, quirks=quirks,
                           recurrence=recurrence)


@pytest.mark.parametrize("quirks,wl,crg", [(True, False, False), (True, True, False),
                                           (False, True, False), (True, True, True),
                                           (False, True, True)])
def test_fused_solve_cg_vs_oracle(ctx, quirks, wl, crg):
    inv, G, aux = small_inversion(ctx, crg=crg)
    WL = {}
    if wl:
        rng = np.random.default_rng(7)
        WL = {k: rng.uniform(0.5, 2.0, inv.Nm) for k in ("s", "x", "y", "z", "tx", "ty", "tz")}
    (dx, err, its), args, n_fused, n_gemv = _solve(inv, aux, quirks, WL, fused=True)
    assert n_fused == 20 and n_gemv == 1          # 1 G*m_k before the loop
    dxo, erro, itso = _oracle(inv, G, aux, args, quirks, WL, crg)
    assert its == itso == 20
    assert rel_norm(dx, dxo) < FUSED_TOL
    assert abs(err - erro) < 1e-8 * abs(erro)
    dxr, _, _ = _oracle(inv, G, aux, args, quirks, WL, crg, recurrence=False)
    assert rel_norm(dx, dxr) < REF_TOL
    (dx2, err2, _), _, _, _ = _solve(inv, aux, quirks, WL, fused=True, args=args)
    assert np.array_equal(dx, dx2) and err == err2


@pytest.mark.parametrize("wgs", [1, 3, 7, 64, 0])
def test_fused_grid_shapes(ctx, wgs):
    """Grids from 13 workgroups (13 row groups of one 4,096-column tile) to
    the default (one per CU, XCD-aligned): groups of one row block and groups
    of many."""
    inv, G, aux = small_inversion(ctx, nx=20, ny=18, nz=9, npt=25, flag=0x3FE)
    try:
        ctx.tune(fused_wgs=wgs * ((inv.Nm + 255) // 256) if wgs else 0)
        (dx, err, its), args, n_fused, _ = _solve(inv, aux, True, {}, fused=True, its=12)
    finally:
        ctx.tune(fused_wgs=0)
    assert n_fused == 12
    dxo, erro, itso = _oracle(inv, G, aux, args, True, {}, False, its=12)
    assert its == itso == 12
    assert rel_norm(dx, dxo) < FUSED_TOL


def test_fused_failed_wait_reruns_two_pass(ctx):
    """A failed wait inside the fused product (forced from the host at row
    block 0 of every row group) makes the solve rerun with the two-pass iteration from the same
    inputs: the result is bit-identical to a plain two-pass solve, the
    context counts the fallback and stays on the two-pass iteration until
    re-enabled."""
    inv, G, aux = small_inversion(ctx)
    args = _cg_inputs(inv, aux)[:4]
    (dx_2, err_2, its_2), _, _, _ = _solve(inv, aux, True, {}, fused=False, args=args)
    try:
        ctx.tune(fused_fail=0)
        (dx_f, err_f, its_f), _, n_fused, _ = _solve(inv, aux, True, {}, fused=True, args=args)
        assert ctx.tuning()["fused_fallbacks"] == 1
        ctx.tune(fused_fail=-1)
        (dx_g, _, _), _, n_fused2, _ = _solve(inv, aux, True, {}, fused=True, args=args)
        assert n_fused2 == 0 and ctx.tuning()["fused_fallbacks"] == 1
    finally:
        ctx.tune(fused_fail=-1, fused_fallbacks=0)
    assert np.array_equal(dx_f, dx_2) and err_f == err_2 and its_f == its_2
    assert np.array_equal(dx_g, dx_2)
    (dx_h, _, _), _, n_fused3, _ = _solve(inv, aux, True, {}, fused=True, args=args)
    assert n_fused3 == 20


@pytest.mark.parametrize("store", ["f32", "bf16"])
def test_fused_reduced_storage_matches_two_pass(ctx, store):
    """With G stored in FP32/BF16 and FP64 products (fused_math 1) both
    iterations read the same stored values and accumulate in FP64: they agree
    to FP64 rounding of the recurrence."""
    inv, G, aux = small_inversion(ctx, nx=20, ny=18, nz=9, npt=15)
    inv.compute_G(store)
    try:
        ctx.tune(fused_math=1)
        (dx_f, err_f, _), args, n_fused, _ = _solve(inv, aux, True, {}, fused=True, its=20)
        (dx_2, err_2, _), _, n_f2, _ = _solve(inv, aux, True, {}, fused=False, its=20, args=args)
    finally:
        ctx.tune(fused_math=0)
        inv.compute_G("f64")
    assert n_fused == 20 and n_f2 == 0
    assert rel_norm(dx_f, dx_2) < REF_TOL


@pytest.mark.parametrize("store", ["f32", "bf16"])
def test_fused_fp32_products_match_two_pass(ctx, store):
    """FP32 products with FP64 sums (fused_math 2, the BF16 default): a
    thread's products of one row (8 for BF16, 8 for FP32 storage) and of one
    column over a block's rows are summed in FP32 with x and the coefficients
    rounded to FP32 (unit roundoff 6e-8), everything else in FP64.  After
    three CG iterations dx agrees with the two-pass FP64-accumulating run
    within 1e-5 (a bound 100x above the FP32 rounding of the products, far
    below the 4e-3 relative rounding of BF16 storage itself), and the run is
    bit-for-bit repeatable."""
    inv, G, aux = small_inversion(ctx, nx=20, ny=18, nz=9, npt=15)
    inv.compute_G(store)
    try:
        ctx.tune(fused_math=2)
        (dx_f, _, _), args, n_fused, _ = _solve(inv, aux, True, {}, fused=True, its=3)
        (dx_g, _, _), _, _, _ = _solve(inv, aux, True, {}, fused=True, its=3, args=args)
        (dx_2, _, _), _, _, _ = _solve(inv, aux, True, {}, fused=False, its=3, args=args)
    finally:
        ctx.tune(fused_math=0)
        inv.compute_G("f64")
    assert n_fused == 3
    assert np.array_equal(dx_f, dx_g)
    assert 0 < rel_norm(dx_f, dx_2) < 1e-5


def test_fused_tolerance_stop(ctx):
    inv, G, aux = small_inversion(ctx)
    mk, mref, Wd, P, _ = _cg_inputs(inv, aux)
    dx, err, its = inv.solve_cg(1e-5, 500, 1.0, mk, mref, Wd, aux["dobs"], P, fused=True)
    dxo, erro, itso = oracle.solve_cg(G, 1e-5, 500, 1.0, mk, mref, Wd, aux["dobs"], P, aux["Ws"],
                                      aux["Wx"], aux["Wy"], aux["Wz"], recurrence=True)
    assert its == itso < 500 and err < 1e-5
    assert rel_norm(dx, dxo) < 1e-9


def test_fused_gauss_newton_l1_irls_vs_oracle(ctx):
    """Three L1 (IRLS) Gauss-Newton steps with the single-read iteration
    against the oracle's recurrence form.  From the second step on, the
    solve's initial residual takes G^T (Wd^2 (G m - dobs)) from the previous
    step's affine single-read misfit pass instead of a G^T product."""
    inv, G, aux = small_inversion(ctx)
    nm = inv.Nm
    lo, hi = np.full(nm, -1000.0), np.full(nm, 1000.0)
    Wd = np.full(inv.Nd, 1.0 / (0.01 * np.abs(aux["dobs"]).max()))
    inv.gn_setup(aux["dobs"], Wd, lo, hi, np.zeros(nm), np.zeros(nm), lp=1, eps2=1e-3,
                 fused=True)
    m = np.zeros(nm)
    WL = {}
    for k in range(3):
        lam = 50.0 * 0.5 ** k
        mis, err, its = inv.gn_step(lam, 0.0, 15)
        P = oracle.barrier_P(m, lo, hi)
        dxo, erro, itso = oracle.solve_cg(G, 0.0, 15, lam, m, np.zeros(nm), Wd, aux["dobs"], P,
                                          aux["Ws"], aux["Wx"], aux["Wy"], aux["Wz"], WL=WL,
                                          quirks=True, recurrence=True)
        m = oracle.barrier_update(m, dxo, lo, hi, m.copy())
        miso = oracle.misfit(G, m, Wd, aux["dobs"])
        WL = {"s": oracle.irls_weights(aux["Ws"], m, 1, 1e-3),
              "x": oracle.irls_weights(aux["Wx"], m, 1, 1e-3),
              "y": oracle.irls_weights(aux["Wy"], m, 1, 1e-3),
              "z": oracle.irls_weights(aux["Wz"], m, 1, 1e-3)}
        assert its == itso == 15
        assert abs(mis - miso) < 1e-10 * miso, k
        assert rel_norm(inv.gn_model(), m) < 1e-10, k


def test_concurrent_single_read_solves_from_several_streams():
    """Single-read solves queued at once from three threads on three contexts
    (three streams): cooperative grids of one device run one at a time, so
    none of their bounded waits fails, and every result equals the same
    solve run alone, bit for bit."""
    import threading
    from graada3d_amd.engine import Context
    ctxs = [Context(0) for _ in range(3)]
    try:
        probs = []
        for k, c in enumerate(ctxs):
            inv, G, aux = small_inversion(c, nx=16 + 4 * k, ny=12, nz=8, npt=14, seed=k)
            probs.append((inv, aux, _cg_inputs(inv, aux)[:4]))
        alone = [inv.solve_cg(0.0, 40, 2.0, *a[:3], aux["dobs"], a[3], fused=True)
                 for inv, aux, a in probs]
        got = [[None] * 6 for _ in probs]
        errs = []

        def run(q):
            inv, aux, a = probs[q]
            try:
                for rep in range(6):
                    got[q][rep] = inv.solve_cg(0.0, 40, 2.0, *a[:3], aux["dobs"], a[3],
                                               fused=True)
            except BaseException as ex:
                errs.append(ex)

        ts = [threading.Thread(target=run, args=(q,)) for q in range(3)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        assert not errs, errs
        for q, (inv, aux, a) in enumerate(probs):
            assert inv.ctx.tuning()["fused_fallbacks"] == 0
            for rep in range(6):
                dx, err, its = got[q][rep]
                assert np.array_equal(dx, alone[q][0]) and (err, its) == alone[q][1:], (q, rep)
    finally:
        for c in ctxs:
            c.close()
