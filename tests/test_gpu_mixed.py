"""Mixed-precision CG (gra_mixed_precision): the iteration's G products on an
FP32 / BF16 copy of the FP64 G, the residual recomputed with the FP64 G every
R iterations, the stop test on those FP64 residuals only.

Pinned by the mathematics of the converged solve: the fixed point of the
reference's iteration satisfies A_r^T (A_t dx - b) = 0 for its operator pair
(A_t the operator of t, A_r that of r; equal with the consistent operator),
computed here by a dense direct solve of the stacked system.  On the
cross-gradient test problem (cond(A_r^T A_t) = 4.6e3) FP64 CG reaches
err < 1e-22 in about 430 iterations, within ~2e-11 of the direct solution;
the mixed-precision solves must reach the same tolerance (it is tested on
FP64 residuals) and land within 1e-9 of the direct solution (1e-8 with the
quirk operator pair, whose A_r^T A_t is not symmetric), a bound set by
the convergence tolerance (relative residual 1e-11 times the conditioning),
not by the copy's rounding (6e-8 FP32, 4e-3 BF16), which the refreshes
remove.  Ill-conditioned systems whose truncated CG iterates are used as they
are (the benchmark inversions) are outside what refreshes can recover: there
reduced-precision G changes the iterates themselves, and bench.py reports
that error.
"""
import numpy as np
import pytest

from graada3d_amd.engine import TIME_REFRESH, GraError
from helpers import rel_norm
from test_cgls_oracle import stacked
from test_gpu_cg import _cg_inputs, small_inversion

pytestmark = pytest.mark.gpu

TOL = 1e-22
LAM = 3.0


def _problem(ctx, wl, seed=0):
    inv, G, aux = small_inversion(ctx, crg=True, seed=seed)
    mk, mref, Wd, P = _cg_inputs(inv, aux)[:4]
    WL = {}
    if wl:
        rng = np.random.default_rng(11)
        WL = {k: rng.uniform(0.8, 1.25, inv.Nm) for k in ("s", "x", "y", "z", "tx", "ty", "tz")}
    Tx, Ty, Tz = aux["T"]
    p = dict(G=G, mk=mk, mref=mref, Wd=Wd, dobs=aux["dobs"], P=P, Ws=aux["Ws"], Wx=aux["Wx"],
             Wy=aux["Wy"], Wz=aux["Wz"], WL=WL,
             crg=dict(a=inv.a_crg, S=inv.S_crg, Tx=Tx, Ty=Ty, Tz=Tz))
    return inv, aux, p


def _direct(p, quirks):
    A_t, b, A_r, _ = stacked(p, LAM, quirks)
    At, Ar = np.vstack(A_t), np.vstack(A_r)
    return np.linalg.solve(Ar.T @ At, Ar.T @ b)


def _solve(inv, p, quirks, fused, maxit=2000):
    kw = {f"WL_{k}": v for k, v in p["WL"].items()}
    return inv.solve_cg(TOL, maxit, LAM, p["mk"], p["mref"], p["Wd"], p["dobs"], p["P"],
                        quirks=quirks, fused=fused, **kw)


@pytest.mark.parametrize("inner", ["f32", "bf16"])
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("quirks,wl", [(False, False), (True, True)])
def test_mixed_converges_to_the_fp64_solution(ctx, inner, fused, quirks, wl):
    inv, aux, p = _problem(ctx, wl)
    sol = _direct(p, quirks)
    dx64, err64, its64 = _solve(inv, p, quirks, fused=False)
    assert its64 < 2000 and err64 < TOL
    R = 10
    inv.mixed_precision(inner, R)
    try:
        inv.ctx.timing(True)
        inv.ctx.timing_reset()
        dx, err, its = _solve(inv, p, quirks, fused)
        _, refresh_passes = inv.ctx.timing_get(TIME_REFRESH)
        inv.ctx.timing(False)
        refreshes = inv.ctx.tuning()["last_refreshes"]
        dx2, err2, its2 = _solve(inv, p, quirks, fused)
    finally:
        inv.ctx.timing(False)
        inv.mixed_precision("f64")
    assert its < 2000 and err < TOL
    assert its % R == 0 and refreshes == its // R      # it stopped on a refreshed residual
    # with the single-read iteration each refresh is one affine single-read
    # pass over the FP64 G (G x and G^T w from one read), otherwise two; the
    # host queues every refresh point up to maxit (2,000) -- those after the
    # device-side stop return at once -- and times each pass
    assert refresh_passes == (1 if fused else 2) * (2000 // R)
    # with the quirk pair A_r^T A_t is not symmetric and CG's attainable
    # accuracy is lower: the FP64 solve itself lands ~2e-9 from the direct one
    bar = 1e-8 if quirks else 1e-9
    assert rel_norm(dx64, sol) < bar
    assert rel_norm(dx, sol) < bar
    # bit-for-bit repeatable
    assert np.array_equal(dx, dx2) and err == err2 and its == its2


def test_mixed_truncated_solve_reports_the_fp64_residual(ctx):
    """A solve stopped by maxit between refresh points still ends on an FP64
    residual: its err equals ||A_r^T (A_t dx - b)||^2 / ||A_r^T b||^2 of the
    returned dx, evaluated densely."""
    inv, aux, p = _problem(ctx, wl=True)
    A_t, b, A_r, _ = stacked(p, LAM, True)
    At, Ar = np.vstack(A_t), np.vstack(A_r)
    inv.mixed_precision("bf16", 7)
    try:
        dx, err, its = _solve(inv, p, True, True, maxit=23)
        refreshes = inv.ctx.tuning()["last_refreshes"]
    finally:
        inv.mixed_precision("f64")
    assert its == 23 and refreshes == 4                # after 7, 14, 21 and the last
    r = Ar.T @ (At @ dx - b)
    atb = Ar.T @ b
    assert abs(err - (r @ r) / (atb @ atb)) < 1e-9 * err


def test_mixed_off_and_invalidation(ctx):
    """gra_mixed_precision(GRA_F64) and any new G return the context to the
    plain FP64 iteration (bit-identical results); the copy needs an FP64 G."""
    inv, aux, p = _problem(ctx, wl=False)
    ref = _solve(inv, p, True, False, maxit=30)
    inv.mixed_precision("bf16", 5)
    inv.mixed_precision("f64")
    off = _solve(inv, p, True, False, maxit=30)
    assert np.array_equal(off[0], ref[0]) and off[1:] == ref[1:]
    inv.mixed_precision("bf16", 5)
    inv.compute_G("f64")                 # a new G: the copy is stale
    again = _solve(inv, p, True, False, maxit=30)
    assert np.array_equal(again[0], ref[0]) and inv.ctx.tuning()["last_refreshes"] == 0
    inv.compute_G("bf16")
    try:
        with pytest.raises(GraError):
            inv.mixed_precision("bf16", 5)
    finally:
        inv.compute_G("f64")
    with pytest.raises(GraError):
        inv.mixed_precision("bf16", 0)
