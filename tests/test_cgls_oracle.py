"""The CGLS oracle (restating GaussNewtonInversion::solve_cg, .cpp:1417-1698)
and the Gauss-Newton helpers, pinned by:

* convergence to the least-squares solution of the stacked system
  [Wd G P; sqrt(l) WL_s Ws P; ...] (numpy lstsq) when the operator is
  consistent (quirks off), with and without the cross-gradient rows;
* an independent dense-matrix restatement of the reference's iteration with
  its operator asymmetries on (quirks on), written from the reference's
  expressions as whole-matrix products;
* quirks on == quirks off when the asymmetric weights are the identity;
* closed forms for the barrier update (m=0, bounds +-1 gives tanh(dx/2)),
  including the reference's sequential overflow fallback;
* numpy for misfit and IRLS weights.
The reference's solve_cg cannot be compiled here (Fwd.h includes GSL headers
the image lacks), so no reference-run vector pins the unconverged iterates
with quirks on: that part is pinned by the dense restatement only.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle


def problem(seed, nd=40, nm=24, crg=False, wl=False):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((nd, nm))
    mk = rng.standard_normal(nm)
    mref = 0.1 * rng.standard_normal(nm)
    Wd = rng.uniform(0.5, 2.0, nd)
    dobs = rng.standard_normal(nd)
    P = rng.uniform(0.2, 1.5, nm)

    def rs(density=0.15):
        m = sp.random(nm, nm, density=density, random_state=rng.integers(1 << 30), format="csr")
        return (m + sp.identity(nm)).tocsr()

    Ws, Wx, Wy, Wz = rs(), rs(), rs(), rs()
    WL = {}
    if wl:
        for k in ("s", "x", "y", "z", "tx", "ty", "tz"):
            WL[k] = rng.uniform(0.5, 2.0, nm)
    c = None
    if crg:
        c = dict(a=0.7, S=rng.uniform(0.5, 1.5, nm), Tx=rs(0.2), Ty=rs(0.2), Tz=rs(0.2))
    return dict(G=G, mk=mk, mref=mref, Wd=Wd, dobs=dobs, P=P, Ws=Ws, Wx=Wx, Wy=Wy, Wz=Wz,
                WL=WL, crg=c)


def stacked(p, lam, quirks=False):
    """Dense A (t operator, P applied) and b of the stacked least-squares system."""
    nm = p["mk"].size
    I = np.ones(nm)
    WL = {k: p["WL"].get(k, I) for k in ("s", "x", "y", "z", "tx", "ty", "tz")}
    sl = np.sqrt(lam)
    D = np.diag
    P = D(p["P"])
    blocks = [D(p["Wd"]) @ p["G"],
              sl * (D(I if quirks else WL["s"]) @ p["Ws"].toarray()),
              sl * D(WL["z"]) @ p["Wz"].toarray(), sl * D(WL["x"]) @ p["Wx"].toarray(),
              sl * D(WL["y"]) @ p["Wy"].toarray()]
    b = [p["Wd"] * (p["dobs"] - p["G"] @ p["mk"]),
         -sl * WL["s"] * (p["Ws"] @ (p["mk"] - p["mref"])),
         -sl * WL["z"] * (p["Wz"] @ p["mk"]), -sl * WL["x"] * (p["Wx"] @ p["mk"]),
         -sl * WL["y"] * (p["Wy"] @ p["mk"])]
    rblocks = [D(p["Wd"]) @ p["G"], sl * D(WL["s"]) @ p["Ws"].toarray(), blocks[2], blocks[3],
               blocks[4]]
    if p["crg"] is not None:
        c = p["crg"]
        T = {k: D(c["a"] * c["S"]) @ c[k].toarray() for k in ("Tx", "Ty", "Tz")}
        tw = (WL["z"], WL["x"], WL["y"]) if quirks else (WL["tz"], WL["tx"], WL["ty"])
        blocks += [D(tw[0]) @ T["Tz"], D(tw[1]) @ T["Tx"], D(tw[2]) @ T["Ty"]]
        b += [-WL["tz"] * (T["Tz"] @ p["mk"]), -WL["tx"] * (T["Tx"] @ p["mk"]),
              -WL["ty"] * (T["Ty"] @ p["mk"])]
        rblocks += [D(WL["tz"]) @ T["Tz"], D(WL["tx"]) @ T["Tx"], D(WL["ty"]) @ T["Ty"]]
        rblocks_init = rblocks[:-1] + [D(WL["y"] if quirks else WL["ty"]) @ T["Ty"]]
    else:
        rblocks_init = rblocks
    return [B @ P for B in blocks], np.concatenate(b), [B @ P for B in rblocks], \
        [B @ P for B in rblocks_init]


def dense_cgls(p, lam, maxit, tol, quirks, recurrence=False):
    """The reference's iteration, restated with dense matrices.  With
    ``recurrence`` the gradient is updated as r + alpha A_r^T t, starting from
    A_r^T s0 (the loop operator) in the first iteration."""
    A_t, b, A_r, A_r0 = stacked(p, lam, quirks)
    At = np.vstack(A_t)
    Ar = np.vstack(A_r)
    Ar0 = np.vstack(A_r0)
    s = -b
    r = Ar0.T @ s
    r_loop0 = Ar.T @ s
    ATb = Ar.T @ b
    ATb2 = ATb @ ATb
    p_ = np.zeros_like(r)
    dx = np.zeros_like(r)
    beta = 0.0
    err, its = np.nan, maxit
    for i in range(maxit):
        p_ = -r + beta * p_
        # t = A (P p): the P is folded into the blocks
        t = At @ p_
        alpha = (r @ r) / (t @ t)
        s = s + alpha * t
        dx = dx + alpha * p_
        rl = r @ r
        if recurrence:
            r = (r_loop0 if i == 0 else r) + alpha * (Ar.T @ t)
        else:
            r = Ar.T @ s
        beta = (r @ r) / rl
        err = (r @ r) / ATb2
        if err < tol:
            its = i + 1
            break
    return dx, err, its


def run_oracle(p, lam, maxit, tol, quirks, recurrence=False):
    return oracle.solve_cg(p["G"], tol, maxit, lam, p["mk"], p["mref"], p["Wd"], p["dobs"],
                           p["P"], p["Ws"], p["Wx"], p["Wy"], p["Wz"], WL=p["WL"], crg=p["crg"],
                           quirks=quirks, recurrence=recurrence)


@pytest.mark.parametrize("crg", [False, True])
def test_converges_to_least_squares(crg):
    p = problem(1, crg=crg, wl=True)
    lam = 0.3
    dx, err, its = run_oracle(p, lam, 400, 1e-28, quirks=False)
    A_t, b, _, _ = stacked(p, lam, quirks=False)
    sol = np.linalg.lstsq(np.vstack(A_t), b, rcond=None)[0]
    assert np.linalg.norm(dx - sol) < 1e-8 * np.linalg.norm(sol)


@pytest.mark.parametrize("crg,wl", [(False, False), (False, True), (True, True)])
def test_matches_dense_restatement_with_quirks(crg, wl):
    p = problem(2, crg=crg, wl=wl)
    lam = 2.5
    dx, err, its = run_oracle(p, lam, 15, 1e-30, quirks=True)
    dx2, err2, its2 = dense_cgls(p, lam, 15, 1e-30, quirks=True)
    assert its == its2 == 15
    assert np.linalg.norm(dx - dx2) < 1e-10 * np.linalg.norm(dx2)
    assert abs(err - err2) < 1e-8 * abs(err2)


@pytest.mark.parametrize("quirks,crg,wl", [(True, False, False), (True, True, True),
                                           (False, True, True)])
def test_recurrence_matches_dense_restatement(quirks, crg, wl):
    """The single-read iteration's update r <- r + alpha A^T t (with the loop
    operator's A^T s0 as the first base) against the dense restatement; with
    quirks and cross-gradient the first base differs from the initial
    residual, so using the latter fails here by O(1)."""
    p = problem(5, crg=crg, wl=wl)
    dx, err, its = run_oracle(p, 2.5, 10, 1e-30, quirks, recurrence=True)
    dx2, err2, its2 = dense_cgls(p, 2.5, 10, 1e-30, quirks, recurrence=True)
    assert its == its2 == 10
    assert np.linalg.norm(dx - dx2) < 1e-11 * np.linalg.norm(dx2)
    assert abs(err - err2) < 1e-11 * abs(err2)


@pytest.mark.parametrize("quirks,crg", [(True, False), (True, True), (False, True)])
def test_recurrence_equals_reference_iteration_to_rounding(quirks, crg):
    """r = A^T s and r + alpha A^T t are the same vector in exact arithmetic:
    a few iterations on a well-conditioned problem agree to rounding."""
    p = problem(6, crg=crg, wl=True)
    a = run_oracle(p, 1.0, 6, 1e-30, quirks)
    b = run_oracle(p, 1.0, 6, 1e-30, quirks, recurrence=True)
    assert np.linalg.norm(a[0] - b[0]) < 1e-12 * np.linalg.norm(a[0])
    assert abs(a[1] - b[1]) < 1e-10 * a[1]


def test_quirks_inert_when_weights_are_identity():
    p = problem(3, crg=True, wl=False)
    a = run_oracle(p, 1.0, 12, 1e-30, quirks=True)
    b = run_oracle(p, 1.0, 12, 1e-30, quirks=False)
    assert np.array_equal(a[0], b[0]) and a[1] == b[1]


@pytest.mark.parametrize("crg", [False, True])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_quirk_iteration_converges_to_its_operator_pair_solution(crg, seed):
    """With quirks on, t uses the operator A_t (t_s without WL_s, cross-gradient
    rows with WL_z/x/y) and r the operator A_r (A^T with WL_s, WL_t*): the
    converged iterate must solve A_r^T (A_t dx - b) = 0, the dense direct
    solution of that square system (not the least-squares solution of either
    operator).  The iteration reaches it to ~1e-11..1e-14 in 200 steps."""
    p = problem(seed, crg=crg, wl=True)
    lam = 0.3
    A_t, b, A_r, _ = stacked(p, lam, quirks=True)
    At, Ar = np.vstack(A_t), np.vstack(A_r)
    sol = np.linalg.solve(Ar.T @ At, Ar.T @ b)
    dx, err, its = run_oracle(p, lam, 200, 1e-30, quirks=True)
    assert np.linalg.norm(dx - sol) < 1e-9 * np.linalg.norm(sol)
    # and it is not the consistent operator's least-squares solution
    ls = np.linalg.lstsq(At, b, rcond=None)[0]
    assert np.linalg.norm(ls - sol) > 1e-6 * np.linalg.norm(sol)


def _coincide(crg):
    """Random weights where the swapped ones coincide: WL_s = 1 (quirk i,
    GN.cpp:1573), WL_tz/tx/ty = WL_z/x/y (quirks ii and iii, :1582-1584,
    :1508 vs :1661)."""
    p = problem(7, crg=crg, wl=True)
    WL = p["WL"]
    WL["s"] = np.ones_like(WL["s"])
    WL["tz"], WL["tx"], WL["ty"] = WL["z"].copy(), WL["x"].copy(), WL["y"].copy()
    return p


@pytest.mark.parametrize("crg", [False, True])
def test_quirks_reduce_to_the_consistent_iteration_where_weights_coincide(crg):
    """Each quirk swaps one weight for another.  Where the swapped weights
    coincide while every other weight stays random, the quirk iteration is
    bit-identical to the consistent one (which lstsq pins); which weight the
    reference uses in each place is fixed only by its source text."""
    p = _coincide(crg)
    a = run_oracle(p, 1.0, 12, 1e-30, quirks=True)
    b = run_oracle(p, 1.0, 12, 1e-30, quirks=False)
    assert np.array_equal(a[0], b[0]) and a[1] == b[1]


@pytest.mark.parametrize("quirk", ["i", "ii", "ii+iii"])
def test_each_quirk_acts_when_its_weights_differ(quirk):
    """From the coinciding weights, changing only the weights a quirk swaps
    makes the quirk iteration differ from the consistent one (and match the
    dense restatement of the reference's expressions)."""
    p = _coincide(crg=quirk != "i")
    WL = p["WL"]
    if quirk == "i":
        WL["s"] = WL["s"] * np.linspace(0.7, 1.3, WL["s"].size)
    elif quirk == "ii":
        WL["tz"] = WL["tz"] * 1.2
    else:
        WL["ty"] = WL["ty"] * 0.8
    a = run_oracle(p, 1.0, 12, 1e-30, quirks=True)
    b = run_oracle(p, 1.0, 12, 1e-30, quirks=False)
    assert np.linalg.norm(a[0] - b[0]) > 1e-6 * np.linalg.norm(b[0])
    d = dense_cgls(p, 1.0, 12, 1e-30, quirks=True)
    assert np.linalg.norm(a[0] - d[0]) < 1e-10 * np.linalg.norm(d[0])


def test_stops_at_tolerance_and_counts_iterations():
    p = problem(4)
    dx, err, its = run_oracle(p, 1.0, 200, 1e-6, quirks=True)
    assert its < 200 and err < 1e-6
    dx0, err0, its0 = run_oracle(p, 1.0, 0, 1e-6, quirks=True)
    assert its0 == 0 and np.isnan(err0) and np.all(dx0 == 0)


def test_barrier_update_closed_form_and_fallback():
    m = np.zeros(3)
    lo, hi = -np.ones(3), np.ones(3)
    dx = np.array([0.1, 1000.0, -0.2])          # cell 1 overflows exp
    mt = oracle.barrier_update(m, dx, lo, hi, np.array([0.5, 0.25, 0.125]))
    t0 = np.tanh(0.05)
    assert abs(mt[0] - t0) < 1e-16
    assert abs(mt[1] - 0.5 * ((1.0 - 1e-6 + 0.25) / 2.0 + t0)) < 1e-16
    assert abs(mt[2] - np.tanh(-0.1)) < 1e-16
    # first cell overflowing averages with the *previous* trial value of cell 1
    mt = oracle.barrier_update(m, np.array([1000.0, 0.1, 0.0]), lo, hi, np.array([0.5, 0.3, 0.0]))
    assert abs(mt[0] - 0.5 * ((1.0 - 1e-6 + 0.5) / 2.0 + 0.3)) < 1e-16


def test_barrier_P_misfit_irls_against_numpy():
    rng = np.random.default_rng(8)
    m = rng.uniform(-1, 1, 50)
    lo, hi = np.full(50, -2.0), np.full(50, 3.0)
    P = oracle.barrier_P(m, lo, hi)
    assert np.allclose(P, (hi - m) * (m - lo) / (hi - lo), rtol=1e-15)
    G = rng.standard_normal((30, 50))
    Wd = rng.uniform(1, 2, 30)
    d = rng.standard_normal(30)
    ref = np.sqrt(np.sum((Wd * (G @ m - d)) ** 2) / 30)
    assert abs(oracle.misfit(G, m, Wd, d) - ref) < 1e-13 * ref
    W = sp.random(50, 50, density=0.1, random_state=1, format="csr") + sp.identity(50)
    for pexp in (0, 1, 2):
        WL = oracle.irls_weights(W.tocsr(), m, pexp, 1e-3)
        x = W @ m
        assert np.allclose(WL, (x * x + 1e-3) ** (-(2 - pexp) / 4.0), rtol=1e-14)
