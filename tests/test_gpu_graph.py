"""CG iterations replayed from a captured hipGraph (GRA_TUNE_CG_GRAPH): the
first iteration is launched as usual, the next k are captured once per solve
and the graph is replayed; the kernels take the iteration index from device
memory.  Every replay must give what the launched iterations give, bit for
bit: the same kernels with the same arguments run in the same order.  The
solves below repeat the capture (one per solve) with the program's own eager
work and synchronisations between them, and check each against the eager
result: maxit not a multiple of k (the tail runs launched), a stop on the
tolerance inside a replay (the later kernels of the replay return at once),
and a resident Gauss-Newton loop."""
import numpy as np
import pytest

from graada3d_amd.engine import TIME_CG_GRAPH
from test_gpu_cg import _cg_inputs, small_inversion

pytestmark = pytest.mark.gpu


def _solve(inv, args, aux, maxit, tol=0.0, graph=0):
    inv.ctx.tune(cg_graph=graph)
    inv.ctx.timing(True)
    inv.ctx.timing_reset()
    try:
        out = inv.solve_cg(tol, maxit, 2.0, *args[:3], aux["dobs"], args[3])
        replays = inv.ctx.timing_get(TIME_CG_GRAPH)[1]
    finally:
        inv.ctx.timing(False)
        inv.ctx.tune(cg_graph=0)
    return out, replays


@pytest.mark.parametrize("maxit,k", [(21, 5), (40, 13), (12, 1), (9, 20)])
def test_graph_replay_bit_identical(ctx, maxit, k):
    inv, G, aux = small_inversion(ctx, crg=True)
    args = _cg_inputs(inv, aux)[:4]
    (dx0, err0, its0), _ = _solve(inv, args, aux, maxit)
    x = np.random.default_rng(1).standard_normal(inv.Nm)
    for rep in range(3):
        inv.G_vec_mul(x)                       # eager work between the solves
        (dx, err, its), replays = _solve(inv, args, aux, maxit, graph=k)
        assert replays == ((maxit - 1) // k if maxit > k else 0), rep
        assert np.array_equal(dx, dx0) and err == err0 and its == its0, rep


def test_graph_stop_inside_a_replay(ctx):
    inv, G, aux = small_inversion(ctx)
    args = _cg_inputs(inv, aux)[:4]
    (dx0, err0, its0), _ = _solve(inv, args, aux, 400, tol=1e-6)
    assert 10 < its0 < 400 and its0 % 10 != 1
    (dx, err, its), replays = _solve(inv, args, aux, 400, tol=1e-6, graph=10)
    assert replays == 39
    assert its == its0 and err == err0 and np.array_equal(dx, dx0)


def test_graph_gauss_newton_steps(ctx):
    inv, G, aux = small_inversion(ctx)
    nm = inv.Nm
    lo, hi = np.full(nm, -1000.0), np.full(nm, 1000.0)
    Wd = np.full(inv.Nd, 1.0 / (0.01 * np.abs(aux["dobs"]).max()))
    out = []
    for graph in (0, 7):
        ctx.tune(cg_graph=graph)
        try:
            inv.gn_setup(aux["dobs"], Wd, lo, hi, np.zeros(nm), np.zeros(nm), lp=1, eps2=1e-3)
            mis = [inv.gn_step(40.0 * 0.5 ** k, 0.0, 25)[0] for k in range(3)]
            out.append((mis, inv.gn_model()))
        finally:
            ctx.tune(cg_graph=0)
    assert out[0][0] == out[1][0] and np.array_equal(out[0][1], out[1][1])
