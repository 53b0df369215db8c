"""Batched Gauss-Newton steps (gra_gn_step_batch) over uneven instances:
every instance's step is queued on its own stream before any is waited for,
fused steps chained one after another, two-pass steps overlapping.  Each
instance's misfit, iteration count and model after the steps must be
bit-identical to the same instance stepped alone (gra_gn_step) on a fresh
context: the same kernels run in the same order on each stream."""
import numpy as np
import pytest

from graada3d_amd import GaussNewtonInversion
from graada3d_amd.engine import Context, gn_step_batch
from graada3d_amd.workloads import add_noise, batch_instances, batch_regularisation

pytestmark = pytest.mark.gpu


def _setup(b, ctx, fused):
    f = GaussNewtonInversion(b.prisms, b.points, b.flag, ctx=ctx)
    f.compute_G("f64")
    W, T = batch_regularisation(b)
    f.set_regularization(*W, *(T or ()))
    ncomp = f.Nd // f.N_obs
    dobs, sigma = add_noise(f.G_vec_mul(b.rho), f.N_obs, ncomp, b.seed)
    nm = f.Nm
    f.gn_setup(dobs, np.repeat(1.0 / sigma, f.N_obs), np.full(nm, -1000.0), np.full(nm, 1000.0),
               np.zeros(nm), np.zeros(nm), lp=b.lp, eps2=1e-2 if b.lp != 2 else 0.0,
               fused=fused, cross_gradient=b.cross_gradient, a_crg=1.0)
    return f


@pytest.fixture(scope="module")
def instances():
    inst = batch_instances(n=64)
    # eight of the smaller instances: L0 and L2, with and without the
    # cross-gradient term, uneven sizes
    order = sorted(range(len(inst)), key=lambda q: inst[q].prisms.shape[0] * inst[q].points.shape[0])
    return [inst[q] for q in order[:8]]


@pytest.mark.parametrize("fused", [False, True])
def test_batched_steps_equal_solo_steps(instances, fused):
    steps, its = 3, 12
    lams = [[30.0 * 0.5 ** k * (1 + 0.1 * q) for q in range(len(instances))] for k in range(steps)]
    ctxs = [Context(0) for _ in instances]
    solo_ctx = Context(0)
    try:
        invs = [_setup(b, c, fused) for b, c in zip(instances, ctxs)]
        got = [gn_step_batch(invs, lams[k], 0.0, its) for k in range(steps)]
        for q, b in enumerate(instances):
            f = _setup(b, solo_ctx, fused)
            for k in range(steps):
                mis, err, n = f.gn_step(lams[k][q], 0.0, its)
                assert n == got[k][2][q] == its
                assert mis == got[k][0][q] and err == got[k][1][q], (q, k)
            assert np.array_equal(f.gn_model(), invs[q].gn_model()), q
    finally:
        for c in ctxs:
            c.close()
        solo_ctx.close()
