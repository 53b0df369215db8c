"""The serving harness: seeded request stream, payload shapes, residency LRU
(CPU), and one short replay on the GPU checked against the oracle."""
import numpy as np
import pytest

from graada3d_amd import serving as S


def test_stream_is_seeded_and_follows_the_mix():
    pool = S.default_cg_pool()
    a = S.request_stream(200.0, 5.0, 11, pool)
    b = S.request_stream(200.0, 5.0, 11, pool)
    assert [(r.t_arrival, r.kind, r.params) for r in a] == \
           [(r.t_arrival, r.kind, r.params) for r in b]
    n = len(a)
    assert 850 < n < 1150                                   # Poisson(1000)
    t = np.array([r.t_arrival for r in a])
    assert np.all(np.diff(t) > 0) and t[-1] < 5.0
    fwd = [r for r in a if r.kind == "forward"]
    assert 0.75 < len(fwd) / n < 0.85
    cells = np.array([r.params["cells"] for r in fwd])
    assert cells.min() >= 5000 and cells.max() <= 130000
    # log-uniform: the median is near the geometric mean sqrt(5e3 * 1.3e5)
    assert 0.7 < np.median(cells) / np.sqrt(5e3 * 1.3e5) < 1.4
    for r in fwd:
        f = r.params["flag"]
        assert f & 1 == 0 and 0 < f <= 0x3FE
        assert 1.0 <= r.params["height"] <= 100.0


def test_forward_payload_shapes():
    p = dict(cells=20000, npts=5000, flag=0x12, height=10.0, seed=3)
    q = S.forward_payload(p)
    prisms, rho, pts = q["prisms"], q["rho"], q["pts"]
    assert abs(prisms.shape[0] - 20000) / 20000 < 0.25
    nx, ny, nz = q["xs"].size - 1, q["ys"].size - 1, q["zs"].size - 1
    assert nx * ny * nz == prisms.shape[0]
    assert np.array_equal(prisms[:, 0], np.repeat(q["xs"][:-1], ny * nz))
    assert np.array_equal(prisms[:, 5], np.tile(q["zs"][1:], nx * ny))
    assert abs(pts.shape[0] - 5000) / 5000 < 0.05
    assert np.all(pts[:, 2] == -10.0) and np.all(prisms[:, 4] >= 0.0)
    assert rho.shape == (prisms.shape[0],) and np.all(np.abs(rho) <= 500.0)


class _FakeCtx:
    def __init__(self, log, name):
        self.log, self.name = log, name

    def close(self):
        self.log.append(self.name)


def test_resident_cache_lru_admission_and_eviction():
    pool = S.default_cg_pool()
    closed = []

    class Cache(S.ResidentCache):
        def _admit(self, prob, need):
            return dict(ctx=_FakeCtx(closed, prob.name), bytes=need)

    sizes = [S.ResidentCache.g_bytes(p) for p in pool]
    budget = sizes[0] + sizes[1] + 1.0           # room for problems 0 and 1 only
    c = Cache(pool, budget)
    info = {}
    c.release(c.get(0, info))
    assert info["admitted"]
    c.release(c.get(1))
    c.release(c.get(0, info))                    # hit; 1 is now least recent
    assert not info["admitted"]
    assert c.admissions == 2 and c.evictions == 0
    c.release(c.get(2))                          # evicts 1 (LRU), maybe 0 too
    assert closed[0] == pool[1].name
    assert c.used <= budget
    assert 2 in c.entries
    with pytest.raises(MemoryError):
        S.ResidentCache(pool, 1.0).get(0)


def test_resident_cache_threads_share_one_admission_and_keep_busy_entries():
    """Two threads asking for the same absent problem admit it once; an
    entry in use is not evicted (the budget is exceeded until release)."""
    import threading
    import time
    pool = S.default_cg_pool()
    closed, admitted = [], []

    class Cache(S.ResidentCache):
        def _admit(self, prob, need):
            admitted.append(prob.name)
            time.sleep(0.05)
            return dict(ctx=_FakeCtx(closed, prob.name), bytes=need)

    sizes = [S.ResidentCache.g_bytes(p) for p in pool]
    c = Cache(pool, max(sizes) + 1.0)            # room for one problem at a time
    got = []
    ts = [threading.Thread(target=lambda: got.append(c.get(0))) for _ in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert admitted == [pool[0].name] and got[0] is got[1] and got[0]["busy"] == 2
    e1 = c.get(1)                                # 0 is busy: kept, budget exceeded
    assert closed == [] and 0 in c.entries and 1 in c.entries
    for e in got + [e1]:
        c.release(e)
    c.release(c.get(2))                          # now the idle 0 and 1 go
    assert set(closed) == {pool[0].name, pool[1].name} and c.used <= c.budget


@pytest.mark.gpu
def test_short_replay_on_gpu():
    import oracle
    from graada3d_amd.engine import Context, forward, forward_rectilinear
    pool = [S.CgProblem("tiny", (10, 10, 5), (1000.0, 1000.0, 500.0), 8, 0x12, 1.0)]
    reqs = S.request_stream(40.0, 0.5, 5, pool)
    for r in reqs:                               # keep the forward requests small
        if r.kind == "forward":
            r.params.update(cells=min(r.params["cells"], 6000), npts=min(r.params["npts"], 1500))
    res = S.serve(reqs, pool, budget_bytes=1e9)
    assert res["requests"] == len(reqs) > 5
    lat, svc = res["latency_ms"], res["service_ms"]
    assert 0 < svc["p50"] <= lat["p50"] and lat["p99"] <= lat["max"]
    assert res["cache"]["admissions"] <= 1
    # the served forward path against the oracle on one request's payload
    r = next(q for q in reqs if q.kind == "forward")
    q = S.forward_payload(r.params)
    ctx = Context(0)
    try:
        got = forward_rectilinear(q["xs"], q["ys"], q["zs"], q["rho"], q["pts"][:50],
                                  r.params["flag"], ctx=ctx)
        got_p = forward(q["prisms"], q["rho"], q["pts"][:50], r.params["flag"], ctx=ctx)
    finally:
        ctx.close()
    ref, st = oracle.forward(q["prisms"], q["rho"], q["pts"][:50], r.params["flag"])
    for k in range(got.shape[1]):
        assert np.abs(got[:, k] - ref[:, k]).max() <= 1e-9 * np.abs(ref[:, k]).max()
        assert np.abs(got_p[:, k] - ref[:, k]).max() <= 1e-10 * np.abs(ref[:, k]).max()


@pytest.mark.gpu
def test_results_alone_equal_in_the_mix():
    """Each request's result in the concurrent mix (3 server streams with
    requests overlapping on the device, problems admitted as memory allows)
    equals its result served alone (one stream, one request on the GPU at a
    time), bit for bit: the forward arrays and the CG solves' dx, error and
    iteration count."""
    pool = [S.CgProblem("tiny", (10, 10, 5), (1000.0, 1000.0, 500.0), 8, 0x12, 0.5),
            S.CgProblem("tiny9", (10, 10, 5), (1000.0, 1000.0, 500.0), 6, 0x3FE, 0.5)]
    reqs = S.request_stream(400.0, 0.1, 11, pool, p_forward=0.6)    # a burst
    for r in reqs:
        if r.kind == "forward":
            r.params.update(cells=min(r.params["cells"], 6000), npts=min(r.params["npts"], 1500))
    kinds = {r.rid: r.kind for r in reqs}
    assert {"forward", "cg"} <= set(kinds.values())
    mix = S.serve(reqs, pool, budget_bytes=1e9, workers=3, outputs=True)
    alone = S.serve(reqs, pool, budget_bytes=1e9, workers=1, outputs=True)
    assert mix["max_in_service"] >= 2 and alone["max_in_service"] == 1
    assert set(mix["outputs"]) == set(alone["outputs"]) == set(kinds)
    for rid, a in alone["outputs"].items():
        b = mix["outputs"][rid]
        if kinds[rid] == "forward":
            assert a.dtype == np.float32 and np.array_equal(a, b), rid
        else:
            assert np.array_equal(a["dx"], b["dx"]), rid
            assert a["err"] == b["err"] and a["iterations"] == b["iterations"] == 20, rid
