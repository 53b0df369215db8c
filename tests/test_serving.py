"""The serving harness: seeded request stream, payload shapes, the default
problem pool, residency with LRU and GDSF eviction (CPU); short replays on
the GPU checked against the oracle and against each request served alone,
and batched forward launches against single ones."""
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


class _Sized(S.ResidentCache):
    """Cache over fake problems of given sizes and admission times."""

    def __init__(self, sizes, budget, policy, admit_s=None, closed=None):
        super().__init__([S.CgProblem(f"p{k}", (1, 1, 1), (1.0, 1.0, 1.0), 1, 0x2, 1.0)
                          for k in range(len(sizes))], budget, policy=policy)
        self.sizes, self.admit_s_of = sizes, admit_s or [0.0] * len(sizes)
        self.closed = [] if closed is None else closed
        self.admitted = []

    def g_bytes(self, prob):
        return float(self.sizes[int(prob.name[1:])])

    def _admit(self, prob, need, stream_of=None, ctx=None):
        import time
        k = int(prob.name[1:])
        self.admitted.append(k)
        time.sleep(self.admit_s_of[k])
        if ctx is None:
            ctx = _FakeCtx(self.closed, f"buf{len(self.admitted)}")
        ctx.holds = prob.name
        return dict(ctx=ctx, bytes=need)


def test_default_pool_covers_the_benchmark_sizes():
    pool = S.default_cg_pool()
    names = [p.name for p in pool]
    assert names[0] == "inversion_small" and names[-1] == "tensor_large"
    assert len(pool) == 66 and abs(sum(p.weight for p in pool) - 1.0) < 1e-12
    assert all(p.store == "f32" and not p.fused for p in pool)
    assert [p.method for p in pool] == ["nodes"] + ["corners"] * 64 + ["nodes"]
    assert [p.name for p in pool if p.pinned] == ["tensor_large"]
    gb = [S.ResidentCache.g_bytes(p) / 1e9 for p in pool]
    assert abs(gb[0] - 0.64) < 1e-9 and abs(gb[-1] - 47.18592) < 1e-9
    assert 56 < sum(gb[1:-1]) < 58                    # the 64 octree instances


def test_resident_cache_lru_admission_and_eviction():
    c = _Sized([3.0, 4.0, 5.0], 3.0 + 4.0 + 1.0, "lru")    # room for 0 and 1 only
    info = {}
    c.release(c.get(0, info))
    assert info["admitted"]
    c.release(c.get(1))
    c.release(c.get(0, info))                    # hit; 1 is now least recent
    assert not info["admitted"]
    assert c.admissions == 2 and c.evictions == 0
    # 2 (5) fits no buffer: 1 is evicted and its buffer freed (3 + 5 fits)
    c.release(c.get(2))
    assert c.closed == ["buf2"] and list(c.entries) == [0, 2] and c.used == 8.0
    # 1 (4) goes into the buffer of a victim that holds it: 2's (5), LRU
    # order 0 (3, too small), then 2; nothing is freed
    c.release(c.get(1))
    assert c.closed == ["buf2"] and list(c.entries) == [0, 1]
    assert c.entries[1]["ctx"].name == "buf3" and c.entries[1]["bytes"] == 5.0
    assert c.reuses == 1 and c.frees == 1 and c.used == 8.0 <= c.budget
    with pytest.raises(MemoryError):
        _Sized([3.0], 1.0, "lru").get(0)
    with pytest.raises(ValueError):
        S.ResidentCache([], 1.0, policy="fifo")


def test_warm_fills_the_budget_most_requested_first():
    c = _Sized([3.0, 4.0, 5.0, 1.0], 8.5, "gdsf")
    for p, w in zip(c.pool, (0.4, 0.1, 0.3, 0.2)):
        p.weight = w
    assert c.warm() == 2                         # 0 (3), then 2 (5); 3 no longer fits
    assert sorted(c.entries) == [0, 2] and c.admitted == [0, 2]
    assert c.admissions == 0 and c.evictions == 0 and c.used == 8.0


def test_pinned_problems_are_warmed_first_and_never_evicted():
    c = _Sized([6.0, 2.0, 2.0, 3.0], 10.0, "lru")
    c.pool[0].pinned = True
    c.pool[0].weight = 0.0
    assert c.warm() == 3 and c.admitted == [0, 1, 2]       # pinned first, then by weight
    c.release(c.get(3))                                    # frees 1's and 2's buffers, not 0's
    assert c.closed == ["buf2", "buf3"] and sorted(c.entries) == [0, 3]
    c.release(c.get(2))                                    # into 3's buffer
    assert sorted(c.entries) == [0, 2] and c.entries[2]["ctx"].name == "buf4"
    assert c.closed == ["buf2", "buf3"] and c.used == 9.0
    c2 = _Sized([6.0, 5.0], 10.0, "gdsf")
    c2.pool[0].pinned = True
    c2.release(c2.get(0))
    with pytest.raises(MemoryError):                       # 6 pinned + 5 > 10
        c2.get(1)
    assert list(c2.entries) == [0] and c2.used == 6.0


def test_gdsf_keeps_a_frequent_problem_that_is_expensive_to_rebuild():
    """A 10-unit problem that takes 40x longer per byte to admit, asked for
    every fifth request between four small ones, with room for it and three
    small ones: LRU evicts it each cycle (it is the least recent when the
    fourth small one comes), GDSF keeps it and rotates the small ones."""
    sizes = [10.0, 1.0, 1.0, 1.0, 1.0]
    admit_s = [0.2, 0.005, 0.005, 0.005, 0.005]
    seq = [0, 1, 2, 3, 4] * 3
    runs = {}
    for policy in ("lru", "gdsf"):
        c = _Sized(sizes, 13.0, policy, admit_s)
        for pid in seq:
            c.release(c.get(pid))
            assert c.used <= c.budget
        runs[policy] = (c.admitted, sum(admit_s[k] for k in c.admitted))
    assert runs["lru"][0].count(0) == 3
    assert runs["gdsf"][0].count(0) == 1
    assert runs["gdsf"][1] < 0.5 * runs["lru"][1]            # admission time spent


def test_resident_cache_threads_share_one_admission_and_keep_busy_entries():
    """Two threads asking for the same absent problem admit it once; an
    entry in use is not evicted (the budget is exceeded until release)."""
    import threading
    c = _Sized([2.0, 2.0, 2.0], 3.0, "gdsf", [0.05, 0.0, 0.0])    # one problem at a time
    got = []
    ts = [threading.Thread(target=lambda: got.append(c.get(0))) for _ in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert c.admitted == [0] and got[0] is got[1] and got[0]["busy"] == 2
    e1 = c.get(1)                                # 0 is busy: kept, budget exceeded
    assert c.closed == [] and 0 in c.entries and 1 in c.entries and c.used == 4.0
    for e in got + [e1]:
        c.release(e)
    # 2 goes into an idle victim's buffer, and the other idle one is freed
    # to bring the held memory back under the budget
    c.release(c.get(2))
    assert list(c.entries) == [2] and len(c.closed) == 1 and c.used == 2.0 <= c.budget


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
    """Each request's result in the concurrent mix (3 forward and 1 CG server
    streams with requests overlapping on the device, forward requests batched into shared
    launches, problems admitted and evicted under a budget smaller than the
    pool) equals its result served alone (one stream, one request on the GPU
    at a time, no batching), bit for bit: the forward arrays and the CG
    solves' dx, error and iteration count."""
    from graada3d_amd.serving import _instances
    inst = _instances()
    size = [b.prisms.shape[0] * b.points.shape[0] * bin(b.flag).count("1") for b in inst]
    two = [int(k) for k in np.argsort(size)[:2]]
    pool = [S.CgProblem("tiny", (10, 10, 5), (1000.0, 1000.0, 500.0), 8, 0x12, 0.3),
            S.CgProblem("tiny9", (10, 10, 5), (1000.0, 1000.0, 500.0), 6, 0x3FE, 0.2)]
    pool += [S.CgProblem(f"batch_{k:02d}", (0, 0, 0), S.EXTENT, 0, 0, 0.25, instance=k)
             for k in two]
    gb = [S.ResidentCache.g_bytes(p) for p in pool]
    budget = sum(gb) - 0.5 * min(gb[2:])          # the two instances do not fit together
    reqs = S.request_stream(400.0, 0.15, 11, pool, p_forward=0.6)    # a burst
    for r in reqs:
        if r.kind == "forward":
            r.params.update(cells=min(r.params["cells"], 6000), npts=min(r.params["npts"], 1500))
    kinds = {r.rid: r.kind for r in reqs}
    probs = {r.params["problem"] for r in reqs if r.kind == "cg"}
    assert {"forward", "cg"} <= set(kinds.values()) and {2, 3} <= probs
    mix = S.serve(reqs, pool, budget_bytes=budget, workers=3, outputs=True)
    alone = S.serve(reqs, pool, budget_bytes=budget, workers=1, cg_workers=0, max_batch=1,
                    outputs=True)
    assert mix["max_in_service"] >= 2 and alone["max_in_service"] == 1
    assert alone["forward_batches"]["launches"] == 0
    print("mix:", mix["forward_batches"], mix["cache"])
    assert mix["cache"]["evictions"] >= 1 and alone["cache"]["evictions"] >= 1
    assert set(mix["outputs"]) == set(alone["outputs"]) == set(kinds)
    for rid, a in alone["outputs"].items():
        b = mix["outputs"][rid]
        if kinds[rid] == "forward":
            assert a.dtype == np.float32 and np.array_equal(a, b), rid
        else:
            assert np.array_equal(a["dx"], b["dx"]), rid
            assert a["err"] == b["err"] and a["iterations"] == b["iterations"] == 20, rid


@pytest.mark.gpu
def test_results_alone_equal_in_the_mix_at_the_benchmark_sizes():
    """The same bit-identity on the serving mix at its specified sizes: forward
    requests of 5k-130k cells on 1k-20k points with 1-9 components (none
    capped), CG requests on the 5,000 x 32,000 survey, on the largest (11.5 GB)
    and a median octree instance and on the pinned 90,000 x 131,072
    nine-component survey (47 GB of FP32 G), under a budget that holds all but
    half of the smallest problem, so that admissions evict."""
    import dataclasses
    S._instances()
    base = S.default_cg_pool()
    gb = [S.ResidentCache.g_bytes(p) for p in base]
    order = np.argsort(gb[1:-1])
    k_big, k_mid = int(order[-1]), int(order[len(order) // 2])
    pool = [dataclasses.replace(base[0], weight=0.3),
            dataclasses.replace(base[1 + k_big], weight=0.25),
            dataclasses.replace(base[1 + k_mid], weight=0.25),
            dataclasses.replace(base[-1], weight=0.2)]
    assert pool[-1].name == "tensor_large" and pool[-1].pinned
    sizes = [S.ResidentCache.g_bytes(p) for p in pool]
    budget = sum(sizes) - 0.5 * min(sizes)
    reqs = S.request_stream(200.0, 0.5, 22, pool)
    kinds = {r.rid: r.kind for r in reqs}
    fwd = [r.params for r in reqs if r.kind == "forward"]
    per_problem = np.bincount([r.params["problem"] for r in reqs if r.kind == "cg"],
                              minlength=len(pool))
    # the stream reaches the ends of the specified ranges and every problem
    assert len(fwd) > 60 and np.all(per_problem >= 1)
    assert max(p["cells"] for p in fwd) > 100000 and max(p["npts"] for p in fwd) > 15000
    assert min(p["cells"] for p in fwd) < 10000
    print("CG requests per problem:", dict(zip([p.name for p in pool], per_problem.tolist())),
          f"budget {budget / 1e9:.2f} GB of {sum(sizes) / 1e9:.2f}", flush=True)
    mix = S.serve(reqs, pool, budget_bytes=budget, workers=4, cg_workers=1, outputs=True)
    print("mix:", mix["forward_batches"], mix["cache"], flush=True)
    alone = S.serve(reqs, pool, budget_bytes=budget, workers=1, cg_workers=0, max_batch=1,
                    outputs=True)
    print("alone:", alone["cache"], flush=True)
    assert mix["max_in_service"] >= 2 and alone["max_in_service"] == 1
    assert alone["forward_batches"]["launches"] == 0
    assert mix["cache"]["evictions"] >= 1 and alone["cache"]["evictions"] >= 1
    assert set(mix["outputs"]) == set(alone["outputs"]) == set(kinds)
    for rid, a in alone["outputs"].items():
        b = mix["outputs"][rid]
        if kinds[rid] == "forward":
            assert a.dtype == np.float32 and np.array_equal(a, b), rid
        else:
            assert np.array_equal(a["dx"], b["dx"]), rid
            assert a["err"] == b["err"] and a["iterations"] == b["iterations"] == 20, rid


@pytest.mark.gpu
def test_batched_forward_equals_single_launches():
    """gra_forward_rectilinear_batch: uneven meshes, point sets and component
    sets in one launch, each result bit-identical to its own launch."""
    from graada3d_amd.engine import Context, forward_rectilinear, forward_rectilinear_batch
    rng = np.random.default_rng(3)
    qs, flags = [], []
    for k in range(6):
        p = dict(cells=int(rng.integers(500, 9000)), npts=int(rng.integers(20, 700)),
                 flag=int(sum(1 << int(b) for b in rng.choice(np.arange(1, 10), 1 + k % 9,
                                                               replace=False))),
                 height=float(rng.uniform(1, 100)), seed=k)
        qs.append(S.forward_payload(p))
        flags.append(p["flag"])
    with Context(0) as ctx:
        single = [forward_rectilinear(q["xs"], q["ys"], q["zs"], q["rho"], q["pts"], f, ctx=ctx)
                  for q, f in zip(qs, flags)]
        batched = forward_rectilinear_batch(
            [(q["xs"], q["ys"], q["zs"], q["rho"], q["pts"], f) for q, f in zip(qs, flags)],
            ctx=ctx)
        again = forward_rectilinear_batch(
            [(q["xs"], q["ys"], q["zs"], q["rho"], q["pts"], f) for q, f in zip(qs[:2], flags[:2])],
            ctx=ctx)
    for a, b in zip(single, batched):
        assert a.shape == b.shape and np.array_equal(a, b)
    for a, b in zip(single, again):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_solves_on_a_borrowed_stream_are_bit_identical():
    """gra_ctx_set_stream: a context's solves queued on another context's
    stream give its own stream's results (two-pass and single-read)."""
    from graada3d_amd.engine import Context
    from test_gpu_cg import _cg_inputs, small_inversion
    with Context(0) as ctx, Context(0) as other:
        inv, G, aux = small_inversion(ctx)
        args = _cg_inputs(inv, aux)[:4]
        for fused in (False, True):
            ref = inv.solve_cg(0.0, 30, 2.0, *args[:3], aux["dobs"], args[3], fused=fused)
            ctx.use_stream_of(other)
            try:
                got = inv.solve_cg(0.0, 30, 2.0, *args[:3], aux["dobs"], args[3], fused=fused)
                y = inv.G_vec_mul(np.ones(inv.Nm))
            finally:
                ctx.use_stream_of(None)
            assert np.array_equal(got[0], ref[0]) and got[1:] == ref[1:], fused
            assert np.array_equal(y, inv.G_vec_mul(np.ones(inv.Nm)))
