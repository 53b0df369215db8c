"""Request serving on one GPU: an open-loop replay of a seeded request stream.

The reference has no server; this module is the harness around the engine's
two request types, the forward model (compute.cpp:246-260) and the CGLS
inner solve (GaussNewtonInversion::solve_cg, GaussNewtonInversion.cpp:
1417-1698), used to measure request latency (arrival to completion).

Stream (seeded numpy PCG64):
* arrivals: Poisson at `rate` requests/s for `duration` s;
* 80 % forward requests: a regular mesh over 4000 x 4000 x 2000 m with a
  log-uniform cell count in [5k, 130k] (nx = ny = 2 nz), a square grid of a
  log-uniform point count in [1k, 20k] at a height uniform in [1, 100] m
  above the mesh top, a uniformly drawn number (1-9) of distinct components,
  density uniform in [-500, 500] kg/m^3 per cell; run matrix-free in FP64
  (by default the node-sum form for rectilinear meshes,
  gra_forward_rectilinear; `forward_method="prisms"` uses the per-prism
  face/edge kernel, gra_forward) and returned as FP32;
* 20 % inversion requests: 20 CGLS iterations (L2, lambda log-uniform in
  [0.1, 100]) on one of a pool of problems (default_cg_pool: the 5,000 x
  32,000 survey, the 64 uneven octree instances, the 90,000 x 131,072
  nine-component survey) whose G stays resident in HBM between requests
  while a byte budget allows (ResidentCache: admission on demand, eviction
  of idle problems by a size- and cost-aware policy).

Request payloads are built ahead by a producer thread (the client's work);
a dispatcher releases each request at its arrival time into the queue of
its kind; `workers` server threads (default 3) take the forward requests
and `cg_workers` (default 1) the CG requests, each thread with its own
context and stream, so up to workers + cg_workers requests are on the GPU at
once; a CG solve runs on its worker's stream, borrowing the resident
problem's context.
"""
from __future__ import annotations

import collections
import functools
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import mesh as M
from .engine import (Context, GaussNewtonInversion, forward, forward_rectilinear,
                     forward_rectilinear_batch)
from .operators import depth_weights_offset, weighting_operators
from .workloads import add_noise

EXTENT = (4000.0, 4000.0, 2000.0)


@dataclass
class Request:
    rid: int
    t_arrival: float
    kind: str                      # "forward" or "cg"
    params: dict = field(default_factory=dict)


@dataclass
class CgProblem:
    name: str
    shape: tuple                   # (nx, ny, nz); unused for a batch instance
    extent: tuple                  # (x1, y1, z1), origin 0
    n_pts: int                     # points per side
    flag: int
    weight: float
    store: str = "f64"             # storage of the resident G
    fused: bool = False            # single-read CG iteration
    instance: int = -1             # >= 0: index into workloads.batch_instances()
    method: str = "faces"          # assembly: "faces", "nodes" or "corners" (compute_G)
    pinned: bool = False           # once admitted, never evicted (warm() admits it first)


def default_cg_pool():
    """The CG problems of the serving mix at the benchmark workloads' sizes,
    each G stored in FP32 (105 GB of G in all, more than the default serving
    budget holds) and assembled by corner sums (the node form on the two
    regular meshes, per-prism corner sums on the octrees):
    * the 5,000 x 32,000 survey (0.64 GB), 45 % of the solves;
    * each of the 64 uneven octree instances (0.02-11.5 GB, 57 GB in all),
      54 % of the solves spread evenly over them;
    * the 90,000 x 131,072 nine-component survey (47 GB), 1 %, pinned:
      admitting or evicting it maps or unmaps 47 GB, which stalled the
      concurrent forward requests for hundreds of milliseconds.
    The solves use the two-pass iteration: the single-read one is a
    cooperative grid that holds every CU for each iteration, which stalls the
    concurrent forward requests."""
    n_inst = len(_instances())
    pool = [CgProblem("inversion_small", (40, 40, 20), (2000.0, 2000.0, 1000.0), 50, 0x12, 0.45,
                      store="f32", method="nodes")]
    for k in range(n_inst):
        pool.append(CgProblem(f"batch_{k:02d}", (0, 0, 0), EXTENT, 0, 0, 0.54 / n_inst,
                              store="f32", instance=k, method="corners"))
    pool.append(CgProblem("tensor_large", (64, 64, 32), EXTENT, 100, 0x3FE, 0.01,
                          store="f32", method="nodes", pinned=True))
    return pool


# single-read solves only where each row group holds at least this many
# row blocks (shorter groups run the two-pass iteration faster)
FUSED_MIN_BLOCKS = 24


_HOST = {}
_HOST_LOCK = threading.Lock()


@functools.lru_cache(maxsize=1)
def _instances():
    """workloads.batch_instances(), generated once per process."""
    from .workloads import batch_instances
    return batch_instances()


def problem_inputs(prob: CgProblem):
    """Host-side inputs of a CG problem (built once per process): prisms,
    points, component flag, true model, and the regularisation operators."""
    with _HOST_LOCK:
        got = _HOST.get(prob.name)
        if got is not None:
            return got
    if prob.instance >= 0:
        from .workloads import batch_regularisation
        b = _instances()[prob.instance]
        W, _ = batch_regularisation(b)
        out = dict(prisms=b.prisms, pts=b.points, flag=b.flag, rho=b.rho, W=W)
    else:
        x1, y1, z1 = prob.extent
        nx, ny, nz = prob.shape
        prisms = M.regular_mesh((0.0, x1), nx, (0.0, y1), ny, (0.0, z1), nz)
        pts = M.grid_points(prob.n_pts, 0.0, x1, prob.n_pts, 0.0, y1, -1.0)
        Z = depth_weights_offset(prisms, 50.0, 3.0)
        W = weighting_operators(prisms, nx, ny, nz, Z, a_s=1e-3)
        c = (0.5 * x1, 0.5 * y1, 0.3 * z1)
        rho = M.model_from_boxes(prisms, [(c[0] - 0.1 * x1, c[0] + 0.1 * x1, c[1] - 0.1 * y1,
                                           c[1] + 0.1 * y1, c[2] - 0.1 * z1, c[2] + 0.1 * z1,
                                           500.0)])
        out = dict(prisms=prisms, pts=pts, flag=prob.flag, rho=rho, W=W)
    with _HOST_LOCK:
        return _HOST.setdefault(prob.name, out)


def request_stream(rate: float, duration: float, seed: int, pool, p_forward: float = 0.8):
    rng = np.random.default_rng(seed)
    out, t, rid = [], 0.0, 0
    probs = np.array([p.weight for p in pool])
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        if rng.random() < p_forward:
            cells = int(round(np.exp(rng.uniform(np.log(5e3), np.log(1.3e5)))))
            npts = int(round(np.exp(rng.uniform(np.log(1e3), np.log(2e4)))))
            ncomp = int(rng.integers(1, 10))
            bits = rng.choice(np.arange(1, 10), ncomp, replace=False)
            flag = int(sum(1 << int(b) for b in bits))
            out.append(Request(rid, t, "forward", dict(
                cells=cells, npts=npts, flag=flag, height=float(rng.uniform(1.0, 100.0)),
                seed=int(rng.integers(1 << 31)))))
        else:
            out.append(Request(rid, t, "cg", dict(
                problem=int(rng.choice(len(pool), p=probs)),
                lam=float(10.0 ** rng.uniform(-1.0, 2.0)), seed=int(rng.integers(1 << 31)))))
        rid += 1
    return out


def forward_payload(p: dict):
    """dict(xs, ys, zs, prisms, rho, pts, evals) of a forward request (evals:
    (node, point) pairs of its node-sum forward)."""
    cells, npts = p["cells"], p["npts"]
    nz = max(1, int(round((cells / 4.0) ** (1.0 / 3.0))))
    nx = max(1, int(round(np.sqrt(cells / nz))))
    prisms = M.regular_mesh((0.0, EXTENT[0]), nx, (0.0, EXTENT[1]), nx, (0.0, EXTENT[2]), nz)
    n = max(2, int(round(np.sqrt(npts))))
    pts = M.grid_points(n, 0.0, EXTENT[0], n, 0.0, EXTENT[1], -p["height"])
    rho = np.random.default_rng(p["seed"]).uniform(-500.0, 500.0, prisms.shape[0])
    return dict(xs=M.linspaced(nx + 1, 0.0, EXTENT[0]), ys=M.linspaced(nx + 1, 0.0, EXTENT[1]),
                zs=M.linspaced(nz + 1, 0.0, EXTENT[2]), prisms=prisms, rho=rho, pts=pts,
                evals=float((nx + 1) * (nx + 1) * (nz + 1) * pts.shape[0]))


class ResidentCache:
    """Byte-budgeted residency of contexts, each holding one problem's G.

    A request for a problem that is not resident admits it (assembles its G,
    counted in that request's latency), first evicting idle problems that
    are not pinned as the budget requires.  Victims, by `policy`:
    * "gdsf" (default; greedy-dual-size-frequency): the idle problem with
      the lowest H = L + hits * cost / bytes, where cost is the problem's
      measured admission time and L (the clock) rises to each victim's H, so
      that a large problem that is expensive to rebuild and asked for often
      stays, and problems not asked for lately age out;
    * "lru": the least recently used idle problem.

    Device memory is recycled, not returned: an admission that needs room
    assembles into the context of the first victim (in policy order) whose
    G buffer holds its G.  Only when no victim's buffer does is a new one
    allocated, after evicting victims and freeing their buffers until the
    budget holds it.  A
    hipMalloc / hipFree of a G of gigabytes stalls the kernels of every
    stream on the device for up to a few hundred milliseconds (measured on
    MI355X: a 12 GB allocate-free cycle held a concurrent 5 ms forward
    request for 200 ms).  The budget bounds the G memory held, resident and
    spare."""

    def __init__(self, pool, budget_bytes: float, device: int = 0, policy: str = "gdsf"):
        if policy not in ("gdsf", "lru"):
            raise ValueError(f"unknown eviction policy {policy!r}")
        self.pool, self.budget, self.device, self.policy = pool, budget_bytes, device, policy
        self.entries = collections.OrderedDict()
        self.spare = []                 # [capacity, ctx] of contexts holding no problem
        self.used = 0.0                 # G capacity held, resident and spare
        self.admissions = 0
        self.evictions = 0
        self.reuses = 0                 # admissions into a spare buffer
        self.frees = 0                  # spare buffers returned to the device
        self.admit_s = 0.0
        self.clock = 0.0
        self.cost = {}                  # pid -> last measured admission time (s)
        self.lock = threading.Lock()
        self.cv = threading.Condition(self.lock)

    @staticmethod
    def g_bytes(prob: CgProblem) -> float:
        """Bytes of the problem's G as the engine lays it out (rows of Nm
        entries padded to a multiple of 64)."""
        b = {"f64": 8.0, "f32": 4.0, "bf16": 2.0}[prob.store]
        if prob.instance >= 0:
            inst = _instances()[prob.instance]
            nm, nd = inst.prisms.shape[0], inst.points.shape[0] * bin(inst.flag).count("1")
        else:
            nx, ny, nz = prob.shape
            nm, nd = nx * ny * nz, prob.n_pts ** 2 * bin(prob.flag).count("1")
        return b * nd * ((nm + 63) // 64 * 64)

    def prepare(self):
        """Builds every problem's host-side inputs (before a replay)."""
        for p in self.pool:
            problem_inputs(p)

    def warm(self, stream_of=None):
        """A server's start-up: admits the pinned problems, then the others,
        most requested first, while they fit in the budget without evictions;
        the admission counters then restart from zero.  Returns the number
        admitted."""
        n = 0
        for pid in sorted(range(len(self.pool)),
                          key=lambda k: (not self.pool[k].pinned, -self.pool[k].weight)):
            if self.used + self.g_bytes(self.pool[pid]) <= self.budget:
                self.release(self.get(pid, stream_of=stream_of))
                n += 1
        self.admissions, self.evictions, self.reuses, self.frees = 0, 0, 0, 0
        self.admit_s = 0.0
        return n

    def _value(self, pid, e):
        cost = self.cost.get(pid, e["bytes"] * 1e-11)     # ~ assembly time before measured
        return self.clock + e["hits"] * cost / e["bytes"]

    def _victims(self):
        idle = [(pid, e) for pid, e in self.entries.items()
                if e.get("ready") and e["busy"] == 0 and not self.pool[pid].pinned]
        if self.policy == "gdsf":
            idle.sort(key=lambda t: t[1]["H"])
        return idle

    def _evict(self, pid):
        """Moves an idle resident problem's context to the spares (lock held)."""
        old = self.entries.pop(pid)
        self.spare.append([old["bytes"], old["ctx"]])
        self.evictions += 1
        if self.policy == "gdsf":
            self.clock = max(self.clock, old["H"])

    def _take_spare(self, need):
        """The smallest spare holding `need` bytes, removed from the list."""
        fit = [s for s in self.spare if s[0] >= need]
        if not fit:
            return None
        best = min(fit, key=lambda s: s[0])
        self.spare.remove(best)
        return best

    def _free_spare(self):
        cap, ctx = self.spare.pop(0)
        ctx.close()
        self.used -= cap
        self.frees += 1

    def _place(self, need):
        """[capacity, ctx or None] for an admission of `need` bytes (lock
        held): the smallest spare that holds it, else new memory if the
        budget has room, else the buffer of a victim that holds it, else new
        memory after evicting and freeing; then spares are freed and victims
        evicted until the budget holds what is kept (in-use problems may
        have pushed it over)."""
        got = self._take_spare(need)
        if got is None and self.used + need <= self.budget:
            got = [need, None]
        if got is None:
            for pid, e in self._victims():
                if e["bytes"] >= need:
                    self._evict(pid)
                    got = self._take_spare(need)
                    break
        if got is None:
            got = [need, None]
        extra = need if got[1] is None else 0.0
        while self.spare and self.used + extra > self.budget:
            self._free_spare()
        for pid, e in self._victims():
            if self.used + extra <= self.budget:
                break
            self._evict(pid)
            self._free_spare()
        return got

    def get(self, pid: int, info: dict | None = None, stream_of=None):
        """The resident entry of problem pid, admitted if needed; the caller
        holds it (it is not evicted) until release().  Safe to call from
        several threads: one admits, the others wait for it.  info["admitted"]
        tells whether this call admitted the problem.  An admission queues
        its GPU work on `stream_of`'s stream (a Context) if given."""
        if info is not None:
            info["admitted"] = False
        with self.lock:
            while True:
                e = self.entries.get(pid)
                if e is None:
                    break
                if e.get("ready"):
                    self.entries.move_to_end(pid)
                    e["busy"] += 1
                    e["hits"] += 1
                    e["H"] = self._value(pid, e)
                    return e
                self.cv.wait()
            prob = self.pool[pid]
            need = self.g_bytes(prob)
            pinned = sum(e["bytes"] for k, e in self.entries.items() if self.pool[k].pinned)
            if need + (0.0 if prob.pinned else pinned) > self.budget:
                raise MemoryError(f"problem {prob.name} ({need / 1e9:.1f} GB) exceeds the budget "
                                  f"beside the pinned problems ({pinned / 1e9:.1f} GB)")
            cap, ctx = self._place(need)
            if ctx is None:
                self.used += cap
            else:
                self.reuses += 1
            placeholder = dict(ready=False, busy=1, bytes=cap, hits=1, H=0.0,
                               lock=threading.Lock())
            if info is not None:
                info["admitted"] = True
            self.entries[pid] = placeholder
            self.admissions += 1
        t0 = time.perf_counter()
        try:
            e = self._admit(prob, need, stream_of, ctx)
        except BaseException:
            with self.lock:
                del self.entries[pid]
                if ctx is not None:
                    self.spare.append([cap, ctx])
                else:
                    self.used -= cap
                self.cv.notify_all()
            raise
        dt = time.perf_counter() - t0
        with self.lock:
            self.admit_s += dt
            self.cost[pid] = dt
            e["bytes"] = cap
            placeholder.update(e, ready=True)
            placeholder["H"] = self._value(pid, placeholder)
            self.cv.notify_all()
        return placeholder

    def release(self, e):
        with self.lock:
            e["busy"] -= 1

    def _admit(self, prob: CgProblem, need: float, stream_of=None, ctx=None):
        """Assembles the problem into `ctx` (a spare context whose G buffer
        holds it) or a new context; returns the entry's fields."""
        h = problem_inputs(prob)
        fresh = ctx is None
        ctx = Context(self.device) if fresh else ctx
        try:
            ctx.tune(fused_min_blocks=FUSED_MIN_BLOCKS)
            if stream_of is not None:
                ctx.use_stream_of(stream_of)
            inv = GaussNewtonInversion(h["prisms"], h["pts"], h["flag"], ctx=ctx)
            inv.compute_G(prob.store, prob.method)
            inv.set_regularization(*h["W"])
            nobs = h["pts"].shape[0]
            dobs, sigma = add_noise(inv.G_vec_mul(h["rho"]), nobs, bin(h["flag"]).count("1"), 7)
            if not np.all(sigma > 0):
                raise ValueError(f"problem {prob.name}: the synthetic anomaly holds no cell "
                                 "(zero data, no noise scale)")
            if stream_of is not None:
                ctx.use_stream_of(None)
        except BaseException:
            if fresh:
                ctx.close()
            raise
        Wd = np.repeat(1.0 / sigma, nobs)
        nm = h["prisms"].shape[0]
        P = np.full(nm, (1500.0 * 1500.0) / 3000.0)     # barrier P at m = 0, bounds +-1500
        return dict(ctx=ctx, inv=inv, bytes=need, dobs=dobs, Wd=Wd, P=P, nm=nm, fused=prob.fused)

    def close(self):
        for e in self.entries.values():
            if e.get("ready"):
                e["ctx"].close()
        for _, ctx in self.spare:
            ctx.close()
        self.entries.clear()
        self.spare = []
        self.used = 0.0


def store_error(prob: CgProblem, lams=(0.1, 1.0, 10.0, 100.0), device: int = 0):
    """The served solve's error from its G assembly, storage and product: dx
    of the 20-iteration solve on the problem as served (prob.method,
    prob.store, prob.fused) against the FP64 two-pass solve on the face/edge
    G from the same inputs (that G's data), max over `lams` of
    ||dx - dx_f64|| / ||dx_f64||, and of the relative difference of the
    returned CG error measure."""
    import dataclasses
    both = [prob, dataclasses.replace(prob, store="f64", fused=False, method="faces")]
    cache = ResidentCache(both, 1e15, device)
    try:
        a, b = cache.get(0), cache.get(1)
        zero = np.zeros(b["nm"])
        e_dx = e_err = 0.0
        for lam in lams:
            ref = b["inv"].solve_cg(0.0, 20, lam, zero, zero, b["Wd"], b["dobs"], b["P"])
            got = a["inv"].solve_cg(0.0, 20, lam, zero, zero, b["Wd"], b["dobs"], b["P"],
                                    fused=a["fused"])
            e_dx = max(e_dx, float(np.linalg.norm(got[0] - ref[0]) / np.linalg.norm(ref[0])))
            e_err = max(e_err, abs(got[1] - ref[1]) / abs(ref[1]))
        return dict(dx_rel_err=e_dx, cg_err_rel_diff=e_err, store=prob.store, fused=prob.fused,
                    method=prob.method, lambdas=list(lams))
    finally:
        cache.close()


def _percentiles(x):
    x = np.asarray(x, float)
    if x.size == 0:
        return None
    return {"p50": float(np.percentile(x, 50)), "p90": float(np.percentile(x, 90)),
            "p95": float(np.percentile(x, 95)), "p99": float(np.percentile(x, 99)),
            "max": float(x.max()), "mean": float(x.mean()), "n": int(x.size)}


def serve(requests, pool, budget_bytes: float = 100e9, device: int = 0, lookahead: int = 32,
          progress=None, forward_method: str = "nodes", records: bool = False,
          workers: int = 3, outputs: bool = False, max_batch: int = 8, policy: str = "gdsf",
          batch_evals: float = 2e8, cg_order: str = "sjf",
          cg_workers: int = 1, prewarm: bool = True):
    """Replays `requests` in arrival order; returns a summary dict (times in ms;
    with `records`, also the per-request arrival / start / finish times; with
    `outputs`, each request's result by request id: the FP32 forward array,
    or dict(dx, err, iterations) of the CG solve).

    `workers` server threads, each with its own engine context (stream),
    take the released requests in arrival order: while one request's kernel
    runs, the next one's inputs are uploaded, and a short request can run
    beside a long one instead of waiting for it to finish.  With
    `cg_workers` > 0, those threads take the forward requests only and
    `cg_workers` more take the CG requests from a queue of their own, so that
    a forward request never waits in the queue behind solves (the two kinds
    still share the device: the forward kernels are FP64-issue bound, the
    solves HBM bound); with 0, every thread takes both kinds.  The CG
    threads take the queued solve with the smallest G first (`cg_order`
    "sjf"; "fifo": arrival order), so that the many solves on small G do not
    wait behind one on a 47 GB G.  A worker that
    takes a small forward request also takes the small forward requests
    queued right behind it whose inputs are ready (up to `max_batch` in all,
    each at most a quarter and all together at most `batch_evals`
    (node, point) evaluations) and serves them in one launch
    (gra_forward_rectilinear_batch; the node-sum form only), each result
    bit-identical to its own launch.  Large requests go alone: a batch ends
    when its largest request does, so batching a large request with small
    ones would hold the small ones' results back.  With `prewarm`,
    the cache is filled before the first arrival (ResidentCache.warm), as a
    server loads its resident problems at start-up."""
    payloads, cv = {}, threading.Condition()
    stop = [False]
    prod_err = []

    def producer():
        try:
            for r in requests:
                if r.kind != "forward":
                    continue
                with cv:
                    while not stop[0] and len(payloads) >= lookahead:
                        cv.wait()
                    if stop[0]:
                        return
                item = forward_payload(r.params)
                with cv:
                    payloads[r.rid] = item
                    cv.notify_all()
        except BaseException as ex:                 # workers waiting on it must not hang
            with cv:
                prod_err.append(ex)
                stop[0] = True
                cv.notify_all()

    def take_payload(rid):
        with cv:
            while rid not in payloads:
                if prod_err:
                    raise RuntimeError("request producer failed") from prod_err[0]
                cv.wait()
            q = payloads.pop(rid)
            cv.notify_all()
            return q

    workers = max(1, int(workers))
    cg_workers = max(0, int(cg_workers))
    if cg_order not in ("sjf", "fifo"):
        raise ValueError(f"unknown CG order {cg_order!r}")
    g_size = [ResidentCache.g_bytes(p) for p in pool]
    batching = forward_method == "nodes" and max_batch > 1
    fctxs = [Context(device) for _ in range(workers + cg_workers)]
    cache = ResidentCache(pool, budget_bytes, device, policy)
    cache.prepare()
    th = threading.Thread(target=producer, daemon=True)
    th.start()
    rec, rec_lock = [], threading.Lock()
    results = {}
    # one queue for both kinds, or a forward queue and a CG queue
    split = cg_workers > 0
    queues = {"forward": collections.deque(), "cg": collections.deque()} if split else \
        {"all": collections.deque()}
    qcv = threading.Condition()
    closing = [False]
    errors = []
    t_start = [0.0]
    n_batched = [0, 0]        # launches with more than one request, requests in them

    def run_forward(fctx, q, flag):
        if forward_method == "nodes":
            return forward_rectilinear(q["xs"], q["ys"], q["zs"], q["rho"], q["pts"], flag,
                                       ctx=fctx)
        return forward(q["prisms"], q["rho"], q["pts"], flag, ctx=fctx)

    def next_work(pending):
        """The head request of `pending` and, for a forward one, the forward
        requests queued right behind it with ready inputs."""
        with qcv:
            while not pending and not closing[0]:
                qcv.wait()
            if not pending:
                return None
            if split and cg_order == "sjf" and pending is queues["cg"]:
                k = min(range(len(pending)),
                        key=lambda q: (g_size[pending[q].params["problem"]], q))
                r = pending[k]
                del pending[k]
                return [r]
            group = [pending.popleft()]
            if group[0].kind == "forward" and batching:
                small = 0.25 * batch_evals
                with cv:
                    head = payloads.get(group[0].rid)
                    total = head["evals"] if head is not None else batch_evals + 1.0
                    while (total <= small and pending and len(group) < max_batch
                           and pending[0].kind == "forward" and pending[0].rid in payloads
                           and payloads[pending[0].rid]["evals"] <= small
                           and total + payloads[pending[0].rid]["evals"] <= batch_evals):
                        total += payloads[pending[0].rid]["evals"]
                        group.append(pending.popleft())
            return group

    def record(r, t0, t1, cold, size, result):
        with rec_lock:
            rec.append(dict(rid=r.rid, kind=r.kind, arrival=r.t_arrival, start=t0, finish=t1,
                            cold=cold, **size))
            if outputs:
                results[r.rid] = result
            n = len(rec)
        if progress and n % progress == 0:
            print(f"[serve] {n}/{len(requests)} requests, t = {t1:.1f} s", flush=True)

    def worker(fctx, pending):
        while True:
            group = next_work(pending)
            if group is None:
                return
            try:
                r = group[0]
                if r.kind == "forward":
                    qs = [take_payload(g.rid) for g in group]
                    t0 = time.perf_counter() - t_start[0]
                    if len(group) == 1:
                        outs = [run_forward(fctx, qs[0], r.params["flag"])]
                    else:
                        outs = forward_rectilinear_batch(
                            [(q["xs"], q["ys"], q["zs"], q["rho"], q["pts"], g.params["flag"])
                             for q, g in zip(qs, group)], ctx=fctx)
                        with rec_lock:
                            n_batched[0] += 1
                            n_batched[1] += len(group)
                    t1 = time.perf_counter() - t_start[0]
                    for g, q, out in zip(group, qs, outs):
                        out = out.astype(np.float32)
                        record(g, t0, t1, False, dict(cells=q["prisms"].shape[0],
                                                      points=q["pts"].shape[0],
                                                      comps=out.shape[1], batch=len(group)), out)
                else:
                    t0 = time.perf_counter() - t_start[0]
                    info = {}
                    e = cache.get(r.params["problem"], info, stream_of=fctx)
                    try:
                        cold = info["admitted"]
                        zero = np.zeros(e["nm"])
                        # the solve runs on this worker's stream (one stream per
                        # worker, whatever the number of resident problems)
                        with e["lock"]:
                            e["ctx"].use_stream_of(fctx)
                            try:
                                dx, err, its = e["inv"].solve_cg(
                                    0.0, 20, r.params["lam"], zero, zero, e["Wd"], e["dobs"],
                                    e["P"], fused=e["fused"])
                            finally:
                                e["ctx"].use_stream_of(None)
                    finally:
                        cache.release(e)
                    t1 = time.perf_counter() - t_start[0]
                    record(r, t0, t1, cold, dict(problem=pool[r.params["problem"]].name),
                           dict(dx=dx, err=err, iterations=its))
            except BaseException as ex:          # reported after the replay
                errors.append(ex)
                with qcv:                        # stop the replay: drain and close
                    closing[0] = True
                    for q in queues.values():
                        q.clear()
                    qcv.notify_all()

    threads = []
    warmed, t_origin_ns = 0, 0
    try:
        # warm each forward context once (first-call module loading is not a request)
        warm = forward_payload(dict(cells=64, npts=16, flag=0x3FE, height=10.0, seed=1))
        for fctx in fctxs:
            run_forward(fctx, warm, 0x3FE)
            if batching:
                forward_rectilinear_batch([(warm["xs"], warm["ys"], warm["zs"], warm["rho"],
                                            warm["pts"], 0x3FE)] * 2, ctx=fctx)
        warmed = cache.warm(fctxs[-1]) if prewarm else 0
        t_start[0] = time.perf_counter()
        t_origin_ns = time.monotonic_ns()
        owner = [queues["forward"]] * workers + [queues["cg"]] * cg_workers if split else \
            [queues["all"]] * workers
        threads = [threading.Thread(target=worker, args=(f, q), daemon=True)
                   for f, q in zip(fctxs, owner)]
        for t in threads:
            t.start()
        for r in requests:
            if errors:
                break
            now = time.perf_counter() - t_start[0]
            if now < r.t_arrival:
                time.sleep(r.t_arrival - now)
            with qcv:
                queues[r.kind if split else "all"].append(r)
                qcv.notify_all()
    finally:
        with qcv:
            closing[0] = True
            qcv.notify_all()
        for t in threads:
            t.join()
        with cv:
            stop[0] = True
            cv.notify_all()
        th.join(timeout=10)
        cache_stats = dict(prewarmed=warmed, admissions=cache.admissions,
                           evictions=cache.evictions, reuses=cache.reuses, frees=cache.frees,
                           admission_s=cache.admit_s, resident_gb=cache.used / 1e9,
                           budget_gb=cache.budget / 1e9, policy=cache.policy,
                           pool_gb=sum(cache.g_bytes(p) for p in pool) / 1e9)
        cache.close()
        for fctx in fctxs:
            fctx.close()
    if errors:
        raise errors[0]
    if prod_err:
        raise RuntimeError("request producer failed") from prod_err[0]
    rec.sort(key=lambda x: x["arrival"])
    lat = [1e3 * (x["finish"] - x["arrival"]) for x in rec]
    svc = [1e3 * (x["finish"] - x["start"]) for x in rec]
    makespan = max(x["finish"] for x in rec) if rec else 0.0
    by = {}
    for kind in ("forward", "cg"):
        sel = [x for x in rec if x["kind"] == kind]
        by[kind] = dict(
            latency_ms=_percentiles([1e3 * (x["finish"] - x["arrival"]) for x in sel]),
            service_ms=_percentiles([1e3 * (x["finish"] - x["start"]) for x in sel]))
    by_problem = {}
    for p in pool:
        sel = [x for x in rec if x["kind"] == "cg" and x.get("problem") == p.name]
        by_problem[p.name] = dict(
            requests=len(sel), cold=sum(x["cold"] for x in sel),
            warm_service_ms=_percentiles([1e3 * (x["finish"] - x["start"]) for x in sel
                                          if not x["cold"]]),
            latency_ms=_percentiles([1e3 * (x["finish"] - x["arrival"]) for x in sel]))
    warm_cg = [1e3 * (x["finish"] - x["start"]) for x in rec if x["kind"] == "cg" and not x["cold"]]
    return dict(
        requests=len(rec), makespan_s=makespan,
        offered_rate=len(rec) / requests[-1].t_arrival if rec else 0.0,
        achieved_rate=len(rec) / makespan if makespan else 0.0,
        workers=workers, cg_workers=cg_workers,
        mean_in_service=sum(svc) / 1e3 / makespan if makespan else 0.0,
        utilization=sum(svc) / 1e3 / makespan / (workers + cg_workers) if makespan else 0.0,
        sustainable_rate=((1e3 / np.mean(svc) if svc else 0.0)
                          if workers + cg_workers == 1 else None),
        latency_ms=_percentiles(lat), service_ms=_percentiles(svc), by_kind=by,
        cg_by_problem=by_problem, cg_warm_service_ms=_percentiles(warm_cg), cache=cache_stats,
        max_in_service=_max_overlap(rec),
        forward_batches=dict(launches=n_batched[0], requests=n_batched[1]),
        **(dict(records=rec, t_origin_ns=t_origin_ns) if records else {}),
        **(dict(outputs=results) if outputs else {}))


def _max_overlap(rec):
    """Largest number of requests in service (start..finish) at one time."""
    ev = sorted([(x["start"], 1) for x in rec] + [(x["finish"], -1) for x in rec])
    n = best = 0
    for _, d in ev:
        n += d
        best = max(best, n)
    return best
