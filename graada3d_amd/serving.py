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
  [0.1, 100]) on one of a pool of problems whose dense G stays resident in
  HBM between requests.  Residency is a byte-budgeted LRU of engine
  contexts: a request for a problem that is not resident assembles its G
  (admission, counted in that request's latency), evicting least recently
  used problems until the budget holds.

Request payloads are built ahead by a producer thread (the client's work);
a dispatcher releases each request at its arrival time and `workers`
server threads (default 3, each with its own context and stream) take them
in that order, so up to `workers` requests are on the GPU at once.
"""
from __future__ import annotations

import collections
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import mesh as M
from .engine import Context, GaussNewtonInversion, forward, forward_rectilinear
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
    shape: tuple                   # (nx, ny, nz)
    extent: tuple                  # (x1, y1, z1), origin 0
    n_pts: int                     # points per side
    flag: int
    weight: float


def default_cg_pool(include_large: bool = False):
    pool = [
        CgProblem("inversion_small", (40, 40, 20), (2000.0, 2000.0, 1000.0), 50, 0x12, 0.5),
        CgProblem("medium_5comp", (30, 30, 15), EXTENT, 40, 0x2 | 0x10 | 0x20 | 0x40 | 0x80, 0.3),
        CgProblem("small_9comp", (24, 24, 12), EXTENT, 30, 0x3FE, 0.2),
    ]
    if include_large:
        pool.append(CgProblem("tensor_large", (64, 64, 32), EXTENT, 100, 0x3FE, 0.05))
    w = np.array([p.weight for p in pool])
    for p, q in zip(pool, w / w.sum()):
        p.weight = float(q)
    return pool


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
    """dict(xs, ys, zs, prisms, rho, pts) of a forward request."""
    cells, npts = p["cells"], p["npts"]
    nz = max(1, int(round((cells / 4.0) ** (1.0 / 3.0))))
    nx = max(1, int(round(np.sqrt(cells / nz))))
    prisms = M.regular_mesh((0.0, EXTENT[0]), nx, (0.0, EXTENT[1]), nx, (0.0, EXTENT[2]), nz)
    n = max(2, int(round(np.sqrt(npts))))
    pts = M.grid_points(n, 0.0, EXTENT[0], n, 0.0, EXTENT[1], -p["height"])
    rho = np.random.default_rng(p["seed"]).uniform(-500.0, 500.0, prisms.shape[0])
    return dict(xs=M.linspaced(nx + 1, 0.0, EXTENT[0]), ys=M.linspaced(nx + 1, 0.0, EXTENT[1]),
                zs=M.linspaced(nz + 1, 0.0, EXTENT[2]), prisms=prisms, rho=rho, pts=pts)


class ResidentCache:
    """Byte-budgeted LRU of contexts, each holding one problem's G."""

    def __init__(self, pool, budget_bytes: float, device: int = 0):
        self.pool, self.budget, self.device = pool, budget_bytes, device
        self.entries = collections.OrderedDict()
        self.used = 0.0
        self.admissions = 0
        self.evictions = 0
        self.lock = threading.Lock()
        self.cv = threading.Condition(self.lock)

    @staticmethod
    def g_bytes(prob: CgProblem) -> float:
        nx, ny, nz = prob.shape
        return 8.0 * nx * ny * nz * prob.n_pts ** 2 * bin(prob.flag).count("1")

    def get(self, pid: int, info: dict | None = None):
        """The resident entry of problem pid, admitted if needed; the caller
        holds it (it is not evicted) until release().  Safe to call from
        several threads: one admits, the others wait for it.  info["admitted"]
        tells whether this call admitted the problem."""
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
                    return e
                self.cv.wait()
            prob = self.pool[pid]
            need = self.g_bytes(prob)
            if need > self.budget:
                raise MemoryError(f"problem {prob.name} ({need / 1e9:.1f} GB) exceeds the budget")
            # evict idle least-recently-used problems; in-use ones stay (the
            # budget may be exceeded until they are released)
            for old_pid in list(self.entries):
                if self.used + need <= self.budget:
                    break
                old = self.entries[old_pid]
                if old.get("ready") and old["busy"] == 0:
                    del self.entries[old_pid]
                    old["ctx"].close()
                    self.used -= old["bytes"]
                    self.evictions += 1
            placeholder = dict(ready=False, busy=1, bytes=need)
            if info is not None:
                info["admitted"] = True
            self.entries[pid] = placeholder
            self.used += need
            self.admissions += 1
        try:
            e = self._admit(prob, need)
        except BaseException:
            with self.lock:
                del self.entries[pid]
                self.used -= need
                self.cv.notify_all()
            raise
        with self.lock:
            placeholder.update(e, ready=True)
            self.cv.notify_all()
        return placeholder

    def release(self, e):
        with self.lock:
            e["busy"] -= 1

    def _admit(self, prob: CgProblem, need: float):
        nx, ny, nz = prob.shape
        x1, y1, z1 = prob.extent
        prisms = M.regular_mesh((0.0, x1), nx, (0.0, y1), ny, (0.0, z1), nz)
        pts = M.grid_points(prob.n_pts, 0.0, x1, prob.n_pts, 0.0, y1, -1.0)
        ctx = Context(self.device)
        inv = GaussNewtonInversion(prisms, pts, prob.flag, ctx=ctx)
        inv.compute_G("f64")
        Z = depth_weights_offset(prisms, 50.0, 3.0)
        inv.set_regularization(*weighting_operators(prisms, nx, ny, nz, Z, a_s=1e-3))
        c = (0.5 * x1, 0.5 * y1, 0.3 * z1)
        rho = M.model_from_boxes(prisms, [(c[0] - 0.1 * x1, c[0] + 0.1 * x1, c[1] - 0.1 * y1,
                                            c[1] + 0.1 * y1, c[2] - 0.1 * z1, c[2] + 0.1 * z1,
                                            500.0)])
        ncomp = bin(prob.flag).count("1")
        dobs, sigma = add_noise(inv.G_vec_mul(rho), pts.shape[0], ncomp, 7)
        if not np.all(sigma > 0):
            ctx.close()
            raise ValueError(f"problem {prob.name}: the synthetic anomaly holds no cell "
                             "(zero data, no noise scale)")
        Wd = np.repeat(1.0 / sigma, pts.shape[0])
        nm = prisms.shape[0]
        P = np.full(nm, (1500.0 * 1500.0) / 3000.0)     # barrier P at m = 0, bounds +-1500
        return dict(ctx=ctx, inv=inv, bytes=need, dobs=dobs, Wd=Wd, P=P, nm=nm)

    def close(self):
        for e in self.entries.values():
            if e.get("ready"):
                e["ctx"].close()
        self.entries.clear()
        self.used = 0.0


def _percentiles(x):
    x = np.asarray(x, float)
    if x.size == 0:
        return None
    return {"p50": float(np.percentile(x, 50)), "p90": float(np.percentile(x, 90)),
            "p99": float(np.percentile(x, 99)), "max": float(x.max()), "mean": float(x.mean()),
            "n": int(x.size)}


def serve(requests, pool, budget_bytes: float = 100e9, device: int = 0, lookahead: int = 32,
          progress=None, forward_method: str = "nodes", records: bool = False,
          workers: int = 3, outputs: bool = False):
    """Replays `requests` in arrival order; returns a summary dict (times in ms;
    with `records`, also the per-request arrival / start / finish times; with
    `outputs`, each request's result by request id: the FP32 forward array,
    or dict(dx, err, iterations) of the CG solve).

    `workers` server threads, each with its own engine context (stream),
    take the released requests in arrival order: while one request's kernel
    runs, the next one's inputs are uploaded, and a short request can run
    beside a long one instead of waiting for it to finish."""
    import queue as _queue
    payloads, cv = {}, threading.Condition()
    stop = [False]

    def producer():
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

    workers = max(1, int(workers))
    fctxs = [Context(device) for _ in range(workers)]
    cache = ResidentCache(pool, budget_bytes, device)
    th = threading.Thread(target=producer, daemon=True)
    th.start()
    rec, rec_lock = [], threading.Lock()
    results = {}
    work = _queue.Queue()
    errors = []
    t_start = [0.0]

    def run_forward(fctx, q, flag):
        if forward_method == "nodes":
            return forward_rectilinear(q["xs"], q["ys"], q["zs"], q["rho"], q["pts"], flag,
                                       ctx=fctx)
        return forward(q["prisms"], q["rho"], q["pts"], flag, ctx=fctx)

    def worker(fctx):
        while True:
            r = work.get()
            if r is None:
                return
            try:
                if r.kind == "forward":
                    with cv:
                        while r.rid not in payloads:
                            cv.wait()
                        q = payloads.pop(r.rid)
                        cv.notify_all()
                    t0 = time.perf_counter() - t_start[0]
                    out = run_forward(fctx, q, r.params["flag"]).astype(np.float32)
                    size = dict(cells=q["prisms"].shape[0], points=q["pts"].shape[0],
                                comps=out.shape[1])
                    cold = False
                    result = out
                else:
                    t0 = time.perf_counter() - t_start[0]
                    info = {}
                    e = cache.get(r.params["problem"], info)
                    try:
                        cold = info["admitted"]
                        zero = np.zeros(e["nm"])
                        dx, err, its = e["inv"].solve_cg(0.0, 20, r.params["lam"], zero, zero,
                                                         e["Wd"], e["dobs"], e["P"])
                    finally:
                        cache.release(e)
                    result = dict(dx=dx, err=err, iterations=its)
                    size = dict(problem=pool[r.params["problem"]].name)
                t1 = time.perf_counter() - t_start[0]
                with rec_lock:
                    rec.append(dict(rid=r.rid, kind=r.kind, arrival=r.t_arrival, start=t0,
                                    finish=t1, cold=cold, **size))
                    if outputs:
                        results[r.rid] = result
                    n = len(rec)
                if progress and n % progress == 0:
                    print(f"[serve] {n}/{len(requests)} requests, t = {t1:.1f} s", flush=True)
            except BaseException as ex:          # reported after the replay
                errors.append(ex)

    threads = []
    try:
        # warm each forward context once (first-call module loading is not a request)
        warm = forward_payload(dict(cells=64, npts=16, flag=0x3FE, height=10.0, seed=1))
        for fctx in fctxs:
            run_forward(fctx, warm, 0x3FE)
        t_start[0] = time.perf_counter()
        threads = [threading.Thread(target=worker, args=(f,), daemon=True) for f in fctxs]
        for t in threads:
            t.start()
        for r in requests:
            now = time.perf_counter() - t_start[0]
            if now < r.t_arrival:
                time.sleep(r.t_arrival - now)
            work.put(r)
    finally:
        for _ in threads:
            work.put(None)
        for t in threads:
            t.join()
        with cv:
            stop[0] = True
            cv.notify_all()
        th.join(timeout=10)
        cache_stats = dict(admissions=cache.admissions, evictions=cache.evictions,
                           resident_gb=cache.used / 1e9)
        cache.close()
        for fctx in fctxs:
            fctx.close()
    if errors:
        raise errors[0]
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
    warm_cg = [1e3 * (x["finish"] - x["start"]) for x in rec if x["kind"] == "cg" and not x["cold"]]
    return dict(
        requests=len(rec), makespan_s=makespan,
        offered_rate=len(rec) / requests[-1].t_arrival if rec else 0.0,
        achieved_rate=len(rec) / makespan if makespan else 0.0,
        workers=workers,
        mean_in_service=sum(svc) / 1e3 / makespan if makespan else 0.0,
        utilization=sum(svc) / 1e3 / makespan / workers if makespan else 0.0,
        sustainable_rate=(1e3 / np.mean(svc) if svc else 0.0) if workers == 1 else None,
        latency_ms=_percentiles(lat), service_ms=_percentiles(svc), by_kind=by,
        cg_warm_service_ms=_percentiles(warm_cg), cache=cache_stats,
        max_in_service=_max_overlap(rec),
        **(dict(records=rec) if records else {}),
        **(dict(outputs=results) if outputs else {}))


def _max_overlap(rec):
    """Largest number of requests in service (start..finish) at one time."""
    ev = sorted([(x["start"], 1) for x in rec] + [(x["finish"], -1) for x in rec])
    n = best = 0
    for _, d in ev:
        n += d
        best = max(best, n)
    return best
