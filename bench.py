#!/usr/bin/env python3
"""Benchmark of the engine's hot path on one GPU.

Workloads
* tensor_large (default): all nine components of the 64x64x32-cell model
  (62.5 m cells) on a 100x100 grid at z = -50 m -- G of 90,000 x 131,072
  (94.4 GB in FP64) -- L1 (IRLS, eps^2 = 1e-3) Gauss-Newton, lambda_k =
  100*0.5^k, 20 CG iterations per outer iteration; default 10 steps.
* inversion_small: the L2 Gauss-Newton inversion of the 40x40x20-cell,
  50x50-point gz+Tzz survey -- dense G of 5,000 x 32,000 resident in HBM,
  Ws/Wx/Wy/Wz regularisation with (z+50)^-1.5 depth weighting, lambda_k =
  100*0.5^k, 50 CG iterations per outer iteration; default 10 steps = the
  whole inversion.

One step = one Gauss-Newton outer iteration on the GPU (barrier P, solve_cg =
CGLS with two passes over G per iteration, barrier model update, misfit, IRLS
weights).  Before the timed steps of each storage type the G-product launch
shapes are chosen by gra_autotune (timed back-to-back product pairs on the
resident G; its time is reported, not counted).  `value` is CG iterations
per second of the plain FP64 path.  The
same steps are then run with G stored in FP32 and BF16 (FP64 accumulation),
with the single-read iteration (one fused pass over G per CG iteration) for
every storage type, and with the single-read iteration over an FP32 / BF16
copy of the FP64 G plus FP64 residual refreshes ("refined"), and reported
as speed-up and error against the plain FP64 run (final model, misfit, and
predicted data per component).  G assembly throughput (entries/s) is
measured for every storage type.  Kernel rooflines use HIP events on the
engine's stream around the G products of every TIMING_STRIDE-th CG
iteration of the timed steps (an event is a queue packet of its own and
slows the launches next to it: timing every iteration cost 3-5 % of the
iteration rate); the CPU baseline is the single-threaded C oracle on a
bounded sample.

* batch: 64 uneven instances (2:1 octree meshes of 2k-60k cells refined at
  seeded box faces, 400-10,000 points, random component subsets; 1.4e10 G
  entries, 114 GB in FP64).  A step assembles every instance's G: once in a
  single batched launch (gra_assemble_batch, `value` = G entries/s) and once
  as 64 separate launches, reported beside it.
* serving: an open-loop replay of a seeded Poisson request stream (default
  200 requests/s for 2 s): 80 % matrix-free forward requests (5k-130k cells,
  1k-20k points, 1-9 components, FP32 out), 20 % 20-iteration CGLS solves on
  resident dense G (LRU admission / eviction), one request at a time on the
  GPU (graada3d_amd/serving.py).  `value` is the p99 request latency.

* karst: the reference's own cave-inversion example (Examples/Karstic_cave,
  interpolated data: 11,374 points, octree refined to 10,490 cells) run end
  to end by the inversion driver; `value` is the wall time of one whole
  inversion, beside the run time the reference logged for it.

Counter probes (--probe, fixed short launch sequences for rocprofv3 passes;
no JSON line):
* gemv: a plain 1 GiB device-to-device copy (the calibration dispatch for
  FETCH_SIZE / WRITE_SIZE), then the 5,000 x 32,000 G of inversion_small in
  FP64 (and FP32, BF16 with --all-dtypes) with 3 launches each of G x and
  G^T w.
* gemv-large: the same copy, then the 90,000 x 131,072 FP64 G of
  tensor_large (assembled by corner sums) with 3 launches each of G x and
  G^T w.
* assemble: G of that survey (gz+Tzz), all nine components on a 20x20x10 mesh
  at 40x40 points, the same two G by corner sums, one node-sum forward
  (FP64 instruction counters).
* octree: the 40x40x20 mesh refined twice at a seeded eighth of its cells
  (158,427 cells): G by face/edge integrals and by per-prism corner sums,
  gz+Tzz at 50x50 points and all nine at 30x30, three times each.
* mat-mul: the matrix-core products with 16 right-hand sides on the
  5,000 x 32,000 G in FP64, FP32 and BF16 (MFMA counters).
* mat-mul-large: the same on the 90,000 x 131,072 G of tensor_large
  (assembled by corner sums).

    python bench.py [--workload tensor_large|inversion_small|batch|serving|karst] [--steps K]
                    [--warmup W]
                    [--stores f64,f32,bf16] [--no-fused] [--no-cpu] [--gpus 1]
                    [--rate R] [--duration S] [--budget-gb B] [--policy gdsf|lru]
                    [--max-batch K] [--workers W]
    python bench.py --probe gemv|gemv-large|assemble|octree|mat-mul|mat-mul-large [--all-dtypes]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "G entries/s, CG iter/s (mixed vs FP64 plain) at stated error; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# FP64 flop per G entry of the assembly kernels (gz+Tzz / all nine components),
# from the SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 counters
FLOP_PER_ENTRY = {True: 1165.0, False: 740.0}
BYTES = {"f64": 8, "f32": 4, "bf16": 2}
TIMING_STRIDE = 10      # CG iterations 0, 10, 20, ... of each solve carry events
REFRESH = 10            # refined variants: FP64 residual refresh every 10 CG iterations
BATCH_MIN_BLOCKS = 24   # batch: single-read CG where each row group gets >= 24 row blocks
# field bits (gs.h:46-57) and component names, in G's row-block order
COMPONENTS = [(1, "gz"), (2, "gx"), (3, "gy"), (4, "Tzz"), (5, "Txz"), (6, "Tyz"), (7, "Txx"),
              (8, "Txy"), (9, "Tyy")]


def build_problem(name):
    from graada3d_amd import mesh as M
    from graada3d_amd.operators import depth_weights_offset, weighting_operators
    from graada3d_amd.workloads import forward_problem_small, tensor_problem_large
    if name == "inversion_small":
        fp = forward_problem_small()
        shape, cg_maxit, lp, eps2, z0 = (40, 40, 20), 50, 2, 0.0, 50.0
        lam0 = 100.0
    else:
        fp = tensor_problem_large()
        shape, cg_maxit, lp, eps2, z0 = (64, 64, 32), 20, 1, 1e-3, 50.0
        lam0 = 100.0
    Z = depth_weights_offset(fp.prisms, z0, 3.0)
    W = weighting_operators(fp.prisms, *shape, Z, a_s=1e-3)
    nm = fp.prisms.shape[0]
    return dict(fp=fp, W=W, cg_maxit=cg_maxit, lp=lp, eps2=eps2,
                lambdas=[lam0 * 0.5 ** k for k in range(64)],
                m_min=np.full(nm, -1500.0), m_max=np.full(nm, 1500.0), mref=np.zeros(nm),
                m_ini=np.zeros(nm))


def run(args):
    from graada3d_amd.engine import (TIME_ASSEMBLE, TIME_FUSED, TIME_GEMV_N, TIME_GEMV_T,
                                     TIME_REFRESH, Context, GaussNewtonInversion)
    from graada3d_amd.workloads import add_noise

    pb = build_problem(args.workload)
    fp = pb["fp"]
    ctx = Context(0)
    inv = GaussNewtonInversion(fp.prisms, fp.points, fp.flag, ctx=ctx)
    nd, nm = inv.Nd, inv.Nm
    ncomp = nd // fp.points.shape[0]
    print(f"[bench] {args.workload}: G {nd} x {nm}", file=sys.stderr, flush=True)

    # synthetic data from the FP64 G
    inv.compute_G("f64")
    d_clean = inv.G_vec_mul(fp.rho)
    dobs, sigma = add_noise(d_clean, fp.points.shape[0], ncomp, fp.seed)
    Wd = np.repeat(1.0 / sigma, fp.points.shape[0])

    # the same G by corner sums over the tensor-product mesh (shared corner
    # values): its rate, and its products against the face/edge G's
    xr = np.random.default_rng(11).standard_normal(nm)
    y_face = inv.G_vec_mul(xr)
    inv.compute_G("f64", method="nodes")          # untimed first call
    ctx.timing(True)
    ctx.timing_reset()
    for _ in range(2):
        inv.compute_G("f64", method="nodes")
    node_ms, _ = ctx.timing_get(TIME_ASSEMBLE)
    ctx.timing(False)
    y_node = inv.G_vec_mul(xr)
    nobs = fp.points.shape[0]
    node_dev = [float(np.linalg.norm(y_node[k * nobs:(k + 1) * nobs] - y_face[k * nobs:(k + 1) * nobs])
                      / np.linalg.norm(y_face[k * nobs:(k + 1) * nobs])) for k in range(ncomp)]
    nodes_info = {"entries_per_s": nd * nm / (node_ms * 1e-3), "avg_ms": node_ms,
                  "kernel": "corner sums over the tensor-product mesh, corner values shared "
                            "through LDS (gra_assemble_rectilinear), FP64",
                  "Gx_rel_dev_vs_face_edge_per_component": node_dev}
    print(f"[bench] node-sum assembly: {nodes_info['entries_per_s']:.3e} entries/s "
          f"(face/edge next), G x deviation {max(node_dev):.1e}", file=sys.stderr, flush=True)

    results = {}
    # (storage, mode): "" two passes over G per CG iteration, "fused" one
    # single-read pass, "refined" the single-read pass over a reduced-precision
    # copy of the FP64 G with FP64 residual refreshes every REFRESH iterations
    variants = [(s, "") for s in args.stores]
    if args.fused:
        variants += [(s, "fused") for s in args.stores]
        variants += [(s, "refined") for s in args.stores if s != "f64"]
    asm = {}
    skipped = []
    for store, mode in variants:
        fused = mode != ""
        key = store + ("_" + mode if mode else "")
        # the context holds one G: assemble it in this variant's storage type
        # (timed the first time, twice on the small workload); the refined
        # variant keeps the FP64 G and makes its reduced-precision copy
        ctx.timing(True)
        ctx.timing_reset()
        reps = 2 if args.workload == "inversion_small" and store not in asm else 1
        for _ in range(reps):
            inv.compute_G("f64" if mode == "refined" else store)
        if store not in asm:
            asm[store], _ = ctx.timing_get(TIME_ASSEMBLE)
        ctx.timing(False)
        asm_ms = asm[store]
        if mode == "refined":
            inv.mixed_precision(store, REFRESH)
        inv.set_regularization(*pb["W"])
        ctx.tune(n_rows=0, n_unroll=0, n_splits=0, t_chunks=0, t_unroll=0, nontemporal=-1, n_block=-1)
        tune_s = 0.0
        if args.autotune and not fused:
            t0 = time.perf_counter()
            ctx.autotune()
            tune_s = time.perf_counter() - t0

        def setup():
            inv.gn_setup(dobs, Wd, pb["m_min"], pb["m_max"], pb["mref"], pb["m_ini"], pb["lp"],
                         pb["eps2"], fused=fused)

        setup()
        for k in range(args.warmup):
            inv.gn_step(pb["lambdas"][k], 0.0, pb["cg_maxit"])
        setup()
        ctx.timing(TIMING_STRIDE)
        ctx.timing_reset()
        ctx.sync()
        t0 = time.perf_counter()
        its, mis = 0, []
        for k in range(args.steps):
            m_, e_, i_ = inv.gn_step(pb["lambdas"][k], 0.0, pb["cg_maxit"])
            its += i_
            mis.append(m_)
        ctx.sync()
        el = time.perf_counter() - t0
        gn_ms, gn_c = ctx.timing_get(TIME_GEMV_N)
        gt_ms, gt_c = ctx.timing_get(TIME_GEMV_T)
        fu_ms, fu_c = ctx.timing_get(TIME_FUSED)
        rf_ms, rf_c = ctx.timing_get(TIME_REFRESH)
        ctx.timing(False)
        if mode == "refined":
            inv.mixed_precision("f64")
        gbytes = float(nd) * nm * BYTES[store]

        def kern(ms, n):
            return dict(avg_ms=ms, launches=n, GBps=gbytes / (ms * 1e-3) / 1e9 if n else None)

        results[key] = dict(
            store=store, fused=fused, mode=mode or "two-pass",
            refresh={"every": REFRESH, "fp64_product_avg_ms": rf_ms, "fp64_products": rf_c}
            if mode == "refined" else None,
            iter_s=its / el, ms_per_step=el / args.steps * 1e3, cg_iters=its,
            misfit=mis[-1], model=inv.gn_model(),
            asm_ms=asm_ms, entries_per_s=nd * nm / (asm_ms * 1e-3),
            gemv_n=kern(gn_ms, gn_c), gemv_t=kern(gt_ms, gt_c), fused_product=kern(fu_ms, fu_c),
            tuning=ctx.tuning(), autotune_s=tune_s)
        if fused and fu_c == 0:
            # the fused grid cannot be co-resident for this G: the solver ran
            # the two-pass iteration; report that instead of a fused figure
            print(f"[bench] {key}: fused grid not co-resident for this G; two-pass ran",
                  file=sys.stderr, flush=True)
            del results[key]
            skipped.append(key)
            continue
        print(f"[bench] {key}: {results[key]['iter_s']:.1f} CG iter/s, assembly "
              f"{results[key]['entries_per_s']:.3e} entries/s", file=sys.stderr, flush=True)
        if not fused:
            results[key]["mat_mul"] = mat_mul_rates(ctx, inv, store, gbytes, gn_ms, gt_ms)

    # predicted data of every final model with the FP64 G, per component
    inv.compute_G("f64")
    comps = [n for b, n in COMPONENTS if fp.flag & (1 << b)]
    y_ref = inv.G_vec_mul(results["f64"]["model"])
    nobs = fp.points.shape[0]
    for key, r in results.items():
        y = inv.G_vec_mul(r["model"])
        r["data_rel_err"] = {c: float(np.linalg.norm(y[k * nobs:(k + 1) * nobs]
                                                     - y_ref[k * nobs:(k + 1) * nobs])
                                      / np.linalg.norm(y_ref[k * nobs:(k + 1) * nobs]))
                             for k, c in enumerate(comps)}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, inv, pb, dobs, Wd)
    ctx.close()
    if skipped:
        results["_skipped"] = skipped
    results["_nodes"] = nodes_info
    return pb, nd, nm, results, cpu


# TFLOP/s of the MFMA flavour used (dense peaks); BF16 G runs on the BF16
# matrix cores with each vector split into three BF16 parts, so it issues
# three MFMA products per useful one
MFMA_PEAK = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0}
MFMA_NAME = {"f64": "f64 16x16x4", "f32": "f32 16x16x4",
             "bf16": "bf16 16x16x32 (G X) / 16x16x16 (G^T R), vectors as 3 BF16 parts"}
MFMA_ISSUE = {"f64": 1.0, "f32": 1.0, "bf16": 3.0}
MM_RHS = 16


def mat_mul_rates(ctx, inv, store, gbytes, gn_ms, gt_ms):
    """The matrix-core products with MM_RHS right-hand sides on the resident G
    (HIP events around each launch, 3 launches each after one untimed):
    G-byte rate, MFMA throughput against the flavour's peak, and the time per
    vector against the single-vector products of the timed CG steps."""
    from graada3d_amd.engine import TIME_MAT_MUL
    rng = np.random.default_rng(17)
    X = rng.standard_normal((inv.Nm, MM_RHS))
    R = rng.standard_normal((inv.Nd, MM_RHS))
    out = {"rhs": MM_RHS, "mfma": MFMA_NAME[store], "peak_TFLOPs": MFMA_PEAK[store]}
    for name, fn, arg, vec_ms in (("G_X", inv.G_mat_mul, X, gn_ms), ("GT_R", inv.GT_mat_mul, R, gt_ms)):
        fn(arg)
        ctx.timing(True)
        ctx.timing_reset()
        for _ in range(3):
            fn(arg)
        ms, _ = ctx.timing_get(TIME_MAT_MUL)
        ctx.timing(False)
        tf = 2.0 * inv.Nd * inv.Nm * MM_RHS / (ms * 1e-3) / 1e12
        out[name] = {"avg_ms": ms, "G_GBps": gbytes / (ms * 1e-3) / 1e9, "TFLOPs": tf,
                     "mfma_frac": tf * MFMA_ISSUE[store] / MFMA_PEAK[store],
                     "hbm_frac": gbytes / (ms * 1e-3) / 8e12,
                     "per_vector_speedup_vs_gemv": vec_ms * MM_RHS / ms if vec_ms else None}
    return out


def all_cores(fn, unit_count, unit):
    """Re-times fn() with the oracle's loops on every host core."""
    import oracle
    n = oracle.max_threads()
    oracle.set_threads(n)
    try:
        c0 = time.perf_counter()
        fn()
        dt = time.perf_counter() - c0
    finally:
        oracle.set_threads(1)
    return {"value": unit_count / dt, "unit": unit, "cores": n, "seconds": dt}


def cpu_baseline(args, inv, pb, dobs, Wd):
    import oracle
    if args.workload == "inversion_small":
        inv.compute_G("f64")
        G = inv.get_G()
        P = oracle.barrier_P(pb["m_ini"], pb["m_min"], pb["m_max"])

        def solve():
            oracle.solve_cg(G, 0.0, args.cpu_iters, pb["lambdas"][0], pb["m_ini"], pb["mref"], Wd,
                            dobs, P, *pb["W"])

        c0 = time.perf_counter()
        solve()
        c1 = time.perf_counter()
        return {"value": args.cpu_iters / (c1 - c0), "unit": "CG iter/s", "cores": 1,
                "kind": "port",
                "sample": f"one solve_cg of {args.cpu_iters} CGLS iterations (plus its setup "
                          f"products) on the same 5000x32000 FP64 G, single-threaded C oracle "
                          f"(sequential GEMV, as the reference's Eigen GEMV runs); {c1 - c0:.1f} s; "
                          f"host has {os.cpu_count()} logical CPUs",
                "all_cores": all_cores(solve, args.cpu_iters, "CG iter/s")}
    # tensor_large: G (94 GB) is not copied to the host; the per-iteration CPU
    # cost is extrapolated from single-threaded oracle products on 2,000 rows.
    fp = pb["fp"]
    rows = 2000
    rng = np.random.default_rng(0)
    Gs = rng.standard_normal((rows, fp.prisms.shape[0]))
    x = rng.standard_normal(fp.prisms.shape[0])
    y = rng.standard_normal(rows)
    reps = 30                       # about 10 s on one core

    def products():
        for _ in range(reps):
            oracle.gemv(Gs, x)
            oracle.gemv(Gs, y, trans=True)

    c0 = time.perf_counter()
    products()
    c1 = time.perf_counter()
    per_iter = (c1 - c0) / reps * (inv.Nd / rows)
    ac = all_cores(products, reps * rows / inv.Nd, "CG iter/s")
    return {"value": 1.0 / per_iter, "unit": "CG iter/s", "cores": 1, "kind": "port",
            "sample": f"single-threaded C oracle G x and G^T w on {rows} rows x {inv.Nm} columns "
                      f"({reps} repetitions, {c1 - c0:.1f} s), extrapolated to {inv.Nd} rows; host "
                      f"has {os.cpu_count()} logical CPUs",
            "all_cores": ac}


def run_batch(args):
    from graada3d_amd.engine import TIME_ASSEMBLE, Context, GaussNewtonInversion, assemble_batch
    from graada3d_amd.workloads import add_noise, batch_instances, batch_regularisation
    steps = args.steps or 2
    warm = 1 if args.warmup is None else args.warmup
    t0 = time.perf_counter()
    inst = batch_instances()
    print(f"[bench] batch: {len(inst)} instances generated in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    fwds = [GaussNewtonInversion(b.prisms, b.points, b.flag, ctx=Context(0)) for b in inst]
    entries = float(sum(f.Nd * f.Nm for f in fwds))
    for _ in range(warm):
        assemble_batch(fwds)
    t_batch, t_seq = [], []
    ctx0 = fwds[0].ctx
    for s in range(steps):
        ctx0.timing(True)
        ctx0.timing_reset()
        c0 = time.perf_counter()
        assemble_batch(fwds)
        t_batch.append(time.perf_counter() - c0)
        k_ms, _ = ctx0.timing_get(TIME_ASSEMBLE)
        ctx0.timing(False)
        c0 = time.perf_counter()
        for f in fwds:
            f.compute_G("f64")
        t_seq.append(time.perf_counter() - c0)
        print(f"[bench] batch step {s}: batched {t_batch[-1]:.3f} s (kernel {k_ms / 1e3:.3f} s), "
              f"sequential {t_seq[-1]:.3f} s", file=sys.stderr, flush=True)
    tb, ts = min(t_batch), min(t_seq)
    # the same batch by per-prism corner sums (gra_assemble_batch_corners),
    # then the face/edge G again for the inversion below
    t_corner = []
    assemble_batch(fwds, method="corners")
    for s in range(steps):
        c0 = time.perf_counter()
        assemble_batch(fwds, method="corners")
        t_corner.append(time.perf_counter() - c0)
    tc = min(t_corner)
    print(f"[bench] batch corner sums: {tc:.3f} s ({entries / tc:.3e} entries/s)",
          file=sys.stderr, flush=True)
    assemble_batch(fwds)
    # batched inversion: per instance, L0 (IRLS, eps^2 = 1e-2) or L2 Gauss-Newton
    # on its octree operators, half of them with the cross-gradient term; the
    # host-side operator build and uploads are not timed
    from graada3d_amd.engine import gn_step_batch
    gn_steps, cg_its = 3, 20
    c0 = time.perf_counter()
    setups = []
    for b, f in zip(inst, fwds):
        W, T = batch_regularisation(b)
        f.set_regularization(*W, *(T or ()))
        ncomp = f.Nd // f.N_obs
        dobs, sigma = add_noise(f.G_vec_mul(b.rho), f.N_obs, ncomp, b.seed)
        nm = f.Nm
        setups.append((dobs, np.repeat(1.0 / sigma, f.N_obs), np.full(nm, -1000.0),
                       np.full(nm, 1000.0), np.zeros(nm), np.zeros(nm)))
    setup_s = time.perf_counter() - c0

    def gn_setup_all(fused):
        for b, f, a in zip(inst, fwds, setups):
            f.gn_setup(*a, lp=b.lp, eps2=1e-2 if b.lp != 2 else 0.0, fused=fused,
                       cross_gradient=b.cross_gradient, a_crg=1.0)
        for f in fwds:
            f.ctx.sync()

    lams = [[100.0 * 0.5 ** k] * len(fwds) for k in range(gn_steps)]
    runs = {}
    # (a) instance after instance, one host synchronisation per step
    # (b) gra_gn_step_batch: each step of all instances in one submission,
    #     two-pass CG (instances overlap) and single-read CG (chained)
    for name, fused, batched in (("sequential_two_pass", False, False),
                                 ("batched_two_pass", False, True),
                                 ("batched_single_read", True, True)):
        gn_setup_all(fused)
        for f in fwds:
            # single-read only where each row group gets enough row blocks
            # (small instances stay two-pass)
            f.ctx.tune(fused_min_blocks=BATCH_MIN_BLOCKS if fused else 0)
        c0 = time.perf_counter()
        total_its, models = 0, None
        if batched:
            for k in range(gn_steps):
                mis, err, its = gn_step_batch(fwds, lams[k], 0.0, cg_its)
                total_its += int(its.sum())
        else:
            for f in fwds:
                for k in range(gn_steps):
                    total_its += f.gn_step(lams[k][0], 0.0, cg_its)[2]
        t_run = time.perf_counter() - c0
        runs[name] = {"seconds": t_run, "cg_iterations": total_its,
                      "cg_iter_per_s": total_its / t_run,
                      "G_TBps": gn_steps * cg_its * (1 if fused else 2) *
                      sum(f.Nd * f.Nm * 8.0 for f in fwds) / t_run / 1e12}
        if name == "sequential_two_pass":
            ref_models = [f.gn_model() for f in fwds]
        elif name == "batched_two_pass":
            runs[name]["bit_identical_to_sequential"] = all(
                np.array_equal(f.gn_model(), m) for f, m in zip(fwds, ref_models))
        print(f"[bench] batch inversion {name}: {total_its} CG iterations in {t_run:.2f} s",
              file=sys.stderr, flush=True)
    t_inv = runs["sequential_two_pass"]["seconds"]
    total_its = runs["sequential_two_pass"]["cg_iterations"]
    cpu = None
    if not args.no_cpu:
        import oracle
        b = inst[0]
        n = max(1, int(2e6 // b.prisms.shape[0]))
        c0 = time.perf_counter()
        oracle.assemble(b.prisms, b.points[:n], b.flag)
        c1 = time.perf_counter()
        e = n * b.prisms.shape[0] * bin(b.flag).count("1")
        ac = all_cores(lambda: oracle.assemble(b.prisms, b.points[:n], b.flag), e, "G entries/s")
        cpu = {"value": e / (c1 - c0), "unit": "G entries/s", "cores": 1, "kind": "port",
               "sample": f"single-threaded C oracle assembly of instance 0's G on {n} of its "
                         f"{b.points.shape[0]} points ({b.prisms.shape[0]} cells, "
                         f"{bin(b.flag).count('1')} components; {c1 - c0:.1f} s)", "all_cores": ac}
    sizes = np.array([[f.Nm, f.N_obs, f.Nd // f.N_obs] for f in fwds])
    out = {
        "metric": METRIC, "value": entries / tb, "unit": "G entries/s", "n_gpus": 1,
        "steps": steps, "warmup": warm, "ms_per_step": tb * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batch: 64 uneven instances, 2:1 octree meshes (2k-60k cells "
                               "refined at seeded box faces), 400-10000 points at 0.5-100 m, "
                               "random component subsets; FP64 G of every instance assembled by "
                               "one batched launch", "entries": entries,
                   "cells": [int(sizes[:, 0].min()), int(sizes[:, 0].max())],
                   "points": [int(sizes[:, 1].min()), int(sizes[:, 1].max())]},
        "batched_s": tb, "sequential_s": ts, "sequential_entries_per_s": entries / ts,
        "speedup_batched_vs_sequential": ts / tb,
        "corner_sums": {"batched_s": tc, "entries_per_s": entries / tc,
                        "speedup_vs_face_edge": tb / tc,
                        "kernel": "per-prism corner sums, one batched launch "
                                  "(gra_assemble_batch_corners), FP64"},
        "batch_inversion": {"instances": len(inst), "gn_steps": gn_steps, "cg_iters_per_step": cg_its,
                            "cg_iterations": total_its, "seconds": t_inv,
                            "cg_iter_per_s": total_its / t_inv,
                            "G_bytes_per_s": 2.0 * 8.0 * sum(f.Nd * f.Nm for f in fwds) *
                            gn_steps * cg_its / t_inv,
                            "L0": int(sum(b.lp == 0 for b in inst)),
                            "cross_gradient": int(sum(b.cross_gradient for b in inst)),
                            "host_setup_s": setup_s, "runs": runs},
        "roofline": {"bound": "fp64-valu", "achieved": None, "peak": 78.6, "unit": "TFLOP/s",
                     "frac": None, "traffic": None,
                     "note": "transcendental-bound FP64 evaluation; HBM writes are "
                             f"{entries * 8 / tb / 1e9:.0f} GB/s of 8000"},
        "cpu_baseline": cpu,
    }
    for f in fwds:
        f.ctx.close()
    print(json.dumps(out))


def run_serving(args):
    from graada3d_amd import serving as S
    pool = S.default_cg_pool()
    reqs = S.request_stream(args.rate, args.duration, args.seed, pool)
    print(f"[bench] serving: {len(reqs)} requests over {args.duration} s at {args.rate}/s, "
          f"{len(pool)} CG problems ({sum(S.ResidentCache.g_bytes(p) for p in pool) / 1e9:.0f} "
          f"GB of G) under a {args.budget_gb:g} GB budget", file=sys.stderr, flush=True)
    res = S.serve(reqs, pool, budget_bytes=args.budget_gb * 1e9,
                  progress=max(50, int(args.rate * 10)), forward_method=args.forward_method,
                  workers=args.workers, max_batch=args.max_batch, policy=args.policy,
                  cg_workers=args.cg_workers)
    cpu = None
    if not args.no_cpu:
        import oracle
        # CPU: the oracle's forward on the first forward request's mesh and
        # components, on a bounded slice of its points, extrapolated per request
        r = next(q for q in reqs if q.kind == "forward")
        q = S.forward_payload(r.params)
        prisms, rho, pts = q["prisms"], q["rho"], q["pts"]
        n = max(1, min(pts.shape[0], int(2e6 // prisms.shape[0])))
        c0 = time.perf_counter()
        oracle.forward(prisms, rho, pts[:n], r.params["flag"])
        c1 = time.perf_counter()
        per_req = (c1 - c0) * pts.shape[0] / n
        ac = all_cores(lambda: oracle.forward(prisms, rho, pts[:n], r.params["flag"]),
                       n / pts.shape[0], "forward requests/s")
        cpu = {"value": 1.0 / per_req, "unit": "forward requests/s", "cores": 1, "kind": "port",
               "sample": f"single-threaded C oracle forward of request {r.rid} ({prisms.shape[0]} "
                         f"cells, {bin(r.params['flag']).count('1')} components) on {n} of its "
                         f"{pts.shape[0]} points ({c1 - c0:.1f} s), extrapolated to the whole "
                         f"request; GPU: {res['by_kind']['forward']['service_ms']['mean']:.1f} ms "
                         f"mean forward service time", "all_cores": ac}
    fwd = res["by_kind"]["forward"]["latency_ms"]
    # the served solves' error from FP32 storage, against FP64 two-pass
    prec = {}
    for name in ["inversion_small"] + sorted(_serving_extremes(pool)) + ["tensor_large"]:
        prob = next(p for p in pool if p.name == name)
        prec[name] = S.store_error(prob)
        print(f"[bench] serving precision {name}: dx {prec[name]['dx_rel_err']:.2e}",
              file=sys.stderr, flush=True)
    out = {
        "metric": METRIC,
        "value": fwd["p99"],
        "unit": "ms (p99 forward request latency)",
        "n_gpus": 1, "steps": res["requests"], "warmup": 0,
        "ms_per_step": res["service_ms"]["mean"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"serving: Poisson {args.rate:g} req/s for {args.duration:g} s "
                               f"(seed {args.seed}), 80% matrix-free forward (5k-130k cells, "
                               f"1k-20k points, 1-9 comps, FP32 out, {args.forward_method} "
                               f"form, up to {args.max_batch} per launch), 20% 20-iteration "
                               f"CGLS on resident FP32 corner-sum G, two-pass products "
                               f"(pool: inversion_small 45%, the 64 octree instances 54%, "
                               f"tensor_large 1%; "
                               f"{res['cache']['pool_gb']:.0f} GB of G under a "
                               f"{args.budget_gb:g} GB {args.policy} budget, tensor_large "
                               f"pinned), {args.workers} forward + {args.cg_workers} CG server "
                               f"streams"},
        "latency_ms": res["latency_ms"], "service_ms": res["service_ms"],
        "by_kind": res["by_kind"], "cg_warm_service_ms": res["cg_warm_service_ms"],
        "cg_by_problem": {k: v for k, v in res["cg_by_problem"].items()
                          if k in {"inversion_small", "tensor_large"} | _serving_extremes(pool)},
        "throughput_req_per_s": res["achieved_rate"],
        "offered_rate": res["offered_rate"], "achieved_rate": res["achieved_rate"],
        "utilization": res["utilization"],
        "mean_in_service": res["mean_in_service"], "workers": res["workers"],
        "cg_workers": res["cg_workers"],
        "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
        "forward_batches": res["forward_batches"], "cache": res["cache"],
        "cg_precision_vs_f64": prec,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def _serving_extremes(pool):
    """Names of the smallest and largest octree instances of the pool."""
    from graada3d_amd import serving as S
    inst = [p for p in pool if p.instance >= 0]
    if not inst:
        return set()
    by = sorted(inst, key=S.ResidentCache.g_bytes)
    return {by[0].name, by[len(by) // 2].name, by[-1].name}


def stage_karst(root, key="interpolated"):
    """The reference's cave-inversion example (configuration files and data
    file, kept as fixtures under tests/golden/karst) laid out under root."""
    import shutil
    folder = {"raw": "invert_raw_data_pivetta2015",
              "interpolated": "invert_interpolated_data_Braitenberg2015"}[key]
    src = os.path.join(ROOT, "tests", "golden", "karst")
    fx = dict(np.load(os.path.join(src, f"{key}.npz"), allow_pickle=False))
    os.makedirs(os.path.join(root, folder), exist_ok=True)
    for name in ("config", "config_data", "config_model", "config_inversion"):
        shutil.copyfile(os.path.join(src, folder, name), os.path.join(root, folder, name))
    os.makedirs(os.path.join(root, "data_preprocessing"), exist_ok=True)
    np.savetxt(os.path.join(root, "data_preprocessing", str(fx["data_name"])), fx["data"],
               fmt="%.17g")
    return os.path.join(root, folder, "config"), fx


def run_karst(args):
    """The whole adaptive inversion of the reference's cave example (11,374
    gz points, 16x13x12 initial octree refined to 10,490 cells, L1
    smallness, 52 lambda trials of up to 100 CG iterations) through the
    driver on the GPU engine, from reading the files to the final model."""
    import tempfile
    from graada3d_amd import driver
    from graada3d_amd.engine import TIME_GEMV_N, TIME_GEMV_T, Context
    steps = args.steps or 3
    warm = 1 if args.warmup is None else args.warmup
    ctx = Context(0)
    tmp = tempfile.mkdtemp(prefix="karst_")
    cfg, fx = stage_karst(tmp)
    times, runs = [], []
    for k in range(warm + steps):
        d, m, inv = driver.read_config(cfg)
        if args.karst_method == 2:
            from dataclasses import replace
            inv = replace(inv, method=2, threshold=args.karst_threshold)
        ctx.timing(k >= warm)
        ctx.timing_reset()
        t0 = time.perf_counter()
        run = driver.AdaptiveInversion(d, m, inv, ctx=ctx, out_dir=tmp)
        res = run.run()
        ctx.sync()
        el = time.perf_counter() - t0
        if k >= warm:
            gn, gt = ctx.timing_get(TIME_GEMV_N), ctx.timing_get(TIME_GEMV_T)
            times.append(el)
            runs.append((run, res, gn, gt))
        print(f"[bench] karst run {k}: {el:.3f} s, misfit {res['misfit']:.9g}, Nm {res['Nm']}",
              file=sys.stderr, flush=True)
        ctx.timing(False)
    run, res, gn, gt = runs[-1]
    nd = run.n_obs
    # G-product bytes: every solve's products at that trial's Nm; the launches
    # beyond the CG iterations (set-up products, misfits) are spread evenly
    trials = run.trials
    its = np.array([t["cg_iterations"] for t in trials], float)
    nms = np.array([t["Nm"] for t in trials], float)
    # bytes of one product: the dense G (8 B per entry) or the compressed
    # rows (8 B value + 4 B index per kept coefficient)
    per = (np.array([12.0 * t["nnz"] for t in trials]) if args.karst_method == 2
           else nd * nms * 8)

    def rate(tm):
        avg_ms, n = tm
        extra = (n - its.sum()) / len(trials)
        byt = float(((its + extra) * per).sum())
        return byt / (avg_ms * n * 1e-3) / 1e9, avg_ms * n
    gbn, msn = rate(gn)
    gbt, mst = rate(gt)
    frac = (gbn * msn + gbt * mst) / (msn + mst) / HBM_PEAK_GBS
    cpu = None
    if not args.no_cpu and args.karst_method == 1:
        cpu = karst_cpu_baseline(cfg, trials, res["Nm"])
    sec = float(np.mean(times))
    out = {
        "metric": "whole adaptive inversion wall time (reference's cave example)",
        "value": sec, "unit": "s", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": sec * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "reference example data (Filtered_anomaly.dat)",
        "config": {"workload": f"cave inversion: {nd} gz points, 16x13x12 octree -> {res['Nm']} "
                               f"cells, {len(trials)} lambda trials, "
                               f"{int(its.sum())} CG iterations, final misfit {res['misfit']:.9g}"
                               + (f", method 2 (wavelet-compressed G, threshold "
                                  f"{args.karst_threshold:g}: {trials[-1]['nnz']} of "
                                  f"{nd * run.wavelet[0]} coefficients kept)"
                                  if args.karst_method == 2 else ", method 1 (dense G)")},
        "roofline": {"bound": "hbm", "achieved": (gbn * msn + gbt * mst) / (msn + mst),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": frac, "traffic": None},
        "gemv_GBps": {"n": gbn, "t": gbt}, "gemv_share_of_wall": (msn + mst) * 1e-3 / sec,
        "reference_logged_s": float(fx["seconds"]),
        "misfit_vs_reference_run": abs(res["misfit"] - float(fx["iters"][-1, 2]))
        / float(fx["iters"][-1, 2]),
        "cpu_baseline": cpu,
    }
    ctx.close()
    print(json.dumps(out))


def karst_cpu_baseline(cfg, trials, final_nm, n_trials=2):
    """The same driver with the oracle behind it, for the first n_trials
    lambda trials (all at the initial 2,496 cells), scaled to the whole run by
    CG work (iterations x Nm per trial)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import tempfile
    import oracle
    from helpers import OracleGN
    from graada3d_amd import driver
    n = oracle.max_threads()
    oracle.set_threads(n)
    try:
        d, m, inv = driver.read_config(cfg)
        c0 = time.perf_counter()
        run = driver.AdaptiveInversion(d, m, inv, out_dir=tempfile.mkdtemp(prefix="karst_cpu_"),
                                       engine_factory=OracleGN)
        c1 = time.perf_counter()
        run.run(max_trials=n_trials)
        c2 = time.perf_counter()
    finally:
        oracle.set_threads(1)

    def work(ts):
        return sum((t["cg_iterations"] + 2) * t["Nm"] for t in ts)
    # columns assembled over the run: the initial mesh plus 8 per split cell
    nm0, nm1 = trials[0]["Nm"], final_nm
    cols = nm0 + 8 * (nm1 - nm0) / 7
    est = (c1 - c0) * cols / nm0 + (c2 - c1) * work(trials) / work(run.trials)
    return {"value": est, "unit": "s (whole inversion, extrapolated)", "cores": n, "kind": "port",
            "sample": f"driver on the C oracle (OpenMP, {n} threads): set-up with the "
                      f"{run.n_obs}x{nm0} G assembly ({c1 - c0:.1f} s, scaled to the run's "
                      f"{cols:.0f} assembled columns) and the first {n_trials} lambda trials "
                      f"({c2 - c1:.1f} s, scaled by CG work = iterations x cells to the GPU run's "
                      f"{len(trials)} trials)"}


def device_copy(nbytes: int) -> None:
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or \
            hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3):      # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def run_probe(args) -> None:
    """Fixed, short launch sequences for rocprofv3 counter passes (the
    --probe modes listed in the module docstring)."""
    from graada3d_amd.engine import Context, Fwd, forward_rectilinear
    from graada3d_amd.workloads import forward_problem_small
    from graada3d_amd import mesh as M

    if args.probe in ("mat-mul", "mat-mul-large"):
        ctx = Context(0)
        if args.probe == "mat-mul":
            fp, method = forward_problem_small(), "faces"
        else:
            from graada3d_amd.workloads import tensor_problem_large
            fp, method = tensor_problem_large(), "nodes"
        f = Fwd(fp.prisms, fp.points, fp.flag, ctx=ctx)
        rng = np.random.default_rng(0)
        X, R = rng.standard_normal((f.Nm, 16)), rng.standard_normal((f.Nd, 16))
        for store in ("f64", "f32", "bf16"):
            f.compute_G(store, method=method)
            for _ in range(4):                         # one warm-up call and three more
                f.G_mat_mul(X)
                f.GT_mat_mul(R)
            print(f"[probe] {store}: G X and G^T R with 16 right-hand sides", flush=True)
        ctx.close()
        return

    if args.probe == "gemv-large":
        # the 90,000 x 131,072 FP64 G of tensor_large (assembled by corner
        # sums, which is quick; the products do not depend on the method)
        device_copy(1 << 30)
        print("[probe] 1 GiB device copy done", flush=True)
        from graada3d_amd.workloads import tensor_problem_large
        ctx = Context(0)
        fp = tensor_problem_large()
        f = Fwd(fp.prisms, fp.points, fp.flag, ctx=ctx)
        f.compute_G("f64", method="nodes")
        rng = np.random.default_rng(0)
        x, y = rng.standard_normal(f.Nm), rng.standard_normal(f.Nd)
        for _ in range(3):
            f.G_vec_mul(x)
            f.GT_vec_mul(y)
        print(f"[probe] f64: 3 x G x, 3 x G^T w on {f.Nd} x {f.Nm}", flush=True)
        ctx.close()
        return

    if args.probe == "octree":
        from graada3d_amd.octree import OctreeMesh
        ctx = Context(0)
        mesh = OctreeMesh((0.0, 2000.0), 40, (0.0, 2000.0), 40, (0.0, 1000.0), 20)
        rng = np.random.default_rng(3)
        for _ in range(2):
            n = mesh.n_cells
            mesh.refine(np.sort(rng.choice(n, n // 8, replace=False)))
        pr = mesh.prisms()
        for flag, npt in ((0x12, 50), (0x3FE, 30)):
            pts = M.grid_points(npt, 0.0, 2000.0, npt, 0.0, 2000.0, -1.0)
            f = Fwd(pr, pts, flag, ctx=ctx)
            for method in ("faces", "corners"):
                for _ in range(3):
                    f.compute_G("f64", method=method)
                print(f"[probe] octree {len(pr)} cells, {f.Nd} x {f.Nm} G, flag {flag:#x}, "
                      f"{method} x3", flush=True)
            del f
        ctx.close()
        return

    if args.probe == "assemble":
        ctx = Context(0)
        fp = forward_problem_small()
        f = Fwd(fp.prisms, fp.points, fp.flag, ctx=ctx)
        f.compute_G("f64")
        print(f"[probe] assembled {f.Nd} x {f.Nm} (gz+Tzz)", flush=True)
        pr = M.regular_mesh((0.0, 2000.0), 20, (0.0, 2000.0), 20, (0.0, 1000.0), 10)
        pts = M.grid_points(40, 0.0, 2000.0, 40, 0.0, 2000.0, -1.0)
        g = Fwd(pr, pts, 0x3FE, ctx=ctx)
        g.compute_G("f64")
        print(f"[probe] assembled {g.Nd} x {g.Nm} (9 components)", flush=True)
        f.compute_G("f64", method="nodes")
        g.compute_G("f64", method="nodes")
        print("[probe] the same two G by corner sums", flush=True)
        xs = M.linspaced(41, 0.0, 2000.0)
        zs = M.linspaced(21, 0.0, 1000.0)
        forward_rectilinear(xs, xs, zs, fp.rho, fp.points, 0x3FE, ctx=ctx)
        print("[probe] node-sum forward 40x40x20 mesh, 2500 points, 9 components", flush=True)
        ctx.close()
        return

    device_copy(1 << 30)
    print("[probe] 1 GiB device copy done", flush=True)
    ctx = Context(0)
    fp = forward_problem_small()
    f = Fwd(fp.prisms, fp.points, fp.flag, ctx=ctx)
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal(f.Nm), rng.standard_normal(f.Nd)
    for store in (("f64", "f32", "bf16") if args.all_dtypes else ("f64",)):
        f.compute_G(store)
        for _ in range(3):
            f.G_vec_mul(x)
            f.GT_vec_mul(y)
        print(f"[probe] {store}: 3 x G x, 3 x G^T w on {f.Nd} x {f.Nm}", flush=True)
    ctx.close()


def mixed_summary(mixed, bar=1e-6):
    """The fastest variant whose final-model and misfit errors are within
    `bar` of the plain FP64 run, and the fastest variant overall."""
    ok = {k: v for k, v in mixed.items()
          if v["model_rel_err"] <= bar and v["misfit_rel_err"] <= bar}
    pick = lambda d: max(d, key=lambda k: d[k]["speedup_vs_f64"]) if d else None
    out = {}
    for name, d in (("fastest_within_1e-6", ok), ("fastest", mixed)):
        k = pick(d)
        out[name] = None if k is None else {
            "variant": k, "speedup_vs_f64": d[k]["speedup_vs_f64"], "iter_s": d[k]["iter_s"],
            "model_rel_err": d[k]["model_rel_err"], "misfit_rel_err": d[k]["misfit_rel_err"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="tensor_large",
                    choices=["inversion_small", "tensor_large", "batch", "serving", "karst"])
    ap.add_argument("--rate", type=float, default=200.0, help="serving: requests/s")
    ap.add_argument("--duration", type=float, default=2.0, help="serving: arrival window, s")
    ap.add_argument("--seed", type=int, default=20240605, help="serving: stream seed")
    ap.add_argument("--workers", type=int, default=4,
                    help="serving: forward server threads, each with its own stream")
    ap.add_argument("--cg-workers", type=int, default=1,
                    help="serving: CG server threads (0: the forward threads take both kinds)")
    ap.add_argument("--budget-gb", type=float, default=80.0,
                    help="serving: bytes of resident G the cache may hold (GB)")
    ap.add_argument("--policy", default="gdsf", choices=["gdsf", "lru"],
                    help="serving: eviction policy of the resident-G cache")
    ap.add_argument("--max-batch", type=int, default=8,
                    help="serving: forward requests per shared launch (1: no batching)")
    ap.add_argument("--forward-method", default="nodes", choices=["nodes", "prisms"],
                    help="serving: node-sum (rectilinear) or per-prism face/edge forward")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--stores", default="f64,f32,bf16")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="use the default G-product launch shapes instead of gra_autotune's")
    ap.add_argument("--no-fused", dest="fused", action="store_false",
                    help="skip the single-read (fused) CG iteration variants")
    ap.add_argument("--cpu-iters", type=int, default=40)
    ap.add_argument("--karst-method", type=int, default=1, choices=[1, 2],
                    help="karst: dense G (1) or wavelet-compressed G (2)")
    ap.add_argument("--karst-threshold", type=float, default=5e-4,
                    help="karst: relative threshold of the wavelet compression (method 2)")
    ap.add_argument("--probe", default=None,
                    choices=["gemv", "gemv-large", "assemble", "octree", "mat-mul",
                             "mat-mul-large"],
                    help="run a fixed launch sequence for rocprofv3 counter passes instead")
    ap.add_argument("--all-dtypes", action="store_true",
                    help="probe gemv: also with G stored in FP32 and BF16")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the engine's "
              "multi-GPU sharding is a design note, not part of this build.", file=sys.stderr)
        sys.exit(2)
    if args.probe:
        run_probe(args)
        return
    if args.workload == "serving":
        # one hardware queue per server stream (five by default; HIP's default
        # is four, and streams that share a queue serialise their work: with
        # four queues the forward p99 tripled); set before the runtime starts
        streams = args.workers + args.cg_workers
        if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < streams:
            os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, max(8, streams)))
        run_serving(args)
        return
    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "karst":
        run_karst(args)
        return
    small = args.workload == "inversion_small"
    if args.steps is None:
        args.steps = 10
    if args.warmup is None:
        args.warmup = 2 if small else 1
    args.stores = [s for s in args.stores.split(",") if s]
    if "f64" not in args.stores:
        args.stores = ["f64"] + args.stores
    args.stores = ["f64"] + [s for s in args.stores if s != "f64"]

    pb, nd, nm, res, cpu = run(args)
    skipped = res.pop("_skipped", [])
    nodes_info = res.pop("_nodes")
    ref = res["f64"]
    mixed = {}
    for s, r in res.items():
        if s == "f64":
            continue
        mixed[s] = {"iter_s": r["iter_s"], "speedup_vs_f64": r["iter_s"] / ref["iter_s"],
                    "model_rel_err": float(np.linalg.norm(r["model"] - ref["model"])
                                           / np.linalg.norm(ref["model"])),
                    "misfit_rel_err": abs(r["misfit"] - ref["misfit"]) / ref["misfit"],
                    "data_rel_err_per_component": r["data_rel_err"],
                    "mode": r["mode"], "g_entries_per_s": r["entries_per_s"]}
        if r["refresh"]:
            mixed[s]["fp64_refresh"] = r["refresh"]
        if r["fused"]:
            mixed[s]["fused_GBps"] = r["fused_product"]["GBps"]
            mixed[s]["fused_avg_ms"] = r["fused_product"]["avg_ms"]
        else:
            mixed[s]["gemv_GBps"] = {"n": r["gemv_n"]["GBps"], "t": r["gemv_t"]["GBps"]}
            mixed[s]["mat_mul"] = r.get("mat_mul")
    dom = "gemv_t" if ref["gemv_t"]["avg_ms"] >= ref["gemv_n"]["avg_ms"] else "gemv_n"
    achieved = ref[dom]["GBps"]
    # HBM bytes per launch of the dominant kernel from rocprofv3 PMC runs
    # (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, calibrated on a 1 GiB device copy;
    # counter passes over --probe gemv and --probe gemv-large)
    if small:
        traffic = {"gemv_t": 1280.6e6 + 0.8e6, "gemv_n": 1282.1e6 + 0.04e6}[dom] / 1e9
    else:
        traffic = {"gemv_t": 94.384e9, "gemv_n": 94.382e9}[dom] / 1e9
    out = {
        "metric": METRIC,
        "value": ref["iter_s"],
        "unit": "CG iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ref["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": ("inversion_small: L2 Gauss-Newton on the 40x40x20 mesh, 50x50 gz+Tzz "
                         "points, dense FP64 G 5000x32000 resident in HBM, 50 CGLS iterations per "
                         "step (plain FP64, 2 passes over G per iteration)") if small else
                        ("tensor_large: L1 IRLS (eps^2 = 1e-3) Gauss-Newton on the 64x64x32 mesh "
                         "(62.5 m cells, 12 seeded random boxes), 100x100 points at z=-50 m, all 9 "
                         "components, dense FP64 G 90000x131072 (94.4 GB) resident in HBM, 20 "
                         "CGLS iterations per step (plain FP64, 2 passes over G per iteration)"),
            "Nd": nd, "Nm": nm, "cg_iters_per_step": pb["cg_maxit"],
            "lambda_schedule": "100*0.5^k"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_unit": "GB per launch"},
        "kernels": {"gemv_n": ref["gemv_n"], "gemv_t": ref["gemv_t"],
                    "events": f"HIP events around the G products of CG iterations 0, "
                              f"{TIMING_STRIDE}, {2 * TIMING_STRIDE}, ... of every timed solve "
                              f"and the set-up products"},
        "g_assembly": {"entries_per_s": ref["entries_per_s"], "avg_ms": ref["asm_ms"],
                       "entries": nd * nm, "store": "f64",
                       "kernel": "face/edge FP64 evaluation, " + ("gz+Tzz" if small else "9 comps"),
                       # FP64 flop per entry from PMC instruction counters
                       "roofline": {"bound": "fp64-valu", "flop_per_entry": FLOP_PER_ENTRY[small],
                                    "achieved": ref["entries_per_s"] * FLOP_PER_ENTRY[small] / 1e12,
                                    "peak": 78.6, "unit": "TFLOP/s",
                                    "frac": ref["entries_per_s"] * FLOP_PER_ENTRY[small] / 78.6e12}},
        "g_assembly_nodes": nodes_info,
        "mat_mul": ref.get("mat_mul"),
        "final_misfit": ref["misfit"],
        "mixed": mixed,
        "mixed_not_run": {k: "fused grid not co-resident for this G (two-pass iteration ran)"
                          for k in skipped},
        "mixed_summary": mixed_summary(mixed),
        "mixed_error_definition": "||m - m_f64|| / ||m_f64|| of the final model, |misfit - "
                                  "misfit_f64| / misfit_f64, and per component ||G m - G m_f64|| "
                                  "/ ||G m_f64|| with the FP64 G, after the same steps, against "
                                  "the plain FP64 GPU run (itself checked against the CPU "
                                  "oracle); modes: two-pass (G stored in that type), fused (one "
                                  "single-read pass per iteration), refined (fused pass over an "
                                  "FP32/BF16 copy of the FP64 G, FP64 residual refresh every "
                                  f"{REFRESH} iterations and after the last)",
        "tuning": ref["tuning"],
        "autotune_s": ref["autotune_s"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
