"""Hardware-counter probe for the G products.

Runs a fixed, short sequence on GPU 0 for ``rocprofv3 --pmc`` passes:

1. a plain device-to-device copy of 1 GiB (the runtime's copy kernel) -- the
   calibration dispatch: its FETCH_SIZE / WRITE_SIZE readings correspond to
   exactly 1 GiB read / written;
2. the 5,000 x 32,000 FP64 G of the small inversion benchmark, then 3 launches
   each of G x and G^T w (and the same with G stored in FP32 and BF16 when
   ``--all-dtypes`` is given).

    rocprofv3 --pmc FETCH_SIZE -d <dir> -o probe -- python3 -m graada3d_amd.probe
    rocprofv3 --pmc WRITE_SIZE -d <dir> -o probe -- python3 -m graada3d_amd.probe

With ``--assemble`` it instead assembles G once for the same survey (gz+Tzz,
1.6e8 entries), once with all nine components on a 20x20x10 mesh and a 40x40
grid (1.44e7 entries), the same two G by corner sums
(gra_assemble_rectilinear) and runs one node-sum forward (40x40x20 mesh,
50x50 points, 9 components), for FP64 instruction counters:

    rocprofv3 --pmc SQ_INSTS_VALU_ADD_F64 SQ_INSTS_VALU_MUL_F64 SQ_INSTS_VALU_FMA_F64 \
        SQ_INSTS_VALU_TRANS_F64 -d <dir> -o asm -- python3 -m graada3d_amd.probe --assemble

With ``--octree`` it assembles G on an octree -- the 40x40x20 mesh with two
rounds of refinement of a seeded eighth of the cells (158,427 cells) -- three
times each by the face/edge kernel (gra_assemble) and by per-prism corner
sums (gra_assemble_corners): gz+Tzz at 50x50 points and all nine components
at 30x30 points, z = -1 m:

    rocprofv3 --kernel-trace --stats -d <dir> -o oct -- python3 -m graada3d_amd.probe --octree

With ``--mat-mul`` it runs the matrix-core products with 16 right-hand sides
(gra_G_mat_mul, gra_GT_mat_mul; 2 launches each) on the same 5,000 x 32,000 G
stored in FP64, FP32 and BF16, for MFMA counters:

    rocprofv3 --pmc SQ_VALU_MFMA_BUSY_CYCLES SQ_BUSY_CYCLES GRBM_GUI_ACTIVE -d <dir> -o mm \
        -- python3 -m graada3d_amd.probe --mat-mul
"""
from __future__ import annotations

import argparse
import ctypes
import sys

import numpy as np


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


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--all-dtypes", action="store_true")
    ap.add_argument("--assemble", action="store_true")
    ap.add_argument("--mat-mul", action="store_true")
    ap.add_argument("--octree", action="store_true")
    args = ap.parse_args(argv)
    from .engine import Context, Fwd, forward_rectilinear
    from .workloads import forward_problem_small
    from . import mesh as M

    if args.mat_mul:
        ctx = Context(0)
        fp = forward_problem_small()
        f = Fwd(fp.prisms, fp.points, fp.flag, ctx=ctx)
        rng = np.random.default_rng(0)
        X, R = rng.standard_normal((f.Nm, 16)), rng.standard_normal((f.Nd, 16))
        for store in ("f64", "f32", "bf16"):
            f.compute_G(store)
            for _ in range(2):
                f.G_mat_mul(X)
                f.GT_mat_mul(R)
            print(f"[probe] {store}: G X and G^T R with 16 right-hand sides", flush=True)
        ctx.close()
        return

    if args.octree:
        from .octree import OctreeMesh
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

    if args.assemble:
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


if __name__ == "__main__":
    main(sys.argv[1:])
