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
    args = ap.parse_args(argv)
    from .engine import Context, Fwd
    from .workloads import forward_problem_small

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
