"""The inversion driver on the GPU engine against the two complete inversion
runs the reference ships (Examples/Karstic_cave): the same lambda trials,
tried-lambda counts and refined octrees, final misfits and models."""
import time

import pytest

from graada3d_amd import driver
from helpers import check_against_run, stage_karst

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("key,nm,its", [("raw", 4239, 3), ("interpolated", 10490, 6)])
def test_driver_matches_reference_run(ctx, tmp_path, key, nm, its):
    cfg, fx = stage_karst(key, tmp_path)
    d, m, inv = driver.read_config(cfg)
    out = tmp_path / "out"
    out.mkdir()
    t0 = time.perf_counter()
    run = driver.AdaptiveInversion(d, m, inv, ctx=ctx, out_dir=str(out))
    res = run.run()
    secs = time.perf_counter() - t0
    print(f"{key}: {res['iterations']} GN iterations, Nm {res['Nm']}, misfit {res['misfit']:.12g}, "
          f"{secs:.2f} s (reference run {float(fx['seconds']):.1f} s)", flush=True)
    assert res["Nm"] == nm and res["iterations"] == its
    check_against_run(run, fx, str(out))
    run.write_result(inv.output, secs)
