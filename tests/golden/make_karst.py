"""Writes the inversion-driver fixtures under tests/golden/karst/ from the two
complete inversion runs the reference ships (run in the build container,
where the reference tree exists):

    python tests/golden/make_karst.py

Source: Examples/Karstic_cave/{invert_raw_data_pivetta2015,
invert_interpolated_data_Braitenberg2015}: the GraAda3D configuration files
the runs used (copied verbatim), the data files they read
(data_preprocessing/*.dat) and what the runs wrote: lambda_misfit_GN and
Iteration_misfit_GN (every trial's lambda and misfit), <output>.txt (final
octree cells and model) and dpredicted_g_z.

Stored: the configuration files as text; per example one npz with the data
file's values (`data`, as read), the trial log (`lam`, `misfit`), the
iteration log (`iters`: iteration, tried lambdas, misfit), the final cells
(`cells`: x0 x1 y0 y1 z0 z1), model (`model`), the predicted data (`dpred`)
and the run's wall time in seconds from its `info` file (`seconds`).
"""
from __future__ import annotations

import os
import shutil

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/Examples/Karstic_cave"
OUT = os.path.join(HERE, "karst")
EXAMPLES = {
    "raw": ("invert_raw_data_pivetta2015", "bouguer_gz_detrended_Pivetta2015.dat"),
    "interpolated": ("invert_interpolated_data_Braitenberg2015", "Filtered_anomaly.dat"),
}


def main():
    os.makedirs(OUT, exist_ok=True)
    for key, (folder, data_name) in EXAMPLES.items():
        src = os.path.join(SRC, folder)
        dst = os.path.join(OUT, folder)
        os.makedirs(dst, exist_ok=True)
        for name in ("config", "config_data", "config_model", "config_inversion"):
            shutil.copyfile(os.path.join(src, name), os.path.join(dst, name))
            os.chmod(os.path.join(dst, name), 0o644)
        data = np.loadtxt(os.path.join(SRC, "data_preprocessing", data_name))
        lm = np.loadtxt(os.path.join(src, "lambda_misfit_GN"), skiprows=1)
        it = np.loadtxt(os.path.join(src, "Iteration_misfit_GN"), skiprows=1, ndmin=2)
        res = np.loadtxt(os.path.join(src, "gz_result.txt"), comments="#")
        dp = np.loadtxt(os.path.join(src, "dpredicted_g_z"))
        with open(os.path.join(src, "info")) as f:
            secs = [float(l.split()[2]) for l in f if l.startswith("Running time")][0]
        np.savez_compressed(os.path.join(OUT, f"{key}.npz"), data=data, data_name=data_name,
                            lam=lm[:, 0], misfit=lm[:, 1], iters=it, cells=res[:, :6],
                            model=res[:, 9], dpred=dp[:, -1], seconds=secs)
        print(key, data.shape, lm.shape, it.shape, res.shape, dp.shape, secs)


if __name__ == "__main__":
    main()
