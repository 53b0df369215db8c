"""Writes the golden fixtures under tests/golden/ (run in the build container,
where the reference tree and oracle/_ref exist):

    python tests/golden/make_golden.py

1. Reference-held golden vectors (data files of the reference's own tests):
   Examples/Forward_modelling/{tests/test_compute/result_no_noise,
   model2/result/data_no_noise, model3/result/data_no_noise}: noise-free
   outputs of the reference `compute` tool (%.15e).  Stored as npz with the
   model definitions their scripts used (anomaly boxes, mesh) and, for the
   40x40x20 models, the cell densities of the committed .vtk files.
2. Vectors from the reference's own prism sources (oracle/_ref, built by
   oracle/Makefile from src/Forward_modelling/*.cpp): all nine components for
   random, grid-aligned and on-face (degenerate-branch) point/prism pairs and
   a small dense G, which the committed files above do not cover (gx, gy,
   Txx/Txy/Tyy on arbitrary geometry, the G layout).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from graada3d_amd import mesh as M  # noqa: E402

REF = "/root/reference/Examples/Forward_modelling"


def read_vtk_density(path, n):
    with open(path) as f:
        lines = f.read().splitlines()
    k = next(i for i, l in enumerate(lines) if l.startswith("LOOKUP_TABLE"))
    vals = []
    for l in lines[k + 1:]:
        vals += [float(x) for x in l.split()]
        if len(vals) >= n:
            break
    return np.array(vals[:n])


def read_boxes(path):
    rows = []
    for l in open(path):
        l = l.split("#")[0].strip()
        if l:
            rows.append([float(x) for x in l.split()[:7]])
    return np.array(rows)


def reference_forward_sets():
    bit = {"g_z": 1, "T_zz": 4, "T_xz": 5, "T_yz": 6, "T_xx": 7, "T_xy": 8, "T_yy": 9}
    sets = {
        # test_make_model/build.sh:1 anomalies, test_compute/run.sh
        "model1": dict(path="tests/test_compute/result_no_noise",
                       boxes=np.array([[600, 900, 550, 850, 200, 500, 500],
                                       [1200, 1500, 550, 850, 200, 500, -500],
                                       [900, 1500, 1300, 1500, 200, 500, 500]], float),
                       mesh=(0, 2000, 40, 0, 2000, 40, 0, 1000, 20),
                       vtk="tests/test_make_model/test_model1.vtk"),
        "model2": dict(path="model2/result/data_no_noise", boxes=read_boxes(REF + "/model2/anomalies.txt"),
                       mesh=(0, 6000, 60, 0, 6000, 60, 0, 3500, 35), vtk=None),
        "model3": dict(path="model3/result/data_no_noise", boxes=read_boxes(REF + "/model3/anomalies.txt"),
                       mesh=(0, 2000, 40, 0, 2000, 40, 0, 1000, 20),
                       vtk="model3/model/test_model.vtk"),
    }
    for name, s in sets.items():
        path = os.path.join(REF, s["path"])
        header = open(path).readline()
        cols = [c for c in header.replace("#", "").split() if c not in ("x", "y", "z", "(m)")]
        names = [c for c in cols if not c.startswith("(")]
        data = np.loadtxt(path)
        flag = 0
        order = []
        for nme in names:
            flag |= 1 << bit[nme]
            order.append(bit[nme])
        x0, x1, nx, y0, y1, ny, z0, z1, nz = s["mesh"]
        out = dict(points=data[:, :3], values=data[:, 3:], bits=np.array(order), flag=flag,
                   boxes=s["boxes"], mesh=np.array(s["mesh"], float))
        if s["vtk"]:
            dens = read_vtk_density(os.path.join(REF, s["vtk"]), nx * ny * nz)
            prisms = M.regular_mesh((x0, x1), nx, (y0, y1), ny, (z0, z1), nz)
            mine = M.model_from_boxes(prisms, s["boxes"])
            assert np.array_equal(mine, dens), f"{name}: box model differs from the .vtk density"
            nzi = np.nonzero(dens)[0]
            out["vtk_nonzero_idx"] = nzi
            out["vtk_nonzero_val"] = dens[nzi]
        np.savez_compressed(os.path.join(HERE, f"ref_forward_{name}.npz"), **out)
        print(name, data.shape, hex(flag))


def pair_cases(rng, n):
    pts, prs = [], []
    while len(pts) < n:
        k = len(pts)
        pr = np.sort(rng.uniform(-500, 1500, (3, 2)), axis=1).ravel()
        if k % 4 == 0:
            pr = np.round(pr / 50) * 50
        if np.any(pr[1::2] <= pr[0::2]):
            continue
        pt = rng.uniform(-1000, 2500, 3)
        if k % 3 == 0:
            pt[0] = pr[rng.integers(0, 2)]
        if k % 5 == 0:
            pt[1] = pr[2 + rng.integers(0, 2)]
        if k % 7 == 0:
            pt[2] = pr[4 + rng.integers(0, 2)] - rng.uniform(0, 100)
        pts.append(pt)
        prs.append(pr)
    return np.array(pts), np.array(prs)


def ref_pairs():
    assert oracle.ref_available(), "oracle/_ref not built (make -C oracle ref)"
    rng = np.random.default_rng(7)
    pts, prs = pair_cases(rng, 1500)
    vals = np.zeros((len(pts), 10))
    status = np.zeros(len(pts), int)
    for q in range(len(pts)):
        v, _ = oracle.prism_field(pts[q], prs[q], 1.0, 0x3FE, use_ref=True)
        vals[q] = v
        _, st = oracle.prism_field(pts[q], prs[q], 1.0, 0x3FE)
        status[q] = st
    keep = status == 0      # degenerate pairs: reference asserts, values unspecified
    np.savez_compressed(os.path.join(HERE, "ref_pairs.npz"), points=pts[keep], prisms=prs[keep],
                        values=vals[keep])
    print("pairs", keep.sum(), "kept of", len(pts))


def ref_G():
    prisms = M.regular_mesh((0.0, 300.0), 6, (0.0, 250.0), 5, (0.0, 200.0), 4)
    pts = M.grid_points(7, -25.0, 325.0, 6, 0.0, 250.0, -10.0)    # y on cell faces
    pts = np.vstack([pts, [[150.0, 100.0, -0.5], [75.0, 125.0, -300.0], [1e4, -3e3, -50.0]]])
    G, st = oracle.assemble(prisms, pts, 0x3FE, use_ref=True)
    np.savez_compressed(os.path.join(HERE, "ref_G_small.npz"), prisms=prisms, points=pts, G=G,
                        flag=0x3FE)
    print("G", G.shape, st)


if __name__ == "__main__":
    oracle.build(with_ref=True)
    reference_forward_sets()
    ref_pairs()
    ref_G()
