"""Gridded model files (the reference / structural models of an inversion) and
their multilinear interpolation (graada3d_amd.gridded), and the driver with
both model files on the oracle.

Pins: every data order x loop order of the file format reads back the grid
that was written; the interpolation reproduces any trilinear polynomial
(a + bx + cy + dz + exy + fxz + gyz + hxyz) on a non-uniform grid, returns the
node values at the nodes, and outside the grid equals its value at the
nearest point of the grid's box (constant extension); a wrong axis order,
weight or extension rule fails one of these."""
import itertools

import numpy as np
import pytest
from dataclasses import replace

from graada3d_amd import driver
from graada3d_amd.gridded import (MultilinearInterpolator, model_from_config,
                                  read_gridded_model)
from helpers import OracleGN, stage_karst

PERMS = ["".join(p) for p in itertools.permutations("xyz")]


def _grid(rng, n):
    return [np.sort(rng.uniform(lo, hi, k)) for (lo, hi), k in
            zip(((0, 100), (-50, 20), (300, 900)), n)]


def _write(path, g, f, order, fast):
    """Records of the file format: the first axis of `fast` innermost."""
    ax = ["xyz".index(ch) for ch in fast]
    lines = []
    idx = [0, 0, 0]
    for s in range(len(g[ax[2]])):
        for m in range(len(g[ax[1]])):
            for q in range(len(g[ax[0]])):
                idx[ax[2]], idx[ax[1]], idx[ax[0]] = s, m, q
                xyz = [g[a][idx[a]] for a in range(3)]
                cols = [xyz["xyz".index(ch)] for ch in order]
                lines.append(" ".join(f"{v:.17g}" for v in cols + [f[tuple(idx)]]))
    path.write_text("\n".join(lines) + "\n")


@pytest.mark.parametrize("order", PERMS)
@pytest.mark.parametrize("fast", PERMS)
def test_every_order_reads_back(tmp_path, order, fast):
    rng = np.random.default_rng(hash((order, fast)) % 2**32)
    n = (4, 3, 5)
    g = _grid(rng, n)
    f = rng.standard_normal(n)
    p = tmp_path / "model.xyz"
    _write(p, g, f, order, fast)
    gx, gy, gz, fr = read_gridded_model(str(p), n, order, fast)
    for a, b in zip((gx, gy, gz), g):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(fr, f)


def test_config_line_gives_sizes_in_data_order(tmp_path):
    rng = np.random.default_rng(3)
    n = (3, 4, 2)
    g = _grid(rng, n)
    f = rng.standard_normal(n)
    _write(tmp_path / "m.xyz", g, f, "zyx", "yxz")
    it = model_from_config(["m.xyz", "zyx", "2", "4", "3", "yxz"], str(tmp_path))
    np.testing.assert_array_equal(it.f, f)
    with pytest.raises(ValueError):
        model_from_config(["m.xyz", "zyx", "2", "4", "4", "yxz"], str(tmp_path))   # too few
    with pytest.raises(ValueError):
        model_from_config(["m.xyz", "zzx", "2", "4", "3", "yxz"], str(tmp_path))
    with pytest.raises(ValueError):
        model_from_config(["m.xyz", "zyx", "2", "4", "3", "01"], str(tmp_path))


def test_trilinear_polynomials_are_reproduced():
    rng = np.random.default_rng(7)
    g = _grid(rng, (6, 5, 7))
    c = rng.standard_normal(8)

    def poly(x, y, z):
        return (c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * y + c[5] * x * z +
                c[6] * y * z + c[7] * x * y * z)

    X, Y, Z = np.meshgrid(*g, indexing="ij")
    it = MultilinearInterpolator(*g, poly(X, Y, Z))
    x = rng.uniform(g[0][0], g[0][-1], 500)
    y = rng.uniform(g[1][0], g[1][-1], 500)
    z = rng.uniform(g[2][0], g[2][-1], 500)
    ref = poly(x, y, z)
    assert np.abs(it(x, y, z) - ref).max() <= 1e-12 * np.abs(ref).max()
    # nodes, the last node included
    np.testing.assert_array_equal(it(X.ravel(), Y.ravel(), Z.ravel()), poly(X, Y, Z).ravel())


def test_outside_takes_the_nearest_box_point():
    rng = np.random.default_rng(8)
    g = _grid(rng, (4, 6, 3))
    it = MultilinearInterpolator(*g, rng.standard_normal((4, 6, 3)))
    x = rng.uniform(-100, 200, 400)
    y = rng.uniform(-150, 120, 400)
    z = rng.uniform(0, 1200, 400)
    cl = [np.clip(v, a[0], a[-1]) for v, a in zip((x, y, z), g)]
    np.testing.assert_array_equal(it(x, y, z), it(*cl))
    assert ((x < g[0][0]) | (x > g[0][-1])).any()


def test_single_node_axis():
    it = MultilinearInterpolator([1.0, 2.0], [5.0], [0.0, 1.0], np.array([[[1.0, 3.0]], [[5.0, 7.0]]]))
    np.testing.assert_allclose(it([1.5, 0.0], [-9.0, 9.0], [0.5, 2.0]), [4.0, 3.0])


def model_files(tmp_path, mesh_cfg):
    """A reference density model and a structural model on grids (node
    spacing not uniform) that cover the mesh; the structural grid extends
    past the mesh only partly, so some cell centres fall outside it."""
    rng = np.random.default_rng(11)
    (x0, x1), (y0, y1), (z0, z1) = mesh_cfg.x, mesh_cfg.y, mesh_cfg.z
    g = [np.sort(np.r_[lo, hi, rng.uniform(lo, hi, k - 2)]) for (lo, hi), k in
         (((x0, x1), 9), ((y0, y1), 8), ((z0, z1), 6))]
    X, Y, Z = np.meshgrid(*g, indexing="ij")
    xc, yc, zc = (x0 + x1) / 2, (y0 + y1) / 2, z0 + 0.4 * (z1 - z0)
    rho = -900.0 * np.exp(-(((X - xc) / 60) ** 2 + ((Y - yc) / 50) ** 2 + ((Z - zc) / 40) ** 2))
    _write(tmp_path / "ref.xyz", g, rho, "xyz", "zyx")
    gs = [g[0][1:], g[1], g[2][:-1]]
    Xs, Ys, Zs = np.meshgrid(*gs, indexing="ij")
    sus = np.tanh((Xs - xc) / 40) + 0.3 * np.sin(Zs / 30)
    _write(tmp_path / "crg.xyz", gs, sus, "zyx", "xyz")
    n, ns = [len(v) for v in g], [len(v) for v in gs]
    petro = [str(tmp_path / "ref.xyz"), "xyz", *map(str, n), "zyx"]
    crg = [str(tmp_path / "crg.xyz"), "zyx", *map(str, ns[::-1]), "xyz"]
    return petro, crg, MultilinearInterpolator(*g, rho), MultilinearInterpolator(*gs, sus)


def model_file_config(tmp_path):
    """The cave example's configuration with both model files, a_crg = 1e-3,
    p_t = 1, three lambdas per Gauss-Newton iteration and a refinement after
    each iteration."""
    cfg, _ = stage_karst("raw", tmp_path)
    d, m, inv = driver.read_config(cfg)
    petro, crg, it_m0, it_m0s = model_files(tmp_path, m)
    inv = replace(inv, use_petro=1, petro_model=petro, use_crg=1, crg_model=crg,
                  a=inv.a[:4] + (1e-3,), p=inv.p[:4] + (1,), n_lambda=3, gn_iterations=3,
                  refine_interval=1)
    return d, m, inv, it_m0, it_m0s


def test_driver_with_model_files_on_the_oracle(tmp_path):
    d, m, inv, it_m0, it_m0s = model_file_config(tmp_path)
    out = tmp_path / "out"
    out.mkdir()
    run = driver.AdaptiveInversion(d, m, inv, out_dir=str(out), engine_factory=OracleGN)
    p = run.mesh.prisms()
    np.testing.assert_array_equal(run.m0, it_m0.at_centres(p))
    np.testing.assert_array_equal(run.m0_s, it_m0s.at_centres(p))
    assert run.engine.use_cross_gradient_constraint and run.engine.a_crg == 1e-3
    assert len(run.T) == 3 and run.T[0].shape == (run.mesh.n_cells,) * 2
    nm0 = run.mesh.n_cells
    res = run.run()
    # the start model was m0 moved inside (m_min, m_max); it was refined
    assert res["Nm"] > nm0 and np.isfinite(res["misfit"])
    p = run.mesh.prisms()
    lv = run.mesh.levels()
    kids = lv > 0
    assert kids.any()
    np.testing.assert_array_equal(run.m0[kids], it_m0.at_centres(p[kids]))
    np.testing.assert_array_equal(run.m0_s[kids], it_m0s.at_centres(p[kids]))
    np.testing.assert_array_equal(run.m0_s[~kids], it_m0s.at_centres(p[~kids]))
    assert len(run.m0) == len(run.m0_s) == res["Nm"]


def test_petrophysical_start_model(tmp_path):
    """GN.cpp:752-766: the start model is m0 with values at or beyond a bound
    put 0.01 inside it (the first CG linearises about it)."""
    d, m, inv, it_m0, _ = model_file_config(tmp_path)
    inv = replace(inv, use_crg=0, m_min=-200.0, m_max=100.0)
    seen = []

    class Recorder(OracleGN):
        def solve_cg(self, tol, maxit, lam, mk, *a, **k):
            seen.append(np.array(mk))
            return super().solve_cg(tol, maxit, lam, mk, *a, **k)

    run = driver.AdaptiveInversion(d, m, inv, out_dir=str(tmp_path), engine_factory=Recorder)
    m0 = run.m0.copy()
    assert (m0 < -200).any() and not run.engine.__dict__.get("use_cross_gradient_constraint")
    run.run(max_trials=1)
    start = np.where(m0 > 100 - 1e-7, 100 - 0.01, m0)
    start = np.where(start < -200 + 1e-7, -200 + 0.01, start)
    np.testing.assert_array_equal(seen[0], start)
