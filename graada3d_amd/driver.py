"""The GraAda3D inversion driver on the engine: the reference's four-file
configuration, its data and model files, and the adaptive Gauss-Newton loop.

    python -m graada3d_amd.driver config          # like the GraAda3D binary

* Configuration (src/GraAda3D.cpp): ``config`` names the data, model and
  inversion parameter files (one per valid line); a valid line is a line with
  ',', ';' and tabs turned into blanks, everything from '#' removed and
  surrounding blanks stripped, non-empty (GraAdaInv::line_process /
  next_valid_line, :183-229).  ``config_data`` (:230-342): data file, number
  of components, their labels (0 V, 1 gz, 2 gx, 3 gy, 4 Tzz, 5 Tzx, 6 Tzy,
  7 Txx, 8 Txy, 9 Tyy), then one "relative absolute" error line per label;
  the data file holds "x y z d_1 .. d_n" per valid line and the components
  are reordered by ascending label (:344-389).  ``config_model``
  (:391-478): x/y/z extents and cell counts, then padding counts and factors.
  ``config_inversion`` (:480-683): method, wavelet threshold, a_s a_z a_x a_y
  a_crg, p_s p_z p_x p_y p_t epsilon2, depth-weighting exponent, lambda_max
  n_lambda decreasing rate, target misfit, stagnation factor, CG tolerance and
  iteration factor, Gauss-Newton iterations, m_min m_max m_ini, maximum
  refinements, minimum cell sizes, refinement interval, refined proportion,
  record flag, cross-gradient flag and model, petrophysical flag and model,
  output name.
* Inversion (GaussNewtonInversion::invert_with_own_CG,
  GaussNewtonInversion.cpp:657-1308, with AdaptiveInversion::refine_mesh,
  AdaptiveInversion.cpp:637-727): per Gauss-Newton iteration a sequence of
  trial lambdas, each a solve_cg, a log-barrier model update and a misfit;
  the best trial is kept; stagnation and misfit-increase tests end the
  trials; IRLS weights are updated from the best model between trials; the
  mesh is refined every `interval` iterations (or on stagnation) at the cells
  with the largest jump indicator, and G grows by expand_G on the GPU.
* Data weighting: Wd_i = 1 / (relative_k |d_i| + absolute_k)
  (InversionBase::set_dobs); depth weighting |mean(z_obs) - z_c|^(-beta/2)
  (set_depth_weighting flag 0).

* Reference and structural models (GaussNewtonInversion.cpp:657-692,
  :752-766, :1131-1150, :1262-1281; AdaptiveInversion.cpp:206-536): a
  petrophysical model file sets m0 at the cell centres by multilinear
  interpolation (graada3d_amd.gridded) and the start model to m0 moved 0.01
  inside (m_min, m_max); a cross-gradient model file sets m0_s the same way,
  and with it the T_x / T_y / T_z operators (InversionBase::set_Tmatrix), the
  weight a_crg, S_crg = I, and the IRLS weights WL_t from T m when p_t != 2.
  After a refinement the children take interpolated m0 / m0_s; with both
  files a cell at the refinement threshold (within 1e-6) is refined too.

Supported subset: methods 1 (own CG on a stored G) and 2 (own CG on the
wavelet-compressed G), with or without padding cells (the level-0 mesh is
then the tensor product of the padded node coordinates), with or without
reference / cross-gradient model files; method 0 raises ValueError.
"""
from __future__ import annotations

import argparse
import os
import sys
import time
from dataclasses import dataclass, field

import numpy as np

LABEL_BIT = {k: 1 << k for k in range(10)}
LABEL_NAME = ["V", "g_z", "g_x", "g_y", "T_zz", "T_xz", "T_yz", "T_xx", "T_xy", "T_yy"]


def line_process(line: str) -> str:
    """GraAdaInv::line_process (GraAda3D.cpp:183-211)."""
    for ch in "\t,;\r\n":
        line = line.replace(ch, " ")
    cut = line.find("#")
    if cut >= 0:
        line = line[:cut]
    return line.strip(" ")


def valid_lines(path: str):
    with open(path) as f:
        for raw in f:
            line = line_process(raw)
            if line:
                yield line


@dataclass
class DataConfig:
    data_file: str
    labels: list
    rel_err: list              # in ascending-label order
    abs_err: list
    flag: int = 0


@dataclass
class ModelConfig:
    x: tuple
    y: tuple
    z: tuple
    n: tuple
    padding: tuple = (0, 0, 0, 0, 0, 0)


@dataclass
class InversionConfig:
    method: int
    threshold: float
    a: tuple                   # a_s, a_z, a_x, a_y, a_crg
    p: tuple                   # p_s, p_z, p_x, p_y, p_t
    eps2: float
    depth_exponent: float
    lambda_max: float
    n_lambda: int
    lambda_rate: float
    target_misfit: float
    stag_tol: float
    cg_tol: float
    cg_factor: float
    gn_iterations: int
    m_min: float
    m_max: float
    m_ini: float
    max_refinements: int
    min_size: tuple
    refine_interval: int
    refine_fraction: float
    record: int
    use_crg: int
    crg_model: list = field(default_factory=list)
    use_petro: int = 0
    petro_model: list = field(default_factory=list)
    output: str = "result"


def read_config(path: str):
    """(DataConfig, ModelConfig, InversionConfig) of a GraAda3D configuration;
    relative file names resolve against the configuration's directory."""
    base = os.path.dirname(os.path.abspath(path))

    def res(p):
        return p if os.path.isabs(p) else os.path.normpath(os.path.join(base, p))

    names = [ln.split()[0] for ln in valid_lines(path)][:3]
    if len(names) < 3:
        raise ValueError(f"{path}: needs the data, model and inversion parameter files")
    d = list(valid_lines(res(names[0])))
    nf = int(d[1].split()[0])
    labels = [int(v) for v in d[2].split()[:nf]]
    errs = [tuple(float(v) for v in d[3 + k].split()[:2]) for k in range(nf)]
    order = sorted(range(nf), key=lambda k: labels[k])
    flag = 0
    for lab in labels:
        flag |= LABEL_BIT[lab]
    data = DataConfig(res(d[0].split()[0]), labels, [errs[k][0] for k in order],
                      [errs[k][1] for k in order], flag)
    m = list(valid_lines(res(names[1])))
    ext = [m[k].split() for k in range(3)]
    pad = tuple(float(m[3 + k].split()[0]) for k in range(6))
    model = ModelConfig((float(ext[0][0]), float(ext[0][1])), (float(ext[1][0]), float(ext[1][1])),
                        (float(ext[2][0]), float(ext[2][1])),
                        (int(ext[0][2]), int(ext[1][2]), int(ext[2][2])), pad)
    v = [ln.split() for ln in valid_lines(res(names[2]))]
    f = [float(x) for x in v[2][:5]]
    p = [float(x) for x in v[3][:6]]
    lam = v[5]
    cg = v[8]
    mm = v[10]
    inv = InversionConfig(
        method=int(v[0][0]), threshold=float(v[1][0]), a=tuple(f), p=tuple(int(x) for x in p[:5]),
        eps2=p[5], depth_exponent=float(v[4][0]), lambda_max=float(lam[0]), n_lambda=int(lam[1]),
        lambda_rate=float(lam[2]), target_misfit=float(v[6][0]), stag_tol=float(v[7][0]),
        cg_tol=float(cg[0]), cg_factor=abs(float(cg[1])), gn_iterations=int(v[9][0]),
        m_min=float(mm[0]), m_max=float(mm[1]), m_ini=float(mm[2]),
        max_refinements=int(v[11][0]), min_size=tuple(float(x) for x in v[12][:3]),
        refine_interval=int(v[13][0]), refine_fraction=float(v[14][0]), record=int(v[15][0]),
        use_crg=int(v[16][0]), crg_model=_model_line(v[17], res),
        use_petro=int(v[18][0]), petro_model=_model_line(v[19], res),
        output=v[20][0] if len(v) > 20 else "result")
    return data, model, inv


def _model_line(tokens, res):
    """A model line with its file name resolved like the other files."""
    return [res(tokens[0])] + list(tokens[1:]) if tokens and tokens[0] != "none" else list(tokens)


def read_data(cfg: DataConfig):
    """Points (N_obs x 3) and dobs (component-blocked, ascending label) of the
    data file (GraAdaInv::read_data_from_file, GraAda3D.cpp:344-389)."""
    rows = [ln.split() for ln in valid_lines(cfg.data_file)]
    nf = len(cfg.labels)
    arr = np.array([[float(x) for x in r[:3 + nf]] for r in rows])
    order = sorted(range(nf), key=lambda k: cfg.labels[k])
    dobs = np.concatenate([arr[:, 3 + k] for k in order])
    return arr[:, :3].copy(), dobs


def data_weights(dobs, n_obs, rel, abs_):
    """Wd_i = 1 / (relative_k |d_i| + absolute_k) per component block."""
    w = np.empty_like(dobs)
    for k in range(len(rel)):
        blk = slice(k * n_obs, (k + 1) * n_obs)
        w[blk] = 1.0 / (rel[k] * np.abs(dobs[blk]) + abs_[k])
    return w


def build_mesh(model: ModelConfig):
    """The initial inversion mesh (GraAda3D.cpp:456-475): regular, or with
    padding cells around the region of interest when any padding count is
    non-zero (Mesh::generate_regular_mesh_with_padding, Mesh.cpp:516-590)."""
    from .mesh import padded_nodes
    from .octree import OctreeMesh
    npx, fx, npy, fy, npz, fz = model.padding
    if int(npx) == 0 and int(npy) == 0 and int(npz) == 0:
        return OctreeMesh(model.x, model.n[0], model.y, model.n[1], model.z, model.n[2])
    nodes = (padded_nodes(model.x, model.n[0], int(npx), fx),
             padded_nodes(model.y, model.n[1], int(npy), fy),
             padded_nodes(model.z, model.n[2], int(npz), fz, both_sides=False))
    return OctreeMesh(None, 0, None, 0, None, 0, nodes=nodes)


class AdaptiveInversion:
    """invert_with_own_CG with adaptive refinement on the engine.  Logs go to
    `out_dir` as lambda_misfit_GN and Iteration_misfit_GN."""

    def __init__(self, data: DataConfig, model: ModelConfig, inv: InversionConfig, ctx=None,
                 out_dir: str = ".", engine_factory=None):
        from . import operators as O
        from .octree import OctreeMesh
        if inv.method not in (1, 2):
            raise ValueError("method must be 1 (stored G) or 2 (wavelet-compressed G); "
                             "method 0 (Eigen's LSCG) is not supported")
        from .gridded import model_from_config
        self.interp_m0 = model_from_config(inv.petro_model) if inv.use_petro else None
        self.interp_m0s = model_from_config(inv.crg_model) if inv.use_crg else None
        self.O = O
        self.cfg, self.dcfg, self.model_cfg = inv, data, model
        self.out_dir = out_dir
        self.points, self.dobs = read_data(data)
        self.n_obs = self.points.shape[0]
        self.Wd = data_weights(self.dobs, self.n_obs, data.rel_err, data.abs_err)
        self.mesh = build_mesh(model)
        if engine_factory is None:
            from .engine import Context, GaussNewtonInversion
            ctx = ctx or Context(0)

            def engine_factory(prisms, points, flag):
                return GaussNewtonInversion(prisms, points, flag, ctx=ctx)
        self.engine = engine_factory(self.mesh.prisms(), self.points, data.flag)
        self._assemble()
        nm = self.mesh.n_cells
        self.m_min = np.full(nm, inv.m_min)
        self.m_max = np.full(nm, inv.m_max)
        prisms = self.mesh.prisms()
        self.m0 = self.interp_m0.at_centres(prisms) if inv.use_petro else np.zeros(nm)
        self.m0_s = self.interp_m0s.at_centres(prisms) if inv.use_crg else None
        # the start model is m0 when a reference model is given (GN.cpp:752-766)
        self.m = self.m0.copy() if inv.use_petro else np.full(nm, inv.m_ini)
        if inv.use_crg:
            self.engine.use_cross_gradient_constraint = True
            self.engine.a_crg = inv.a[4]
        self._build_operators()

    def _assemble(self):
        """compute_G, or compute_G_wavelet over the reordered cells for
        method 2 (GaussNewtonInversion::invert, GaussNewtonInversion.cpp:2-29)."""
        if self.cfg.method == 2:
            self.wavelet = self.engine.compute_G_wavelet(self.cfg.threshold,
                                                         self.mesh.reordered())
        else:
            self.engine.compute_G("f64")

    # -- operators (InversionBase::init_matrices; GN.cpp:711-718) ------------
    def _build_operators(self):
        O, c = self.O, self.cfg
        prisms = self.mesh.prisms()
        D = self.mesh.difference_matrices()
        Z = O.depth_weights_reference(prisms, self.points, c.depth_exponent)
        a_s, a_z, a_x, a_y = c.a[:4]
        self.W = O.weighting_operators_from(prisms, D, Z, a_s=a_s, a_x=a_x, a_y=a_y, a_z=a_z)
        self.T = ()
        if c.use_crg:
            # InversionBase::set_Tmatrix; S_crg = I (InversionBase.cpp:113-130)
            self.T = O.cross_gradient_operators_from(prisms, D, self.m0_s)
            self.engine.S_crg = np.ones(prisms.shape[0])
        self.engine.set_regularization(*self.W, *self.T)

    def _irls(self, m_opt):
        """WL_s from Ws (m - m0), WL_x/y/z from Wx/Wy/Wz m where the exponent
        is not 2 (GN.cpp:962-1022)."""
        c = self.cfg
        p_s, p_z, p_x, p_y = c.p[:4]
        Ws, Wx, Wy, Wz = self.W
        WL = {}
        e = self.engine
        if p_s != 2:
            WL["s"] = e.irls_weights(Ws, m_opt - self.m0, p_s, c.eps2)
        if p_x != 2:
            WL["x"] = e.irls_weights(Wx, m_opt, p_x, c.eps2)
        if p_y != 2:
            WL["y"] = e.irls_weights(Wy, m_opt, p_y, c.eps2)
        if p_z != 2:
            WL["z"] = e.irls_weights(Wz, m_opt, p_z, c.eps2)
        if c.use_crg and c.p[4] != 2:
            # GN.cpp:1010-1021, :1266-1280
            for k, T in zip(("tx", "ty", "tz"), self.T):
                WL[k] = e.irls_weights(T, m_opt, c.p[4], c.eps2)
        return WL

    def _refine(self):
        """AdaptiveInversion::refine_mesh (AdaptiveInversion.cpp:637-727; with
        model files :206-388 and :389-536)."""
        c = self.cfg
        nm = self.mesh.n_cells
        a = c.refine_fraction
        if a < 0 or abs(a * nm) < 1:
            return 0
        ind = self.mesh.indicator(self.m)
        tol = np.sort(ind)[::-1][int(a * nm)]
        p = self.mesh.prisms()
        d = p[:, 1::2] - p[:, 0::2]
        small = np.zeros(nm, bool)
        for k in range(3):
            small |= (d[:, k] < c.min_size[k]) | (np.abs(d[:, k] - c.min_size[k]) < 1e-7)
        hit = ind > tol
        if c.use_crg and c.use_petro:
            hit |= np.abs(ind - tol) < 1e-6
        marked = np.flatnonzero(hit & ~small)
        if marked.size == 0:
            return 0
        split, kids = self.mesh.refine(marked)
        if self.cfg.method == 2:
            # the compressed G is built again over the new reordered cells
            # (AdaptiveInversion.cpp:690-700)
            self.engine.set_mesh(self.mesh.prisms())
            self._assemble()
        else:
            self.engine.expand_G(split, kids)
        # children inherit the parent's values (Mesh.cpp:1591-1597)
        reps = np.ones(nm, np.int64)
        reps[split] = 8
        self.m, self.m0, self.m_min, self.m_max = (np.repeat(v, reps) for v in
                                                   (self.m, self.m0, self.m_min, self.m_max))
        # children of a split cell take the model files' values at their centres
        kid_rows = np.flatnonzero(np.repeat(reps == 8, reps))
        kid_prisms = self.mesh.prisms()[kid_rows]
        if c.use_petro:
            self.m0[kid_rows] = self.interp_m0.at_centres(kid_prisms)
        if c.use_crg:
            self.m0_s = np.repeat(self.m0_s, reps)
            self.m0_s[kid_rows] = self.interp_m0s.at_centres(kid_prisms)
        self._build_operators()
        return int(split.size)

    def run(self, max_trials: int | None = None, verbose: bool = False):
        """The Gauss-Newton loop.  `max_trials` stops after that many lambda
        trials in total (a bounded sample of the run); `verbose` prints the
        reference's per-trial line."""
        c, e = self.cfg, self.engine
        self.trials = []
        lam_log = open(os.path.join(self.out_dir, "lambda_misfit_GN"), "w")
        it_log = open(os.path.join(self.out_dir, "Iteration_misfit_GN"), "w")
        lam_log.write(f"{'lambda':>25}{'misfit':>25}\n")
        it_log.write(f"{'Iteranation number':>25}{'number of tried regulartion parameters':>25}"
                     f"{'misfit':>25}\n")
        # start model rule (GN.cpp:752-780)
        hi = (self.m > self.m_max) | (np.abs(self.m - self.m_max) < 1e-7)
        if c.use_petro:
            self.m = np.where(hi, self.m_max - 0.01, self.m)
            lo = (self.m < self.m_min) | (np.abs(self.m - self.m_min) < 1e-7)
            self.m = np.where(lo, self.m_min + 0.01, self.m)
        else:
            self.m = np.where(hi, 0.5 * (self.m_max + self.m_min), self.m)
            lo = (self.m < self.m_min) | (np.abs(self.m - self.m_min) < 1e-7)
            self.m = np.where(lo, 0.5 * (self.m_max + self.m_min), self.m)
        misfit = e.misfit(self.m, self.Wd, self.dobs)
        lam = c.lambda_max
        lambda_opt, misfit_opt, m_opt = lam, misfit, self.m.copy()
        WL = {}
        c_ref, i_between = 0, 0
        history = []
        for i in range(c.gn_iterations):
            nm = self.mesh.n_cells
            m_trial = self.m.copy()
            misfit_last1, misfit_last2 = 10 * misfit, 100 * misfit
            misfit_last_iteration = misfit
            if i == 0 or c.n_lambda == 1 or abs(c.lambda_rate - 1) < 1e-6:
                lam = c.lambda_max
            else:
                lam = lambda_opt * (1.0 / c.lambda_rate) ** min(3, c.n_lambda - 2)
                if lam < min(c.lambda_max * c.lambda_rate ** 20, 1.0e-5):
                    lam = 1.0
            P = e.barrier_P(self.m, self.m_min, self.m_max)
            cg_maxit = int(c.cg_factor * nm) if c.cg_factor < 1 else int(c.cg_factor)
            j = 0
            while j < c.n_lambda:
                kw = {f"WL_{k}": v for k, v in WL.items()}
                dx, cg_err, cg_its = e.solve_cg(c.cg_tol, cg_maxit, lam, self.m, self.m0, self.Wd,
                                                self.dobs, P, quirks=True, **kw)
                m_trial = e.barrier_update(self.m, dx, self.m_min, self.m_max, m_trial)
                misfit_last2, misfit_last1 = misfit_last1, misfit
                misfit = e.misfit(m_trial, self.Wd, self.dobs)
                lam_log.write(f"{lam:25.14e}{misfit:25.14e}\n")
                self.trials.append(dict(lam=lam, misfit=misfit, cg_iterations=int(cg_its),
                                        cg_error=float(cg_err), Nm=nm,
                                        nnz=self.wavelet[1] if c.method == 2 else None))
                if verbose:
                    print(f"  #{j} lambda={lam:g}, misfit={misfit:g} (CG iterations: {cg_its}, "
                          f"error: {cg_err:g})", flush=True)
                if j == 0 or misfit < misfit_opt:
                    misfit_opt, m_opt = misfit, m_trial.copy()
                    if i > 0:
                        self.m = m_opt.copy()
                    lambda_opt = lam
                if max_trials is not None and len(self.trials) >= max_trials:
                    break
                if j > 1 and (abs(misfit - misfit_last1) / min(misfit, misfit_last1) < c.stag_tol
                              and abs(misfit_last1 - misfit_last2) /
                              min(misfit_last1, misfit_last2) < c.stag_tol):
                    break
                if misfit > misfit_last1 and misfit_last1 > misfit_last2 and j > 1:
                    break
                lam *= c.lambda_rate
                WL = self._irls(m_opt)
                if misfit < c.target_misfit:
                    break
                j += 1
            self.m, misfit = m_opt.copy(), misfit_opt
            tried = c.n_lambda if j == c.n_lambda else j + 1
            it_log.write(f"{i + 1:25d}{tried:25d}{misfit:25.14e}\n")
            history.append(dict(iteration=i + 1, lambdas=tried, misfit=misfit, Nm=nm))
            if max_trials is not None and len(self.trials) >= max_trials:
                break
            if misfit < c.target_misfit or i == c.gn_iterations - 1:
                break
            flag1 = abs(misfit - misfit_last_iteration) / min(misfit, misfit_last_iteration) < \
                c.stag_tol
            if c_ref >= c.max_refinements and flag1:
                break
            i_between += 1
            if c_ref < c.max_refinements and (i_between % c.refine_interval == 0 or flag1):
                i_between = 0
                c_ref += 1
                # W and the IRLS weights are rebuilt from m after every
                # refinement attempt (GN.cpp:1137-1262)
                self._refine()
                WL = self._irls(self.m)
                m_opt = self.m.copy()
        lam_log.close()
        it_log.close()
        self.misfit, self.history = misfit, history
        return dict(misfit=misfit, Nm=self.mesh.n_cells, iterations=len(history), history=history)

    def write_result(self, name: str, seconds: float | None = None):
        """<name>.txt (Mesh::out_model_txt, Mesh.cpp:1416-1460), dobs_* and
        dpredicted_* (InversionBase::out_data, InversionBase.cpp:743-783) and
        `info` (GraAda3D.cpp:759-782; the final element count is the number of
        leaf cells)."""
        p = self.mesh.prisms()
        c = 0.5 * (p[:, 0::2] + p[:, 1::2])
        with open(os.path.join(self.out_dir, name + ".txt"), "w") as f:
            f.write("#The first 6 parameters give the dimensions of a cell. The 7-9 parameters are "
                    "the cell center. The 10th parameter is the value within the cell.\n")
            f.write("".join(f"{h:<22}" for h in ("#x0(m)", "x1(m)", "#y0(m)", "y1(m)", "#z0(m)",
                                                 "z1(m)", "xc(m)", "yc(m)", "zc(m)", "value")) + "\n")
            for k in range(p.shape[0]):
                f.write("".join(f"{v:<22.14e}" for v in (*p[k], *c[k], self.m[k])) + "\n")
        dpre = self.engine.G_vec_mul(self.m)
        labels = sorted(self.dcfg.labels)
        for tag, d in (("dobs", self.dobs), ("dpredicted", dpre)):
            for j, lab in enumerate(labels):
                with open(os.path.join(self.out_dir, f"{tag}_{LABEL_NAME[lab]}"), "w") as f:
                    for i in range(self.n_obs):
                        x, y, z = self.points[i]
                        f.write(f"{x:<30.15e}{y:<30.15e}{z:<30.15e}{d[i + j * self.n_obs]:<30.15e}\n")
        m = self.model_cfg
        with open(os.path.join(self.out_dir, "info"), "w") as f:
            f.write("1. Initial inverion mesh\n")
            f.write(f"X range  {m.x[0]:g}, {m.x[1]:g}\nY range: {m.y[0]:g}, {m.y[1]:g}\n"
                    f"Z range {m.z[0]:g}, {m.z[1]:g}\n")
            f.write(f"Number of elements along z axis in the initial mesh {m.n[2]}\n"
                    f"Number of elements along x axis in the initial mesh {m.n[0]}\n"
                    f"Number of elements along y axis in the initial mesh {m.n[1]}\n"
                    f"Total number of elements in the initial mesh {m.n[0] * m.n[1] * m.n[2]}\n\n")
            f.write(f"2. Final mesh\nElement number {self.mesh.n_cells}\n\n")
            f.write(f"3. Evaluation of the inversion\nFinal misfit {self.misfit:g}\n")
            if seconds is not None:
                f.write(f"Running time {seconds:g} seconds\n")


def main(argv=None):
    ap = argparse.ArgumentParser(description="GraAda3D inversion on the MI355X engine")
    ap.add_argument("config")
    ap.add_argument("--out", default=".", help="directory for the logs and results")
    args = ap.parse_args(argv)
    data, model, inv = read_config(args.config)
    os.makedirs(args.out, exist_ok=True)
    t0 = time.perf_counter()
    run = AdaptiveInversion(data, model, inv, out_dir=args.out)
    res = run.run()
    run.write_result(inv.output, time.perf_counter() - t0)
    print(f"final misfit {res['misfit']:.6g}, {res['Nm']} cells, {res['iterations']} iterations")
    return 0


if __name__ == "__main__":
    sys.exit(main())
