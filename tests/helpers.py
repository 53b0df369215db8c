"""Shared helpers for the tests (fixture loading, error measures)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


def rel_norm(a, b):
    """||a - b||_2 / ||b||_2 (b the reference)."""
    a = np.asarray(a, float).ravel()
    b = np.asarray(b, float).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def max_rel_to_scale(a, b):
    """max|a - b| / max|b|."""
    b = np.asarray(b, float)
    s = np.abs(b).max()
    return np.abs(np.asarray(a, float) - b).max() / (s if s > 0 else 1.0)


def model_from_fixture(fx):
    from graada3d_amd import mesh as M
    x0, x1, nx, y0, y1, ny, z0, z1, nz = fx["mesh"]
    prisms = M.regular_mesh((x0, x1), int(nx), (y0, y1), int(ny), (z0, z1), int(nz))
    rho = M.model_from_boxes(prisms, fx["boxes"])
    return prisms, rho


class OracleGN:
    """The oracle behind the interface of engine.GaussNewtonInversion
    (compute_G, expand_G, G_vec_mul, set_regularization, solve_cg, barrier_P,
    barrier_update, misfit, irls_weights), so the Gauss-Newton driver can run
    on the CPU as a checker."""

    def __init__(self, prisms, points, field_flag):
        self.prisms = np.asarray(prisms, float).copy()
        self.points = np.asarray(points, float).copy()
        self.flag = int(field_flag)
        self.G = None
        self.W = None

    @property
    def Nm(self):
        return self.prisms.shape[0]

    def set_mesh(self, prisms):
        self.prisms = np.asarray(prisms, float).copy()
        self.G = None

    def compute_G_wavelet(self, relative_threshold, order):
        """The oracle's compressed G; the products and the CG then use the
        operator it represents as a dense matrix."""
        import oracle
        self.wav = oracle.WaveletG(self.prisms, self.points, self.flag, order, relative_threshold)
        assert self.wav.status == 0
        self.G = self.wav.dense()
        return self.wav.n_dy, self.wav.nnz

    def compute_G(self, store="f64"):
        import oracle
        self.G, st = oracle.assemble(self.prisms, self.points, self.flag)
        assert st == 0, st

    def expand_G(self, split_cols, child_prisms):
        import oracle
        parts_p, parts_g = [], []
        kids = np.asarray(child_prisms, float).reshape(-1, 8, 6)
        Gk, st = oracle.assemble(kids.reshape(-1, 6), self.points, self.flag)
        assert st == 0, st
        pos = {int(c): k for k, c in enumerate(split_cols)}
        for j in range(self.Nm):
            if j in pos:
                k = pos[j]
                parts_p.append(kids[k])
                parts_g.append(Gk[:, 8 * k:8 * k + 8])
            else:
                parts_p.append(self.prisms[j:j + 1])
                parts_g.append(self.G[:, j:j + 1])
        self.prisms = np.concatenate(parts_p)
        self.G = np.ascontiguousarray(np.concatenate(parts_g, axis=1))

    def G_vec_mul(self, m):
        import oracle
        return oracle.gemv(self.G, m)

    def set_regularization(self, Ws, Wx, Wy, Wz, Tx=None, Ty=None, Tz=None):
        self.W = (Ws, Wx, Wy, Wz)
        self.T = None if Tx is None else (Tx, Ty, Tz)

    def solve_cg(self, tol, maxit, lam, mk, mref, Wd, dobs, P, WL_s=None, WL_x=None, WL_y=None,
                 WL_z=None, WL_tx=None, WL_ty=None, WL_tz=None, quirks=True, **_):
        import oracle
        WL = {k: v for k, v in dict(s=WL_s, x=WL_x, y=WL_y, z=WL_z, tx=WL_tx, ty=WL_ty,
                                    tz=WL_tz).items() if v is not None}
        crg = None
        if getattr(self, "use_cross_gradient_constraint", False):
            crg = dict(Tx=self.T[0], Ty=self.T[1], Tz=self.T[2], a=self.a_crg, S=self.S_crg)
        return oracle.solve_cg(self.G, tol, maxit, lam, mk, mref, Wd, dobs, P, *self.W, WL=WL,
                               crg=crg, quirks=quirks)

    def barrier_P(self, m, m_min, m_max):
        import oracle
        return oracle.barrier_P(m, m_min, m_max)

    def barrier_update(self, m, dx, m_min, m_max, m_trial):
        import oracle
        return oracle.barrier_update(m, dx, m_min, m_max, m_trial)

    def misfit(self, m, Wd, dobs):
        import oracle
        return oracle.misfit(self.G, m, Wd, dobs)

    def irls_weights(self, W, m, p, eps2):
        import oracle
        return oracle.irls_weights(W, m, p, eps2)


KARST_FOLDERS = {"raw": "invert_raw_data_pivetta2015",
                 "interpolated": "invert_interpolated_data_Braitenberg2015"}


def stage_karst(key, root):
    """Lay out one reference cave-inversion example under `root` as the
    reference ships it (configuration files in <folder>/, the data file in
    data_preprocessing/); returns (config path, fixture dict)."""
    import shutil
    fx = load(os.path.join("karst", f"{key}.npz"))
    folder = KARST_FOLDERS[key]
    dst = os.path.join(str(root), folder)
    os.makedirs(dst, exist_ok=True)
    for name in ("config", "config_data", "config_model", "config_inversion"):
        shutil.copyfile(os.path.join(GOLDEN, "karst", folder, name), os.path.join(dst, name))
    os.makedirs(os.path.join(str(root), "data_preprocessing"), exist_ok=True)
    np.savetxt(os.path.join(str(root), "data_preprocessing", str(fx["data_name"])), fx["data"],
               fmt="%.17g")
    return os.path.join(dst, "config"), fx


def check_against_run(run, fx, out_dir, misfit_rtol=1e-3, final_rtol=1e-6, model_rtol=1e-5):
    """Compare a finished driver run with a reference run's logs and outputs."""
    lm = np.loadtxt(os.path.join(out_dir, "lambda_misfit_GN"), skiprows=1, ndmin=2)
    it = np.loadtxt(os.path.join(out_dir, "Iteration_misfit_GN"), skiprows=1, ndmin=2)
    assert lm.shape[0] == fx["lam"].shape[0], (lm.shape, fx["lam"].shape)
    np.testing.assert_allclose(lm[:, 0], fx["lam"], rtol=1e-13)
    np.testing.assert_allclose(lm[:, 1], fx["misfit"], rtol=misfit_rtol)
    np.testing.assert_array_equal(it[:, :2], fx["iters"][:, :2])
    np.testing.assert_allclose(it[:, 2], fx["iters"][:, 2], rtol=final_rtol)
    cells = run.mesh.prisms()
    assert cells.shape == fx["cells"].shape
    np.testing.assert_array_equal(cells, fx["cells"])
    assert rel_norm(run.m, fx["model"]) < model_rtol
    assert rel_norm(run.engine.G_vec_mul(run.m), fx["dpred"]) < final_rtol


# ---- 60-digit evaluation of a prism's field by its corner sum ---------------
# The mathematics that pins both the reference's face/edge values and the
# corner-sum (node) assembly where the two disagree: the prism integral of
# component `bit` equals c_bit * sum over the 8 corners of s * F_bit(corner - p)
# (s = +1 at (x1, y1, z1), alternating), evaluated here in decimal arithmetic
# at 60 digits with the same limits at x = 0 / y = 0 as the device code.

def _dec_pi():
    from decimal import Decimal
    return 16 * _dec_atan_small(Decimal(1) / 5) - 4 * _dec_atan_small(Decimal(1) / 239)


def _dec_atan_small(v):
    """atan(v) for |v| <= 1: halve the argument until |v| < 0.05
    (atan v = 2 atan(v / (1 + sqrt(1 + v^2)))), then the Taylor series."""
    from decimal import Decimal, getcontext
    n = 0
    while abs(v) > Decimal("0.05"):
        v = v / (1 + (1 + v * v).sqrt())
        n += 1
    eps = Decimal(10) ** (-getcontext().prec - 2)
    s, term, k, v2 = Decimal(0), v, 0, v * v
    while True:
        t = term / (2 * k + 1)
        if abs(t) < eps:
            break
        s += -t if k % 2 else t
        term *= v2
        k += 1
    return s * (2 ** n)


def _dec_atan_ratio(num, den, half_pi):
    """atan(num / den) with den = +0 giving sgn(num) * pi/2 (the IEEE limit of
    num / +0) and num = 0 giving 0."""
    from decimal import Decimal
    if num == 0:
        return Decimal(0)
    if den == 0:
        return half_pi if num > 0 else -half_pi
    q = num / den
    if abs(q) <= 1:
        return _dec_atan_small(q)
    return (half_pi if q > 0 else -half_pi) - _dec_atan_small(1 / q)


def corner_exact(prism, point, bit, digits=60):
    """Field component `bit` (1..9, the reference's field bits) of the unit
    density prism [x0, x1, y0, y1, z0, z1] at `point`, from its corner sum at
    `digits` decimal digits; the point must lie above the prism (z < z0)."""
    from decimal import Decimal, localcontext
    G0 = Decimal("6.673e-11")
    with localcontext() as ctx:
        ctx.prec = digits
        half_pi = _dec_pi() / 2
        tot = Decimal(0)
        for a in (0, 1):
            for b in (0, 1):
                for c in (0, 1):
                    x = Decimal(float(prism[a])) - Decimal(float(point[0]))
                    y = Decimal(float(prism[2 + b])) - Decimal(float(point[1]))
                    z = Decimal(float(prism[4 + c])) - Decimal(float(point[2]))
                    r = (x * x + y * y + z * z).sqrt()
                    axis = x == 0 and y == 0            # on the corner's vertical line

                    def lx(): return (x + r).ln()
                    def ly(): return (y + r).ln()
                    def lz(): return (z + r).ln()
                    def axy(): return _dec_atan_ratio(x * y, z * r, half_pi)
                    def ayz():
                        return _dec_atan_small(z / r) if axis else _dec_atan_ratio(y * z, x * r,
                                                                                  half_pi)
                    def axz():
                        return _dec_atan_small(z / r) if axis else _dec_atan_ratio(x * z, y * r,
                                                                                  half_pi)
                    F = {1: lambda: x * ly() + y * lx() - z * axy(),
                         2: lambda: y * lz() + z * ly() - x * ayz(),
                         3: lambda: z * lx() + x * lz() - y * axz(),
                         4: axy, 5: ly, 6: lx, 7: ayz, 8: lz, 9: axz}[bit]()
                    s = (1 if a else -1) * (1 if b else -1) * (1 if c else -1)
                    tot += s * F
        scale = {1: -G0 * Decimal(10) ** 5, 2: -G0 * Decimal(10) ** 5,
                 3: -G0 * Decimal(10) ** 5}.get(bit)
        if scale is None:
            scale = (-1 if bit in (4, 7, 9) else 1) * G0 * Decimal(10) ** 9
        return float(tot * scale)
