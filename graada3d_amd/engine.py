"""Host-side mirror of the reference's forward / inversion interfaces.

``Fwd`` mirrors ``class Fwd`` (src/Forward_modelling/Fwd.h:14-105):
``compute_G``, ``G_vec_mul``, ``GT_vec_mul``, ``get_G``, ``compute_gobs``.
``GaussNewtonInversion`` adds ``solve_cg`` with the argument list of
``GaussNewtonInversion::solve_cg`` (src/Inversion/GaussNewtonInversion.h:70-93)
and the barrier / misfit / IRLS helpers of ``invert_with_own_CG``.
``forward`` is the matrix-free loop of the ``compute`` tool
(src/compute.cpp:246-260).

Every numerical call goes through libgraeng's C ABI onto the GPU; there is no
CPU fallback.  Errors raise ``GraError`` (the reference asserts or aborts).
"""
from __future__ import annotations

import atexit
import ctypes
import sys
import weakref

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import GraError, GraCsr, GraCgArgs, GraGnSetup, c_dp, c_i32p, c_i64p

# Field bits (gs.h:46-57)
GZ, GX, GY, TZZ, TXZ, TYZ, TXX, TXY, TYY = (1 << 1, 1 << 2, 1 << 3, 1 << 4, 1 << 5, 1 << 6,
                                            1 << 7, 1 << 8, 1 << 9)
FIELD_NAMES = {1: "gz", 2: "gx", 3: "gy", 4: "Tzz", 5: "Txz", 6: "Tyz", 7: "Txx", 8: "Txy",
               9: "Tyy"}
NAME_TO_BIT = {v: k for k, v in FIELD_NAMES.items()}


def flag_from_names(names) -> int:
    """'gz/Tzz' or ['gz','Tzz'] -> field flag (aliases as compute.cpp:168-219)."""
    if isinstance(names, str):
        names = names.split("/")
    alias = {"gxx": "Txx", "gxy": "Txy", "gyx": "Txy", "Tyx": "Txy", "gxz": "Txz", "gzx": "Txz",
             "Tzx": "Txz", "gyy": "Tyy", "gyz": "Tyz", "gzy": "Tyz", "Tzy": "Tyz", "gzz": "Tzz"}
    f = 0
    for n in names:
        n = alias.get(n, n)
        if n not in NAME_TO_BIT:
            raise ValueError(f"invalid component name {n!r}")
        f |= 1 << NAME_TO_BIT[n]
    return f


def component_bits(flag: int):
    return [b for b in range(1, 10) if flag & (1 << b)]


def _f64(a, n=None, name="array"):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError(f"{name}: expected {n} values, got {a.size}")
    return a


def _p(a):
    return None if a is None else a.ctypes.data_as(c_dp)


# Contexts still open; closed by an atexit hook while the GPU runtime is
# intact (a context left to the garbage collector during interpreter teardown
# would call into the runtime after its own exit handlers have run).
_LIVE = weakref.WeakSet()


@atexit.register
def _close_all():
    for c in list(_LIVE):
        c.close()


class Context:
    """One engine context (device buffers + its own GPU stream)."""

    def __init__(self, device: int = 0):
        self.L = _lib.load()
        h = ctypes.c_void_p()
        st = self.L.gra_ctx_create(int(device), ctypes.byref(h))
        if st != 0:
            raise GraError(st, "gra_ctx_create", "(is a GPU visible?)")
        self.h = h
        _LIVE.add(self)

    def check(self, st: int, where: str):
        if st != 0:
            msg = self.L.gra_last_error(self.h).decode(errors="replace")
            raise GraError(st, where, msg)

    def close(self):
        if getattr(self, "h", None):
            self.L.gra_ctx_destroy(self.h)
            self.h = None
        _LIVE.discard(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        # no runtime calls once the interpreter is finalising: the atexit
        # hook has closed every live context before that point
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self.check(self.L.gra_sync(self.h), "gra_sync")

    def use_stream_of(self, other: "Context | None"):
        """Queue this context's later work on `other`'s stream (None: back on
        its own).  Waits for the work queued on the stream in use first.
        `other` must stay open while this context uses its stream."""
        s = self.L.gra_ctx_stream(other.h) if other is not None else None
        self.check(self.L.gra_ctx_set_stream(self.h, s), "gra_ctx_set_stream")

    # timing (GPU events on the context stream)
    def timing(self, on):
        """False/True, or a stride n: CG-loop launches timed on every n-th
        iteration (others on every launch)."""
        self.check(self.L.gra_timing_enable(self.h, int(on)), "timing")

    def timing_reset(self):
        self.check(self.L.gra_timing_reset(self.h), "timing_reset")

    def timing_get(self, which: int):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self.check(self.L.gra_timing_get(self.h, which, ctypes.byref(ms), ctypes.byref(n)),
                   "timing_get")
        return ms.value, n.value


    def tune(self, **kw):
        """Override G-product launch shapes (n_rows, n_unroll, n_splits,
        t_chunks, t_unroll, nontemporal, n_block); 0 (-1 for nontemporal and
        n_block) = auto."""
        for k, v in kw.items():
            self.check(self.L.gra_tune_set(self.h, _lib.TUNE[k], int(v)), f"tune {k}")

    def autotune(self, reps: int = 0):
        """Pick the G-product launch shapes for the current G by timing
        back-to-back G x / G^T w pairs; returns (ms G x, ms G^T w)."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self.check(self.L.gra_autotune(self.h, int(reps), ctypes.byref(a), ctypes.byref(b)),
                   "autotune")
        return a.value, b.value

    def tuning(self):
        out = {}
        for k, p in _lib.TUNE.items():
            v = ctypes.c_int64()
            self.check(self.L.gra_tune_get(self.h, p, ctypes.byref(v)), "tune_get")
            out[k] = v.value
        return out


TIME_GEMV_N, TIME_GEMV_T, TIME_ASSEMBLE, TIME_CG_ITER, TIME_FUSED, TIME_MAT_MUL = 0, 1, 2, 3, 4, 5
TIME_REFRESH, TIME_CG_GRAPH, TIME_AFFINE = 6, 7, 8


def csr_struct(m, keep: list) -> GraCsr:
    """scipy sparse (square) -> GraCsr (arrays appended to `keep`)."""
    m = sp.csr_matrix(m)
    m.sort_indices()
    rp = np.ascontiguousarray(m.indptr, dtype=np.int64)
    ci = np.ascontiguousarray(m.indices, dtype=np.int32)
    v = np.ascontiguousarray(m.data, dtype=np.float64)
    keep += [rp, ci, v]
    return GraCsr(m.shape[0], v.size, rp.ctypes.data_as(c_i64p), ci.ctypes.data_as(c_i32p),
                  v.ctypes.data_as(c_dp))


class Fwd:
    """Dense forward operator G on the GPU (mirror of the reference's Fwd)."""

    def __init__(self, prisms, points, field_flag: int = GZ, ctx: Context | None = None):
        self.ctx = ctx or Context()
        self.prisms = _f64(prisms).reshape(-1, 6)
        self.points = _f64(points).reshape(-1, 3)
        self.field_flag = int(field_flag)
        self.Nm = self.prisms.shape[0]
        self.N_obs = self.points.shape[0]
        self.Nd = len(component_bits(self.field_flag)) * self.N_obs
        self._assembled = False

    def compute_G(self, store: str = "f64", method: str = "faces"):
        """Fwd::compute_G (Fwd.cpp:47-79); G stays in device memory.

        method "faces": the reference's per-prism face/edge formulation
        (gra_assemble); "nodes": corner sums shared between the cells of a
        tensor-product mesh (gra_assemble_rectilinear; the prisms must be a
        regular_mesh-ordered rectilinear mesh and the points above it);
        "corners": corner sums per prism for any mesh (gra_assemble_corners;
        the points above every prism)."""
        c = self.ctx
        if method == "faces":
            st = c.L.gra_assemble(c.h, _p(self.prisms), self.Nm, _p(self.points), self.N_obs,
                                  self.field_flag, _lib.DTYPES[store])
            c.check(st, "gra_assemble")
        elif method == "corners":
            st = c.L.gra_assemble_corners(c.h, _p(self.prisms), self.Nm, _p(self.points),
                                          self.N_obs, self.field_flag, _lib.DTYPES[store])
            c.check(st, "gra_assemble_corners")
        elif method == "nodes":
            from .mesh import rectilinear_axes
            axes = rectilinear_axes(self.prisms)
            if axes is None:
                raise ValueError("method 'nodes' needs a tensor-product mesh in regular_mesh order")
            xs, ys, zs = axes
            st = c.L.gra_assemble_rectilinear(c.h, _p(xs), xs.size - 1, _p(ys), ys.size - 1,
                                              _p(zs), zs.size - 1, _p(self.points), self.N_obs,
                                              self.field_flag, _lib.DTYPES[store])
            c.check(st, "gra_assemble_rectilinear")
        else:
            raise ValueError(f"unknown assembly method {method!r}")
        self._assembled = True

    def set_mesh(self, prisms):
        """Fwd::set_mesh (Fwd.cpp:221-225): new cells; G must be assembled
        again."""
        self.prisms = _f64(prisms).reshape(-1, 6).copy()
        self.Nm = self.prisms.shape[0]
        self._assembled = False

    def compute_G_wavelet(self, relative_threshold: float, order):
        """Fwd::compute_G_wavelet (Fwd.cpp:148-219): G over the cells in
        `order` (reordered position -> cell, e.g. OctreeMesh.reordered()),
        D4-transformed per row and thresholded at relative_threshold times
        the row's largest coefficient; the products use it until the next
        compute_G.  Returns (n_dy, kept coefficients)."""
        c = self.ctx
        o = np.ascontiguousarray(order, dtype=np.int64)
        if o.shape != (self.Nm,):
            raise ValueError(f"order must have {self.Nm} entries")
        st = c.L.gra_assemble_wavelet(c.h, _p(self.prisms), self.Nm, _p(self.points), self.N_obs,
                                      self.field_flag, o.ctypes.data_as(_lib.c_i64p),
                                      float(relative_threshold))
        c.check(st, "gra_assemble_wavelet")
        self._assembled = True
        return self.wavelet_info()

    def wavelet_info(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self.ctx.check(self.ctx.L.gra_wavelet_info(self.ctx.h, ctypes.byref(a), ctypes.byref(b)),
                       "wavelet_info")
        return a.value, b.value

    def _need_G(self):
        if not self._assembled:
            raise GraError(1, "Fwd", "compute_G() has not been called")

    def G_vec_mul(self, vec):
        self._need_G()
        v = _f64(vec, self.Nm, "vec")
        out = np.empty(self.Nd)
        self.ctx.check(self.ctx.L.gra_G_vec_mul(self.ctx.h, _p(v), _p(out)), "G_vec_mul")
        return out

    def GT_vec_mul(self, vec):
        self._need_G()
        v = _f64(vec, self.Nd, "vec")
        out = np.empty(self.Nm)
        self.ctx.check(self.ctx.L.gra_GT_vec_mul(self.ctx.h, _p(v), _p(out)), "GT_vec_mul")
        return out

    def G_mat_mul(self, X):
        """G @ X for the k columns of X (Nm x k) on the matrix cores
        (gra_G_mat_mul); returns Nd x k."""
        self._need_G()
        X = np.asarray(X, dtype=np.float64)
        if X.ndim != 2 or X.shape[0] != self.Nm:
            raise ValueError(f"X: expected shape ({self.Nm}, k)")
        k = X.shape[1]
        Xc = np.asfortranarray(X)
        out = np.empty((self.Nd, k), order="F")
        self.ctx.check(self.ctx.L.gra_G_mat_mul(self.ctx.h, Xc.ctypes.data_as(c_dp), k,
                                                out.ctypes.data_as(c_dp)), "G_mat_mul")
        return out

    def GT_mat_mul(self, R):
        """G^T @ R for the k columns of R (Nd x k) on the matrix cores
        (gra_GT_mat_mul); returns Nm x k."""
        self._need_G()
        R = np.asarray(R, dtype=np.float64)
        if R.ndim != 2 or R.shape[0] != self.Nd:
            raise ValueError(f"R: expected shape ({self.Nd}, k)")
        k = R.shape[1]
        Rc = np.asfortranarray(R)
        out = np.empty((self.Nm, k), order="F")
        self.ctx.check(self.ctx.L.gra_GT_mat_mul(self.ctx.h, Rc.ctypes.data_as(c_dp), k,
                                                 out.ctypes.data_as(c_dp)), "GT_mat_mul")
        return out

    def compute_gobs(self, rho):
        return self.G_vec_mul(rho)

    def expand_G(self, split_cols, child_prisms):
        """AdaptiveInversion::expand_G (AdaptiveInversion.cpp:23-106): replace
        the columns of the split cells (old column indices, ascending) by their
        8 children's columns, computed on the GPU.  Returns the new Nm."""
        self._need_G()
        sc = np.ascontiguousarray(split_cols, dtype=np.int64)
        ch = _f64(child_prisms).reshape(-1, 6)
        if ch.shape[0] != 8 * sc.size:
            raise ValueError("need 8 child prisms per split cell")
        st = self.ctx.L.gra_expand_cols(self.ctx.h, sc.ctypes.data_as(_lib.c_i64p), sc.size,
                                        _p(ch))
        self.ctx.check(st, "expand_G")
        parts, last = [], 0
        for s, col in enumerate(sc):
            parts += [self.prisms[last:col], ch[8 * s:8 * s + 8]]
            last = col + 1
        parts.append(self.prisms[last:])
        self.prisms = np.ascontiguousarray(np.vstack(parts))
        self.Nm = self.prisms.shape[0]
        return self.Nm

    def col_norms(self):
        """||G(:, j)||_2 per column (the reads of set_depth_weighting(beta, 1),
        InversionBase.cpp:519-526)."""
        self._need_G()
        out = np.empty(self.Nm)
        self.ctx.check(self.ctx.L.gra_col_norms(self.ctx.h, _p(out)), "col_norms")
        return out

    def get_G(self, col_major: bool = False):
        """Host copy of G as float64 (Nd x Nm)."""
        self._need_G()
        out = np.empty((self.Nd, self.Nm)) if not col_major else np.empty((self.Nm, self.Nd))
        self.ctx.check(self.ctx.L.gra_copy_G(self.ctx.h, _p(out), 1 if col_major else 0),
                       "copy_G")
        return out if not col_major else out.T


class GaussNewtonInversion(Fwd):
    """solve_cg and the Gauss-Newton helpers on the GPU-resident G."""

    def __init__(self, prisms, points, field_flag: int = GZ, ctx: Context | None = None):
        super().__init__(prisms, points, field_flag, ctx)
        self.use_cross_gradient_constraint = False
        self.a_crg = 1.0
        self.S_crg = None
        self._ops_key = None

    def set_regularization(self, Ws, Wx, Wy, Wz, Tx=None, Ty=None, Tz=None):
        keep = []
        W = [csr_struct(m, keep) for m in (Ws, Wx, Wy, Wz)]
        if Tx is not None:
            T = [ctypes.byref(csr_struct(m, keep)) for m in (Tx, Ty, Tz)]
        else:
            T = [None, None, None]
        st = self.ctx.L.gra_set_regularization(self.ctx.h, *[ctypes.byref(w) for w in W], *T)
        self.ctx.check(st, "set_regularization")

    def solve_cg(self, tol, maxit, lam, mk, mref, Wd, dobs, P, WL_s=None, WL_x=None, WL_y=None,
                 WL_z=None, WL_tx=None, WL_ty=None, WL_tz=None, quirks: bool = True,
                 fused: bool = False):
        """GaussNewtonInversion::solve_cg; diagonal operators as vectors (None =
        identity).  Regularisation operators come from set_regularization.
        Returns (dx, err, iterations)."""
        self._need_G()
        nm, nd = self.Nm, self.Nd
        arr = dict(mk=_f64(mk, nm, "mk"), mref=_f64(mref, nm, "mref"), Wd=_f64(Wd, nd, "Wd"),
                   dobs=_f64(dobs, nd, "dobs"), P=_f64(P, nm, "P"))
        wl = {k: (None if v is None else _f64(v, nm, k)) for k, v in
              dict(WL_s=WL_s, WL_x=WL_x, WL_y=WL_y, WL_z=WL_z, WL_tx=WL_tx, WL_ty=WL_ty,
                   WL_tz=WL_tz).items()}
        S = None if self.S_crg is None else _f64(self.S_crg, nm, "S_crg")
        a = GraCgArgs()
        a.tol, a.maxit, a.lambda_ = float(tol), int(maxit), float(lam)
        for k, v in arr.items():
            setattr(a, k, _p(v))
        for k, v in wl.items():
            setattr(a, k, _p(v))
        a.use_cross_gradient = 1 if self.use_cross_gradient_constraint else 0
        a.a_crg = float(self.a_crg)
        a.S_crg = _p(S)
        a.quirk_mode = 1 if quirks else 0
        a.fused = 1 if fused else 0
        dx = np.empty(nm)
        err = ctypes.c_double(np.nan)
        its = ctypes.c_int(0)
        st = self.ctx.L.gra_solve_cg(self.ctx.h, ctypes.byref(a), _p(dx), ctypes.byref(err),
                                     ctypes.byref(its))
        self.ctx.check(st, "solve_cg")
        return dx, err.value, its.value

    def mixed_precision(self, inner: str = "bf16", refresh_every: int = 10):
        """Mixed-precision CG on this context's FP64 G (gra_mixed_precision):
        the iteration's products on an FP32 / BF16 copy, FP64 residual
        refreshes every `refresh_every` iterations; inner "f64" switches it
        off."""
        self._need_G()
        st = self.ctx.L.gra_mixed_precision(self.ctx.h, _lib.DTYPES[inner], int(refresh_every))
        self.ctx.check(st, "mixed_precision")

    def barrier_P(self, m, m_min, m_max):
        nm = self.Nm
        out = np.empty(nm)
        st = self.ctx.L.gra_barrier_P(self.ctx.h, _p(_f64(m, nm)), _p(_f64(m_min, nm)),
                                      _p(_f64(m_max, nm)), _p(out))
        self.ctx.check(st, "barrier_P")
        return out

    def barrier_update(self, m, dx, m_min, m_max, m_trial):
        nm = self.Nm
        mt = _f64(m_trial, nm).copy()
        st = self.ctx.L.gra_barrier_update(self.ctx.h, _p(_f64(m, nm)), _p(_f64(dx, nm)),
                                           _p(_f64(m_min, nm)), _p(_f64(m_max, nm)), _p(mt))
        self.ctx.check(st, "barrier_update")
        return mt

    def misfit(self, m, Wd, dobs):
        out = ctypes.c_double()
        st = self.ctx.L.gra_misfit(self.ctx.h, _p(_f64(m, self.Nm)), _p(_f64(Wd, self.Nd)),
                                   _p(_f64(dobs, self.Nd)), ctypes.byref(out))
        self.ctx.check(st, "misfit")
        return out.value

    def irls_weights(self, W, m, p: int, eps2: float):
        keep = []
        c = csr_struct(W, keep)
        out = np.empty(W.shape[0])
        st = self.ctx.L.gra_irls_weights(self.ctx.h, ctypes.byref(c), _p(_f64(m, W.shape[0])),
                                         int(p), float(eps2), _p(out))
        self.ctx.check(st, "irls_weights")
        return out

    # resident Gauss-Newton loop (model and CG state stay on the GPU)
    def gn_setup(self, dobs, Wd, m_min, m_max, mref, m_ini, lp: int = 2, eps2: float = 0.0,
                 fused: bool = False, cross_gradient: bool = False, a_crg: float = 0.0,
                 S_crg=None):
        """Resident Gauss-Newton state; with cross_gradient the T operators of
        set_regularization enter every CG (weight a_crg, mask S_crg)."""
        nm, nd = self.Nm, self.Nd
        self._gn_keep = [_f64(dobs, nd), _f64(Wd, nd), _f64(m_min, nm), _f64(m_max, nm),
                         _f64(mref, nm), _f64(m_ini, nm)]
        S = _f64(S_crg, nm, "S_crg") if S_crg is not None else None
        self._gn_keep.append(S)
        s = GraGnSetup(*[_p(a) for a in self._gn_keep[:6]], int(lp), float(eps2),
                       1 if fused else 0, 1 if cross_gradient else 0, float(a_crg),
                       _p(S) if S is not None else None)
        self.ctx.check(self.ctx.L.gra_gn_setup_problem(self.ctx.h, ctypes.byref(s)), "gn_setup")

    def gn_step(self, lam, cg_tol, cg_maxit):
        mis = ctypes.c_double()
        err = ctypes.c_double()
        its = ctypes.c_int()
        st = self.ctx.L.gra_gn_step(self.ctx.h, float(lam), float(cg_tol), int(cg_maxit),
                                    ctypes.byref(mis), ctypes.byref(err), ctypes.byref(its))
        self.ctx.check(st, "gn_step")
        return mis.value, err.value, its.value

    def gn_model(self):
        m = np.empty(self.Nm)
        self.ctx.check(self.ctx.L.gra_gn_get_model(self.ctx.h, _p(m)), "gn_model")
        return m


def gn_step_batch(invs, lambdas, cg_tol: float, cg_maxit: int):
    """One resident Gauss-Newton step on each inversion (each with its own
    context, one device) in a single batched submission
    (gra_gn_step_batch): returns (misfits, cg_errs, cg_iterations)."""
    n = len(invs)
    hs = (ctypes.c_void_p * n)(*[f.ctx.h.value for f in invs])
    lam = np.ascontiguousarray(lambdas, dtype=np.float64)
    if lam.size != n:
        raise ValueError("one lambda per inversion")
    mis, err = np.empty(n), np.empty(n)
    its = (ctypes.c_int * n)()
    L = invs[0].ctx.L
    st = L.gra_gn_step_batch(hs, n, _p(lam), float(cg_tol), int(cg_maxit), _p(mis), _p(err), its)
    if st != 0:
        msg = "; ".join(L.gra_last_error(f.ctx.h).decode(errors="replace") for f in invs
                        if L.gra_last_error(f.ctx.h))
        raise GraError(st, "gn_step_batch", msg)
    return mis, err, np.array(its[:], dtype=int)


def assemble_batch(fwds, store: str = "f64", method: str = "faces"):
    """Assemble the G of every Fwd in `fwds` (each with its own context, one
    device) in a single batched launch: the face/edge values
    (gra_assemble_batch) or, with method "corners", per-prism corner sums
    (gra_assemble_batch_corners; points above every prism of their item)."""
    if method not in ("faces", "corners"):
        raise ValueError(f"unknown assembly method {method!r}")
    n = len(fwds)
    if n == 0:
        return
    L = fwds[0].ctx.L
    hs = (ctypes.c_void_p * n)(*[f.ctx.h.value for f in fwds])
    items = (_lib.GraBatchItem * n)()
    for b, f in enumerate(fwds):
        items[b] = _lib.GraBatchItem(_p(f.prisms), f.Nm, _p(f.points), f.N_obs, f.field_flag)
    fn = L.gra_assemble_batch if method == "faces" else L.gra_assemble_batch_corners
    st = fn(hs, n, items, _lib.DTYPES[store])
    fwds[0].ctx.check(st, "assemble_batch")
    for f in fwds:
        f._assembled = True


def forward_rectilinear(xs, ys, zs, rho, points, field_flag: int, ctx: Context | None = None):
    """Forward of a rectilinear mesh (node coordinates xs, ys, zs; cell order
    i*ny*nz + j*nz + k) by the node-sum form: (N_obs, n_comp) array.  Points
    must lie above the mesh."""
    ctx = ctx or Context()
    xs, ys, zs = _f64(xs), _f64(ys), _f64(zs)
    nx, ny, nz = xs.size - 1, ys.size - 1, zs.size - 1
    points = _f64(points).reshape(-1, 3)
    rho = _f64(rho, nx * ny * nz, "rho")
    nc = len(component_bits(field_flag))
    out = np.empty((points.shape[0], nc))
    st = ctx.L.gra_forward_rectilinear(ctx.h, _p(xs), nx, _p(ys), ny, _p(zs), nz, _p(rho),
                                       _p(points), points.shape[0], int(field_flag), _p(out))
    ctx.check(st, "forward_rectilinear")
    return out


def forward_rectilinear_batch(requests, ctx: Context | None = None):
    """Several forward_rectilinear requests in one launch; `requests` is a
    list of (xs, ys, zs, rho, points, field_flag); returns one (N_obs, n_comp)
    array per request, bit-identical to forward_rectilinear of each."""
    ctx = ctx or Context()
    n = len(requests)
    items = (_lib.GraFwdRectItem * n)()
    keep, outs = [], []
    for it, (xs, ys, zs, rho, points, flag) in zip(items, requests):
        xs, ys, zs = _f64(xs), _f64(ys), _f64(zs)
        nx, ny, nz = xs.size - 1, ys.size - 1, zs.size - 1
        points = _f64(points).reshape(-1, 3)
        rho = _f64(rho, nx * ny * nz, "rho")
        out = np.empty((points.shape[0], len(component_bits(flag))))
        keep += [xs, ys, zs, points, rho]
        outs.append(out)
        it.xs, it.nx, it.ys, it.ny, it.zs, it.nz = _p(xs), nx, _p(ys), ny, _p(zs), nz
        it.rho, it.pts, it.N_obs, it.field_flag, it.out = (_p(rho), _p(points), points.shape[0],
                                                           int(flag), _p(out))
    ctx.check(ctx.L.gra_forward_rectilinear_batch(ctx.h, n, items), "forward_rectilinear_batch")
    return outs


def forward(prisms, rho, points, field_flag: int, ctx: Context | None = None):
    """compute.cpp:246-260 matrix-free forward: (N_obs, n_comp) array."""
    ctx = ctx or Context()
    prisms = _f64(prisms).reshape(-1, 6)
    points = _f64(points).reshape(-1, 3)
    rho = _f64(rho, prisms.shape[0], "rho")
    nc = len(component_bits(field_flag))
    out = np.empty((points.shape[0], nc))
    st = ctx.L.gra_forward(ctx.h, _p(prisms), _p(rho), prisms.shape[0], _p(points),
                           points.shape[0], int(field_flag), _p(out))
    ctx.check(st, "forward")
    return out
