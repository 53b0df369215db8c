"""CPU oracle (test infrastructure only).

ctypes bindings for ``oracle/liboracle.so`` (the C restatement of the
reference's prism kernels and CGLS solver) and, where it has been built,
``oracle/_ref/libref_prism.so`` (the reference's own prism sources).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only as the checker.  The product
(``graada3d_amd``) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_prism.so")

_dp = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def build(with_ref: bool | None = None) -> None:
    """Compile liboracle.so, and oracle/_ref when the reference tree exists."""
    subprocess.check_call(["make", "-s", "-C", HERE])
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference/src/Forward_modelling")
    if with_ref:
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(with_ref=False)
        L = ctypes.CDLL(LIB_PATH)
        L.ora_prism_field.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_uint32, _dp]
        L.ora_assemble.argtypes = [_dp, _i64, _dp, _i64, ctypes.c_uint32, _i64, _i64, _dp]
        L.ora_forward.argtypes = [_dp, _dp, _i64, _dp, _i64, ctypes.c_uint32, _dp]
        L.ora_gemv_n.argtypes = [_dp, _i64, _i64, _dp, _dp]
        L.ora_gemv_t.argtypes = [_dp, _i64, _i64, _dp, _dp]
        L.ora_barrier_P.argtypes = [_dp, _dp, _dp, _i64, _dp]
        L.ora_barrier_update.argtypes = [_dp, _dp, _dp, _dp, _i64, _dp]
        L.ora_misfit.argtypes = [_dp, _i64, _i64, _dp, _dp, _dp]
        L.ora_misfit.restype = ctypes.c_double
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    """Threads of the oracle's point / row / column loops (results do not
    depend on it; 1 = the reference's single-threaded GEMV)."""
    lib().ora_set_threads(int(n))


def max_threads() -> int:
    return int(lib().ora_max_threads())


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own prism kernel, or None when it was not built."""
    global _ref
    if _ref is None and ref_available():
        R = ctypes.CDLL(REF_PATH)
        R.ref_prism_field.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_uint32, _dp]
        R.ref_assemble.argtypes = [_dp, _i64, _dp, _i64, ctypes.c_uint32, _i64, _i64, _dp]
        R.ref_forward.argtypes = [_dp, _dp, _i64, _dp, _i64, ctypes.c_uint32, _dp]
        _ref = R
    return _ref


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def n_components(flag: int) -> int:
    return bin(flag & 0x3FE).count("1")


def prism_field(point, prism, rho, flag, use_ref=False):
    """All requested components of one (point, prism) pair, indexed by field bit."""
    out = np.zeros(10)
    p, q = _c(point), _c(prism)
    L = ref() if use_ref else lib()
    st = (L.ref_prism_field if use_ref else L.ora_prism_field)(_ptr(p), _ptr(q), float(rho), flag, _ptr(out))
    return out, st


def assemble(prisms, pts, flag, row0=0, nrows=None, use_ref=False):
    """Row-major G (n_comp*N_obs x Nm); only rows of points [row0,row0+nrows) filled."""
    prisms, pts = _c(prisms), _c(pts)
    nm, nobs = prisms.shape[0], pts.shape[0]
    nrows = nobs - row0 if nrows is None else nrows
    G = np.zeros((n_components(flag) * nobs, nm))
    L = ref() if use_ref else lib()
    fn = L.ref_assemble if use_ref else L.ora_assemble
    st = fn(_ptr(prisms), nm, _ptr(pts), nobs, flag, row0, nrows, _ptr(G))
    return G, st


def forward(prisms, rho, pts, flag, use_ref=False):
    prisms, rho, pts = _c(prisms), _c(rho), _c(pts)
    out = np.zeros((pts.shape[0], n_components(flag)))
    L = ref() if use_ref else lib()
    fn = L.ref_forward if use_ref else L.ora_forward
    st = fn(_ptr(prisms), _ptr(rho), prisms.shape[0], _ptr(pts), pts.shape[0], flag, _ptr(out))
    return out, st


def gemv(G, x, trans=False):
    G, x = _c(G), _c(x)
    nd, nm = G.shape
    y = np.zeros(nm if trans else nd)
    (lib().ora_gemv_t if trans else lib().ora_gemv_n)(_ptr(G), nd, nm, _ptr(x), _ptr(y))
    return y


# ---------------------------------------------------------------- CGLS ----

class _CSR(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("rp", ctypes.POINTER(ctypes.c_int64)),
                ("ci", ctypes.POINTER(ctypes.c_int32)), ("v", _dp)]


def _csr(m, keep):
    """scipy CSR -> C struct (arrays kept alive in `keep`)."""
    m = m.tocsr()
    m.sort_indices()
    rp = np.ascontiguousarray(m.indptr, dtype=np.int64)
    ci = np.ascontiguousarray(m.indices, dtype=np.int32)
    v = np.ascontiguousarray(m.data, dtype=np.float64)
    keep += [rp, ci, v]
    return _CSR(m.shape[0], rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                ci.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _ptr(v))


def solve_cg(G, tol, maxit, lam, mk, mref, Wd, dobs, P, Ws, Wx, Wy, Wz,
             WL=None, crg=None, quirks=True, recurrence=False):
    """Oracle ``solve_cg``.  G dense row-major; W* scipy sparse; WL a dict of
    diagonal vectors with keys s,x,y,z,tx,ty,tz (missing = identity); crg
    None or dict(a=..., S=vec or None, Tx=, Ty=, Tz= scipy sparse).
    ``recurrence`` selects the r <- r + alpha A^T t update of the
    single-read iteration instead of the reference's r = A^T s.
    Returns (dx, err, iterations)."""
    L = lib()
    keep = []
    WL = WL or {}
    n = mk.shape[0]
    ops = [Ws, Wx, Wy, Wz, Ws.T, Wx.T, Wy.T, Wz.T]
    cs = [_csr(o, keep) for o in ops]
    if crg is not None:
        tops = [crg["Tx"], crg["Ty"], crg["Tz"], crg["Tx"].T, crg["Ty"].T, crg["Tz"].T]
        ts = [_csr(o, keep) for o in tops]
        a_crg = float(crg["a"])
        S = crg.get("S")
        S = _c(S) if S is not None else None
    else:
        import scipy.sparse as sp
        z = sp.csr_matrix((n, n))
        ts = [_csr(z, keep) for _ in range(6)]
        a_crg, S = 0.0, None

    def v(key):
        a = WL.get(key)
        return _c(a) if a is not None else None

    wl = [v(k) for k in ("s", "x", "y", "z", "tx", "ty", "tz")]
    dx = np.zeros(n)
    err = ctypes.c_double(np.nan)
    its = ctypes.c_int(-1)
    fn = L.ora_solve_cg
    CP = ctypes.POINTER(_CSR)
    fn.argtypes = ([ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                    ctypes.POINTER(ctypes.c_int), ctypes.c_double, _dp, _dp, _dp, _dp, _dp]
                   + [CP] * 8 + [ctypes.c_int, ctypes.c_double, _dp] + [CP] * 6 + [_dp] * 7
                   + [ctypes.c_int, ctypes.c_int, _dp, _i64, _i64, _dp])
    G = _c(G)
    arrs = [_c(mk), _c(mref), _c(Wd), _c(dobs), _c(P)]
    fn(float(tol), int(maxit), ctypes.byref(err), ctypes.byref(its), float(lam),
       *[_ptr(a) for a in arrs], *[ctypes.byref(c) for c in cs],
       1 if crg is not None else 0, a_crg, _ptr(S), *[ctypes.byref(c) for c in ts],
       *[_ptr(a) for a in wl], 1 if quirks else 0, 1 if recurrence else 0, _ptr(G), G.shape[0],
       G.shape[1], _ptr(dx))
    return dx, err.value, its.value


def barrier_P(m, mmin, mmax):
    m, mmin, mmax = _c(m), _c(mmin), _c(mmax)
    P = np.zeros_like(m)
    lib().ora_barrier_P(_ptr(m), _ptr(mmin), _ptr(mmax), m.shape[0], _ptr(P))
    return P


def barrier_update(m, dx, mmin, mmax, m_trial):
    m, dx, mmin, mmax = _c(m), _c(dx), _c(mmin), _c(mmax)
    mt = _c(m_trial).copy()
    lib().ora_barrier_update(_ptr(m), _ptr(dx), _ptr(mmin), _ptr(mmax), m.shape[0], _ptr(mt))
    return mt


def misfit(G, m, Wd, dobs):
    G = _c(G)
    return lib().ora_misfit(_ptr(G), G.shape[0], G.shape[1], _ptr(_c(m)), _ptr(_c(Wd)), _ptr(_c(dobs)))


def irls_weights(W, m, p, eps2):
    keep = []
    c = _csr(W, keep)
    out = np.zeros(W.shape[0])
    fn = lib().ora_irls_weights
    fn.argtypes = [ctypes.POINTER(_CSR), _dp, ctypes.c_int, ctypes.c_double, _dp]
    fn(ctypes.byref(c), _ptr(_c(m)), int(p), float(eps2), _ptr(out))
    return out


# ------------------------------------------------------------- wavelets ----

_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)


def d4_filters():
    h, g = np.zeros(4), np.zeros(4)
    lib().ora_d4_filters(_ptr(h), _ptr(g))
    return h, g


def d4_transform(a, forward=True):
    """The full periodic D4 transform of a (length a power of two)."""
    a = _c(a).copy()
    fn = lib().ora_d4_transform
    fn.argtypes = [_dp, ctypes.c_int64, ctypes.c_int]
    if fn(_ptr(a), a.shape[0], 1 if forward else -1):
        raise ValueError("length must be a power of two >= 2")
    return a


def wavelet_length(n):
    fn = lib().ora_wavelet_length
    fn.argtypes = [ctypes.c_int64]
    fn.restype = ctypes.c_int64
    return int(fn(int(n)))


class WaveletG:
    """compute_G_wavelet's compressed rows: CSR (rowptr, ci, v) over the
    n_dy wavelet coefficients, with the cell order it was built for."""

    def __init__(self, prisms, pts, flag, order, rel):
        prisms, pts = _c(prisms), _c(pts)
        self.order = np.ascontiguousarray(order, dtype=np.int64)
        self.Nm = prisms.shape[0]
        self.n_dy = wavelet_length(self.Nm)
        self.n_rows = n_components(flag) * pts.shape[0]
        fn = lib().ora_wavelet_rows
        fn.argtypes = [_dp, ctypes.c_int64, _dp, ctypes.c_int64, ctypes.c_uint32, _i64p,
                       ctypes.c_double, ctypes.c_int64, _i64p, _i64p, _i32p, _dp]
        counts = np.zeros(self.n_rows, np.int64)
        args = (_ptr(prisms), self.Nm, _ptr(pts), pts.shape[0], flag,
                self.order.ctypes.data_as(_i64p), float(rel), self.n_dy)
        st = fn(*args, counts.ctypes.data_as(_i64p), None, None, None)
        self.rowptr = np.zeros(self.n_rows + 1, np.int64)
        np.cumsum(counts, out=self.rowptr[1:])
        self.ci = np.zeros(self.rowptr[-1], np.int32)
        self.v = np.zeros(self.rowptr[-1])
        fn(*args, counts.ctypes.data_as(_i64p), self.rowptr.ctypes.data_as(_i64p),
           self.ci.ctypes.data_as(_i32p), _ptr(self.v))
        self.status = st

    @property
    def nnz(self):
        return int(self.rowptr[-1])

    def _prod(self, name, x, n_out):
        fn = getattr(lib(), name)
        fn.argtypes = [_i64p, _i32p, _dp, ctypes.c_int64, ctypes.c_int64, _i64p, ctypes.c_int64,
                       _dp, _dp]
        y = np.zeros(n_out)
        fn(self.rowptr.ctypes.data_as(_i64p), self.ci.ctypes.data_as(_i32p), _ptr(self.v),
           self.n_rows, self.n_dy, self.order.ctypes.data_as(_i64p), self.Nm, _ptr(_c(x)), _ptr(y))
        return y

    def G_vec_mul(self, x):
        return self._prod("ora_wavelet_gemv_n", x, self.n_rows)

    def GT_vec_mul(self, r):
        return self._prod("ora_wavelet_gemv_t", r, self.Nm)

    def dense(self):
        """The operator as a dense (n_rows x Nm) matrix in the original cell
        order: row i = W^T c_i mapped back through the cell order."""
        out = np.zeros((self.n_rows, self.Nm))
        for i in range(self.n_rows):
            a = np.zeros(self.n_dy)
            sl = slice(self.rowptr[i], self.rowptr[i + 1])
            a[self.ci[sl]] = self.v[sl]
            out[i, self.order] = d4_transform(a, forward=False)[:self.Nm]
        return out
