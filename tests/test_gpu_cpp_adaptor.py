"""The C++ host side: tests/cpp/adaptor_check drives the engine only through
include/graeng.hpp (Fwd::compute_G / G_vec_mul / GT_vec_mul and
GaussNewtonInversion::solve_cg call shapes) and its results match the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from graada3d_amd import mesh as M
from graada3d_amd import operators as O
from helpers import rel_norm

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "adaptor_check")


def _csr_bytes(m):
    m = m.tocsr()
    m.sort_indices()
    return (np.array([m.shape[0], m.nnz], np.int64).tobytes() + m.indptr.astype(np.int64).tobytes()
            + m.indices.astype(np.int32).tobytes() + m.data.astype(np.float64).tobytes())


def test_cpp_adaptor_matches_oracle(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/adaptor_check not built (run __graft_entry__.build())")
    prisms = M.regular_mesh((0.0, 400.0), 8, (0.0, 300.0), 6, (0.0, 200.0), 4)
    pts = M.grid_points(8, -20.0, 420.0, 7, -20.0, 320.0, -2.0)
    flag = 0x12
    nm, nobs = prisms.shape[0], pts.shape[0]
    nd = 2 * nobs
    rng = np.random.default_rng(12)
    x, y = rng.standard_normal(nm), rng.standard_normal(nd)
    mk, mref = rng.uniform(-20, 20, nm), np.zeros(nm)
    G, _ = oracle.assemble(prisms, pts, flag)
    dobs = G @ rng.uniform(-300, 300, nm)
    Wd = np.full(nd, 1.0 / np.abs(dobs).max())
    P = oracle.barrier_P(mk, np.full(nm, -800.0), np.full(nm, 800.0))
    Z = O.depth_weights_offset(prisms, 50.0, 3.0)
    Ws, Wx, Wy, Wz = O.weighting_operators(prisms, 8, 6, 4, Z)
    lam, tol, maxit = 2.0, 0.0, 25
    blob = (np.array([nm, nobs, flag, maxit], np.int64).tobytes()
            + np.array([lam, tol], np.float64).tobytes()
            + b"".join(np.ascontiguousarray(a, np.float64).tobytes()
                       for a in (prisms, pts, x, y, mk, mref, Wd, dobs, P))
            + b"".join(_csr_bytes(m) for m in (Ws, Wx, Wy, Wz)))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    fin.write_bytes(blob)
    r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = np.frombuffer(fout.read_bytes(), np.float64)
    Gx, GTy, dx = out[:nd], out[nd:nd + nm], out[nd + nm:nd + 2 * nm]
    norms = out[nd + 2 * nm:nd + 3 * nm]
    Gx_corners = out[nd + 3 * nm:2 * nd + 3 * nm]
    assert out.size == 2 * nd + 3 * nm + 2
    err, its = out[-2], int(out[-1])
    # the per-prism corner-sum G (compute_G_corners) times x
    assert rel_norm(Gx_corners, G @ x) < 1e-10
    assert rel_norm(Gx, G @ x) < 1e-12
    assert rel_norm(GTy, G.T @ y) < 1e-12
    assert np.abs(norms / np.linalg.norm(G, axis=0) - 1.0).max() < 1e-14
    dxo, erro, itso = oracle.solve_cg(G, tol, maxit, lam, mk, mref, Wd, dobs, P, Ws, Wx, Wy, Wz)
    assert its == itso == maxit
    assert rel_norm(dx, dxo) < 1e-10
    assert abs(err - erro) < 1e-8 * erro
