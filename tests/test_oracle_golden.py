"""The CPU oracle against the reference's own golden vectors.

* Committed noise-free outputs of the reference `compute` tool
  (Examples/Forward_modelling/{tests/test_compute/result_no_noise,
  model2/result/data_no_noise, model3/result/data_no_noise}), printed %.15e,
  so the bar is a few units of the 16th digit.
* Values written by the reference's own prism sources (oracle/_ref):
  all nine components on random, grid-aligned and on-face geometry, and a
  small dense G in the row-major component-blocked layout.  The oracle keeps
  the reference's operation order and libm, so these must match bit for bit.
"""
import numpy as np
import pytest

import oracle
from helpers import load, model_from_fixture


def _forward_check(name, stride=1):
    fx = load(f"ref_forward_{name}.npz")
    prisms, rho = model_from_fixture(fx)
    nz = rho != 0
    pts = fx["points"][::stride]
    gold = fx["values"][::stride]
    out, st = oracle.forward(prisms[nz], rho[nz], pts, int(fx["flag"]))
    assert st == 0
    bits = sorted(fx["bits"].tolist())
    cols = [bits.index(b) for b in fx["bits"]]     # file column order -> bit order
    out = out[:, cols]
    rel = np.abs(out - gold) / np.abs(gold)
    return rel


@pytest.mark.parametrize("name,stride", [("model1", 1), ("model3", 2), ("model2", 9)])
def test_forward_matches_reference_golden(name, stride):
    rel = _forward_check(name, stride)
    # %.15e output: 16 significant digits
    assert rel.max() < 2e-15, rel.max()


def test_box_models_match_committed_vtk():
    """The density models rebuilt from the anomaly boxes with the reference's
    full-containment rule equal the densities of the committed .vtk files."""
    for name in ("model1", "model3"):
        fx = load(f"ref_forward_{name}.npz")
        _, rho = model_from_fixture(fx)
        idx = np.nonzero(rho)[0]
        assert np.array_equal(idx, fx["vtk_nonzero_idx"])
        assert np.array_equal(rho[idx], fx["vtk_nonzero_val"])


def test_pairs_bit_exact_vs_reference_vectors():
    fx = load("ref_pairs.npz")
    for q in range(fx["points"].shape[0]):
        v, st = oracle.prism_field(fx["points"][q], fx["prisms"][q], 1.0, 0x3FE)
        assert st == 0
        assert np.array_equal(v[1:], fx["values"][q][1:], equal_nan=True), q


def test_dense_G_bit_exact_vs_reference():
    fx = load("ref_G_small.npz")
    G, st = oracle.assemble(fx["prisms"], fx["points"], int(fx["flag"]))
    assert st == 0
    assert np.array_equal(G, fx["G"])


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
def test_oracle_vs_reference_build_live():
    rng = np.random.default_rng(11)
    for k in range(400):
        pr = np.sort(rng.uniform(-300, 900, (3, 2)), axis=1).ravel()
        if k % 2:
            pr = np.round(pr / 25) * 25
            if np.any(pr[1::2] <= pr[0::2]):
                continue
        pt = rng.uniform(-600, 1500, 3)
        if k % 3 == 0:
            pt[0] = pr[1]
        if k % 5 == 0:
            pt[1] = pr[2]
        a, s1 = oracle.prism_field(pt, pr, 1.0, 0x3FE)
        if s1:
            continue
        b, _ = oracle.prism_field(pt, pr, 1.0, 0x3FE, use_ref=True)
        assert np.array_equal(a, b, equal_nan=True), (pt, pr)
