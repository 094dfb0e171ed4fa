"""Uniform-grid evaluation with the points generated on the GPU
(ct_eval_cubic; reference evaluate_electron_density_on_cubic,
src/evaluate_density.cu:425-660) against the same quantities at the same
points passed explicitly, and against the oracle."""
import os

import numpy as np
import pytest

import fchk_py
from chemtoolscuda_amd.cube import UniformGrid
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ct():
    import chemtoolscuda_amd

    return chemtoolscuda_amd


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("name", ["h2o.fchk", "ch4.fchk", "atom_c_g_pure_orbital.fchk"])
def test_cubic_density_matches_points_and_oracle(ct, name):
    path = os.path.join(fchk_py.FCHK_DIR, name)
    d = fchk_py.load(path)
    mol = ct.Molecule(path)
    g = UniformGrid.from_molecule(mol.numbers, mol.coordinates, spacing=0.5)
    rho = mol.compute_electron_density_on_cubic_grid(g.origin, g.axes, g.shape, False)
    assert rho.shape == g.shape
    pts = g.points
    ref = mol.compute_electron_density(pts)
    assert _rel(rho.reshape(-1), ref) < 1e-12
    sub = np.arange(0, pts.shape[0], max(1, pts.shape[0] // 400))
    orc = oracle.density(d["basis"].flat(), d["P"], pts[sub])["rho"]
    assert _rel(rho.reshape(-1)[sub], orc) < 1e-12


def test_cubic_several_quantities_skewed_axes_chunks_and_mixed(ct):
    from chemtoolscuda_amd import synthetic

    w = synthetic.water(10)
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    axes = np.array([[0.21, 0.02, 0.0], [-0.03, 0.19, 0.01], [0.0, 0.04, 0.23]])
    g = UniformGrid(origin=[-3.0, -2.5, -3.2], axes=axes, shape=(23, 19, 29))
    names = ["rho", "grad", "lap", "tau_pd", "rdg"]
    cub = mol.compute_on_cubic_grid(g.origin, g.axes, g.shape, names)
    pts = g.points
    ref = mol.compute(pts, names)
    for k in names:
        assert cub[k].shape[:3] == g.shape
        assert _rel(cub[k].reshape(ref[k].shape), ref[k]) < 1e-12, k
    mol.set_workspace_limit(256 * 25 * 4 * 40)  # many chunks
    cub2 = mol.compute_on_cubic_grid(g.origin, g.axes, g.shape, names)
    for k in names:
        assert np.array_equal(cub2[k], cub[k]), k
    mix = mol.compute_on_cubic_grid(g.origin, g.axes, g.shape, ["rho"], precision="mixed")
    m = cub["rho"] >= 1e-8
    assert np.max(np.abs(mix["rho"] - cub["rho"])[m] / cub["rho"][m]) <= 1e-6
    with pytest.raises(Exception):
        mol.compute_on_cubic_grid(g.origin, g.axes, (0, 2, 2), ["rho"])
