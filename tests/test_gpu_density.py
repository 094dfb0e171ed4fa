"""GPU parity of the FP64 density path against the pinned FP64 oracle.

Tolerance: the FP64 GPU path must agree with the oracle to <= 1e-12
relative.  Relative error is measured per point against the scale of the
quantity at that point, floored at 1e-10 of the largest value so that
near-zero components (gradient components that cross zero, rho in the far
tail) are compared in absolute terms at that level.
"""
import os

import numpy as np
import pytest

import fchk_py
from oracle import descriptors as D
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


def relerr(got, ref):
    got = np.asarray(got, dtype=np.float64).reshape(ref.shape[0], -1)
    ref = np.asarray(ref, dtype=np.float64).reshape(ref.shape[0], -1)
    scale = np.maximum(np.linalg.norm(ref, axis=1), 1e-10 * np.abs(ref).max())
    return float(np.max(np.linalg.norm(got - ref, axis=1) / scale))


@pytest.fixture(scope="module")
def ct():
    import chemtoolscuda_amd

    return chemtoolscuda_amd


def _points(basis, n, seed, spread=3.0):
    rng = np.random.default_rng(seed)
    c = basis.atom_xyz[rng.integers(0, basis.natom, size=n)]
    return c + rng.normal(0.0, spread / 2, size=(n, 3))


FIXTURES = ["atom_he.fchk", "atom_be.fchk", "atom_kr.fchk", "atom_c_g_pure_orbital.fchk",
            "atom_be_f_pure_orbital.fchk", "atom_08_O_N09_M2_ub3lyp_ccpvtz_g09.fchk", "ch4.fchk", "h2o.fchk",
            "qm9_000092_HF_cc-pVDZ.fchk", "test.fchk"]


@pytest.mark.parametrize("name", FIXTURES)
def test_fchk_loader_matches_python_reader(ct, name):
    path = os.path.join(fchk_py.FCHK_DIR, name)
    d = fchk_py.load(path)
    mol = ct.Molecule(path)
    assert mol.nbasis == d["basis"].nbasis
    assert np.array_equal(mol.numbers, d["numbers"])
    assert np.allclose(mol.coordinates, d["basis"].atom_xyz, rtol=0, atol=0)
    assert np.allclose(mol.charges, d["charges"], rtol=0, atol=0)
    assert np.allclose(mol.density_matrix, d["P"], rtol=1e-14, atol=1e-14)
    assert mol.get_file_path() == path


@pytest.mark.parametrize("name", FIXTURES)
def test_fchk_density_family_fp64(ct, name):
    path = os.path.join(fchk_py.FCHK_DIR, name)
    d = fchk_py.load(path)
    mol = ct.Molecule(path)
    pts = _points(d["basis"], 1500, 11)
    ref = oracle.density(d["basis"].flat(), d["P"], pts, rho=True, grad=True, hess=True, lap=True, tau=True)
    assert relerr(mol.compute_density(pts), ref["rho"]) < TOL
    assert relerr(mol.compute_electron_density(pts), ref["rho"]) < TOL
    assert relerr(mol.compute_gradient(pts), ref["grad"]) < TOL
    assert relerr(mol.compute_electron_density_gradient(pts), ref["grad"]) < TOL
    assert relerr(mol.compute_hessian(pts), ref["hess"]) < TOL
    assert relerr(mol.compute_laplacian(pts), ref["lap"]) < TOL
    assert relerr(mol.compute_positive_definite_kinetic_energy_density(pts), ref["tau"]) < TOL


def test_water_all_quantities_fp64(ct):
    from chemtoolscuda_amd import synthetic

    w = synthetic.water(4000)
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    pts = w.points
    ref = oracle.density(w.basis.flat(), w.P, pts, rho=True, grad=True, hess=True, lap=True, tau=True)
    rho, grad, lap, tau = ref["rho"], ref["grad"], ref["lap"], ref["tau"]
    got = mol.compute(pts, ["rho", "grad", "hess", "lap", "tau_pd", "rdg", "shannon", "weizsacker",
                            "thomas_fermi", "gge", "general_ked", "ked_alpha"], a=0.3, b=0.7, alpha=0.25)
    assert relerr(got["rho"], rho) < TOL
    assert relerr(got["grad"], grad) < TOL
    assert relerr(got["hess"], ref["hess"]) < TOL
    assert relerr(got["lap"], lap) < TOL
    assert relerr(got["tau_pd"], tau) < TOL
    assert relerr(got["rdg"], D.reduced_density_gradient(rho, grad)) < 1e-11
    assert relerr(got["shannon"], D.shannon(rho)) < 1e-11
    assert relerr(got["weizsacker"], D.weizsacker(rho, grad)) < 1e-11
    assert relerr(got["thomas_fermi"], D.thomas_fermi(rho)) < 1e-11
    # the GGE combination cancels; compare against the scale of its terms
    gge_ref = D.gradient_expansion_general(rho, grad, lap, 0.3, 0.7)
    assert relerr(got["gge"], gge_ref) < 1e-11
    assert relerr(got["general_ked"], D.general_ked(tau, lap, 0.25)) < 1e-11
    assert relerr(got["ked_alpha"], D.general_kinetic_energy_density(tau, lap, 0.25)) < 1e-11
    # single-quantity methods and reference spellings
    assert relerr(mol.compute_reduced_density_gradient(pts), D.reduced_density_gradient(rho, grad)) < 1e-11
    assert relerr(mol.compute_shannon_information_density(pts), D.shannon(rho)) < 1e-11
    assert relerr(mol.compute_weizsacker_ked(pts), D.weizsacker(rho, grad)) < 1e-11
    assert relerr(mol.compute_thomas_fermi_ked(pts), D.thomas_fermi(rho)) < 1e-11
    assert relerr(mol.compute_gradient_expansion_ked(pts), D.gradient_expansion_general(rho, grad, lap, 1 / 9, 1 / 6)) < 1e-11
    assert relerr(mol.compute_empirical_gradient_expansion_ked(pts),
                  D.gradient_expansion_general(rho, grad, lap, 1 / 5, 1 / 6)) < 1e-11
    assert relerr(mol.compute_general_gradient_expansion_ked(pts, 0.3, 0.7), gge_ref) < 1e-11
    assert relerr(mol.compute_general_ked(pts, 0.4), D.general_ked(tau, lap, 0.4)) < 1e-11
    assert relerr(mol.compute_general_kinetic_energy_density(pts, 0.4), D.general_ked(tau, lap, 0.4)) < 1e-11
    assert relerr(mol.compute_schrodinger_ked(pts), D.general_ked(tau, lap, 0.0)) < 1e-11
    assert relerr(mol.compute_laplacian_electron_density(pts), lap) < TOL
    assert relerr(mol.compute_electron_density_hessian(pts), ref["hess"]) < TOL
    v = np.random.default_rng(3).normal(size=(100, 3))
    assert np.allclose(mol.compute_norm_of_vector(v), D.norm(v), rtol=1e-15)


def test_edge_cases(ct):
    from chemtoolscuda_amd import synthetic

    w = synthetic.water(10)
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    assert mol.compute_density(np.zeros((0, 3))).shape == (0,)
    assert mol.compute_gradient(np.zeros((0, 3))).shape == (0, 3)
    one = mol.compute_density(w.points[:1])
    ref = oracle.density(w.basis.flat(), w.P, w.points[:1])["rho"]
    assert relerr(one, ref) < TOL
    with pytest.raises(Exception):
        mol.compute_density(np.zeros((5, 2)))
    with pytest.raises(Exception):
        mol.compute_density(w.points, precision="half")


def test_chunking_matches_single_pass(ct):
    """Forcing many small chunks gives identical results (chunk boundaries at 256 points)."""
    from chemtoolscuda_amd import synthetic

    w = synthetic.water(5000)
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    a = mol.compute(w.points, ["rho", "grad", "lap"])
    mol.set_workspace_limit(256 * 5 * 64 * 8 * 3)
    b = mol.compute(w.points, ["rho", "grad", "lap"])
    assert mol.last_report()["n_chunks"] > 3
    for k in a:
        assert np.array_equal(a[k], b[k])
