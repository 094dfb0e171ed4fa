"""GPU ESP (Hermite pre-contraction + polynomial form + table Boys function)
against the FP64 oracle (per-point McMurchie-Davidson V_ij, pinned by the
helium closed form, the Poisson equation and the far-field monopole).

Tolerance: 1e-12 relative to the scale of the two terms, |V_nuc| + |V_el|,
because the ESP crosses zero where they cancel.
"""
import os

import numpy as np
import pytest

import fchk_py
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ct():
    import chemtoolscuda_amd

    return chemtoolscuda_amd


def _points(xyz, n, seed, spread=2.5):
    rng = np.random.default_rng(seed)
    c = xyz[rng.integers(0, xyz.shape[0], size=n)]
    pts = c + rng.normal(0.0, spread, size=(n, 3))
    d = np.min(np.linalg.norm(pts[:, None, :] - xyz[None], axis=2), axis=1)
    return pts[d > 0.05]


def _check(got, ref, xyz, Z, pts, tol=1e-12):
    nuc = np.sum(Z[None, :] / np.linalg.norm(pts[:, None, :] - xyz[None], axis=2), axis=1)
    scale = np.abs(nuc) + np.abs(nuc - ref)
    err = np.abs(got - ref) / scale
    assert err.max() < tol, err.max()
    return err.max()


@pytest.mark.parametrize("name,npts", [("atom_he.fchk", 300), ("atom_be.fchk", 200), ("h2o.fchk", 120),
                                       ("atom_c_g_pure_orbital.fchk", 60), ("atom_be_f_pure_orbital.fchk", 80),
                                       ("ch4.fchk", 60), ("test.fchk", 12)])
def test_esp_fchk(ct, name, npts):
    path = os.path.join(fchk_py.FCHK_DIR, name)
    d = fchk_py.load(path)
    mol = ct.Molecule(path)
    xyz = d["basis"].atom_xyz
    pts = _points(xyz, npts, 3)
    got = mol.compute_electrostatic_potential(pts)
    ref = oracle.esp(d["basis"].flat(), d["P"], d["charges"], pts)
    _check(got, ref, xyz, d["charges"], pts)


def test_esp_synthetic_cartesian_and_surface(ct):
    from chemtoolscuda_amd import synthetic

    w = synthetic.water(200)
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    pts = _points(w.basis.atom_xyz, 200, 4)
    _check(mol.compute_electrostatic_potential(pts), oracle.esp(w.basis.flat(), w.P, w.charges, pts),
           w.basis.atom_xyz, w.charges, pts)
    e = synthetic.esp30(2000)
    mol = ct.molecule_from_basis(e.basis, e.P, e.charges)
    pts = e.points[:24]
    got = mol.compute_electrostatic_potential(e.points)[:24]
    _check(got, oracle.esp(e.basis.flat(), e.P, e.charges, pts), e.basis.atom_xyz, e.charges, pts)


def test_esp_cache_invalidation_and_edge_cases(ct):
    from chemtoolscuda_amd import synthetic

    w = synthetic.water(50)
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    pts = _points(w.basis.atom_xyz, 50, 5)
    a = mol.compute_electrostatic_potential(pts)
    P2 = 0.5 * w.P
    mol.set_density_matrix(P2)
    b = mol.compute_electrostatic_potential(pts)
    _check(b, oracle.esp(w.basis.flat(), P2, w.charges, pts), w.basis.atom_xyz, w.charges, pts)
    xyz2 = w.basis.atom_xyz + np.array([0.3, -0.2, 0.1])
    mol.set_coordinates(xyz2)
    c = mol.compute_electrostatic_potential(pts)
    _check(c, oracle.esp(w.basis.with_coordinates(xyz2).flat(), P2, w.charges, pts), xyz2, w.charges, pts)
    assert not np.allclose(a, b)
    assert mol.compute_electrostatic_potential(np.zeros((0, 3))).shape == (0,)


@pytest.mark.parametrize("case", ["water", "esp30", "atom_c_g_pure_orbital.fchk", "ch4.fchk"])
def test_esp_mixed_against_fp64(ct, case):
    """Mixed ESP (Boys function, geometry and accumulation in FP64, Hermite
    polynomial in FP32) against the FP64 kernel, relative to |V_nuc| + |V_el|."""
    from chemtoolscuda_amd import synthetic

    if case.endswith(".fchk"):
        path = os.path.join(fchk_py.FCHK_DIR, case)
        d = fchk_py.load(path)
        mol = ct.Molecule(path)
        xyz, Z = d["basis"].atom_xyz, d["charges"]
        pts = _points(xyz, 400, 6)
    else:
        w = synthetic.WORKLOADS[case]()
        mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
        xyz, Z = w.basis.atom_xyz, w.charges
        pts = w.points[:4000] if case == "esp30" else _points(xyz, 2000, 6)
    ref = mol.compute_electrostatic_potential(pts)
    got = mol.compute_electrostatic_potential(pts, precision="mixed")
    err = _check(got, ref, xyz, Z, pts, tol=1e-6)
    print(f"{case}: mixed ESP max error / (|V_nuc| + |V_el|) = {err:.2e}")
