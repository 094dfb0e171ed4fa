"""GPU ESP (Hermite pre-contraction + polynomial form + table Boys function)
against the FP64 oracle (per-point McMurchie-Davidson V_ij, pinned by the
helium closed form, the Poisson equation and the far-field monopole), and the
device Boys function against 40-digit mpmath vectors.

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


@pytest.mark.parametrize("lmax_pairs", [4, 8])
def test_device_boys_against_mpmath_golden(lmax_pairs):
    """F_0..F_m(T) as the ESP kernels evaluate them (both table grids, every
    order), relative error against tests/golden/boys_mp.npz (mpmath hyp1f1 at
    40 digits): T on and between the grid points, across both table ends
    (T = 50 and 60), from 0 to 1e6."""
    from chemtoolscuda_amd import capi

    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "boys_mp.npz"))
    worst = 0.0
    for m in range(lmax_pairs + 1):
        F = capi.boys(g["T"], m, lmax_pairs)
        err = np.abs(F - g["F"][:, : m + 1]) / g["F"][:, : m + 1]
        worst = max(worst, float(err.max()))
    print(f"device Boys (pairs up to L = {lmax_pairs}): max rel. error {worst:.2e}")
    assert worst < 4e-15


@pytest.mark.parametrize("name,npts", [("atom_he.fchk", 300), ("atom_be.fchk", 200), ("h2o.fchk", 120),
                                       ("atom_c_g_pure_orbital.fchk", 60), ("atom_be_f_pure_orbital.fchk", 80),
                                       ("atom_c_g_cartesian_orbital.fchk", 60),
                                       ("atom_be_f_cartesian_orbital.fchk", 80),
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


def test_esp_cartesian_f_and_g_synthetic(ct):
    """Four atoms of the 300-atom workload's recipe (Cartesian s, p, d, f, g on
    heavy atoms and f on H: pair L up to 8), FP64 against the oracle, mixed
    against FP64."""
    from chemtoolscuda_amd import synthetic
    from chemtoolscuda_amd.basis import BasisSet

    rng = np.random.default_rng(11)
    xyz = np.array([[0.0, 0.0, 0.0], [2.6, 0.3, -0.4], [-1.1, 2.2, 0.5], [0.4, -0.9, 2.3]])
    els = ["C", "O", "H", "N"]
    shells = []
    for a, el in enumerate(els):
        shells += synthetic._tz_builder(rng, a, el)
    basis = BasisSet(shells, xyz, np.array([6, 8, 1, 7]))
    assert max(abs(s.l) for s in basis.shells) == 4
    P = synthetic._density_matrix(rng, basis.nbasis, 11)
    Z = np.array([6.0, 8.0, 1.0, 7.0])
    mol = ct.molecule_from_basis(basis, P, Z)
    pts = _points(xyz, 120, 12)
    ref = oracle.esp(basis.flat(), P, Z, pts)
    got = mol.compute_electrostatic_potential(pts)
    print(f"Cartesian f/g: FP64 max error {_check(got, ref, xyz, Z, pts):.2e}")
    mixed = mol.compute_electrostatic_potential(pts, precision="mixed")
    print(f"Cartesian f/g: mixed max error {_check(mixed, ref, xyz, Z, pts, tol=1e-6):.2e}")


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
    """Mixed ESP (pair geometry, Boys function F_0..F_L and the sum over pairs
    in FP64; each pair's Hermite polynomial and its sum over N in FP32) against
    the FP64 kernel, relative to |V_nuc| + |V_el|."""
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
