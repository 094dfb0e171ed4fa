"""Pins the FP64 host oracle before anything is compared against it.

Every oracle function is checked against something it does not share code
with: the wavefunction itself (C^T S C = I, Tr PS = N_e, Gaussian's own
"Total SCF Density"), closed forms, finite differences, quadrature, the
Poisson equation, and scipy's hyp1f1 (golden Boys vectors).
"""
import os

import numpy as np
import pytest

import fchk_py
from chemtoolscuda_amd import synthetic
from chemtoolscuda_amd.basis import BasisSet, Shell
from oracle import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("path", fchk_py.fixture_paths(), ids=os.path.basename)
def test_fchk_orthonormal_mos_and_scf_density(path):
    """Normalisation, component order and pure transforms: C^T S C = I."""
    d = fchk_py.load(path)
    flat = d["basis"].flat()
    S = oracle.overlap(flat)
    for C in (d["C_alpha"], d["C_beta"]):
        assert np.abs(C.T @ S @ C - np.eye(C.shape[1])).max() < 1e-7
    assert abs(np.trace(d["P"] @ S) - d["n_electrons"]) < 1e-6
    assert np.abs(d["P"] - d["scf_density"]).max() < 1e-7


def test_cartesian_fg_components_unit_normalised():
    """Every Cartesian f and g component of a one-primitive shell has S_ii = 1."""
    basis = BasisSet([Shell(0, 3, [0.7], [1.0]), Shell(0, 4, [1.3], [1.0]), Shell(0, -4, [0.9], [1.0])],
                     np.zeros((1, 3)))
    S = oracle.overlap(basis.flat())
    assert np.allclose(np.diag(S), 1.0, atol=1e-13)


def test_boys_against_hyp1f1_golden():
    g = np.load(os.path.join(GOLD, "boys.npz"))
    F = oracle.boys(16, g["T"])
    assert np.allclose(F[:, :9], g["F"][:, :9], rtol=1e-12, atol=0)
    assert np.allclose(F, g["F"], rtol=1e-11, atol=0)
    assert np.allclose(oracle.boys(8, [0.0])[0], 1.0 / (2 * np.arange(9) + 1), rtol=0, atol=1e-15)


def test_boys_against_mpmath_golden():
    """40-digit mpmath vectors (tests/golden/make_boys_mp.py): on and between
    the 0.125 grid points, around T = 50 and 60, tiny and large T."""
    g = np.load(os.path.join(GOLD, "boys_mp.npz"))
    F = oracle.boys(8, g["T"])
    err = np.abs(F - g["F"]) / g["F"]
    assert err.max() < 4e-15, err.max()


def test_ao_closed_forms_s_and_p():
    a, c = 0.83, 0.6
    basis = BasisSet([Shell(0, 0, [a], [c]), Shell(0, 1, [a], [c])], np.array([[0.1, -0.2, 0.3]]))
    rng = np.random.default_rng(0)
    pts = rng.uniform(-2, 2, size=(50, 3))
    ao = oracle.ao(basis.flat(), pts, 0)[0]
    d = pts - basis.atom_xyz[0]
    e = np.exp(-a * np.sum(d * d, axis=1))
    ns = (2 * a / np.pi) ** 0.75
    npp = np.sqrt((2 * a / np.pi) ** 1.5 * 4 * a)
    assert np.allclose(ao[0], c * ns * e, rtol=1e-14)
    for k in range(3):
        assert np.allclose(ao[1 + k], c * npp * d[:, k] * e, rtol=1e-13, atol=1e-300)


def test_pure_d_equals_cartesian_combination():
    a = 1.1
    basis = BasisSet([Shell(0, 2, [a], [1.0]), Shell(0, -2, [a], [1.0])], np.zeros((1, 3)))
    pts = np.random.default_rng(1).uniform(-1.5, 1.5, size=(40, 3))
    ao = oracle.ao(basis.flat(), pts, 0)[0]
    xx, yy, zz, xy, xz, yz = ao[:6]
    c0, c1, s1, c2, s2 = ao[6:]
    assert np.allclose(c0, (2 * zz - xx - yy) / 2, atol=1e-14)
    assert np.allclose(c1, xz, atol=1e-14)
    assert np.allclose(s1, yz, atol=1e-14)
    assert np.allclose(c2, np.sqrt(3) / 2 * (xx - yy), atol=1e-14)
    assert np.allclose(s2, xy, atol=1e-14)


def _mixed_basis():
    rng = np.random.default_rng(7)
    shells = []
    for l in (0, 1, 2, -2, 3, -3, 4, -4):
        n = 2
        shells.append(Shell(int(rng.integers(0, 2)), l, rng.uniform(0.3, 2.0, n), rng.uniform(0.2, 1.0, n)))
    return BasisSet(shells, np.array([[0.0, 0.0, 0.0], [0.4, -0.3, 0.9]]))


def test_ao_derivatives_match_finite_differences():
    basis = _mixed_basis()
    flat = basis.flat()
    pts = np.random.default_rng(2).uniform(-1.5, 1.5, size=(30, 3))
    ao = oracle.ao(flat, pts, 2)
    h = 1e-5
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        fp = oracle.ao(flat, pts + e, 1)
        fm = oracle.ao(flat, pts - e, 1)
        fd = (fp[0] - fm[0]) / (2 * h)
        assert np.allclose(ao[1 + k], fd, atol=2e-8 * np.abs(ao[1 + k]).max())
        # second derivatives column k: d/dk of (d/dx, d/dy, d/dz)
        fd2 = (fp[1:4] - fm[1:4]) / (2 * h)
        hidx = {0: (4, 5, 6), 1: (5, 7, 8), 2: (6, 8, 9)}[k]
        for j in range(3):
            ref = ao[hidx[j]]
            assert np.allclose(ref, fd2[j], atol=2e-7 * np.abs(ref).max())


def _sphere_quadrature(nr=160, nth=36, nph=72, rm=1.0):
    x, wx = np.polynomial.legendre.leggauss(nr)
    r = rm * (1 + x) / (1 - x)
    wr = wx * 2 * rm / (1 - x) ** 2 * r ** 2
    ct, wt = np.polynomial.legendre.leggauss(nth)
    ph = np.arange(nph) * 2 * np.pi / nph
    st = np.sqrt(1 - ct ** 2)
    R, CT, PH = np.meshgrid(r, ct, ph, indexing="ij")
    ST = np.sqrt(1 - CT ** 2)
    pts = np.stack([R * ST * np.cos(PH), R * ST * np.sin(PH), R * CT], axis=-1).reshape(-1, 3)
    w = (wr[:, None, None] * wt[None, :, None] * (2 * np.pi / nph)).repeat(nph, axis=2).reshape(-1)
    del st
    return pts, w


@pytest.mark.parametrize("name", ["atom_he.fchk", "atom_be.fchk", "atom_o.fchk",
                                  "atom_c_g_pure_orbital.fchk", "atom_08_O_N09_M2_ub3lyp_ccpvtz_g09.fchk"])
def test_density_integrates_to_electron_count(name):
    d = fchk_py.load(os.path.join(fchk_py.FCHK_DIR, name))
    flat = d["basis"].flat()
    pts, w = _sphere_quadrature()
    pts = pts + d["basis"].atom_xyz[0]
    rho = oracle.density(flat, d["P"], pts)["rho"]
    assert abs(np.dot(w, rho) - d["n_electrons"]) < 2e-6 * d["n_electrons"]


def test_helium_density_is_p11_phi_squared_and_derivatives_consistent():
    d = fchk_py.load(os.path.join(fchk_py.FCHK_DIR, "atom_he.fchk"))
    flat = d["basis"].flat()
    pts = np.random.default_rng(3).uniform(-2, 2, size=(64, 3))
    phi = oracle.ao(flat, pts, 0)[0, 0]
    out = oracle.density(flat, d["P"], pts, rho=True, grad=True, hess=True, lap=True, tau=True)
    assert np.allclose(out["rho"], d["P"][0, 0] * phi ** 2, rtol=1e-14)
    assert np.allclose(out["lap"], np.trace(out["hess"], axis1=1, axis2=2), rtol=1e-12, atol=1e-300)


def test_density_derivatives_against_finite_differences():
    w = synthetic.water(64)
    flat = w.basis.flat()
    pts = w.points
    out = oracle.density(flat, w.P, pts, rho=True, grad=True, hess=True, lap=True, tau=True)
    h = 1e-5
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        op = oracle.density(flat, w.P, pts + e, rho=True, grad=True)
        om = oracle.density(flat, w.P, pts - e, rho=True, grad=True)
        fd = (op["rho"] - om["rho"]) / (2 * h)
        assert np.allclose(out["grad"][:, k], fd, rtol=1e-6, atol=1e-9 * np.abs(fd).max())
        fdg = (op["grad"] - om["grad"]) / (2 * h)
        assert np.allclose(out["hess"][:, :, k], fdg, rtol=1e-5, atol=1e-8 * np.abs(fdg).max())
    # tau_PD = 1/2 sum_k |grad psi_k|^2 with P = 2 C C^T  (positive, orbital form)
    ao = oracle.ao(flat, pts, 1)
    evals, evecs = np.linalg.eigh(w.P)
    keep = evals > 1e-10
    psi_grad = np.einsum("mk,dmn->dkn", evecs[:, keep], ao[1:4])
    tau = 0.5 * np.einsum("k,dkn->n", evals[keep], psi_grad ** 2)
    assert np.allclose(out["tau"], tau, rtol=1e-10, atol=1e-14)


def test_esp_helium_closed_form():
    from scipy.special import erf
    d = fchk_py.load(os.path.join(fchk_py.FCHK_DIR, "atom_he.fchk"))
    sh = d["basis"].shells[0]
    pts = np.random.default_rng(4).uniform(-3, 3, size=(40, 3)) + d["basis"].atom_xyz[0]
    esp = oracle.esp(d["basis"].flat(), d["P"], d["charges"], pts)
    r = np.linalg.norm(pts - d["basis"].atom_xyz[0], axis=1)
    N = (2 * sh.exps / np.pi) ** 0.75
    ref = d["charges"][0] / r
    for p in range(sh.exps.size):
        for q in range(sh.exps.size):
            g = sh.exps[p] + sh.exps[q]
            ref -= d["P"][0, 0] * sh.coefs[p] * sh.coefs[q] * N[p] * N[q] * (np.pi / g) ** 1.5 * erf(np.sqrt(g) * r) / r
    assert np.allclose(esp, ref, rtol=1e-12, atol=1e-13)


def test_esp_poisson_equation():
    """Laplacian of the ESP equals 4 pi rho away from the nuclei."""
    w = synthetic.water(8)
    flat = w.basis.flat()
    rng = np.random.default_rng(5)
    dirs = rng.standard_normal((8, 3))
    pts = 1.2 * dirs / np.linalg.norm(dirs, axis=1, keepdims=True) + np.array([0.0, 0.0, 0.6])
    h = 2e-3
    lap = -6 * oracle.esp(flat, w.P, w.charges, pts)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        lap += oracle.esp(flat, w.P, w.charges, pts + e) + oracle.esp(flat, w.P, w.charges, pts - e)
    lap /= h * h
    rho = oracle.density(flat, w.P, pts)["rho"]
    assert np.allclose(lap, 4 * np.pi * rho, rtol=2e-4, atol=1e-4)


def test_esp_far_field_monopole():
    """r * ESP -> Z_tot - N_e; Richardson over r and 2r removes the dipole term."""
    w = synthetic.water(4)
    flat = w.basis.flat()
    u = np.array([[1.0, 0, 0], [0, 0.6, 0.8], [-0.48, 0.64, -0.6]])
    r1, r2 = 400.0, 800.0
    e1 = oracle.esp(flat, w.P, w.charges, r1 * u)
    e2 = oracle.esp(flat, w.P, w.charges, r2 * u)
    q = 2.0 * r2 * e2 - r1 * e1
    n_el = np.trace(w.P @ oracle.overlap(flat))
    assert np.allclose(q, w.charges.sum() - n_el, rtol=0, atol=1e-3)


def test_adjoints_against_finite_differences():
    w = synthetic.water(200)
    flat = w.basis.flat()
    pts = w.points
    wt = np.random.default_rng(6).uniform(0, 1, size=pts.shape[0])
    dP = oracle.drho_dP(flat, pts, wt)
    rng = np.random.default_rng(8)
    for _ in range(3):
        i, j = rng.integers(0, w.basis.nbasis, size=2)
        E = np.zeros_like(w.P)
        E[i, j] += 1
        E[j, i] += 1
        h = 1e-3
        Lp = np.dot(wt, oracle.density(flat, w.P + h * E, pts)["rho"])
        Lm = np.dot(wt, oracle.density(flat, w.P - h * E, pts)["rho"])
        fd = (Lp - Lm) / (2 * h)
        assert np.isclose(fd, dP[i, j] + dP[j, i] if i != j else 2 * dP[i, i], rtol=1e-9)
    dR = oracle.drho_dR(flat, w.P, pts, wt)
    h = 1e-5
    for A in range(w.basis.natom):
        for k in range(3):
            xyz = w.basis.atom_xyz.copy()
            xyz[A, k] += h
            Lp = np.dot(wt, oracle.density(w.basis.with_coordinates(xyz).flat(), w.P, pts)["rho"])
            xyz[A, k] -= 2 * h
            Lm = np.dot(wt, oracle.density(w.basis.with_coordinates(xyz).flat(), w.P, pts)["rho"])
            assert np.isclose((Lp - Lm) / (2 * h), dR[A, k], rtol=1e-6, atol=1e-9)
