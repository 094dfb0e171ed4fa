"""Pins the numpy restatement of the density-based descriptors.

The descriptor formulas (reference src/evaluate_densbased.cu:44-223) are
checked against facts they do not share code with:
  * Weizsaecker: tau_W = |grad rho|^2 / (8 rho) equals tau_PD exactly for a
    one-orbital density (rho = n psi^2), a dropped factor of 2 fails it;
  * general KED at a = 0: tau_PD - lap/4 is the Schroedinger KED
    -1/2 sum_k n_k psi_k lap psi_k, built here from AO second derivatives;
  * Shannon: the integral of rho ln rho for a normalised Gaussian density has
    the closed form N ln(N (2a/pi)^{3/2}) - 3N/2 (a sign slip fails it);
  * Thomas-Fermi and RDG: their published constants C_F = 2.871234 and
    1/(2 (3 pi^2)^{1/3}) = 0.161620, the homogeneity tau_TF(l rho) =
    l^{5/3} tau_TF(rho), and the invariance of s under the coordinate scaling
    rho_g(r) = g^3 rho(g r);
  * masked divide: values at and below the 1e-20 mask.
"""
import numpy as np

from chemtoolscuda_amd import synthetic
from chemtoolscuda_amd.basis import BasisSet, Shell
from oracle import descriptors as D
from oracle import oracle


def _water_orbitals(n=200):
    w = synthetic.water(n)
    evals, evecs = np.linalg.eigh(w.P)
    return w, evals, evecs


def test_weizsacker_equals_tau_for_one_orbital_density():
    w, evals, evecs = _water_orbitals()
    flat = w.basis.flat()
    c = evecs[:, -1]
    P1 = 2.0 * np.outer(c, c)
    out = oracle.density(flat, P1, w.points, rho=True, grad=True, tau=True)
    keep = out["rho"] > 1e-12
    tw = D.weizsacker(out["rho"], out["grad"])
    assert np.allclose(tw[keep], out["tau"][keep], rtol=1e-9, atol=0)


def test_general_ked_at_zero_is_schroedinger_ked():
    w, evals, evecs = _water_orbitals()
    flat = w.basis.flat()
    out = oracle.density(flat, w.P, w.points, rho=True, lap=True, tau=True)
    ao = oracle.ao(flat, w.points, 2)
    keep = evals > 1e-10
    C = evecs[:, keep]
    psi = C.T @ ao[0]
    lap_psi = C.T @ (ao[4] + ao[7] + ao[9])
    tau_s = -0.5 * np.einsum("k,kn->n", evals[keep], psi * lap_psi)
    got = D.general_ked(out["tau"], out["lap"], 0.0)
    scale = np.abs(out["tau"]).max()
    assert np.allclose(got, tau_s, rtol=1e-10, atol=1e-12 * scale)
    # a = 1 leaves tau_PD; the C++ t_alpha at alpha = -1/4 is the same KED
    assert np.array_equal(D.general_ked(out["tau"], out["lap"], 1.0), out["tau"])
    assert np.allclose(D.general_kinetic_energy_density(out["tau"], out["lap"], -0.25), tau_s,
                       rtol=1e-10, atol=1e-12 * scale)


def test_shannon_integral_of_gaussian_density():
    a, nel = 0.7, 2.0
    basis = BasisSet([Shell(0, 0, [a], [1.0])], np.zeros((1, 3)))
    r = np.linspace(1e-6, 12.0, 120001)
    pts = np.stack([r, np.zeros_like(r), np.zeros_like(r)], axis=1)
    rho = oracle.density(basis.flat(), np.array([[nel]]), pts)["rho"]
    integrand = 4 * np.pi * r ** 2 * D.shannon(rho)
    val = np.sum(0.5 * (integrand[1:] + integrand[:-1]) * np.diff(r))
    exact = nel * np.log(nel * (2 * a / np.pi) ** 1.5) - 1.5 * nel
    assert abs(val - exact) < 1e-8 * abs(exact)


def test_thomas_fermi_and_rdg_constants_and_scaling():
    assert abs(D.thomas_fermi(np.array([1.0]))[0] - 2.871234000188191) < 1e-12
    g = np.array([[1.0, 0.0, 0.0]])
    assert abs(D.reduced_density_gradient(np.array([1.0]), g)[0] - 0.16162045967399548) < 1e-14
    rho = np.array([0.3, 1.7, 4e-3])
    assert np.allclose(D.thomas_fermi(2.5 * rho), 2.5 ** (5 / 3) * D.thomas_fermi(rho), rtol=1e-14)
    # coordinate scaling: a Gaussian with exponent a g^2 and the same electron
    # count is rho_g(r) = g^3 rho(g r); s_g(r) = s(g r)
    a, gam = 0.9, 1.6
    pts = np.random.default_rng(5).uniform(-1.5, 1.5, size=(40, 3))
    s = []
    for alpha, p in ((a, gam * pts), (a * gam ** 2, pts)):
        basis = BasisSet([Shell(0, 0, [alpha], [1.0])], np.zeros((1, 3)))
        out = oracle.density(basis.flat(), np.array([[2.0]]), p, rho=True, grad=True)
        s.append(D.reduced_density_gradient(out["rho"], out["grad"]))
    assert np.allclose(s[0], s[1], rtol=1e-12)


def test_masked_divide_mask_value():
    num = np.array([1.0, 2.0, 3.0, 4.0])
    den = np.array([2.0, 1e-20, 0.0, -1.0])
    assert np.array_equal(D.masked_divide(num, den), np.array([0.5, 1e-20, 1e-20, 1e-20]))
