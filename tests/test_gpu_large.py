"""The 300-atom workload: M = 5400 Cartesian functions with f on H and f+g on
the heavy atoms, points drawn per atom with radius ~ Gamma(3, 1).

FP64 path against the oracle (1e-12, norm-wise per point as in
test_gpu_density.py) for rho, grad, Hessian, Laplacian, tau_PD, RDG and
Shannon on a subsample; the split-BF16 mixed path against the FP64 path for
the Hessian, RDG and Shannon outputs, each error taken against the per-point
condition scale of its quantity (the magnitude of the terms it is summed
from, computed from |P| and |AO| with the oracle), at 1e-6.
"""
import numpy as np
import pytest

from oracle import descriptors as D
from oracle import oracle

pytestmark = pytest.mark.gpu

PAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]  # oracle second-derivative slots 4..9


@pytest.fixture(scope="module")
def large():
    import chemtoolscuda_amd
    from chemtoolscuda_amd import synthetic

    w = synthetic.large300(6000)
    mol = chemtoolscuda_amd.molecule_from_basis(w.basis, w.P, w.charges)
    names = ["rho", "grad", "hess", "lap", "tau_pd", "rdg", "shannon"]
    f64 = mol.compute(w.points, names, precision="fp64")
    mix = mol.compute(w.points, names, precision="mixed")
    return w, f64, mix, mol.last_report()


def _normwise(got, ref, floor):
    a = np.asarray(got).reshape(ref.shape[0], -1)
    b = np.asarray(ref).reshape(ref.shape[0], -1)
    den = np.maximum(np.linalg.norm(b, axis=1), floor)
    return float(np.max(np.linalg.norm(a - b, axis=1) / den))


def test_large_fp64_matches_oracle(large):
    w, f64, _, _ = large
    sub = np.arange(0, w.points.shape[0], 30)
    ref = oracle.density(w.basis.flat(), w.P, w.points[sub], rho=True, grad=True, hess=True, lap=True, tau=True)
    for k, rk in (("rho", "rho"), ("grad", "grad"), ("hess", "hess"), ("lap", "lap"), ("tau_pd", "tau")):
        r = ref[rk]
        floor = 1e-10 * np.abs(r).max()
        assert _normwise(f64[k][sub], r, floor) < 1e-12, k
    rho, grad = ref["rho"], ref["grad"]
    ok = rho > 1e-10
    rdg = D.reduced_density_gradient(rho, grad)
    assert np.max(np.abs(f64["rdg"][sub] - rdg)[ok] / np.abs(rdg[ok])) < 1e-11
    sh = D.shannon(rho)
    assert np.max(np.abs(f64["shannon"][sub] - sh)[ok] / (np.abs(sh[ok]) + rho[ok])) < 1e-11


def test_large_mixed_hessian_rdg_shannon(large):
    w, f64, mix, rep = large
    sub = np.arange(0, w.points.shape[0], 20)
    pts = w.points[sub]
    ao = oracle.ao(w.basis.flat(), pts, 2)
    aP = np.abs(w.P)
    a = np.abs(ao)
    y = [aP @ a[d] for d in range(4)]
    s_rho = np.sum(a[0] * y[0], axis=0)
    s_grad = np.sqrt(sum((2 * np.sum(a[1 + k] * y[0], axis=0)) ** 2 for k in range(3)))
    s_hess = np.zeros((len(sub), 3, 3))
    for slot, (k, l) in enumerate(PAIRS):
        s = 2 * np.sum(a[4 + slot] * y[0], axis=0) + 2 * np.sum(a[1 + k] * y[1 + l], axis=0)
        s_hess[:, k, l] = s_hess[:, l, k] = s
    rho = f64["rho"][sub]
    gnorm = np.linalg.norm(f64["grad"][sub], axis=1)
    mask = (rho >= 1e-8) & (gnorm > 0)

    err_rho = np.abs(mix["rho"][sub] - rho) / rho
    err_h = np.max(np.abs(mix["hess"][sub] - f64["hess"][sub]).reshape(-1, 9) / s_hess.reshape(-1, 9), axis=1)
    s = f64["rdg"][sub]
    s_scale = s * (s_grad / gnorm + 4.0 / 3.0 * s_rho / rho)
    err_s = np.abs(mix["rdg"][sub] - s) / s_scale
    sh_scale = np.abs(np.log(rho) + 1.0) * s_rho
    err_sh = np.abs(mix["shannon"][sub] - f64["shannon"][sub]) / sh_scale
    stats = {k: float(v[mask].max()) for k, v in
             (("rho", err_rho), ("hess", err_h), ("rdg", err_s), ("shannon", err_sh))}
    print(f"large300 mixed vs fp64 (max over {int(mask.sum())} points): {stats}; "
          f"re-evaluated {rep['n_reevaluated']} of {w.points.shape[0]}")
    for k, v in stats.items():
        assert v <= 1e-6, (k, v)
    # rho itself, over every point
    full = f64["rho"] >= 1e-8
    assert np.max(np.abs(mix["rho"] - f64["rho"])[full] / f64["rho"][full]) <= 1e-6
