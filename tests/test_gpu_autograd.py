"""Adjoint kernels and the autograd wrapper.

dL/dP and dL/dR_A from the HIP adjoint are checked against the oracle's
adjoint (itself pinned by finite differences in test_oracle.py) to 1e-10
relative, and the autograd gradients of L = sum_g w_g (rho_g - rho_t,g)^2
against central finite differences of the GPU FP64 forward pass to 1e-6
relative: h = 1e-4 for P (L is quadratic in P, so the central difference is
exact up to rounding) and, for the centres, Richardson-extrapolated central
differences at h = 2e-5 and 1e-5 (tight exponents up to 5000 make the plain
O(h^2) truncation error of the order of 1e-5 relative).

Mixed forward pass: for dL/dP the residual is formed as rho(P - P_t) (rho is
linear in P, so for fixed centres this is the same function of P): the
FP32-quality contraction is then accurate to ~1e-7 of the residual itself and
the per-component 1e-6 criterion holds.  For dL/dR_A (rho_t fixed, so no such
rewrite) the mixed forward's error, |d rho_g| <= 1e-6 rho_g, enters
dL/drho_g = 2 w_g (rho_g - rho_t,g) directly; the backward kernels are FP64,
so the gradient can differ from the FP64 finite difference by at most
sum_g 2 w_g 1e-6 rho_g |d rho_g / d R_A,k|, which the test computes per
component (d rho_g / d R_A = -2 sum_{i on A} grad phi_i (P phi)_i with the
oracle's AOs) and adds to the 1e-6 finite-difference tolerance.
"""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ct():
    import chemtoolscuda_amd

    return chemtoolscuda_amd


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_backward_matches_oracle(ct):
    from chemtoolscuda_amd import synthetic

    w = synthetic.diff40(3000)
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    wt = w.extra["weights"]
    dP, dR = mol.backward(w.points, wt)
    flat = w.basis.flat()
    dP_ref = oracle.drho_dP(flat, w.points, wt)
    dR_ref = oracle.drho_dR(flat, w.P, w.points, wt)
    assert _rel(dP, dP_ref) < 1e-10
    assert _rel(dR, dR_ref) < 1e-10
    assert np.allclose(dP, dP.T, rtol=0, atol=1e-12 * np.abs(dP).max())


def test_backward_chunked_and_host_device_inputs(ct):
    import torch
    from chemtoolscuda_amd import synthetic

    w = synthetic.water(5000)
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    wt = np.random.default_rng(1).uniform(size=5000)
    a = mol.backward(w.points, wt)
    mol.set_workspace_limit(256 * 4 * 128 * 8 * 3)
    b = mol.backward(w.points, wt)
    assert _rel(b[0], a[0]) < 1e-13 and _rel(b[1], a[1]) < 1e-12
    from chemtoolscuda_amd import capi
    h = capi.Handle(w.basis.flat(), w.P, w.charges)
    pts = torch.from_numpy(w.points).cuda()
    wt_d = torch.from_numpy(wt).cuda()
    dP = torch.empty((w.basis.nbasis,) * 2, dtype=torch.float64, device="cuda")
    dR = torch.empty((w.basis.natom, 3), dtype=torch.float64, device="cuda")
    h.backward(pts.data_ptr(), 5000, wt_d.data_ptr(), dP.data_ptr(), dR.data_ptr())
    assert _rel(dP.cpu().numpy(), a[0]) < 1e-13 and _rel(dR.cpu().numpy(), a[1]) < 1e-12


@pytest.mark.parametrize("workload,npts", [("water", 3000), ("diff40", 3000)])
def test_mixed_backward_against_fp64(ct, workload, npts):
    """Mixed adjoint (FP32 AOs, split-bf16 Gram matrix and adjoint contraction,
    FP64 partial sums) against the FP64 adjoint, per element relative to the
    condition scale of each sum: sum_g |w_g| |phi_i| |phi_j| for dL/dP and
    sum_g |w_g| 2 sum_{i on A} |d_k phi_i| (|P| |phi|)_i for dL/dR (floored at
    1e-14 of the largest), at 1e-6."""
    from chemtoolscuda_amd import synthetic

    w = synthetic.WORKLOADS[workload](npts)
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    wt = np.random.default_rng(3).uniform(-1.0, 1.0, size=w.points.shape[0])
    dP64, dR64 = mol.backward(w.points, wt)
    dPm, dRm = mol.backward(w.points, wt, precision="mixed")
    ao = oracle.ao(w.basis.flat(), w.points, 1)
    a0 = np.abs(ao[0])
    sP = (a0 * np.abs(wt)) @ a0.T
    # pairs of far-apart functions have a zero or negligible scale (their
    # products underflow): floor the scale at 1e-14 of the largest
    eP = np.max(np.abs(dPm - dP64) / np.maximum(sP, 1e-14 * sP.max()))
    y = np.abs(w.P) @ a0
    atom_of = np.repeat([s.atom for s in w.basis.shells], [s.nfunc for s in w.basis.shells])
    sR = np.zeros_like(dR64)
    for k in range(3):
        per_fn = 2.0 * np.abs(ao[1 + k]) * y @ np.abs(wt)
        np.add.at(sR[:, k], atom_of, per_fn)
    eR = np.max(np.abs(dRm - dR64) / np.maximum(sR, 1e-14 * sR.max()))
    print(f"{workload}: mixed adjoint max error / condition scale: dP {eP:.2e}, dR {eR:.2e}")
    assert eP <= 1e-6 and eR <= 1e-6, (eP, eR)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autograd_against_finite_differences(ct, precision):
    import torch
    from chemtoolscuda_amd import synthetic
    from chemtoolscuda_amd.autograd import DensityModel

    w = synthetic.diff40(4000)
    dev = torch.device("cuda")
    model = DensityModel(w.basis, w.charges, precision=precision)
    ref_model = DensityModel(w.basis, w.charges, precision="fp64")
    pts = torch.from_numpy(w.points).to(dev)
    wts = torch.from_numpy(w.extra["weights"]).to(dev)
    Pt = torch.from_numpy(w.extra["P_target"]).to(dev)
    R0 = torch.from_numpy(w.basis.atom_xyz.copy()).to(dev)
    with torch.no_grad():
        rho_t = ref_model.density(Pt, R0, pts)

    def loss(m, P, R):
        # the mixed forward holds its error to 1e-6 of the residual rho - rho_t
        tgt = rho_t if m.precision == "mixed" else None
        return torch.sum(wts * (m.density(P, R, pts, rho_target=tgt) - rho_t) ** 2)

    P = torch.from_numpy(w.P.copy()).to(dev).requires_grad_(True)
    R = R0.clone().requires_grad_(True)
    L = loss(model, P, R)
    L.backward()
    gP = P.grad.cpu().numpy()
    gR = R.grad.cpu().numpy()
    # central finite differences of the FP64 forward pass
    rng = np.random.default_rng(2)
    M = w.basis.nbasis
    with torch.no_grad():
        hP = 1e-4
        for _ in range(4):
            i, j = (int(v) for v in rng.integers(0, M, size=2))
            E = torch.zeros((M, M), dtype=torch.float64, device=dev)
            E[i, j] += 1.0
            E[j, i] += 1.0
            fd = (loss(ref_model, P + hP * E, R) - loss(ref_model, P - hP * E, R)).item() / (2 * hP)
            an = gP[i, j] + gP[j, i] if i != j else 2 * gP[i, i]
            assert abs(fd - an) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(gP).max()), (i, j, fd, an)
        def central(E, h):
            return (loss(ref_model, P, R + h * E) - loss(ref_model, P, R - h * E)).item() / (2 * h)

        for A in rng.choice(w.basis.natom, size=4, replace=False):
            for k in range(3):
                E = torch.zeros_like(R)
                E[A, k] = 1.0
                fd = (4 * central(E, 1e-5) - central(E, 2e-5)) / 3
                an = gR[A, k]
                assert abs(fd - an) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(gR).max()), (A, k, fd, an)


