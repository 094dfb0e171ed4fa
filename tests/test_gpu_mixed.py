"""Mixed-precision path (FP32 AOs, three-piece bf16 split MFMA contraction,
FP64 epilogue, FP64 re-evaluation of flagged points) against the FP64 path
and the oracle.

Tolerances (the target of the mixed path): rho to <= 1e-6 relative where
rho >= 1e-8.  Gradient, Laplacian, tau_PD and Hessian are compared against
the magnitude of the terms they are summed from (the per-point condition
scale), also at 1e-6: |d rho| against 2 sum_i |d phi_i (P phi)_i| and so on,
because the gradient vanishes at critical points and the Laplacian changes
sign, where a pointwise relative error has no meaning.
"""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ct():
    import chemtoolscuda_amd

    return chemtoolscuda_amd


def _condition_scales(flat, P, pts):
    ao = oracle.ao(flat, pts, 2)
    aP = np.abs(P)
    phi = ao[0]
    y0 = aP @ np.abs(phi)
    s_rho = np.sum(np.abs(phi) * y0, axis=0)
    s_grad = np.sqrt(sum((2 * np.sum(np.abs(ao[1 + k]) * y0, axis=0)) ** 2 for k in range(3)))
    lapphi = ao[4] + ao[7] + ao[9]
    s_tau = 0.5 * sum(np.sum(np.abs(ao[1 + k]) * (aP @ np.abs(ao[1 + k])), axis=0) for k in range(3))
    s_lap = 2 * np.sum(np.abs(lapphi) * y0, axis=0) + 4 * s_tau
    return s_rho, s_grad, s_lap, s_tau


def _check(got, ref, scale, tol, mask):
    a = np.asarray(got).reshape(len(scale), -1)
    b = np.asarray(ref).reshape(len(scale), -1)
    err = np.linalg.norm(a - b, axis=1) / scale
    worst = float(err[mask].max())
    assert worst <= tol, worst
    return worst


@pytest.mark.parametrize("workload,npts", [("water", 6000), ("organic60", 20000)])
def test_mixed_matches_fp64_and_oracle(ct, workload, npts):
    from chemtoolscuda_amd import synthetic

    w = synthetic.WORKLOADS[workload]()
    rng = np.random.default_rng(21)
    pts = w.points[rng.choice(w.points.shape[0], size=min(npts, w.points.shape[0]), replace=False)]
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    names = ["rho", "grad", "lap", "tau_pd"]
    mix = mol.compute(pts, names, precision="mixed")
    rep = mol.last_report()
    f64 = mol.compute(pts, names, precision="fp64")
    rho = f64["rho"]
    mask = rho >= 1e-8
    rel = np.abs(mix["rho"] - rho) / np.abs(rho)
    assert rel[mask].max() <= 1e-6, rel[mask].max()
    # condition-scaled errors on a subsample (the oracle's AO arrays are large)
    sub = np.arange(0, pts.shape[0], max(1, pts.shape[0] // 1500))
    s_rho, s_grad, s_lap, s_tau = _condition_scales(w.basis.flat(), w.P, pts[sub])
    m = mask[sub]
    _check(mix["grad"][sub], f64["grad"][sub], s_grad, 1e-6, m)
    _check(mix["lap"][sub], f64["lap"][sub], s_lap, 1e-6, m)
    _check(mix["tau_pd"][sub], f64["tau_pd"][sub], s_tau, 1e-6, m)
    # and the FP64 path agrees with the oracle on the same points, every quantity
    ref = oracle.density(w.basis.flat(), w.P, pts[sub], rho=True, grad=True, lap=True, tau=True)
    assert np.max(np.abs(f64["rho"][sub] - ref["rho"]) / np.abs(ref["rho"])) < 1e-12
    # (condition scales: the Laplacian's two terms cancel, both sides round
    # relative to their size |lap - 4 tau| + 4 tau; grad against
    # max(|grad|, sqrt(8 rho tau)))
    tau = np.abs(ref["tau"])
    scales = {"grad": np.maximum(np.linalg.norm(ref["grad"], axis=1), np.sqrt(8 * np.abs(ref["rho"]) * tau)),
              "lap": np.abs(ref["lap"] - 4 * tau) + 4 * tau, "tau_pd": tau}
    for k, kr in (("grad", "grad"), ("lap", "lap"), ("tau_pd", "tau")):
        d = np.abs(f64[k][sub] - ref[kr]).reshape(len(sub), -1).max(axis=1)
        r = d / (scales[k] + 1e-10 * scales[k].max())
        assert r.max() < 1e-12, (workload, k, r.max())
    print(f"{workload}: max rel err rho {rel[mask].max():.2e}, re-evaluated {rep['n_reevaluated']} "
          f"of {pts.shape[0]}, max estimate {rep['max_error_estimate']:.2e}")


def test_mixed_hessian_and_rho_modes(ct):
    from chemtoolscuda_amd import synthetic

    w = synthetic.water(3000)
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    pts = w.points
    f64 = mol.compute(pts, ["rho", "grad", "hess"], precision="fp64")
    mix = mol.compute(pts, ["rho", "grad", "hess"], precision="mixed")
    mask = f64["rho"] >= 1e-8
    assert (np.abs(mix["rho"] - f64["rho"]) / f64["rho"])[mask].max() <= 1e-6
    hscale = np.abs(f64["hess"]).reshape(len(pts), -1).max(axis=1) + np.abs(f64["rho"])
    assert (np.abs(mix["hess"] - f64["hess"]).reshape(len(pts), -1).max(axis=1) / hscale)[mask].max() <= 2e-6
    for names in (["rho"], ["rho", "grad"]):
        m2 = mol.compute(pts, names, precision="mixed")
        assert (np.abs(m2["rho"] - f64["rho"]) / f64["rho"])[mask].max() <= 1e-6


def test_forced_reevaluation_reproduces_fp64(ct):
    """A zero-width tolerance flags every point: the mixed call then returns the
    FP64 results (to rounding: few re-evaluated points use the row-block-split
    FP64 contraction, whose partial sums are added in another order)."""
    from chemtoolscuda_amd import synthetic

    w = synthetic.water(2500)
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    names = ["rho", "grad", "lap", "tau_pd"]
    f64 = mol.compute(w.points, names, precision="fp64")
    mix = mol.compute(w.points, names, precision="mixed", mixed_tol=1e-300)
    rep = mol.last_report()
    floor_ok = np.abs(f64["rho"]) >= 1e-10
    assert rep["n_reevaluated"] == int(floor_ok.sum())
    for k in names:
        a = np.asarray(mix[k][floor_ok]).reshape(int(floor_ok.sum()), -1)
        b = np.asarray(f64[k][floor_ok]).reshape(int(floor_ok.sum()), -1)
        den = np.maximum(np.linalg.norm(b, axis=1), 1e-10 * np.abs(b).max())
        assert np.max(np.linalg.norm(a - b, axis=1) / den) < 1e-13, k


def test_mixed_error_model_is_conservative(ct):
    """Without re-evaluation, every point's actual error stays below the estimate
    gamma * |phi|^T|P||phi| that drives the re-evaluation (checked through the
    reported maximum estimate/|rho| over accepted points)."""
    from chemtoolscuda_amd import synthetic

    w = synthetic.organic60()
    rng = np.random.default_rng(5)
    pts = w.points[rng.choice(w.points.shape[0], size=30000, replace=False)]
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    raw = mol.compute(pts, ["rho"], precision="mixed", mixed_tol=1e30)
    rep = mol.last_report()
    f64 = mol.compute(pts, ["rho"], precision="fp64")
    mask = np.abs(f64["rho"]) >= 1e-10
    rel = np.abs(raw["rho"] - f64["rho"])[mask] / np.abs(f64["rho"][mask])
    print(f"no re-evaluation: max rel err {rel.max():.2e}, p99.9 {np.quantile(rel, 0.999):.2e}, "
          f"max estimate {rep['max_error_estimate']:.2e}, reevaluated {rep['n_reevaluated']}")
    assert rep["n_reevaluated"] == 0
    assert rel.max() <= rep["max_error_estimate"]


@pytest.mark.parametrize("workload,npts", [("water", 3000), ("organic60", 3000), ("large300", 400)])
def test_mixed_error_model_pointwise(ct, workload, npts):
    """Without re-evaluation, the error of every quantity stays below
    gamma * (its condition scale), gamma = 6u for these (FP64 running-sum)
    modes -- the estimate the finaliser applies to rho.  Printed: the largest
    ratio error / (u * scale), u = 2^-24, per quantity."""
    from chemtoolscuda_amd import synthetic

    w = synthetic.large300(npts) if workload == "large300" else synthetic.WORKLOADS[workload]()
    rng = np.random.default_rng(8)
    pts = w.points[rng.choice(w.points.shape[0], size=min(npts, w.points.shape[0]), replace=False)]
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    names = ["rho", "grad", "lap", "tau_pd"]
    raw = mol.compute(pts, names, precision="mixed", mixed_tol=1e30, route="density")
    assert mol.last_report()["n_reevaluated"] == 0
    f64 = mol.compute(pts, names, precision="fp64")
    scales = _condition_scales(w.basis.flat(), w.P, pts)
    u = 2.0 ** -24
    ratios = {}
    for k, sc in zip(names, scales):
        a = np.asarray(raw[k]).reshape(len(pts), -1)
        b = np.asarray(f64[k]).reshape(len(pts), -1)
        ratios[k] = float(np.max(np.linalg.norm(a - b, axis=1) / (u * sc)))
    # density-only mode: its own tile (triangle of P, doubled right of each
    # wave's diagonal block); the FP64 density-only path uses the same shortcut
    raw_r = mol.compute(pts, ["rho"], precision="mixed", mixed_tol=1e30, route="density")["rho"]
    f64_r = mol.compute(pts, ["rho"], precision="fp64")["rho"]
    ratios["rho (rho-only mode)"] = float(np.max(np.abs(raw_r - f64_r) / (u * scales[0])))
    assert np.max(np.abs(f64_r - f64["rho"]) / scales[0]) < 1e-14
    print(f"{workload} (M={w.basis.nbasis}): max error / (u * condition scale): "
          + ", ".join(f"{k} {v:.2f}" for k, v in ratios.items()))
    for k, v in ratios.items():
        assert v <= 6.0, (k, v)


def _pivoted_cholesky(P, rel_tol=1e-12):
    """numpy restatement of the library's factorisation of P (host code,
    chemtoolscuda_amd/csrc/mol.cpp): P = L L^T, columns in pivot order (largest
    remaining diagonal first), stop at rel_tol of the largest diagonal."""
    n = P.shape[0]
    d = np.diag(P).astype(np.float64).copy()
    tol = rel_tol * np.abs(d).max()
    used = np.zeros(n, dtype=bool)
    cols = []
    while not used.all():
        p = int(np.argmax(np.where(used, -np.inf, d)))
        if d[p] <= tol:
            break
        L = np.array(cols).T if cols else np.zeros((n, 0))
        c = (P[:, p] - L @ L[p]) / np.sqrt(d[p])
        c[used] = 0.0
        c[p] = np.sqrt(d[p])
        d[~used] -= c[~used] ** 2
        used[p] = True
        cols.append(c)
    return np.array(cols).T


def _factor_scales(flat, L, pts):
    """Condition scales of the factor route (psi = L^T phi): the first-order
    propagation of per-column bounds b = |L|^T |x| of every psi column into
    rho = psi.psi, grad = 2 psi.d psi, tau = 1/2 sum_k d_k psi.d_k psi and
    lap = 2 (psi.lap psi + sum_k d_k psi.d_k psi)."""
    ao = oracle.ao(flat, pts, 2)
    aL = np.abs(L)
    lapphi = ao[4] + ao[7] + ao[9]
    psi = [L.T @ ao[k] for k in range(4)]
    psil = L.T @ lapphi
    b = [aL.T @ np.abs(ao[k]) for k in range(4)]
    bl = aL.T @ np.abs(lapphi)
    s_rho = 2 * np.sum(np.abs(psi[0]) * b[0], axis=0)
    s_grad = np.sqrt(sum((2 * np.sum(np.abs(psi[k + 1]) * b[0] + np.abs(psi[0]) * b[k + 1], axis=0)) ** 2
                         for k in range(3)))
    s_tau = sum(np.sum(np.abs(psi[k + 1]) * b[k + 1], axis=0) for k in range(3))
    s_lap = 2 * np.sum(np.abs(psil) * b[0] + np.abs(psi[0]) * bl, axis=0) + 4 * s_tau
    return s_rho, s_grad, s_lap, s_tau


@pytest.mark.parametrize("workload,npts", [("water", 3000), ("organic60", 12000), ("large300", 400)])
def test_factor_route_error_model_pointwise(ct, workload, npts):
    """Factor route (P = L L^T, psi = L^T phi on the split-bf16 engine):
    without re-evaluation the error of rho stays below gamma = 3u times the
    bound the kernel computes and the finaliser applies, 2 sum_j |psi_j|
    (|L|^T |phi|)_j, and the other quantities below 6u times their
    factor-route condition scales.  Also the rho-only and grad modes, and the
    reported rank.  Spatial screening off: this is the error model of the
    split-bf16 engine at every point (screened tiles add a bound of their own,
    tests/test_gpu_screening.py)."""
    from chemtoolscuda_amd import synthetic

    w = synthetic.large300(npts) if workload == "large300" else synthetic.WORKLOADS[workload]()
    rng = np.random.default_rng(9)
    pts = w.points[rng.choice(w.points.shape[0], size=min(npts, w.points.shape[0]), replace=False)]
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    mol.set_screening(0.0)
    P = 0.5 * (w.P + w.P.T)
    L = _pivoted_cholesky(P)
    names = ["rho", "grad", "lap", "tau_pd"]
    raw = mol.compute(pts, names, precision="mixed", mixed_tol=1e30, route="factor")
    rep = mol.last_report()
    assert rep["n_reevaluated"] == 0 and rep["factor_rank"] == L.shape[1]
    f64 = mol.compute(pts, names, precision="fp64")
    scales = _factor_scales(w.basis.flat(), L, pts)
    u = 2.0 ** -24
    ratios = {}
    for k, sc in zip(names, scales):
        a = np.asarray(raw[k]).reshape(len(pts), -1)
        b = np.asarray(f64[k]).reshape(len(pts), -1)
        ratios[k] = float(np.max(np.linalg.norm(a - b, axis=1) / (u * sc)))
    for names2 in (["rho"], ["rho", "grad"]):
        r2 = mol.compute(pts, names2, precision="mixed", mixed_tol=1e30, route="factor")
        assert mol.last_report()["factor_rank"] == L.shape[1]
        for k in names2:
            i = names.index(k)
            a = np.asarray(r2[k]).reshape(len(pts), -1)
            b = np.asarray(f64[k]).reshape(len(pts), -1)
            ratios[f"{k} ({'+'.join(names2)} mode)"] = float(np.max(np.linalg.norm(a - b, axis=1) / (u * scales[i])))
    print(f"{workload} (M={w.basis.nbasis}, rank {L.shape[1]}): factor route, max error / (u * condition scale): "
          + ", ".join(f"{k} {v:.2f}" for k, v in ratios.items()))
    for k, v in ratios.items():
        assert v <= (3.0 if k.startswith("rho") else 6.0), (k, v)
    # with re-evaluation (and screening): the 1e-6 target on rho
    mol.set_screening(1e-15)
    mix = mol.compute(pts, ["rho"], precision="mixed", route="factor")["rho"]
    m = f64["rho"] >= 1e-8
    assert np.max(np.abs(mix - f64["rho"])[m] / f64["rho"][m]) <= 1e-6


def test_route_selection_and_errors(ct):
    from chemtoolscuda_amd import synthetic

    w = synthetic.water(500)
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    mol.compute(w.points, ["rho"], precision="mixed")
    assert mol.last_report()["factor_rank"] == -1  # rank 5 in a 128-row padding: the density route is cheaper
    mol.compute(w.points, ["rho"], precision="fp64")
    assert mol.last_report()["factor_rank"] == -1  # the FP64 path contracts with P unless asked
    mol.compute(w.points, ["rho"], precision="fp64", route="factor")
    assert mol.last_report()["factor_rank"] == 5
    # an indefinite P has no factor: the factor route is refused, auto falls back
    P2 = w.P.copy()
    P2[0, 0] -= 10.0
    mol.set_density_matrix(P2)
    with pytest.raises(RuntimeError, match="not positive semidefinite"):
        mol.compute(w.points, ["rho"], precision="mixed", route="factor")
    a = mol.compute(w.points, ["rho"], precision="mixed")["rho"]
    b = mol.compute(w.points, ["rho"], precision="fp64")["rho"]
    m = np.abs(b) >= 1e-8
    assert np.max(np.abs(a - b)[m] / np.abs(b)[m]) <= 1e-6


def test_indefinite_P_with_zero_diagonal_has_no_factor(ct):
    """A pivoted Cholesky sees only the diagonal: a symmetric P with a zero
    diagonal and non-zero off-diagonal elements (indefinite: trace 0) would
    pass as "rank 0" and give rho = 0.  The handle checks the residual
    P - L L^T as well, so such a P has no factor: the factor route is refused
    in both precisions and the mixed AUTO route contracts with P."""
    from chemtoolscuda_amd import synthetic

    w = synthetic.water(300)
    P = np.zeros_like(w.P)
    P[0, 1] = P[1, 0] = 0.6
    P[1, 2] = P[2, 1] = 0.3
    mol = ct.molecule_from_basis(w.basis, P, w.charges)
    for prec in ("fp64", "mixed"):
        with pytest.raises(RuntimeError, match="not positive semidefinite"):
            mol.compute(w.points, ["rho"], precision=prec, route="factor")
    o = synthetic.organic60()
    rng = np.random.default_rng(51)
    E = rng.normal(0.0, 1e-2, size=o.P.shape)
    P = E + E.T
    np.fill_diagonal(P, 0.0)
    mol = ct.molecule_from_basis(o.basis, P, o.charges)
    pts = o.points[rng.choice(o.points.shape[0], 4000, replace=False)]
    auto = mol.compute(pts, ["rho", "grad"], precision="mixed")
    assert mol.last_report()["factor_rank"] == -1  # (a rank-0 factor would have been the cheaper route)
    dens = mol.compute(pts, ["rho", "grad"], precision="mixed", route="density")
    for q in ("rho", "grad"):
        assert np.array_equal(auto[q], dens[q])
    ref = mol.compute(pts, ["rho"], precision="fp64")["rho"]
    assert np.max(np.abs(ref)) > 1e-3  # rho is far from the zero a rank-0 factor gives


def test_device_resident_chunks_match_host_path(ct):
    """Device-resident points and outputs take the batched path (flagged
    points of all chunks re-evaluated in one FP64 pass at the end); host arrays
    re-evaluate chunk by chunk through pinned staging.  Per point both run the
    same kernels, so the results agree bit for bit, also with many small
    chunks (spatial screening off: with it, the point tiles, and so the
    skipped blocks, follow the chunking)."""
    import torch
    from chemtoolscuda_amd import capi, synthetic

    w = synthetic.organic60()
    rng = np.random.default_rng(12)
    pts = np.ascontiguousarray(w.points[rng.choice(w.points.shape[0], size=20000, replace=False)])
    n = len(pts)
    names = ["rho", "grad", "lap", "tau_pd"]
    h = capi.Handle(w.basis.flat(), w.P, w.charges)
    h.set_screening(0.0)
    host = {q: np.zeros((n, capi.WIDTH.get(q, 1))) for q in names}
    h.eval(pts.ctypes.data, n, {q: a.ctypes.data for q, a in host.items()}, "mixed")
    dev_pts = torch.from_numpy(pts).cuda()
    for limit_mb in (0, 24):  # library default (one chunk), then ~1 800 points per chunk
        if limit_mb:
            h.set_workspace_limit(limit_mb << 20)
        dev = {q: torch.empty((n, capi.WIDTH.get(q, 1)), dtype=torch.float64, device="cuda") for q in names}
        torch.cuda.synchronize()
        rep = h.eval(dev_pts.data_ptr(), n, {q: t.data_ptr() for q, t in dev.items()}, "mixed")
        if limit_mb:
            assert rep["n_chunks"] >= 6
        for q in names:
            assert np.array_equal(dev[q].cpu().numpy(), host[q]), (limit_mb, q)


def _factor_hess_scale(flat, L, pts):
    """Condition scale of the factor route's Hessian, H_kl = 2 sum_j (psi_j
    d_kl psi_j + d_k psi_j d_l psi_j): 2 sum_j (|d_kl psi_j| b_j + |psi_j|
    b_kl,j + |d_k psi_j| b_l,j + |d_l psi_j| b_k,j) with b_x = |L|^T |x|,
    largest over kl."""
    ao = oracle.ao(flat, pts, 2)
    aL = np.abs(L)
    psi = [L.T @ a for a in ao]
    b = [aL.T @ np.abs(a) for a in ao]
    idx = {(0, 0): 4, (0, 1): 5, (0, 2): 6, (1, 1): 7, (1, 2): 8, (2, 2): 9}
    out = np.zeros(pts.shape[0])
    for (k, l), i in idx.items():
        s = 2 * np.sum(np.abs(psi[i]) * b[0] + np.abs(psi[0]) * b[i] + np.abs(psi[1 + k]) * b[1 + l]
                       + np.abs(psi[1 + l]) * b[1 + k], axis=0)
        out = np.maximum(out, s)
    return out


@pytest.mark.parametrize("workload,npts", [("water", 2000), ("organic60", 2000), ("large300", 300)])
def test_factor_route_hessian_pointwise(ct, workload, npts):
    """Hessian mode on the factor route (three launches: {phi, d phi} for rho,
    grad and t_kl; {phi, d_kl phi} twice for h_kl): without re-evaluation rho
    within 3u of its bound, the Hessian within 6u of its factor-route scale;
    with re-evaluation rho to 1e-6 and rho ln rho to 1e-6 of |rho ln rho| + rho.
    Spatial screening off for the pointwise model, on for the 1e-6 targets."""
    from chemtoolscuda_amd import synthetic

    w = synthetic.large300(npts) if workload == "large300" else synthetic.WORKLOADS[workload]()
    rng = np.random.default_rng(13)
    pts = w.points[rng.choice(w.points.shape[0], size=min(npts, w.points.shape[0]), replace=False)]
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    mol.set_screening(0.0)
    L = _pivoted_cholesky(0.5 * (w.P + w.P.T))
    names = ["rho", "grad", "hess"]
    raw = mol.compute(pts, names, precision="mixed", mixed_tol=1e30, route="factor")
    assert mol.last_report()["factor_rank"] == L.shape[1]
    f64 = mol.compute(pts, names + ["rdg", "shannon"], precision="fp64")
    s_rho, s_grad, _, _ = _factor_scales(w.basis.flat(), L, pts)
    s_hess = _factor_hess_scale(w.basis.flat(), L, pts)
    u = 2.0 ** -24
    r_rho = float(np.max(np.abs(raw["rho"] - f64["rho"]) / (u * s_rho)))
    r_grad = float(np.max(np.linalg.norm(raw["grad"] - f64["grad"], axis=1) / (u * s_grad)))
    r_hess = float(np.max(np.abs(raw["hess"] - f64["hess"]).reshape(len(pts), 9).max(axis=1) / (u * s_hess)))
    print(f"{workload} (M={w.basis.nbasis}, rank {L.shape[1]}): factor route Hessian mode, max error / (u * scale): "
          f"rho {r_rho:.2f}, grad {r_grad:.2f}, hess {r_hess:.2f}")
    assert r_rho <= 3.0 and r_grad <= 6.0 and r_hess <= 6.0, (r_rho, r_grad, r_hess)
    mol.set_screening(1e-15)
    mix = mol.compute(pts, ["rho", "hess", "rdg", "shannon"], precision="mixed", route="factor")
    m = f64["rho"] >= 1e-8
    rho = f64["rho"][m]
    assert np.max(np.abs(mix["rho"][m] - rho) / rho) <= 1e-6
    sh = f64["shannon"][m]
    assert np.max(np.abs(mix["shannon"][m] - sh) / (np.abs(sh) + rho)) <= 1e-6


@pytest.mark.parametrize("workload,npts", [("water", 3000), ("organic60", 4000)])
def test_fp64_factor_route_matches_density_route(ct, workload, npts):
    """FP64 on the factor route (route="factor"; also what re-evaluates the
    mixed factor route's flagged points) against FP64 on P, every mode: the
    same quantities summed in another order -- rho to 1e-12 relative, the
    others to 1e-11 of the magnitude of their terms (the Laplacian against
    |lap - 4 tau| + 4 tau, the rest against their norm floored at 1e-10 of the
    largest)."""
    from chemtoolscuda_amd import synthetic

    w = synthetic.WORKLOADS[workload]()
    rng = np.random.default_rng(14)
    pts = w.points[rng.choice(w.points.shape[0], size=min(npts, w.points.shape[0]), replace=False)]
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    for names in (["rho"], ["rho", "grad"], ["rho", "grad", "lap", "tau_pd"], ["rho", "grad", "hess"]):
        a = mol.compute(pts, names, precision="fp64", route="factor")
        assert mol.last_report()["factor_rank"] > 0
        b = mol.compute(pts, names, precision="fp64")
        for k in names:
            x = np.asarray(a[k]).reshape(len(pts), -1)
            y = np.asarray(b[k]).reshape(len(pts), -1)
            den = np.maximum(np.linalg.norm(y, axis=1), 1e-10 * np.abs(y).max())
            if k == "lap":
                tau = b["tau_pd"]
                den = np.abs(b["lap"] - 4 * tau) + 4 * tau
            tol = 1e-12 if k == "rho" else 1e-11
            assert np.max(np.linalg.norm(x - y, axis=1) / den) < tol, (names, k)
