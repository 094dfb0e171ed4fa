"""ct_eval_noalloc (the allocation-free variant of ct_eval) against ct_eval.

Through the C ABI (ctypes) on device arrays: two calls in a row, the second
with fewer points, FP64 and mixed (screened factor route, flagged points
re-evaluated chunk by chunk in the scratch), rows and columns.  FP64 on the
density route (the reference's formulation) must be the same bits as
ct_eval; on the screened factor route, whose point tiles follow the chunks,
within 1e-13 of each quantity's scale; mixed must meet the mixed path's bar
(1e-6 of each quantity's scale against FP64 where rho >= 1e-8).  The library's own device
buffers must not grow during the calls, and host pointers, too many points
and unknown layouts are rejected with a message.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    return torch


def _outputs(torch, qs, n):
    from chemtoolscuda_amd import capi

    return {q: torch.full((n, capi.WIDTH.get(q, 1)), np.nan, dtype=torch.float64, device="cuda") for q in qs}


def _run(torch, h, fn, pts, qs, prec, params=None):
    outs = _outputs(torch, qs, pts.shape[0])
    torch.cuda.synchronize()
    rep = fn(pts.data_ptr(), pts.shape[0], {q: t.data_ptr() for q, t in outs.items()}, prec, params)
    torch.cuda.synchronize()
    return outs, rep


def _scale(torch, q, ref):
    """Error scale of each quantity (the mixed path's bar): rho and tau_PD
    themselves; grad against max(|grad|, sqrt(8 rho tau)); the Laplacian
    against |lap - 4 tau| + 4 tau, the size of its two cancelling terms."""
    rho = ref["rho"][:, 0].abs()
    tau = ref["tau_pd"][:, 0].abs()
    if q == "rho":
        return rho
    if q == "tau_pd":
        return tau
    if q == "grad":
        return torch.maximum(torch.linalg.vector_norm(ref["grad"], dim=1), torch.sqrt(8 * rho * tau))
    return (ref["lap"][:, 0] - 4 * tau).abs() + 4 * tau


@pytest.mark.parametrize("case", ["water", "organic60"])
def test_noalloc_matches_ct_eval_on_two_calls(torch, case):
    from chemtoolscuda_amd import capi, synthetic

    w = synthetic.water(40000) if case == "water" else synthetic.organic60()
    rng = np.random.default_rng(41)
    sel = rng.choice(w.points.shape[0], min(60000, w.points.shape[0]), replace=False)
    pts_all = torch.from_numpy(np.ascontiguousarray(w.points[sel])).cuda()
    h = capi.Handle(w.basis.flat(), w.P, w.charges)
    qs = ["rho", "grad", "lap", "tau_pd"]
    for n in (pts_all.shape[0], pts_all.shape[0] // 3 + 7):  # the second call smaller
        pts = pts_all[:n].contiguous()
        dens = {"route": 1}
        refd, _ = _run(torch, h, h.eval, pts, qs, "fp64", dens)
        gotd, _ = _run(torch, h, h.eval_noalloc, pts, qs, "fp64", dens)
        for q in qs:
            assert torch.equal(gotd[q], refd[q]), (case, n, q)
        ref64, _ = _run(torch, h, h.eval, pts, qs, "fp64")
        got64, rep = _run(torch, h, h.eval_noalloc, pts, qs, "fp64")
        for q in qs:
            sc = _scale(torch, q, ref64).clamp(min=1e-300)
            d = (torch.linalg.vector_norm(got64[q] - ref64[q], dim=1) / sc)[ref64["rho"][:, 0] >= 1e-8].max().item()
            assert d < 1e-13, (case, n, q, d)
        gotm = _outputs(torch, qs, n)
        ptrs = {q: t.data_ptr() for q, t in gotm.items()}
        torch.cuda.synchronize()
        held = capi.device_allocation_held()
        h.eval_noalloc(pts.data_ptr(), n, ptrs, "mixed")
        torch.cuda.synchronize()
        # the library's buffers did not grow; and, once the runtime has set
        # up the mixed kernels' first launches in this process, a call leaves
        # the device's free memory exactly as it was
        assert capi.device_allocation_held() == held
        free0 = torch.cuda.mem_get_info()[0]
        repm = h.eval_noalloc(pts.data_ptr(), n, ptrs, "mixed")
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] == free0
        assert capi.device_allocation_held() == held
        refm, _ = _run(torch, h, h.eval, pts, qs, "mixed")
        mask = ref64["rho"][:, 0] >= 1e-8
        for q in qs:
            sc = _scale(torch, q, ref64).clamp(min=1e-300)
            e_alloc_free = (torch.linalg.vector_norm(gotm[q] - ref64[q], dim=1) / sc)[mask].max().item()
            e_between = (torch.linalg.vector_norm(gotm[q] - refm[q], dim=1) / sc)[mask].max().item()
            print(f"{case} n={n} {q}: no-alloc mixed vs FP64 {e_alloc_free:.2e}, vs ct_eval mixed {e_between:.2e}")
            assert e_alloc_free < 1e-6 and e_between < 2e-6, (case, n, q)
        assert repm["n_reevaluated"] >= 0 and repm["n_chunks"] >= 1


def test_noalloc_hessian_mode_on_two_calls(torch):
    """The Hessian mode in the scratch: the factor route's psi (FP64 path:
    stored by its first launch and read by the two second-derivative
    launches; mixed path: the same, and the FP64 re-evaluation's slices hold
    their own) carved from the caller-sized region, two calls in a row, the
    second smaller: FP64 within 1e-13 and mixed within 1e-6 of each
    quantity's scale against ct_eval's FP64 (the Hessian against
    max|H_kl| + 8 tau)."""
    from chemtoolscuda_amd import capi, synthetic

    w = synthetic.organic60()
    rng = np.random.default_rng(43)
    sel = rng.choice(w.points.shape[0], 30000, replace=False)
    pts_all = torch.from_numpy(np.ascontiguousarray(w.points[sel])).cuda()
    h = capi.Handle(w.basis.flat(), w.P, w.charges)
    qs = ["rho", "grad", "hess", "tau_pd", "lap"]
    for n in (pts_all.shape[0], pts_all.shape[0] // 3 + 5):
        pts = pts_all[:n].contiguous()
        ref64, _ = _run(torch, h, h.eval, pts, qs, "fp64", {"route": 2})
        got64, rep = _run(torch, h, h.eval_noalloc, pts, qs, "fp64", {"route": 2})
        assert rep["factor_rank"] > 0
        gotm, repm = _run(torch, h, h.eval_noalloc, pts, qs, "mixed", {"route": 2})
        assert repm["n_reevaluated"] > 0
        mask = ref64["rho"][:, 0] >= 1e-8
        tau = ref64["tau_pd"][:, 0].abs()
        for q in ("rho", "grad", "hess"):
            if q == "hess":
                sc = ref64["hess"].abs().amax(dim=1) + 8 * tau
            else:
                sc = _scale(torch, q, ref64)
            sc = sc.clamp(min=1e-300)
            e64 = (torch.linalg.vector_norm(got64[q] - ref64[q], dim=1) / sc)[mask].max().item()
            em = (torch.linalg.vector_norm(gotm[q] - ref64[q], dim=1) / sc)[mask].max().item()
            print(f"hessian mode n={n} {q}: no-alloc FP64 {e64:.2e}, no-alloc mixed {em:.2e}")
            assert e64 < 1e-13 and em < 1e-6, (n, q, e64, em)


def test_noalloc_columns_and_rejections(torch):
    from chemtoolscuda_amd import capi, synthetic

    w = synthetic.water(5000)
    h = capi.Handle(w.basis.flat(), w.P, w.charges)
    rows = torch.from_numpy(np.ascontiguousarray(w.points)).cuda()
    cols = rows.t().contiguous()
    n = rows.shape[0]
    a = _outputs(torch, ["rho", "grad"], n)
    b = {"rho": torch.empty(n, dtype=torch.float64, device="cuda"),
         "grad": torch.empty((3, n), dtype=torch.float64, device="cuda")}
    torch.cuda.synchronize()
    h.eval_noalloc(rows.data_ptr(), n, {q: t.data_ptr() for q, t in a.items()}, "fp64")
    h.eval_noalloc(cols.data_ptr(), n, {q: t.data_ptr() for q, t in b.items()}, "fp64", {"out_layout": 1}, "col")
    torch.cuda.synchronize()
    assert torch.equal(a["rho"][:, 0], b["rho"]) and torch.equal(a["grad"].t(), b["grad"])
    host = np.ascontiguousarray(w.points)
    rho = np.empty(n)
    with pytest.raises(RuntimeError, match="device memory"):
        h.eval_noalloc(host.ctypes.data, n, {"rho": rho.ctypes.data}, "fp64")
    with pytest.raises(RuntimeError, match="device memory"):
        h.eval_noalloc(rows.data_ptr(), 1 << 31, {"rho": a["rho"].data_ptr()}, "fp64")
    with pytest.raises(RuntimeError, match="layout"):
        h.eval_noalloc(rows.data_ptr(), n, {"rho": a["rho"].data_ptr()}, "fp64", {"out_layout": 5})
