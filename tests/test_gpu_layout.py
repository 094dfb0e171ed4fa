"""Points and output layouts at the C ABI.

The reference's free functions read column-major (3, N) points
(src/evaluate_density.cu:663-664, 745-752; its C++ tests pass them so,
tests/test_electron_density.cu:177) and the gradient / Hessian can come back
column-major (return_row = false, src/evaluate_gradient.cu:1399-1413,
src/evaluate_hessian.cu:2244-2279), while the Molecule methods and the ESP use
(N, 3) rows.  ct_eval / ct_esp / ct_backward take the points layout as an
argument and ct_eval_params.out_layout selects the grad / hess layout: the
same points in either layout give the same bits, and column outputs are the
exact transposes of row outputs.  All calls go through the C ABI (ctypes),
with host and device arrays, FP64 and mixed, one chunk and many.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    return torch


def _handle(w):
    from chemtoolscuda_amd import capi

    return capi.Handle(w.basis.flat(), w.P, w.charges)


def _eval(h, pts_rows, quantities, precision, layout, out_layout, device, torch, params=None):
    """Runs ct_eval with the points in `layout` and outputs in `out_layout`;
    returns numpy arrays in the layout they were written."""
    from chemtoolscuda_amd import capi

    n = pts_rows.shape[0]
    pts = np.ascontiguousarray(pts_rows if layout == "row" else pts_rows.T)
    par = dict(params or {})
    par["out_layout"] = capi.LAYOUT[out_layout]
    shapes = {}
    for q in quantities:
        wdt = capi.WIDTH.get(q, 1)
        shapes[q] = (n,) if wdt == 1 else ((n, wdt) if out_layout == "row" else (wdt, n))
    if device:
        tp = torch.from_numpy(pts).cuda()
        outs = {q: torch.full(shapes[q], np.nan, dtype=torch.float64, device="cuda") for q in quantities}
        rep = h.eval(tp.data_ptr(), n, {q: t.data_ptr() for q, t in outs.items()}, precision, par, layout)
        torch.cuda.synchronize()
        return {q: t.cpu().numpy() for q, t in outs.items()}, rep
    outs = {q: np.full(shapes[q], np.nan) for q in quantities}
    rep = h.eval(pts.ctypes.data, n, {q: a.ctypes.data for q, a in outs.items()}, precision, par, layout)
    return outs, rep


def _same_bits(a, b):
    assert a.shape == b.shape
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("device", [False, True])
def test_eval_layouts_give_the_same_bits(torch, precision, device):
    from chemtoolscuda_amd import synthetic

    w = synthetic.water(3000)
    h = _handle(w)
    qs = ["rho", "grad", "lap", "tau_pd", "rdg"]
    ref, rep = _eval(h, w.points, qs, precision, "row", "row", device, torch)
    assert not any(np.isnan(v).any() for v in ref.values())
    for layout in ("row", "col"):
        for out_layout in ("row", "col"):
            got, _ = _eval(h, w.points, qs, precision, layout, out_layout, device, torch)
            for q in qs:
                g = got[q].T if (out_layout == "col" and got[q].ndim == 2) else got[q]
                _same_bits(np.ascontiguousarray(g), ref[q])


@pytest.mark.parametrize("device", [False, True])
def test_hessian_column_layout_and_many_chunks(torch, device):
    """Hessian (9, n) rows 3k+l; screened mixed factor route on the 624-function
    system in many small chunks (host: per-chunk re-evaluation through staged
    columns; device: one batched re-evaluation gathering from the columns)."""
    from chemtoolscuda_amd import synthetic

    w = synthetic.organic60()
    rng = np.random.default_rng(21)
    pts = w.points[rng.choice(w.points.shape[0], 6000, replace=False)]
    h = _handle(w)
    h.set_workspace_limit(64 << 20)
    qs = ["rho", "grad", "hess", "lap"]
    for precision in ("fp64", "mixed"):
        ref, rep = _eval(h, pts, qs, precision, "row", "row", device, torch)
        assert rep["n_chunks"] > 1
        got, rep2 = _eval(h, pts, qs, precision, "col", "col", device, torch)
        assert rep2["n_reevaluated"] == rep["n_reevaluated"]
        _same_bits(np.ascontiguousarray(got["grad"].T), ref["grad"])
        _same_bits(np.ascontiguousarray(got["hess"].T), ref["hess"])
        for q in ("rho", "lap"):
            _same_bits(got[q], ref[q])


@pytest.mark.parametrize("device", [False, True])
def test_esp_and_backward_layouts(torch, device):
    from chemtoolscuda_amd import synthetic

    w = synthetic.water(2000)
    h = _handle(w)
    n = w.points.shape[0]
    wt = np.random.default_rng(3).uniform(0, 1, n)
    M, na = w.basis.nbasis, w.basis.natom
    res = {}
    for layout in ("row", "col"):
        pts = np.ascontiguousarray(w.points if layout == "row" else w.points.T)
        for prec in ("fp64", "mixed"):
            if device:
                tp, tw = torch.from_numpy(pts).cuda(), torch.from_numpy(wt).cuda()
                esp = torch.empty(n, dtype=torch.float64, device="cuda")
                dP = torch.empty((M, M), dtype=torch.float64, device="cuda")
                dR = torch.empty((na, 3), dtype=torch.float64, device="cuda")
                h.esp(tp.data_ptr(), n, esp.data_ptr(), prec, layout)
                h.backward(tp.data_ptr(), n, tw.data_ptr(), dP.data_ptr(), dR.data_ptr(), prec, layout)
                torch.cuda.synchronize()
                res[layout, prec] = (esp.cpu().numpy(), dP.cpu().numpy(), dR.cpu().numpy())
            else:
                esp, dP, dR = np.empty(n), np.empty((M, M)), np.empty((na, 3))
                h.esp(pts.ctypes.data, n, esp.ctypes.data, prec, layout)
                h.backward(pts.ctypes.data, n, wt.ctypes.data, dP.ctypes.data, dR.ctypes.data, prec, layout)
                res[layout, prec] = (esp, dP, dR)
    for prec in ("fp64", "mixed"):
        (e1, P1, R1), (e2, P2, R2) = res["row", prec], res["col", prec]
        _same_bits(e1, e2)
        _same_bits(P1, P2)
        # dL/dR: per-row sums land in FP64 atomics, whose order varies from call
        # to call (the last bits only)
        assert np.abs(R1 - R2).max() <= 1e-13 * np.abs(R1).max()


def test_bad_layout_is_rejected():
    import ctypes

    from chemtoolscuda_amd import capi, synthetic

    w = synthetic.water(10)
    h = _handle(w)
    rho = np.empty(10)
    o = capi.Outputs()
    o.rho = rho.ctypes.data
    st = capi.lib().ct_eval(h.ptr, w.points.ctypes.data, 10, 7, capi.Q["rho"], 0, None, ctypes.byref(o), None)
    assert st == -1 and b"layout" in capi.lib().ct_last_error()
