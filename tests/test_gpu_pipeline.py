"""Host arrays at the C ABI: the pipelined path (chunks of up to 2^20 points
through three slots of pinned and device staging, uploads, downloads and the
FP64 re-evaluation on their own streams) gives the bits of the
device-resident path, for every combination of
host and device points, outputs and density targets, both layouts, FP64 and
mixed, over several chunks with a ragged last one.  The reference's Molecule
methods take and return host arrays (src/pymolecule.cu:46-56), copying points
in and results out per chunk (src/evaluate_density.cu:745-752, 830-832).

The density route is used for the bit comparisons: with spatial screening the
tiles of 16 points depend on the chunk boundaries, so the screened factor
route is compared within its error bound instead."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

QS = ["rho", "grad", "lap", "tau_pd", "hess"]
N = 2 * (1 << 19) + 70_001  # six pipelined chunks (2^16 points growing to 2^18 and back), the last one ragged


@pytest.fixture(scope="module")
def torch():
    import torch

    return torch


@pytest.fixture(scope="module")
def water():
    from chemtoolscuda_amd import synthetic

    return synthetic.water(N)


def _same_bits(a, b, what):
    assert a.shape == b.shape, what
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), what


def _run(torch, h, pts_rows, precision, params, layout="row", out_layout="row", pts_dev=True, out_dev=True,
         tgt=None, tgt_dev=True):
    from chemtoolscuda_amd import capi

    n = pts_rows.shape[0]
    pts = np.ascontiguousarray(pts_rows if layout == "row" else pts_rows.T)
    par = dict(params)
    par["out_layout"] = capi.LAYOUT[out_layout]
    keep = []
    if tgt is not None:
        t = torch.from_numpy(tgt).cuda() if tgt_dev else np.ascontiguousarray(tgt)
        keep.append(t)
        par["rho_target"] = t.data_ptr() if tgt_dev else t.ctypes.data
    shapes = {}
    for q in QS:
        wdt = capi.WIDTH.get(q, 1)
        shapes[q] = (n,) if wdt == 1 else ((n, wdt) if out_layout == "row" else (wdt, n))
    if pts_dev:
        tp = torch.from_numpy(pts).cuda()
        pp = tp.data_ptr()
    else:
        tp = pts
        pp = pts.ctypes.data
    if out_dev:
        outs = {q: torch.full(shapes[q], np.nan, dtype=torch.float64, device="cuda") for q in QS}
        ptrs = {q: t.data_ptr() for q, t in outs.items()}
    else:
        outs = {q: np.full(shapes[q], np.nan) for q in QS}
        ptrs = {q: a.ctypes.data for q, a in outs.items()}
    rep = h.eval(pp, n, ptrs, precision, par, layout)
    torch.cuda.synchronize()
    got = {q: (t.cpu().numpy() if out_dev else t) for q, t in outs.items()}
    del tp, keep
    return got, rep


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_host_arrays_match_device_bits(torch, water, precision):
    from chemtoolscuda_amd import capi

    w = water
    h = capi.Handle(w.basis.flat(), w.P, w.charges)
    par = {"route": 1}
    ref, rep0 = _run(torch, h, w.points, precision, par)
    assert not any(np.isnan(v).any() for v in ref.values())
    assert rep0["n_chunks"] == 1
    cases = [
        dict(pts_dev=False, out_dev=False),
        dict(pts_dev=True, out_dev=False),
        dict(pts_dev=False, out_dev=True),
        dict(pts_dev=False, out_dev=False, layout="col", out_layout="col"),
        dict(pts_dev=True, out_dev=False, layout="col", out_layout="row"),
    ]
    for case in cases:
        got, rep = _run(torch, h, w.points, precision, par, **case)
        print(precision, case, "chunks", rep["n_chunks"], "reevaluated", rep["n_reevaluated"], flush=True)
        assert rep["n_chunks"] == 6, case
        assert rep["n_reevaluated"] == rep0["n_reevaluated"], case
        for q in QS:
            g = got[q]
            if case.get("out_layout") == "col" and g.ndim == 2:
                g = np.ascontiguousarray(g.T)
            _same_bits(g, ref[q], (q, case))


def test_host_targets_match_device_bits(torch, water):
    """rho_target from host memory (uploaded with the points) gives the bits and
    the flag decisions of the same targets in device memory."""
    from chemtoolscuda_amd import capi

    w = water
    h = capi.Handle(w.basis.flat(), w.P, w.charges)
    par = {"route": 1, "mixed_tol": 1e-7}
    base, _ = _run(torch, h, w.points, "fp64", par)
    tgt = base["rho"] * (1.0 + 1e-3 * np.cos(np.arange(N)))
    ref, rep0 = _run(torch, h, w.points, "mixed", par, tgt=tgt)
    for case in (dict(tgt_dev=False), dict(tgt_dev=False, pts_dev=False, out_dev=False),
                 dict(tgt_dev=True, pts_dev=False, out_dev=True)):
        got, rep = _run(torch, h, w.points, "mixed", par, tgt=tgt, **case)
        assert rep["n_chunks"] == 6, case
        assert rep["n_reevaluated"] == rep0["n_reevaluated"], case
        for q in QS:
            _same_bits(got[q], ref[q], (q, case))


def test_screened_host_arrays_within_bound(torch):
    """The default (screened factor) route with host arrays: chunk boundaries
    move the screening tiles, so the values agree with the device-resident
    call within the mixed tolerance of the FP64 result rather than bit for
    bit."""
    from chemtoolscuda_amd import capi, synthetic

    w = synthetic.organic60()
    rng = np.random.default_rng(5)
    pts = w.points[np.sort(rng.choice(w.points.shape[0], (1 << 19) + 100_003, replace=False))]
    h = capi.Handle(w.basis.flat(), w.P, w.charges)
    f64, _ = _run(torch, h, pts, "fp64", {})
    dev, rep_d = _run(torch, h, pts, "mixed", {})
    host, rep_h = _run(torch, h, pts, "mixed", {}, pts_dev=False, out_dev=False)
    assert rep_d["n_chunks"] == 1 and rep_h["n_chunks"] > 2
    rho = f64["rho"]
    m = rho >= 1e-10
    for got in (dev, host):
        assert np.max(np.abs(got["rho"] - rho)[m] / rho[m]) <= 1e-6
    # gradients against the scale the error bound uses (|grad| or sqrt(8 rho tau))
    sc = np.maximum(np.linalg.norm(f64["grad"], axis=1), np.sqrt(8 * np.maximum(rho, 0) * f64["tau_pd"]))
    mm = m & (sc > 0)
    for got in (dev, host):
        err = np.max(np.abs(got["grad"] - f64["grad"]), axis=1)
        assert np.max(err[mm] / sc[mm]) <= 2e-6


def test_pipeline_many_small_chunks(torch, water):
    """A workspace limit that forces chunks far below 2^20 points: the slots
    wrap round many times."""
    from chemtoolscuda_amd import capi

    w = water
    sub = w.points[:200_000]
    h = capi.Handle(w.basis.flat(), w.P, w.charges)
    par = {"route": 1}
    ref, _ = _run(torch, h, sub, "mixed", par)
    h.set_workspace_limit(8 << 20)
    got, rep = _run(torch, h, sub, "mixed", par, pts_dev=False, out_dev=False, layout="col", out_layout="col")
    assert rep["n_chunks"] > 10
    for q in QS:
        g = got[q]
        if g.ndim == 2:
            g = np.ascontiguousarray(g.T)
        _same_bits(g, ref[q], q)
