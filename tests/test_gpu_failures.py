"""Failure handling forced from the host.

* Allocation failure: ct_set_device_allocation_limit caps the device memory
  the library's buffers may hold; a chunk workspace beyond the cap fails as a
  failed hipMalloc does, and evaluation halves the chunk and retries (the
  reference asserts instead, src/evaluate_density.cu:690-705).  The results
  must not change: the FP64 density route (the reference's formulation) the
  same bits; the screened FP64 factor route, whose point tiles follow the
  chunks, within 1e-13 of each quantity's scale; mixed within its bar.  A
  cap too small for even the smallest chunk must end in CT_E_NOMEM with a
  message.
* FP64 re-evaluation in slices: with host arrays, a small workspace and a
  mixed tolerance no point meets, every point of every chunk is re-evaluated
  in FP64 in workspace-sized slices; the results are the FP64 path's.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture()
def capped():
    from chemtoolscuda_amd import capi

    yield capi
    capi.set_device_allocation_limit(-1)


def _organic_sample(n, seed):
    from chemtoolscuda_amd import synthetic

    w = synthetic.organic60()
    rng = np.random.default_rng(seed)
    return w, np.ascontiguousarray(w.points[rng.choice(w.points.shape[0], n, replace=False)])


def test_allocation_failure_halves_the_chunk(capped):
    import chemtoolscuda_amd as ct

    w, pts = _organic_sample(20000, 31)
    mol = ct.molecule_from_basis(w.basis, w.P, w.charges)
    qs = ["rho", "grad", "lap", "tau_pd"]
    # (host arrays: the pipelined chunks, up to 2^20 points, 2^16 first and last)
    ref = mol.compute(pts, qs)
    n0 = mol.last_report()["n_chunks"]
    assert mol.last_report()["factor_rank"] > 0
    ref_dens = mol.compute(pts, qs, route="density")
    ref_mixed = mol.compute(pts, qs, precision="mixed")
    # fresh handle (its buffers have not grown yet); the cap (on everything
    # the process's handles hold) leaves 192 MiB more: room for chunks of a
    # few thousand points, not for one of 20 000 (512 MB in FP64)
    mol2 = ct.molecule_from_basis(w.basis, w.P, w.charges)
    capped.set_device_allocation_limit(capped.device_allocation_held() + (192 << 20))
    got = mol2.compute(pts, qs)
    rep = mol2.last_report()
    assert rep["n_chunks"] > n0, rep
    for q in qs:
        scale = np.maximum(np.abs(ref[q]).reshape(len(pts), -1).max(axis=1), 1e-8)
        d = np.abs(got[q] - ref[q]).reshape(len(pts), -1).max(axis=1)
        assert (d / scale).max() < 1e-13, q
    got_dens = mol2.compute(pts, qs, route="density")
    assert mol2.last_report()["n_chunks"] > n0
    for q in qs:
        assert np.array_equal(got_dens[q], ref_dens[q]), q
    got_mixed = mol2.compute(pts, qs, precision="mixed")
    # (the mixed path's smaller workspace may fit the chunks of the ramp)
    assert mol2.last_report()["n_chunks"] >= n0
    for q in qs:
        scale = np.maximum(np.abs(ref[q]).reshape(len(pts), -1).max(axis=1), 1e-8)
        d = np.abs(got_mixed[q] - ref_mixed[q]).reshape(len(pts), -1).max(axis=1)
        assert (d / scale).max() < 2e-6, q
    # too small for any chunk: a clean CT_E_NOMEM
    mol3 = ct.molecule_from_basis(w.basis, w.P, w.charges)
    capped.set_device_allocation_limit(capped.device_allocation_held())
    with pytest.raises(RuntimeError, match="cannot allocate"):
        mol3.compute(pts, qs)
    capped.set_device_allocation_limit(-1)
    assert np.array_equal(mol3.compute(pts, qs, route="density")["rho"], ref_dens["rho"])


def test_host_arrays_reevaluate_every_point_in_slices():
    from chemtoolscuda_amd import capi

    w, pts = _organic_sample(12000, 32)
    h = capi.Handle(w.basis.flat(), w.P, w.charges)
    h.set_workspace_limit(48 << 20)
    qs = ["rho", "grad", "lap"]
    n = len(pts)

    def run(prec, params):
        out = {q: np.empty((n, capi.WIDTH.get(q, 1))) for q in qs}
        rep = h.eval(pts.ctypes.data, n, {q: a.ctypes.data for q, a in out.items()}, prec, params)
        return out, rep

    ref, _ = run("fp64", {"route": 2})  # FP64 on the factor route, as the re-evaluation runs
    got, rep = run("mixed", {"mixed_tol": 1e-15, "rho_floor": 1e-300})
    assert rep["n_chunks"] > 1 and rep["n_reevaluated"] == n, rep
    for q in qs:
        scale = np.maximum(np.abs(ref[q]).max(axis=1), 1e-300)
        assert (np.abs(got[q] - ref[q]).max(axis=1) / scale).max() < 1e-12, q
