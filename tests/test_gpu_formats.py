"""Molecule(path) on .wfn / .wfx / Molden files: the GPU results equal the FP64
oracle's on the fchk file the wavefunction was written from
(tests/wfn_writers.py; the oracle reads the basis and P through the test-side
fchk parser, not the library's readers), for the density, its gradient and
Laplacian; and the same as the GPU results of that fchk file; mixed precision
within 1e-6."""
import os

import numpy as np
import pytest

import fchk_py
import wfn_writers as W
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fixture", ["h2o.fchk", "atom_c_g_cartesian_orbital.fchk", "qm9_000092_HF_cc-pVDZ.fchk"])
def test_formats_match_fchk(tmp_path, fixture):
    import chemtoolscuda_amd as ct

    path = os.path.join(fchk_py.FCHK_DIR, fixture)
    d = fchk_py.load(path)
    rng = np.random.default_rng(11)
    pts = d["basis"].atom_xyz[rng.integers(0, d["basis"].natom, 2000)] + rng.normal(size=(2000, 3))
    ref_mol = ct.Molecule(path)
    names = ["rho", "grad", "lap"]
    ref = ref_mol.compute(pts, names, precision="fp64")
    orc = oracle.density(d["basis"].flat(), d["P"], pts, rho=True, grad=True, lap=True)
    for fmt, write in (("wfn", W.write_wfn), ("wfx", W.write_wfx), ("molden", W.write_molden)):
        out = os.path.join(tmp_path, fixture.replace(".fchk", "." + fmt))
        write(out, d)
        mol = ct.Molecule(out)
        assert mol.get_file_path() == out
        got = mol.compute(pts, names, precision="fp64")
        assert np.max(np.abs(got["rho"] - orc["rho"]) / orc["rho"]) < 1e-11, fmt
        for k in ("grad", "lap"):
            assert np.max(np.abs(got[k] - orc[k])) < 1e-11 * np.abs(orc[k]).max(), (fmt, k, "oracle")
        assert np.max(np.abs(got["rho"] - ref["rho"]) / ref["rho"]) < 1e-11, fmt
        for k in ("grad", "lap"):
            assert np.max(np.abs(got[k] - ref[k])) < 1e-11 * np.abs(ref[k]).max(), (fmt, k)
        mix = mol.compute(pts, ["rho"], precision="mixed")["rho"]
        m = ref["rho"] >= 1e-8
        assert np.max(np.abs(mix - ref["rho"])[m] / ref["rho"][m]) <= 1e-6, fmt
