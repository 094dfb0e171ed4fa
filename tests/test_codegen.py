"""The generated AO polynomial code (codegen/gen_ao.py) against the oracle.

The generator builds solid harmonics from the general closed form and the
derivatives from the S0/S1/S2 radial-sum identities; the oracle uses the
explicit Helgaker table and per-primitive 1-D derivative factors.  Agreement
for every shell type and derivative slot pins the generator on the CPU.
"""
import numpy as np
import pytest
import sympy as sp

from chemtoolscuda_amd.basis import BasisSet, Shell
from chemtoolscuda_amd.codegen import gen_ao
from oracle import oracle


@pytest.mark.parametrize("l", gen_ao.SHELL_TYPES)
@pytest.mark.parametrize("dset", ["val", "grad", "lap", "hess"])
def test_generated_components_match_oracle(l, dset):
    L = abs(l)
    exps = np.array([1.7, 0.45])
    coefs = np.array([0.6, 0.5])
    centre = np.array([[0.2, -0.1, 0.3]])
    basis = BasisSet([Shell(0, l, exps, coefs)], centre)
    pts = np.random.default_rng(abs(l) * 7 + len(dset)).uniform(-1.5, 1.5, size=(25, 3))
    ref = oracle.ao(basis.flat(), pts, 2)  # (10, nfunc, n)
    d = pts - centre
    r2 = np.sum(d * d, axis=1)
    radial = coefs * np.sqrt((2 * exps / np.pi) ** 1.5 * (4 * exps) ** L)
    e = np.exp(-np.outer(r2, exps))
    s0 = e @ radial
    s1 = e @ (radial * -2 * exps)
    s2 = e @ (radial * 4 * exps ** 2)
    x, y, z, S0, S1, S2 = gen_ao.x, gen_ao.y, gen_ao.z, gen_ao.s0, gen_ao.s1, gen_ao.s2
    comps = gen_ao.components(l)
    slot_to_ref = {"val": [0], "grad": [0, 1, 2, 3], "lap": [0, 1, 2, 3, None], "hess": list(range(10))}[dset]
    for ic, q in enumerate(comps):
        exprs = gen_ao.derivative_exprs(q, L, dset)
        for slot, ex in enumerate(exprs):
            f = sp.lambdify((x, y, z, S0, S1, S2), ex, "numpy")
            got = np.broadcast_to(f(d[:, 0], d[:, 1], d[:, 2], s0, s1, s2), (pts.shape[0],))
            r = slot_to_ref[slot]
            want = ref[4, ic] + ref[7, ic] + ref[9, ic] if r is None else ref[r, ic]
            assert np.allclose(got, want, rtol=1e-12, atol=1e-13 * np.abs(want).max() + 1e-300), (l, dset, ic, slot)


def test_generated_header_is_current():
    import os
    hdr = os.path.join(gen_ao.OUT_DIR, "ao_poly.h")
    assert os.path.exists(hdr)
    src = open(hdr).read()
    for dset in gen_ao.DERIV_SETS:
        assert f"ao_shell_dispatch_{dset}" in src
