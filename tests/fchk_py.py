"""Independent Python reader for Gaussian formatted checkpoint files.

Test-side twin of the product's C++ loader (chemtoolscuda_amd/csrc/fchk.cpp):
both are checked against each other and, through C^T S C = I and the stored
"Total SCF Density", against the wavefunction itself.  Conventions follow the
reference loader (src/iodata.cpp:15-211): shell type -1 (SP) is split into an
s and a p segment sharing exponents; negative L means pure; P is the sum over
spin channels of C occ C^T (src/iodata.cpp:154-160).
"""
from __future__ import annotations

import os
import re

import numpy as np

from chemtoolscuda_amd.basis import BasisSet, Shell

FCHK_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fchk")

_HEADER = re.compile(r"^(\S.{0,40}?)\s+([IRC])\s+(N=)?\s*(\S+)\s*$")


def read_fields(path):
    fields = {}
    with open(path) as f:
        lines = f.read().splitlines()
    i = 2
    while i < len(lines):
        m = _HEADER.match(lines[i].rstrip())
        i += 1
        if not m:
            continue
        name, kind, is_arr, val = m.group(1).strip(), m.group(2), m.group(3), m.group(4)
        if not is_arr:
            fields[name] = int(val) if kind == "I" else (float(val) if kind == "R" else val)
            continue
        n = int(val)
        vals = []
        while len(vals) < n and i < len(lines):
            vals += lines[i].split()
            i += 1
        if kind == "I":
            fields[name] = np.array([int(v) for v in vals[:n]])
        elif kind == "R":
            fields[name] = np.array([float(v) for v in vals[:n]])
        else:
            fields[name] = vals
    return fields


def load(path):
    """Returns dict(basis, P, charges, numbers, C_alpha, C_beta, occ_alpha, occ_beta, scf_density)."""
    f = read_fields(path)
    nbasis = f["Number of basis functions"]
    xyz = f["Current cartesian coordinates"].reshape(-1, 3)
    types = f["Shell types"]
    nprims = f["Number of primitives per shell"]
    s2a = f["Shell to atom map"] - 1
    exps = f["Primitive exponents"]
    cc = f["Contraction coefficients"]
    spc = f.get("P(S=P) Contraction coefficients")
    shells = []
    off = 0
    for t, n, a in zip(types, nprims, s2a):
        e = exps[off:off + n]
        if t == -1:
            shells.append(Shell(int(a), 0, e, cc[off:off + n]))
            shells.append(Shell(int(a), 1, e, spc[off:off + n]))
        else:
            shells.append(Shell(int(a), int(t), e, cc[off:off + n]))
        off += n
    basis = BasisSet(shells, xyz, f["Atomic numbers"])
    assert basis.nbasis == nbasis
    nmo = f["Alpha MO coefficients"].size // nbasis
    Ca = f["Alpha MO coefficients"].reshape(nmo, nbasis).T
    na, nb = f["Number of alpha electrons"], f["Number of beta electrons"]
    occa = np.zeros(nmo)
    occa[:na] = 1.0
    occb = np.zeros(nmo)
    occb[:nb] = 1.0
    if "Beta MO coefficients" in f:
        Cb = f["Beta MO coefficients"].reshape(nmo, nbasis).T
        P = (Ca * occa) @ Ca.T + (Cb * occb) @ Cb.T
    else:
        Cb = Ca
        P = (Ca * (occa + occb)) @ Ca.T
    scf = None
    if "Total SCF Density" in f:
        tri = f["Total SCF Density"]
        scf = np.zeros((nbasis, nbasis))
        il = np.tril_indices(nbasis)
        # fchk stores the lower triangle row by row
        scf[il[0], il[1]] = tri
        scf = scf + scf.T - np.diag(np.diag(scf))
    return {"basis": basis, "P": P, "charges": f["Nuclear charges"], "numbers": f["Atomic numbers"],
            "C_alpha": Ca, "C_beta": Cb, "occ_alpha": occa, "occ_beta": occb, "scf_density": scf,
            "n_electrons": na + nb}


def fixture_paths():
    return sorted(os.path.join(FCHK_DIR, p) for p in os.listdir(FCHK_DIR) if p.endswith(".fchk"))
