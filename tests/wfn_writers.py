"""Test-side writers of AIM .wfn / .wfx and Molden files from an fchk fixture.

They turn the reference's own fchk fixtures into the other formats so the
native readers (chemtoolscuda_amd/csrc/wavefunction.cpp) can be checked
against the fchk reader on identical wavefunctions.  The format rules used
here are stated independently of the readers:

* wfn / wfx list uncontracted Cartesian primitives x^a y^b z^c exp(-a r^2)
  with unnormalised coefficients and the AIM type codes 1..35;
* Molden lists contracted shells and MO coefficients of normalised
  contracted functions, Cartesian g in Molden's own order.

Pure shells are written for wfn / wfx as their Cartesian expansion, whose
coefficients are fitted from the oracle's AO values (pure = T cart, exact).
"""
from __future__ import annotations

import numpy as np

from chemtoolscuda_amd.basis import BasisSet, Shell
from oracle import oracle

# AIM primitive type codes -> (a, b, c)
AIM_TYPES = {
    1: (0, 0, 0), 2: (1, 0, 0), 3: (0, 1, 0), 4: (0, 0, 1),
    5: (2, 0, 0), 6: (0, 2, 0), 7: (0, 0, 2), 8: (1, 1, 0), 9: (1, 0, 1), 10: (0, 1, 1),
    11: (3, 0, 0), 12: (0, 3, 0), 13: (0, 0, 3), 14: (2, 1, 0), 15: (2, 0, 1), 16: (0, 2, 1),
    17: (1, 2, 0), 18: (1, 0, 2), 19: (0, 1, 2), 20: (1, 1, 1),
    21: (4, 0, 0), 22: (0, 4, 0), 23: (0, 0, 4), 24: (3, 1, 0), 25: (3, 0, 1), 26: (1, 3, 0),
    27: (0, 3, 1), 28: (1, 0, 3), 29: (0, 1, 3), 30: (2, 2, 0), 31: (2, 0, 2), 32: (0, 2, 2),
    33: (2, 1, 1), 34: (1, 2, 1), 35: (1, 1, 2),
}
TYPE_OF = {v: k for k, v in AIM_TYPES.items()}

# component order of the fchk-convention basis (the order the reference evaluates)
CART = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
    3: [(3, 0, 0), (0, 3, 0), (0, 0, 3), (1, 2, 0), (2, 1, 0), (2, 0, 1), (1, 0, 2), (0, 1, 2), (0, 2, 1), (1, 1, 1)],
    4: [(0, 0, 4), (0, 1, 3), (0, 2, 2), (0, 3, 1), (0, 4, 0), (1, 0, 3), (1, 1, 2), (1, 2, 1), (1, 3, 0), (2, 0, 2),
        (2, 1, 1), (2, 2, 0), (3, 0, 1), (3, 1, 0), (4, 0, 0)],
}
# Molden's Cartesian g order: xxxx yyyy zzzz xxxy xxxz yyyx yyyz zzzx zzzy xxyy xxzz yyzz xxyz yyxz zzxy
MOLDEN_G = [(4, 0, 0), (0, 4, 0), (0, 0, 4), (3, 1, 0), (3, 0, 1), (1, 3, 0), (0, 3, 1), (1, 0, 3), (0, 1, 3),
            (2, 2, 0), (2, 0, 2), (0, 2, 2), (2, 1, 1), (1, 2, 1), (1, 1, 2)]


def _dfact(n):
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def cart_norm(alpha, a, b, c):
    L = a + b + c
    return np.sqrt((2 * alpha / np.pi) ** 1.5 * (4 * alpha) ** L / (_dfact(2 * a - 1) * _dfact(2 * b - 1) * _dfact(2 * c - 1)))


_T = {}


def pure_to_cart(L):
    """T with pure_m = sum_k T[m, k] cart_k for normalised primitives of one exponent."""
    if L not in _T:
        basis = BasisSet([Shell(0, -L, [1.0], [1.0]), Shell(0, L, [1.0], [1.0])], np.zeros((1, 3)))
        pts = np.random.default_rng(L).normal(size=(400, 3))
        ao = oracle.ao(basis.flat(), pts, 0)[0]
        npure = 2 * L + 1
        T, *_ = np.linalg.lstsq(ao[npure:].T, ao[:npure].T, rcond=None)
        assert np.allclose(ao[npure:].T @ T, ao[:npure].T, atol=1e-12)
        _T[L] = T.T
    return _T[L]


def _orbitals(d):
    """(C (nbasis, nmo), occupations, spin labels) of the occupied orbitals."""
    if d["C_beta"] is d["C_alpha"]:
        occ = d["occ_alpha"] + d["occ_beta"]
        keep = occ > 0
        return d["C_alpha"][:, keep], occ[keep], ["Alpha"] * int(keep.sum())
    ka, kb = d["occ_alpha"] > 0, d["occ_beta"] > 0
    C = np.hstack([d["C_alpha"][:, ka], d["C_beta"][:, kb]])
    occ = np.concatenate([d["occ_alpha"][ka], d["occ_beta"][kb]])
    return C, occ, ["Alpha"] * int(ka.sum()) + ["Beta"] * int(kb.sum())


def _primitives(d):
    """Uncontracted Cartesian primitives: centres (1-based), types, exponents, coefficients (nmo, nprim)."""
    basis = d["basis"]
    C, occ, spin = _orbitals(d)
    centers, types, exps, cols = [], [], [], []
    f = 0
    for s in basis.shells:
        L = abs(s.l)
        block = C[f:f + s.nfunc]
        f += s.nfunc
        cart = pure_to_cart(L).T @ block if s.l < 0 else block
        for a, c in zip(s.exps, s.coefs):
            for k, abc in enumerate(CART[L]):
                centers.append(s.atom + 1)
                types.append(TYPE_OF[abc])
                exps.append(a)
                cols.append(cart[k] * c * cart_norm(a, *abc))
    return centers, types, exps, np.array(cols).T, occ, spin


def _d(x):
    return f"{x:.14E}".replace("E", "D")


def write_wfn(path, d, title="from fchk"):
    centers, types, exps, coef, occ, _ = _primitives(d)
    basis = d["basis"]
    nmo, nprim, nnuc = coef.shape[0], len(exps), basis.natom
    out = [title, f"GAUSSIAN      {nmo:10d} MOL ORBITALS {nprim:6d} PRIMITIVES {nnuc:8d} NUCLEI"]
    for i, (xyz, z, q) in enumerate(zip(basis.atom_xyz, d["numbers"], d["charges"])):
        out.append(f"  X{i + 1:<4d}{i + 1:4d}    (CENTRE{i + 1:3d}) {xyz[0]:20.14f}{xyz[1]:20.14f}{xyz[2]:20.14f}  "
                   f"CHARGE = {q:4.1f}")
    for label, vals in (("CENTRE ASSIGNMENTS  ", centers), ("TYPE ASSIGNMENTS    ", types)):
        for s in range(0, nprim, 20):
            out.append(label + "".join(f"{v:3d}" for v in vals[s:s + 20]))
    for s in range(0, nprim, 5):
        out.append("EXPONENTS " + "".join(f"{_d(v):>22s}" for v in exps[s:s + 5]))
    for j in range(nmo):
        out.append(f"MO{j + 1:5d}     MO 0.0        OCC NO = {occ[j]:13.7f}  ORB. ENERGY =    0.000000")
        for s in range(0, nprim, 5):
            out.append("".join(f"{_d(v):>22s}" for v in coef[j, s:s + 5]))
    out.append("END DATA")
    out.append(" THE SCF ENERGY =      0.000000000000 THE VIRIAL(-V/T)=   2.00000000")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def write_wfx(path, d, title="from fchk"):
    centers, types, exps, coef, occ, spin = _primitives(d)
    basis = d["basis"]

    def sec(name, vals, per=5, fmt="{: .14E}"):
        lines = [f"<{name}>"]
        vals = list(vals)
        for s in range(0, len(vals), per):
            lines.append(" ".join(fmt.format(v) for v in vals[s:s + per]))
        lines.append(f"</{name}>")
        return lines

    nel = float(np.sum(occ))
    out = ["<Title>", title, "</Title>", "<Keywords>", "GTO", "</Keywords>",
           "<Number of Nuclei>", str(basis.natom), "</Number of Nuclei>",
           "<Number of Occupied Molecular Orbitals>", str(len(occ)), "</Number of Occupied Molecular Orbitals>",
           "<Number of Perturbations>", "0", "</Number of Perturbations>",
           "<Net Charge>", "0", "</Net Charge>",
           "<Number of Electrons>", f"{nel:.0f}", "</Number of Electrons>",
           "<Nuclear Names>"] + [f"X{i + 1}" for i in range(basis.natom)] + ["</Nuclear Names>"]
    out += sec("Atomic Numbers", d["numbers"], fmt="{:d}")
    out += sec("Nuclear Charges", d["charges"])
    out += sec("Nuclear Cartesian Coordinates", basis.atom_xyz.ravel(), per=3)
    out += ["<Number of Primitives>", str(len(exps)), "</Number of Primitives>"]
    out += sec("Primitive Centers", centers, per=10, fmt="{:d}")
    out += sec("Primitive Types", types, per=10, fmt="{:d}")
    out += sec("Primitive Exponents", exps)
    out += sec("Molecular Orbital Occupation Numbers", occ)
    out += sec("Molecular Orbital Energies", np.zeros(len(occ)))
    out += ["<Molecular Orbital Spin Types>"] + [f" {s}" for s in spin] + ["</Molecular Orbital Spin Types>"]
    out.append("<Molecular Orbital Primitive Coefficients>")
    for j in range(coef.shape[0]):
        out.append(f"<MO Number> {j + 1} </MO Number>")
        for s in range(0, coef.shape[1], 4):
            out.append(" ".join(f"{v: .14E}" for v in coef[j, s:s + 4]))
    out.append("</Molecular Orbital Primitive Coefficients>")
    out += ["<Energy>", " 0.0", "</Energy>", "<Virial Ratio (-V/T)>", " 2.0", "</Virial Ratio (-V/T)>"]
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def write_molden(path, d, angstrom=False, coef_scale=1.7):
    """Contraction coefficients are written times coef_scale (Molden readers
    renormalise contractions); MO coefficients refer to normalised contracted
    functions."""
    basis = d["basis"]
    C, occ, spin = _orbitals(d)
    pure = {abs(s.l) for s in basis.shells if s.l < 0}
    unit = 0.52917721092 if angstrom else 1.0
    out = ["[Molden Format]", "[Title]", " from fchk", "[Atoms] " + ("Angs" if angstrom else "AU")]
    for i, (xyz, z) in enumerate(zip(basis.atom_xyz, d["numbers"])):
        out.append(f"X{i + 1} {i + 1} {int(z)} {xyz[0] * unit:.14f} {xyz[1] * unit:.14f} {xyz[2] * unit:.14f}")
    out.append("[GTO]")
    starts = np.concatenate([[0], np.cumsum([s.nfunc for s in basis.shells])])
    order = []  # file function n -> (fchk function index, norm of its contraction)
    for atom in range(basis.natom):
        out.append(f"  {atom + 1} 0")
        for si, s in enumerate(basis.shells):
            if s.atom != atom:
                continue
            L = abs(s.l)
            out.append(f" {'spdfg'[L]} {len(s.exps)} 1.00")
            for a, c in zip(s.exps, s.coefs):
                out.append(f"  {a:.14E} {c * coef_scale:.14E}")
            ee = np.sqrt(np.outer(s.exps, s.exps))
            S = (2 * ee / np.add.outer(s.exps, s.exps)) ** (L + 1.5)
            norm = np.sqrt(s.coefs @ S @ s.coefs)
            start = int(starts[si])
            if s.l == 4:
                perm = [CART[4].index(abc) for abc in MOLDEN_G]
            else:
                perm = list(range(s.nfunc))
            for k in perm:
                order.append((start + k, norm))
        out.append("")
    if 2 in pure and 3 in pure:
        out.append("[5D]")
    elif 2 in pure:
        out.append("[5D10F]")
    elif 3 in pure:
        out.append("[7F]")
    if 4 in pure:
        out.append("[9G]")
    out.append("[MO]")
    for j in range(C.shape[1]):
        out += [" Sym= A", " Ene= 0.0", f" Spin= {spin[j]}", f" Occup= {occ[j]:.10f}"]
        for n, (src, norm) in enumerate(order):
            out.append(f"{n + 1:6d} {C[src, j] * norm: .14E}")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")
