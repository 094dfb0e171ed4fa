"""Native .wfn / .wfx / Molden readers (chemtoolscuda_amd/csrc/wavefunction.cpp),
host side only: no GPU needed.

The reference reads every format through iodata.load_one (src/iodata.cpp:17-20);
it ships fchk fixtures only.  The readers are pinned by
  * closed forms from the format definitions: one-primitive orbitals of every
    AIM type code 1..35 (wfn and wfx) and of every Molden Cartesian g
    component, whose densities are (c x^a y^b z^c exp(-a r^2))^2 exactly;
  * the reference's 16 fchk fixtures written out as wfn / wfx / Molden by the
    test-side writers (tests/wfn_writers.py) and read back: same density and
    gradient as the fchk reader to rounding;
  * on every file read: Tr(P S) = N_electrons with the oracle's analytic
    overlap of the loaded basis, and P S P = 2 P for closed-shell SCF
    wavefunctions (the density of doubly occupied orthonormal orbitals).
"""
import os

import numpy as np
import pytest

import fchk_py
import wfn_writers as W
from oracle import oracle


@pytest.fixture(scope="module")
def ct():
    import chemtoolscuda_amd

    return chemtoolscuda_amd


def _points(basis, n, seed):
    rng = np.random.default_rng(seed)
    return basis.atom_xyz[rng.integers(0, basis.natom, n)] + rng.normal(size=(n, 3))


@pytest.mark.parametrize("fmt", ["wfn", "wfx", "molden"])
def test_round_trip_of_every_fchk_fixture(ct, tmp_path, fmt):
    write = {"wfn": W.write_wfn, "wfx": W.write_wfx, "molden": W.write_molden}[fmt]
    for path in fchk_py.fixture_paths():
        d = fchk_py.load(path)
        out = os.path.join(tmp_path, os.path.basename(path).replace(".fchk", "." + fmt))
        write(out, d)
        w = ct.read_wavefunction(out)
        b = ct.basis_from_wavefunction(w)
        assert np.allclose(w["atom_xyz"], d["basis"].atom_xyz, rtol=0, atol=1e-12)
        assert np.array_equal(w["numbers"], d["numbers"])
        pts = _points(d["basis"], 80, 3)
        ref = oracle.density(d["basis"].flat(), d["P"], pts, grad=True)
        got = oracle.density(b.flat(), w["P"], pts, grad=True)
        assert np.max(np.abs(got["rho"] - ref["rho"]) / ref["rho"]) < 1e-12, path
        assert np.max(np.abs(got["grad"] - ref["grad"])) < 1e-12 * np.abs(ref["grad"]).max(), path
        S = oracle.overlap(b.flat())
        PS = w["P"] @ S
        assert abs(np.trace(PS) - d["n_electrons"]) < 1e-7, path
        if d["C_beta"] is d["C_alpha"]:
            assert np.max(np.abs(PS @ w["P"] - 2 * w["P"])) < 1e-7 * np.abs(w["P"]).max(), path


def _aim_closed_form(centre, alpha, coefs, pts):
    r = pts - centre
    rho = np.zeros(len(pts))
    for t, c in coefs.items():
        a, b, cc = W.AIM_TYPES[t]
        rho += (c * r[:, 0] ** a * r[:, 1] ** b * r[:, 2] ** cc * np.exp(-alpha * np.sum(r * r, axis=1))) ** 2
    return rho


def test_aim_type_codes_closed_form(ct, tmp_path):
    """35 orbitals, orbital t = c_t x^a y^b z^c exp(-alpha r^2) for AIM type t,
    occupation 1: rho = sum_t of their squares, for .wfn and .wfx."""
    centre, alpha = np.array([0.3, -0.2, 0.1]), 0.7
    coefs = {t: 0.1 * t * (-1) ** t for t in range(1, 36)}
    n = 35
    wfn = ["closed form", f"GAUSSIAN      {n:10d} MOL ORBITALS {n:6d} PRIMITIVES {1:8d} NUCLEI",
           f"  C    1    (CENTRE  1) {centre[0]:12.8f}{centre[1]:12.8f}{centre[2]:12.8f}  CHARGE =  6.0"]
    for label, vals in (("CENTRE ASSIGNMENTS  ", [1] * n), ("TYPE ASSIGNMENTS    ", list(range(1, 36)))):
        for s in range(0, n, 20):
            wfn.append(label + "".join(f"{v:3d}" for v in vals[s:s + 20]))
    for s in range(0, n, 5):
        wfn.append("EXPONENTS " + "".join(" 0.7000000D+00" for _ in range(min(5, n - s))))
    for j in range(n):
        wfn.append(f"MO{j + 1:5d}     MO 0.0        OCC NO =    1.00000000  ORB. ENERGY =   -1.000000")
        row = [coefs[j + 1] if p == j else 0.0 for p in range(n)]
        for s in range(0, n, 5):
            # D16.8 fields written back to back, as fixed-width Fortran output does
            wfn.append("".join(f"{v:16.8E}".replace("E", "D") for v in row[s:s + 5]))
    wfn.append("END DATA")
    p1 = os.path.join(tmp_path, "types.wfn")
    with open(p1, "w") as f:
        f.write("\n".join(wfn) + "\n")
    wfx = ["<Number of Nuclei>", "1", "</Number of Nuclei>", "<Atomic Numbers>", "6", "</Atomic Numbers>",
           "<Nuclear Charges>", "6.0", "</Nuclear Charges>",
           "<Nuclear Cartesian Coordinates>", " ".join(str(v) for v in centre), "</Nuclear Cartesian Coordinates>",
           "<Number of Primitives>", str(n), "</Number of Primitives>",
           "<Primitive Centers>", " ".join(["1"] * n), "</Primitive Centers>",
           "<Primitive Types>", " ".join(str(t) for t in range(1, 36)), "</Primitive Types>",
           "<Primitive Exponents>", " ".join(["0.7"] * n), "</Primitive Exponents>",
           "<Molecular Orbital Occupation Numbers>", " ".join(["1.0"] * n), "</Molecular Orbital Occupation Numbers>",
           "<Molecular Orbital Primitive Coefficients>"]
    for j in range(n):
        wfx += [f"<MO Number> {j + 1} </MO Number>", " ".join(repr(coefs[j + 1] if p == j else 0.0) for p in range(n))]
    wfx.append("</Molecular Orbital Primitive Coefficients>")
    p2 = os.path.join(tmp_path, "types.wfx")
    with open(p2, "w") as f:
        f.write("\n".join(wfx) + "\n")
    pts = centre + np.random.default_rng(4).normal(size=(200, 3))
    expect = _aim_closed_form(centre, alpha, coefs, pts)
    for p in (p1, p2):
        w = ct.read_wavefunction(p)
        assert list(w["shell_l"]) == [0, 1, 2, 3, 4]  # one Cartesian shell per L, same centre and exponent
        got = oracle.density(ct.basis_from_wavefunction(w).flat(), w["P"], pts)["rho"]
        assert np.max(np.abs(got - expect) / expect) < 1e-13, p
    w = ct.read_wavefunction(p1)
    assert list(w["numbers"]) == [6] and list(w["charges"]) == [6.0]


def test_molden_cartesian_g_order_closed_form(ct, tmp_path):
    """One Cartesian g shell (one primitive); orbital k has coefficient 1 on
    the k-th function of Molden's g list (xxxx yyyy zzzz xxxy xxxz yyyx yyyz
    zzzx zzzy xxyy xxzz yyzz xxyz yyxz zzxy) and occupation w_k: rho =
    sum_k w_k (N_k x^a y^b z^c exp(-alpha r^2))^2 with N_k the L2
    normalisation of that component."""
    names = "xxxx yyyy zzzz xxxy xxxz yyyx yyyz zzzx zzzy xxyy xxzz yyzz xxyz yyxz zzxy".split()
    alpha = 1.3
    lines = ["[Molden Format]", "[Atoms] AU", "Ne 1 10 0.0 0.0 0.0", "[GTO]", "  1 0", " g 1 1.00",
             f"  {alpha} 1.0", "", "[MO]"]
    occ = [0.5 + 0.1 * k for k in range(15)]
    for k in range(15):
        lines += [" Sym= A", " Ene= -1.0", " Spin= Alpha", f" Occup= {occ[k]}"]
        lines += [f"{m + 1:4d} {1.0 if m == k else 0.0}" for m in range(15)]
    p = os.path.join(tmp_path, "g.molden")
    with open(p, "w") as f:
        f.write("\n".join(lines) + "\n")
    w = ct.read_wavefunction(p)
    pts = np.random.default_rng(5).normal(size=(150, 3))
    expect = np.zeros(len(pts))
    for k, nm in enumerate(names):
        a, b, c = nm.count("x"), nm.count("y"), nm.count("z")
        g = W.cart_norm(alpha, a, b, c) * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c * np.exp(-alpha * (pts ** 2).sum(1))
        expect += occ[k] * g ** 2
    got = oracle.density(ct.basis_from_wavefunction(w).flat(), w["P"], pts)["rho"]
    assert np.max(np.abs(got - expect) / expect) < 1e-13


def test_molden_angstrom_units(ct, tmp_path):
    d = fchk_py.load(os.path.join(fchk_py.FCHK_DIR, "h2o.fchk"))
    p = os.path.join(tmp_path, "w.molden")
    W.write_molden(p, d, angstrom=True, coef_scale=0.3)
    w = ct.read_wavefunction(p)
    assert np.allclose(w["atom_xyz"], d["basis"].atom_xyz, rtol=0, atol=1e-12)
    pts = _points(d["basis"], 50, 6)
    ref = oracle.density(d["basis"].flat(), d["P"], pts)["rho"]
    got = oracle.density(ct.basis_from_wavefunction(w).flat(), w["P"], pts)["rho"]
    assert np.max(np.abs(got - ref) / ref) < 1e-12


def test_reader_errors(ct, tmp_path):
    p = os.path.join(tmp_path, "x.xyz")
    with open(p, "w") as f:
        f.write("1\n\nH 0 0 0\n")
    with pytest.raises(RuntimeError, match="unknown wavefunction file type"):
        ct.read_wavefunction(p)
    with pytest.raises(RuntimeError, match="cannot open"):
        ct.read_wavefunction(os.path.join(tmp_path, "missing.wfn"))
    d = fchk_py.load(os.path.join(fchk_py.FCHK_DIR, "atom_he.fchk"))
    full = os.path.join(tmp_path, "he.wfn")
    W.write_wfn(full, d)
    with open(full) as f:
        lines = f.read().splitlines()
    cut = os.path.join(tmp_path, "cut.wfn")
    with open(cut, "w") as f:
        f.write("\n".join(lines[:-4]) + "\n")  # last orbital cut short
    with pytest.raises(RuntimeError, match="short last orbital|mismatch"):
        ct.read_wavefunction(cut)
    fchk = ct.read_wavefunction(os.path.join(fchk_py.FCHK_DIR, "atom_he.fchk"))
    assert np.allclose(fchk["P"], d["P"], rtol=0, atol=1e-14)


def _fchk_lines(name):
    with open(os.path.join(fchk_py.FCHK_DIR, name)) as f:
        return f.read().splitlines()


def _field_span(lines, key):
    """(first, end) line indices of a field's header and values."""
    i = next(k for k, ln in enumerate(lines) if ln.startswith(key))
    j = i + 1
    while j < len(lines) and lines[j].startswith(" "):
        j += 1
    return i, j


@pytest.mark.parametrize("case", ["truncated", "missing_shell_types", "short_shell_map", "zero_basis",
                                  "bad_centre", "short_beta", "shell_type_7", "negative_exponent",
                                  "bad_scalar"])
def test_malformed_fchk_is_rejected(ct, tmp_path, case):
    """Malformed .fchk files end in a readable error (CT_E_IO / CT_E_UNSUPPORTED
    through the host-only reader; ct_mol_create_from_fchk parses the same way
    before any device work), never in a crash or an out-of-range read."""
    lines = _fchk_lines("ch4.fchk" if case != "short_beta" else "atom_08_O_N09_M2_ub3lyp_ccpvtz_g09.fchk")
    if case == "truncated":
        i, j = _field_span(lines, "Alpha MO coefficients")
        lines = lines[: i + 3]
        match = "truncated array"
    elif case == "missing_shell_types":
        i, j = _field_span(lines, "Shell types")
        lines = lines[:i] + lines[j:]
        match = "Shell types"
    elif case == "short_shell_map":
        i, j = _field_span(lines, "Shell to atom map")
        n = int(lines[i].split("N=")[1])
        vals = " ".join(lines[i + 1:j]).split()[:-1]
        lines = lines[:i] + [lines[i].split("N=")[0] + f"N={n - 1:12d}", " ".join(vals)] + lines[j:]
        match = "length mismatch"
    elif case == "zero_basis":
        i, _ = _field_span(lines, "Number of basis functions")
        lines[i] = lines[i][:43] + "I" + " " * 15 + "0"
        match = "basis size"
    elif case == "bad_centre":
        i, j = _field_span(lines, "Shell to atom map")
        lines[i + 1] = " ".join(["99"] + lines[i + 1].split()[1:])
        match = "shell centre"
    elif case == "short_beta":
        i, j = _field_span(lines, "Beta MO coefficients")
        lines = lines[:i] + [lines[i].split("N=")[0] + "N=           5", " ".join(["1.0"] * 5)] + lines[j:]
        match = "beta MO coefficient count"
    elif case == "shell_type_7":
        i, j = _field_span(lines, "Shell types")
        lines[i + 1] = " ".join(["7"] + lines[i + 1].split()[1:])
        match = "unsupported"
    elif case == "negative_exponent":
        i, j = _field_span(lines, "Primitive exponents")
        lines[i + 1] = " ".join(["-1.0"] + lines[i + 1].split()[1:])
        match = "exponent"
    else:  # bad_scalar: a non-integer electron count
        i, _ = _field_span(lines, "Number of alpha electrons")
        lines[i] = lines[i][:43] + "I" + " " * 13 + "2.5"
        match = "electron count"
    p = os.path.join(tmp_path, f"{case}.fchk")
    with open(p, "w") as f:
        f.write("\n".join(lines) + "\n")
    with pytest.raises(RuntimeError, match=match):
        ct.read_wavefunction(p)
