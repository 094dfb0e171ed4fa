"""Build-time code generator for the AO kernels (replaces the reference's
sympy scripts ``generate/generate_derivs_contraction.py``,
``generate/generate_sec_derivs_cont.py`` and ``generate/generate_hessian_cont.py``,
whose output was pasted by hand into ``src/evaluate_{gradient,laplacian,hessian}.cu``).

Model.  A contracted shell on centre A evaluated at r, with d = r - A, is

    phi_mu(r) = q_mu(d) * S0,   S0 = sum_p c_p N_p(L) exp(-a_p |d|^2)

where q_mu is a homogeneous polynomial of degree L (a Cartesian monomial with
its per-component normalisation, or a real solid harmonic) and N_p(L) the
radial part of the primitive normalisation.  With S1 = sum_p c_p N_p (-2 a_p) e_p
and S2 = sum_p c_p N_p (4 a_p^2) e_p the derivatives are

    d_k phi   = q_k S0 + d_k q S1
    d_kl phi  = q_kl S0 + (d_l q_k + d_k q_l + delta_kl q) S1 + d_k d_l q S2
    lap phi   = (lap q) S0 + (2L + 3) q S1 + |d|^2 q S2

so one kernel pass evaluates the primitives' exponentials once per shell and
all components/derivatives come from the polynomial code emitted here
(common-subexpression eliminated with sympy).

Component order is the fchk/iodata order the reference evaluates
(reference src/evaluate_density.cu:80,114,143,196,235-236,314).  Real solid
harmonics are built from the closed form of Helgaker, Jorgensen & Olsen,
"Molecular Electronic-Structure Theory", eqs. 6.4.47-6.4.50 (the oracle uses
the explicit table instead, so the two are checked against each other).

Output: ``csrc/generated/ao_poly.h`` plus a JSON file with static flop counts
per (shell type, derivative set), used for the AO roofline.

Run:  python chemtoolscuda_amd/codegen/gen_ao.py
"""
from __future__ import annotations

import json
import os
import re
from math import comb, factorial

import sympy as sp

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(os.path.dirname(HERE), "csrc", "generated")

x, y, z, s0, s1, s2 = sp.symbols("x y z s0 s1 s2", real=True)
X = (x, y, z)

# Shell type codes used on the device: index into this list.
SHELL_TYPES = [0, 1, 2, 3, 4, -2, -3, -4]

CART_ORDER = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
    3: [(3, 0, 0), (0, 3, 0), (0, 0, 3), (1, 2, 0), (2, 1, 0), (2, 0, 1), (1, 0, 2), (0, 1, 2),
        (0, 2, 1), (1, 1, 1)],
    4: [(0, 0, 4), (0, 1, 3), (0, 2, 2), (0, 3, 1), (0, 4, 0), (1, 0, 3), (1, 1, 2), (1, 2, 1),
        (1, 3, 0), (2, 0, 2), (2, 1, 1), (2, 2, 0), (3, 0, 1), (3, 1, 0), (4, 0, 0)],
}


def dfact(n: int) -> int:
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def real_solid_harmonic(l: int, m: int):
    """S_lm(x,y,z) (m>0 cosine-like, m<0 sine-like), Helgaker eqs. 6.4.47-6.4.50."""
    am = abs(m)
    vm = sp.Rational(0) if m >= 0 else sp.Rational(1, 2)
    N = sp.sqrt(sp.Rational(2 * factorial(l + am) * factorial(l - am), 2 ** (1 if m == 0 else 0))) \
        / (2 ** am * factorial(l))
    expr = 0
    for t in range((l - am) // 2 + 1):
        for u in range(t + 1):
            v = vm
            vmax = sp.floor(sp.Rational(am, 2) - vm) + vm
            while v <= vmax:
                c = (-1) ** int(t + v - vm) * sp.Rational(1, 4) ** t * comb(l, t) * comb(l - t, am + t) \
                    * comb(t, u) * comb(am, int(2 * v))
                expr += c * x ** int(2 * t + am - 2 * (u + v)) * y ** int(2 * (u + v)) * z ** (l - 2 * t - am)
                v += 1
    return sp.expand(N * expr)


def components(l: int):
    """Polynomials q_mu (including the angular part of the normalisation)."""
    if l >= 0:
        out = []
        for a, b, c in CART_ORDER[l]:
            h = 1 / sp.sqrt(dfact(2 * a - 1) * dfact(2 * b - 1) * dfact(2 * c - 1))
            out.append(h * x ** a * y ** b * z ** c)
        return out
    L = -l
    h = 1 / sp.sqrt(dfact(2 * L - 1))
    ms = [0]
    for k in range(1, L + 1):
        ms += [k, -k]
    return [sp.expand(h * real_solid_harmonic(L, m)) for m in ms]


# Derivative sets: name -> list of output slots per component.
#   val : value
#   grad: value, d/dx, d/dy, d/dz
#   lap : value, d/dx, d/dy, d/dz, laplacian
#   hess: value, d/dx, d/dy, d/dz, xx, xy, xz, yy, yz, zz
DERIV_SETS = {"val": 1, "grad": 4, "lap": 5, "hess": 10}


def derivative_exprs(q, L, dset):
    out = [q * s0]
    if dset == "val":
        return out
    qk = [sp.diff(q, X[k]) for k in range(3)]
    for k in range(3):
        out.append(qk[k] * s0 + X[k] * q * s1)
    if dset == "grad":
        return out
    if dset == "lap":
        lapq = sum(sp.diff(q, X[k], 2) for k in range(3))
        r2 = x * x + y * y + z * z
        out.append(lapq * s0 + (2 * L + 3) * q * s1 + r2 * q * s2)
        return out
    for k, l in [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]:
        qkl = sp.diff(q, X[k], X[l])
        d = q if k == l else 0
        out.append(qkl * s0 + (X[l] * qk[k] + X[k] * qk[l] + d) * s1 + X[k] * X[l] * q * s2)
    return out


def _num(expr):
    """Replace exact constants by double literals so the emitted code is plain."""
    return expr.xreplace({n: sp.Float(n, 20) for n in expr.atoms(sp.Number) if not n.is_Integer})


def gen_function(l, dset):
    L = abs(l)
    comps = components(l)
    nslot = DERIV_SETS[dset]
    exprs = []
    for q in comps:
        exprs += [sp.expand(e) for e in derivative_exprs(q, L, dset)]
    repl, reduced = sp.cse([_num(e) for e in exprs], symbols=sp.numbered_symbols("t"), optimizations="basic")
    ops = sum(sp.count_ops(v) for _, v in repl) + sum(sp.count_ops(e) for e in reduced)
    name = f"ao_shell_{'c' if l >= 0 else 'p'}{L}_{dset}"
    lines = [f"// l = {l}: {len(comps)} functions x {nslot} slots, {int(ops)} ops after CSE",
             "template <typename T, typename St>",
             f"__device__ __forceinline__ void {name}(T x, T y, T z, T s0, T s1, T s2, St&& st) {{"]
    used = set()
    for e in list(reduced) + [v for _, v in repl]:
        used |= {str(s) for s in e.free_symbols}
    for sym in ("x", "y", "z", "s0", "s1", "s2"):
        if sym not in used:
            lines.append(f"  (void){sym};")
    for sym, val in repl:
        lines.append(f"  const T {sym} = {sp.ccode(val)};")
    for ic in range(len(comps)):
        for d in range(nslot):
            e = reduced[ic * nslot + d]
            lines.append(f"  st({d}, {ic}, {sp.ccode(e)});")
    lines.append("}")
    src = "\n".join(lines)
    # literal doubles must stay doubles for T=double and become floats for T=float:
    src = src.replace("pow(", "ct_ipow(")
    # keep arithmetic in T: wrap floating literals so T=float stays single precision
    src = re.sub(r"(?<![\w.])(\d+\.\d+(?:e[-+]?\d+)?)", r"T(\1)", src)
    return name, src, int(ops), len(comps)


def main():
    os.makedirs(OUT_DIR, exist_ok=True)
    parts = ["// GENERATED by chemtoolscuda_amd/codegen/gen_ao.py -- do not edit.",
             "#pragma once",
             "#include <cstdint>",
             "",
             "template <typename T> __device__ __forceinline__ T ct_ipow(T v, int n) {",
             "  T r = v;",
             "  for (int k = 1; k < n; ++k) r *= v;",
             "  return r;",
             "}",
             ""]
    flops = {}
    for dset in DERIV_SETS:
        for l in SHELL_TYPES:
            name, src, ops, nc = gen_function(l, dset)
            parts.append(src)
            parts.append("")
            flops[f"{l}:{dset}"] = {"ops": ops, "nfunc": nc}
    # dispatch helper: shell type code (index into SHELL_TYPES) -> generated function
    for dset in DERIV_SETS:
        parts.append("// st(slot, component, value) stores one AO value of the shell; shell types")
        parts.append("// with |l| > LMAX are compiled out (the kernel is instantiated per basis max L,")
        parts.append("// so a basis without f or g shells does not pay their register pressure)")
        parts.append("template <typename T, int LMAX, typename St>")
        parts.append(f"__device__ __forceinline__ void ao_shell_dispatch_{dset}(int code, T x, T y, T z, T s0, T s1,")
        parts.append("    T s2, St&& st) {")
        parts.append("  switch (code) {")
        for code, l in enumerate(SHELL_TYPES):
            fn = f"ao_shell_{'c' if l >= 0 else 'p'}{abs(l)}_{dset}"
            parts.append(f"    case {code}: if constexpr ({abs(l)} <= LMAX) {fn}<T>(x, y, z, s0, s1, s2, st); break;")
        parts.append("    default: break;")
        parts.append("  }")
        parts.append("}")
        parts.append("")
    with open(os.path.join(OUT_DIR, "ao_poly.h"), "w") as f:
        f.write("\n".join(parts))
    write_shell_monomials()
    with open(os.path.join(OUT_DIR, "ao_flops.json"), "w") as f:
        json.dump({"shell_types": SHELL_TYPES, "flops": flops}, f, indent=1)
    print("wrote", os.path.join(OUT_DIR, "ao_poly.h"))


def write_shell_monomials():
    """Host table: each shell component as a sum of Cartesian monomials (with
    the angular normalisation factor), used by the ESP pre-contraction."""
    rows, offsets = [], [0]
    for code, l in enumerate(SHELL_TYPES):
        for ic, q in enumerate(components(l)):
            poly = sp.Poly(sp.expand(q), x, y, z)
            for (a, b, c), coef in zip(poly.monoms(), poly.coeffs()):
                rows.append(f"    {{{float(coef)!r}, {a}, {b}, {c}, {ic}}},")
        offsets.append(len(rows))
    text = ["// GENERATED by chemtoolscuda_amd/codegen/gen_ao.py -- do not edit.",
            "// Shell components as sums of Cartesian monomials x^a y^b z^c, including the",
            "// angular part of the L2 normalisation (the same polynomials the AO kernels use).",
            "#pragma once",
            "#include <cstdint>",
            "namespace ct {",
            "struct MonoTerm {",
            "  double c;",
            "  int8_t a, b, z, comp;",
            "};",
            f"constexpr int kShellMonoOffset[{len(offsets)}] = {{{', '.join(map(str, offsets))}}};",
            "constexpr MonoTerm kShellMonos[] = {",
            *rows,
            "};",
            "}  // namespace ct",
            ""]
    with open(os.path.join(OUT_DIR, "shell_monomials.h"), "w") as f:
        f.write("\n".join(text))


if __name__ == "__main__":
    main()
