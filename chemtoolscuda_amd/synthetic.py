"""Synthetic workloads: molecules, density matrices and grids.

There is no network and no wavefunction file for the benchmark systems, so
each workload is generated from a fixed recipe and seed
(``numpy.random.default_rng(index)``, index = 1..5).  The recipes:

1. ``water``  -- O at the origin, H at 1.81 bohr, H-O-H 104.5 deg.  O 3s2p1d
   Cartesian (15 functions), each H 2s1p (5): M = 25.  3-6 primitives per
   shell, exponents log-uniform in [0.1, 5000] (s) / [0.1, 50] (p, d),
   coefficients U(0.1, 1).  P = 2 C C^T with C the first 5 columns of
   Q from QR of a seeded N(0,1) M x M matrix.  10^4 points uniform in
   [-4, 4]^3 bohr.
2. ``organic60`` -- 60 atoms C30 N3 O3 H24 at random positions in an 18 bohr
   box with >= 2.2 bohr separation; cc-pVDZ-like pure shells (heavy 3s2p1d =
   14 functions, H 2s1p = 5): M = 624.  n_occ = floor(249 / 2) = 124.
   Cubic grid, 0.2 bohr spacing, 5 bohr padding (about 2.8M points).
3. ``large300`` -- 300 atoms C55 N10 O10 H225 in a 31 bohr box; Cartesian
   shells with f+g on heavy atoms (2s1p1d1f1g = 36 functions) and f on H
   (2s1f = 12): M = 5400.  n_occ = 352.  10^6 points, drawn per atom with
   radius ~ Gamma(3, 1) (density r^2 e^{-r}) and isotropic direction.
4. ``diff40`` -- 40 atoms (C20 O4 H16), DZ-like pure shells up to d (heavy
   3s2p1d = 14, H 2s1p = 5): M = 416, 5*10^5 points drawn as in 3.  Also
   returns weights w ~ U(0,1) and a target density matrix P + sym N(0, 1e-3).
5. ``esp30`` -- 30 atoms (C12 O6 H12), pure shells up to d (M = 312), 2*10^5
   points on the 1.4 x van der Waals surface.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from .basis import BasisSet, Shell

# Atomic numbers and van der Waals radii (Bondi, converted to bohr).
_Z = {"H": 1, "C": 6, "N": 7, "O": 8}
_ANG = 1.0 / 0.529177210903
_VDW = {"H": 1.20 * _ANG, "C": 1.70 * _ANG, "N": 1.55 * _ANG, "O": 1.52 * _ANG}


@dataclass
class Workload:
    name: str
    basis: BasisSet
    P: np.ndarray
    charges: np.ndarray
    points: np.ndarray
    n_occ: int
    extra: Optional[dict] = None


def _shell(rng, atom, l, nprim, lo, hi):
    exps = np.sort(np.exp(rng.uniform(np.log(lo), np.log(hi), size=nprim)))[::-1]
    coefs = rng.uniform(0.1, 1.0, size=nprim)
    return Shell(atom, l, exps, coefs)


def _density_matrix(rng, M, n_occ):
    A = rng.standard_normal((M, M))
    Q, _ = np.linalg.qr(A[:, :n_occ])
    C = Q[:, :n_occ]
    return 2.0 * C @ C.T


def _random_positions(rng, n, box, min_sep):
    pos = np.empty((0, 3))
    while pos.shape[0] < n:
        cand = rng.uniform(0.0, box, size=3)
        if pos.shape[0] == 0 or np.min(np.linalg.norm(pos - cand, axis=1)) >= min_sep:
            pos = np.vstack([pos, cand])
    return pos


def _per_atom_radial_points(rng, centres, n_total):
    nat = centres.shape[0]
    counts = np.full(nat, n_total // nat)
    counts[: n_total - counts.sum()] += 1
    out = []
    for a in range(nat):
        r = rng.gamma(3.0, 1.0, size=counts[a])
        v = rng.standard_normal((counts[a], 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        out.append(centres[a] + r[:, None] * v)
    return np.concatenate(out)


def _elements(counts):
    els = []
    for sym, n in counts:
        els += [sym] * n
    return els


def water(n_points: int = 10_000) -> Workload:
    rng = np.random.default_rng(1)
    half = np.deg2rad(104.5 / 2.0)
    xyz = np.array([[0.0, 0.0, 0.0],
                    [1.81 * np.sin(half), 0.0, 1.81 * np.cos(half)],
                    [-1.81 * np.sin(half), 0.0, 1.81 * np.cos(half)]])
    shells = []
    for l in (0, 0, 0, 1, 1, 2):
        lo, hi = (0.1, 5000.0) if l == 0 else (0.1, 50.0)
        shells.append(_shell(rng, 0, l, int(rng.integers(3, 7)), lo, hi))
    for a in (1, 2):
        for l in (0, 0, 1):
            lo, hi = (0.1, 5000.0) if l == 0 else (0.1, 50.0)
            shells.append(_shell(rng, a, l, int(rng.integers(3, 7)), lo, hi))
    basis = BasisSet(shells, xyz, np.array([8, 1, 1]))
    M = basis.nbasis
    P = _density_matrix(rng, M, 5)
    pts = rng.uniform(-4.0, 4.0, size=(n_points, 3))
    return Workload("water", basis, P, np.array([8.0, 1.0, 1.0]), pts, 5)


def _heavy_dz(rng, atom, pure=True):
    d = -2 if pure else 2
    return [_shell(rng, atom, 0, 6, 0.5, 5000.0), _shell(rng, atom, 0, 2, 0.3, 3.0),
            _shell(rng, atom, 0, 1, 0.1, 0.3), _shell(rng, atom, 1, 3, 0.3, 20.0),
            _shell(rng, atom, 1, 1, 0.1, 0.3), _shell(rng, atom, d, 1, 0.3, 1.0)]


def _hydrogen_dz(rng, atom):
    return [_shell(rng, atom, 0, 3, 0.3, 30.0), _shell(rng, atom, 0, 1, 0.1, 0.3),
            _shell(rng, atom, 1, 1, 0.5, 1.2)]


def _molecule(rng, elements, box, builder):
    pos = _random_positions(rng, len(elements), box, 2.2)
    shells = []
    for a, el in enumerate(elements):
        shells += builder(rng, a, el)
    nums = np.array([_Z[e] for e in elements])
    return BasisSet(shells, pos, nums), nums.astype(np.float64)


def _dz_builder(rng, a, el):
    return _hydrogen_dz(rng, a) if el == "H" else _heavy_dz(rng, a)


def organic60(spacing: float = 0.2, padding: float = 5.0) -> Workload:
    rng = np.random.default_rng(2)
    els = _elements([("C", 30), ("N", 3), ("O", 3), ("H", 24)])
    rng.shuffle(els)
    basis, Z = _molecule(rng, els, 18.0, _dz_builder)
    n_occ = int(Z.sum()) // 2
    P = _density_matrix(rng, basis.nbasis, n_occ)
    lo = basis.atom_xyz.min(axis=0) - padding
    hi = basis.atom_xyz.max(axis=0) + padding
    axes = [np.arange(lo[k], hi[k] + 1e-9, spacing) for k in range(3)]
    g = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)
    return Workload("organic60", basis, P, Z, np.ascontiguousarray(g), n_occ,
                    {"grid_shape": tuple(len(a) for a in axes), "origin": lo, "spacing": spacing})


def _tz_builder(rng, a, el):
    if el == "H":
        return [_shell(rng, a, 0, 3, 0.3, 30.0), _shell(rng, a, 0, 1, 0.1, 0.3),
                _shell(rng, a, 3, 1, 0.8, 1.2)]
    return [_shell(rng, a, 0, 6, 0.5, 5000.0), _shell(rng, a, 0, 1, 0.1, 0.3),
            _shell(rng, a, 1, 3, 0.15, 20.0), _shell(rng, a, 2, 1, 0.3, 1.5),
            _shell(rng, a, 3, 1, 0.5, 1.2), _shell(rng, a, 4, 1, 0.8, 1.5)]


def large300(n_points: int = 1_000_000) -> Workload:
    rng = np.random.default_rng(3)
    els = _elements([("C", 55), ("N", 10), ("O", 10), ("H", 225)])
    rng.shuffle(els)
    basis, Z = _molecule(rng, els, 31.0, _tz_builder)
    n_occ = int(Z.sum()) // 2
    P = _density_matrix(rng, basis.nbasis, n_occ)
    pts = _per_atom_radial_points(rng, basis.atom_xyz, n_points)
    return Workload("large300", basis, P, Z, pts, n_occ)


def diff40(n_points: int = 500_000) -> Workload:
    rng = np.random.default_rng(4)
    els = _elements([("C", 20), ("O", 4), ("H", 16)])
    rng.shuffle(els)
    basis, Z = _molecule(rng, els, 14.0, _dz_builder)
    n_occ = int(Z.sum()) // 2
    P = _density_matrix(rng, basis.nbasis, n_occ)
    pts = _per_atom_radial_points(rng, basis.atom_xyz, n_points)
    w = rng.uniform(0.0, 1.0, size=n_points)
    E = rng.normal(0.0, 1e-3, size=P.shape)
    P_target = P + 0.5 * (E + E.T)
    return Workload("diff40", basis, P, Z, pts, n_occ, {"weights": w, "P_target": P_target})


def surface_points(rng, xyz, elements, n_points, scale=1.4):
    radii = np.array([_VDW[e] * scale for e in elements])
    out = []
    need = n_points
    while need > 0:
        a = rng.integers(0, len(elements), size=2 * need)
        v = rng.standard_normal((2 * need, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        cand = xyz[a] + radii[a, None] * v
        d = np.linalg.norm(cand[:, None, :] - xyz[None, :, :], axis=2)
        inside = (d < radii[None, :] * (1.0 - 1e-9)).any(axis=1)
        keep = cand[~inside][:need]
        out.append(keep)
        need -= keep.shape[0]
    return np.concatenate(out)


def esp30(n_points: int = 200_000) -> Workload:
    rng = np.random.default_rng(5)
    els = _elements([("C", 12), ("O", 6), ("H", 12)])
    rng.shuffle(els)
    basis, Z = _molecule(rng, els, 12.0, _dz_builder)
    n_occ = int(Z.sum()) // 2
    P = _density_matrix(rng, basis.nbasis, n_occ)
    pts = surface_points(rng, basis.atom_xyz, els, n_points)
    return Workload("esp30", basis, P, Z, pts, n_occ)


WORKLOADS = {"water": water, "organic60": organic60, "large300": large300, "diff40": diff40,
             "esp30": esp30}


def morton_order(points: np.ndarray, bits: int = 10) -> np.ndarray:
    """Permutation that sorts points along a 3-D Morton (Z-order) curve."""
    p = np.asarray(points, dtype=np.float64)
    lo = p.min(axis=0)
    span = np.maximum(p.max(axis=0) - lo, 1e-12)
    q = np.minimum(((p - lo) / span * ((1 << bits) - 1)).astype(np.uint64), (1 << bits) - 1)
    code = np.zeros(p.shape[0], dtype=np.uint64)
    for b in range(bits):
        for k in range(3):
            code |= ((q[:, k] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + k)
    return np.argsort(code, kind="stable")
