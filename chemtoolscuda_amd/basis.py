"""Basis-set data model (host side).

Mirrors the reference's ``GeneralizedContractionShell`` / ``MolecularBasis``
(reference ``include/contracted_shell.h:10-77``) with generalised contractions
already split into segments that share exponents (an SP shell becomes an s
shell and a p shell with the same exponents).

Conventions (reference ``src/iodata.cpp:57-133``):
  * ``l`` is the signed angular momentum: 0..4 Cartesian, -2..-4 pure.
  * ``coefs`` are contraction coefficients of L2-normalised primitives.
  * Basis functions are numbered shell by shell in the order given, with the
    fchk component order inside a shell (see ``n_functions``).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

import numpy as np

_NFUNC = {0: 1, 1: 3, 2: 6, -2: 5, 3: 10, -3: 7, 4: 15, -4: 9}


def n_functions(l: int) -> int:
    """Number of basis functions in a shell (reference include/contracted_shell.h:10-37)."""
    try:
        return _NFUNC[int(l)]
    except KeyError:
        raise ValueError(f"unsupported angular momentum {l}") from None


@dataclass
class Shell:
    atom: int
    l: int
    exps: np.ndarray
    coefs: np.ndarray

    def __post_init__(self):
        self.exps = np.asarray(self.exps, dtype=np.float64).ravel()
        self.coefs = np.asarray(self.coefs, dtype=np.float64).ravel()
        if self.exps.shape != self.coefs.shape or self.exps.size == 0:
            raise ValueError("shell needs matching, non-empty exponent and coefficient arrays")
        n_functions(self.l)

    @property
    def nfunc(self) -> int:
        return n_functions(self.l)


@dataclass
class BasisSet:
    shells: List[Shell]
    atom_xyz: np.ndarray
    atom_numbers: np.ndarray = field(default=None)

    def __post_init__(self):
        self.atom_xyz = np.ascontiguousarray(self.atom_xyz, dtype=np.float64).reshape(-1, 3)
        if self.atom_numbers is None:
            self.atom_numbers = np.zeros(self.atom_xyz.shape[0], dtype=np.int32)
        self.atom_numbers = np.asarray(self.atom_numbers, dtype=np.int32)
        for s in self.shells:
            if not 0 <= s.atom < self.atom_xyz.shape[0]:
                raise ValueError(f"shell centre index {s.atom} out of range")

    @property
    def nbasis(self) -> int:
        return int(sum(s.nfunc for s in self.shells))

    @property
    def natom(self) -> int:
        return int(self.atom_xyz.shape[0])

    def flat(self) -> dict:
        """Flat arrays: the layout passed across the C ABI (include/chemtools_amd.h)."""
        nprim = np.array([s.exps.size for s in self.shells], dtype=np.int32)
        off = np.zeros(len(self.shells), dtype=np.int32)
        if len(self.shells) > 1:
            off[1:] = np.cumsum(nprim)[:-1]
        return {
            "shell_atom": np.array([s.atom for s in self.shells], dtype=np.int32),
            "shell_l": np.array([s.l for s in self.shells], dtype=np.int32),
            "shell_nprim": nprim,
            "shell_prim_off": off,
            "exps": np.concatenate([s.exps for s in self.shells]) if self.shells else np.zeros(0),
            "coefs": np.concatenate([s.coefs for s in self.shells]) if self.shells else np.zeros(0),
            "atom_xyz": self.atom_xyz.copy(),
        }

    def with_coordinates(self, atom_xyz) -> "BasisSet":
        return BasisSet(self.shells, np.asarray(atom_xyz, dtype=np.float64), self.atom_numbers)
