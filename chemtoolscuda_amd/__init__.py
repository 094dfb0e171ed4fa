"""MI355X-native evaluator of Gaussian-basis electron-density descriptors.

The product path is ``Molecule`` (pybind11 over the C ABI in
``include/chemtools_amd.h``), whose methods run hand-written HIP kernels for
gfx950.  Importing this package fails loudly when the native library has not
been built: there is no CPU fallback.
"""
from __future__ import annotations

import os

_HERE = os.path.dirname(os.path.abspath(__file__))

# One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64 (same
# soname, linked by file name).  Loading it first makes the native library
# bind to that copy; the other order would map two runtimes and PyTorch's
# would then find no device.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - the library itself does not need torch
    pass

try:
    from . import _chemtools_amd as _native  # noqa: F401
except ImportError as exc:  # pragma: no cover - exercised only on an unbuilt tree
    raise ImportError(
        "chemtoolscuda_amd: native extension not built; run `python chemtoolscuda_amd/build_ext.py` "
        f"(or __graft_entry__.build()) first ({exc})") from exc

from .basis import BasisSet, Shell, n_functions  # noqa: E402

Molecule = _native.Molecule

LIBRARY_PATH = os.path.join(_HERE, "libchemtools_amd.so")


def molecule_from_basis(basis: BasisSet, P, charges=None, numbers=None) -> "Molecule":
    """Builds a ``Molecule`` from an in-memory ``BasisSet`` and density matrix."""
    import numpy as np

    f = basis.flat()
    if charges is None:
        charges = np.asarray(basis.atom_numbers, dtype=np.float64)
    if numbers is None:
        numbers = basis.atom_numbers
    return Molecule(f["shell_atom"], f["shell_l"], f["shell_nprim"], f["shell_prim_off"], f["exps"], f["coefs"],
                    f["atom_xyz"], np.asarray(charges, dtype=np.float64), np.asarray(P, dtype=np.float64),
                    np.asarray(numbers, dtype=np.int32))


def read_wavefunction(path: str) -> dict:
    """Host-side contents of a .fchk / .wfn / .wfx / Molden file as the native
    readers produce them (shell table, P, coordinates, charges, numbers)."""
    from .capi import read_wavefunction as _read

    return _read(path)


def basis_from_wavefunction(w: dict) -> BasisSet:
    """``BasisSet`` of a ``read_wavefunction`` result."""
    shells = [Shell(int(a), int(l), w["exps"][o:o + n], w["coefs"][o:o + n])
              for a, l, n, o in zip(w["shell_atom"], w["shell_l"], w["shell_nprim"], w["shell_prim_off"])]
    return BasisSet(shells, w["atom_xyz"], w["numbers"])


__all__ = ["Molecule", "BasisSet", "Shell", "n_functions", "molecule_from_basis", "read_wavefunction",
           "basis_from_wavefunction", "LIBRARY_PATH"]
