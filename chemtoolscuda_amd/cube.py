"""Uniform grids and Gaussian cube files (host side).

``UniformGrid`` describes the grid ``origin + i axes[0] + j axes[1] + k axes[2]``
(0 <= i < shape[0], ...; k fastest) that ``Molecule.compute_on_cubic_grid``
and ``compute_electron_density_on_cubic_grid`` evaluate on the GPU without
uploading points.  ``UniformGrid.from_molecule`` builds the grid the way the
reference does (``get_grid_from_coordinates_charges``, src/utils.cpp:127-257,
and ``UniformGrid``, include/utils.h:17-65): the box is aligned with the
principal axes of the nuclear-charge "inertia" tensor and padded around each
atom by an element-dependent distance plus an extension.

``write_cube`` / ``read_cube`` handle the Gaussian cube format (values with
the last index fastest, six per line).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# Padding around an atom of nuclear charge Z (bohr), the branch chain of
# get_grid_from_coordinates_charges (src/utils.cpp:163-188): upper bounds of
# Z, tested in order; Z above 53 gets no padding.
_PAD = [(1, 1, 2.6), (None, 6, 2.8), (7, 7, 3.0), (8, 8, 3.3), (9, 9, 2.9), (None, 15, 3.3), (16, 16, 3.4),
        (None, 35, 3.5), (None, 52, 3.7), (53, 53, 3.8)]


def atom_padding(z) -> float:
    """Padding of an atom of nuclear charge z: the first matching branch
    (equal to z, or z at most the bound), 0 beyond 53."""
    for eq, hi, pad in _PAD:
        if (eq is not None and z == eq) or (eq is None and z <= hi):
            return pad
    return 0.0


@dataclass
class UniformGrid:
    origin: np.ndarray  # (3,)
    axes: np.ndarray    # (3, 3): row r is the step along index r
    shape: tuple        # (n0, n1, n2)

    def __post_init__(self):
        self.origin = np.asarray(self.origin, dtype=np.float64).reshape(3)
        self.axes = np.asarray(self.axes, dtype=np.float64).reshape(3, 3)
        self.shape = tuple(int(v) for v in self.shape)
        if len(self.shape) != 3 or min(self.shape) <= 0:
            raise ValueError("shape must be three positive integers")

    @property
    def size(self) -> int:
        return int(np.prod(self.shape))

    @property
    def points(self) -> np.ndarray:
        """All points, (n0 n1 n2, 3), last index fastest."""
        idx = np.indices(self.shape).reshape(3, -1).T.astype(np.float64)
        return self.origin + idx @ self.axes

    @classmethod
    def from_molecule(cls, numbers, coordinates, spacing: float = 0.2, extension: float = 0.0) -> "UniformGrid":
        """Box around the molecule in its principal-axes frame (src/utils.cpp:127-257).

        The frame is the eigenbasis v of I = sum_A Z_A (|r_A - c|^2 1 - (r_A - c)(r_A - c)^T),
        c the charge-weighted centre; in that frame each atom spans
        r_A v -/+ pad(Z_A) and the box is their union, widened by `extension`.
        The axes are the rotated unit vectors times `spacing`; the shape is
        ceil(extent / spacing) along each axis (include/utils.h:39-47)."""
        Z = np.asarray(numbers, dtype=np.float64)
        R = np.asarray(coordinates, dtype=np.float64).reshape(-1, 3)
        com = Z @ R / Z.sum()
        itensor = np.zeros((3, 3))
        for z, xyz in zip(Z, R - com):
            itensor += z * (np.dot(xyz, xyz) * np.eye(3) - np.outer(xyz, xyz))
        _, v = np.linalg.eigh(itensor)
        rc = R @ v
        pad = np.array([atom_padding(z) for z in Z])
        lo = np.min(rc - pad[:, None], axis=0) - extension
        hi = np.max(rc + pad[:, None], axis=0) + extension
        shape = np.ceil(np.abs(hi - lo) / spacing).astype(int)
        return cls(origin=lo @ v.T, axes=v.T * spacing, shape=tuple(shape))


def write_cube(path: str, grid: UniformGrid, values, numbers, coordinates, charges=None,
               title: str = "chemtoolscuda_amd", comment: str = "") -> None:
    """Gaussian cube file; `values` of shape grid.shape (or flat, last index fastest)."""
    vals = np.asarray(values, dtype=np.float64).reshape(-1)
    if vals.size != grid.size:
        raise ValueError("values do not match the grid")
    numbers = np.asarray(numbers, dtype=int)
    coordinates = np.asarray(coordinates, dtype=np.float64).reshape(-1, 3)
    charges = numbers.astype(np.float64) if charges is None else np.asarray(charges, dtype=np.float64)
    lines = [title, comment,
             f"{len(numbers):5d} {grid.origin[0]:12.6f} {grid.origin[1]:12.6f} {grid.origin[2]:12.6f}"]
    for n, ax in zip(grid.shape, grid.axes):
        lines.append(f"{n:5d} {ax[0]:12.6f} {ax[1]:12.6f} {ax[2]:12.6f}")
    for z, q, xyz in zip(numbers, charges, coordinates):
        lines.append(f"{z:5d} {q:12.6f} {xyz[0]:12.6f} {xyz[1]:12.6f} {xyz[2]:12.6f}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
        rows = vals.reshape(-1, grid.shape[2])
        for row in rows:  # one (i, j) line block at a time, six values per line
            for s in range(0, row.size, 6):
                f.write(" ".join(f"{v:13.5E}" for v in row[s:s + 6]) + "\n")


def read_cube(path: str):
    """Returns (grid, values of shape grid.shape, numbers, coordinates)."""
    with open(path) as f:
        f.readline()
        f.readline()
        head = f.readline().split()
        natom = int(head[0])
        origin = np.array([float(v) for v in head[1:4]])
        shape, axes = [], []
        for _ in range(3):
            t = f.readline().split()
            shape.append(int(t[0]))
            axes.append([float(v) for v in t[1:4]])
        numbers, coords = [], []
        for _ in range(abs(natom)):
            t = f.readline().split()
            numbers.append(int(t[0]))
            coords.append([float(v) for v in t[2:5]])
        data = np.array(f.read().split(), dtype=np.float64)
    grid = UniformGrid(origin, np.array(axes), tuple(shape))
    return grid, data.reshape(grid.shape), np.array(numbers), np.array(coords)
