"""Uniform grids and cube files (host side, CPU)."""
import numpy as np
import pytest

from chemtoolscuda_amd import synthetic
from chemtoolscuda_amd.cube import UniformGrid, atom_padding, read_cube, write_cube


def test_grid_from_molecule_encloses_every_padded_atom():
    w = synthetic.water(10)
    numbers = np.array([8, 1, 1])
    g = UniformGrid.from_molecule(numbers, w.basis.atom_xyz, spacing=0.25, extension=1.0)
    # axes: orthogonal, |axis| = spacing
    assert np.allclose(g.axes @ g.axes.T, 0.25 ** 2 * np.eye(3), atol=1e-12)
    # every atom, padded by its radius + extension, lies inside the box
    unit = g.axes / 0.25
    local = (w.basis.atom_xyz - g.origin) @ unit.T  # coordinates along the grid axes
    ext = np.array(g.shape) * 0.25
    for z, x in zip(numbers, local):
        pad = atom_padding(int(z)) + 1.0
        assert np.all(x - pad >= -1e-9) and np.all(x + pad <= ext + 1e-9)
    assert g.points.shape == (g.size, 3)
    # last index fastest
    p = g.points
    assert np.allclose(p[1] - p[0], g.axes[2]) and np.allclose(p[g.shape[2]] - p[0], g.axes[1])


def test_cube_file_round_trip(tmp_path):
    g = UniformGrid(origin=[-1.0, 0.5, 2.0], axes=np.diag([0.3, 0.2, 0.1]), shape=(3, 4, 7))
    vals = np.random.default_rng(0).uniform(-2, 2, size=g.shape) * 10.0 ** np.random.default_rng(1).integers(-8, 3, size=g.shape)
    path = tmp_path / "x.cube"
    write_cube(str(path), g, vals, [8, 1], [[0, 0, 0], [0, 0, 1.8]])
    g2, v2, nums, xyz = read_cube(str(path))
    assert g2.shape == g.shape and np.allclose(g2.origin, g.origin) and np.allclose(g2.axes, g.axes)
    assert np.allclose(v2, vals, rtol=1e-5, atol=0)
    assert list(nums) == [8, 1] and np.allclose(xyz[1], [0, 0, 1.8])


def test_padding_follows_the_reference_branch_chain():
    """src/utils.cpp:163-188: Z == 1 2.6, Z <= 6 2.8 (ghost atoms, Z = 0,
    too), 7 3.0, 8 3.3, 9 2.9, <= 15 3.3, 16 3.4, <= 35 3.5, <= 52 3.7, 53 3.8,
    and no padding above 53 (the chain has no branch there)."""
    expect = {0: 2.8, 1: 2.6, 2: 2.8, 6: 2.8, 7: 3.0, 8: 3.3, 9: 2.9, 10: 3.3, 15: 3.3, 16: 3.4, 17: 3.5,
              35: 3.5, 36: 3.7, 51: 3.7, 52: 3.7, 53: 3.8, 54: 0.0, 92: 0.0}
    for z, pad in expect.items():
        assert atom_padding(z) == pad, z
    assert atom_padding(7.5) == 3.3  # a non-integer core charge takes the first bound it meets
