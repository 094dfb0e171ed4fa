"""Cutoff radii of the mixed factor route's spatial screening
(chemtoolscuda_amd/csrc/screen_host.cpp), host side only: no GPU needed.

The reference has no screening of the density path (it evaluates every basis
function at every point, src/evaluate_density.cu:25-411), so the radius is
pinned against the oracle's AO values and derivatives (oracle/oracle.py ao(),
an independent FP64 restatement of the AO formulas): for shells of every type
with random contractions, no value, first or second derivative or Laplacian
of any function of the shell exceeds eps at any sampled point beyond the
radius; and for one-primitive s shells, whose bound is exact up to the
Laplacian's |-6a| term, the radius is tight (something exceeds eps a little
inside it).
"""
import numpy as np
import pytest

from oracle import oracle


@pytest.fixture(scope="module")
def capi():
    from chemtoolscuda_amd import capi as c

    c.lib()
    return c


def _single_shell(l, exps, coefs):
    return {"shell_atom": np.array([0], np.int32), "shell_l": np.array([l], np.int32),
            "shell_nprim": np.array([len(exps)], np.int32), "shell_prim_off": np.array([0], np.int32),
            "exps": np.asarray(exps, float), "coefs": np.asarray(coefs, float), "atom_xyz": np.zeros((1, 3))}


def _max_abs(flat, pts):
    ao = oracle.ao(flat, pts, 2)  # value, d_x, d_y, d_z, d_xx, d_xy, d_xz, d_yy, d_yz, d_zz
    lap = ao[4] + ao[7] + ao[9]
    return max(np.abs(ao).max(), np.abs(lap).max())


def _shell_points(rng, radii, ndir=300):
    d = rng.normal(size=(ndir, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    # the coordinate axes and diagonals too (where monomials peak)
    extra = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [1, -1, 0], [0, 1, -1]], float)
    d = np.vstack([d, extra / np.linalg.norm(extra, axis=1)[:, None]])
    return np.vstack([r * d for r in radii])


@pytest.mark.parametrize("l", [0, 1, 2, 3, 4, -2, -3, -4])
@pytest.mark.parametrize("eps", [1e-10, 1e-15])
def test_radius_bounds_every_derivative(capi, l, eps):
    rng = np.random.default_rng(100 + 10 * abs(l) + (l < 0))
    for trial in range(6):
        nprim = int(rng.integers(1, 7))
        exps = np.exp(rng.uniform(np.log(0.05), np.log(5e3), nprim))
        coefs = rng.uniform(-1.0, 1.0, nprim)
        R = capi.shell_screening_radius(l, exps, coefs, eps)
        assert np.isfinite(R) and R > 0
        pts = _shell_points(rng, [R, R * 1.02, R + 0.5, 1.3 * R, 2.0 * R])
        got = _max_abs(_single_shell(l, exps, coefs), pts)
        assert got <= eps, (l, trial, R, got)


@pytest.mark.parametrize("alpha", [0.1, 1.0, 30.0])
def test_radius_is_tight_for_s_primitives(capi, alpha):
    eps = 1e-12
    R = capi.shell_screening_radius(0, [alpha], [1.0], eps)
    rng = np.random.default_rng(7)
    flat = _single_shell(0, [alpha], [1.0])
    assert _max_abs(flat, _shell_points(rng, [R])) <= eps
    assert _max_abs(flat, _shell_points(rng, [0.97 * R])) > eps


def test_radius_errors(capi):
    with pytest.raises(RuntimeError, match="positive"):
        capi.shell_screening_radius(0, [1.0], [1.0], 0.0)
    with pytest.raises(RuntimeError, match="unsupported"):
        capi.shell_screening_radius(5, [1.0], [1.0], 1e-12)
    with pytest.raises(RuntimeError, match="exponents"):
        capi.shell_screening_radius(0, [-1.0], [1.0], 1e-12)
