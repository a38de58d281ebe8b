"""Potential interpolation and file formats (multiscat_amd/formats.py) against closed forms,
the reference's data files (tests/data of the reference, kept as golden fixtures) and
round trips."""

from __future__ import annotations

import numpy as np
import pytest

from multiscat_amd import basis as B
from multiscat_amd import formats as FM
from tests.conftest import GOLDEN, golden, wolken_expected_intensities


def _grid(nx, ny, nzs, z0, z1):
    x = np.arange(nx)[:, None, None] / nx
    y = np.arange(ny)[None, :, None] / ny
    z = np.linspace(z0, z1, nzs)
    return x, y, z


def test_interpolate_known_modes_exactly():
    """V = (a + b z) + c(z) cos(2 pi x) + d(z) sin(2 pi (x + y)) with linear c, d: the
    components are exact (linear in z, so the interpolation is exact too)."""
    nx, ny = 12, 10
    x, y, z = _grid(nx, ny, 41, -1.0, 7.0)
    zz = z[None, None, :]
    V = (2.0 + 0.5 * zz) + (1.0 - 0.1 * zz) * np.cos(2 * np.pi * x) \
        + (0.3 + 0.2 * zz) * np.sin(2 * np.pi * (x + y))
    shape = (5, 5, 30)
    out = FM.interpolate_potential(V, z, shape=shape, z_range=(-1.0, 7.0))
    nodes, _ = B.lobatto_rule(30, -1.0, 7.0)
    exp = np.zeros(shape, complex)
    exp[0, 0] = 2.0 + 0.5 * nodes
    exp[1, 0] = exp[-1, 0] = 0.5 * (1.0 - 0.1 * nodes)          # cos -> (e^{i} + e^{-i}) / 2
    exp[1, 1] = (0.3 + 0.2 * nodes) / 2j                        # sin -> (e^{i} - e^{-i}) / 2i
    exp[-1, -1] = -(0.3 + 0.2 * nodes) / 2j
    np.testing.assert_allclose(out, exp, atol=1e-12)


def test_interpolate_weighted_form_carries_the_lobatto_basis_weights():
    """weighted=True is the reference's storage (multiscat/interpolate.py:84-92): node k times
    the Lobatto basis weight 1/sqrt(w_k), with sqrt(w_k) the reference Fortran helper's
    get_lobatto_weights output (golden fixture helpers.npz) on the same rule."""
    zmin, zmax, n = golden("helpers.npz")["lobatto0"]
    sqrt_w = golden("helpers.npz")["lobatto0_w"]
    nz = int(n)
    x, y, z = _grid(6, 6, 17, zmin, zmax)
    V = np.exp(-z)[None, None, :] * (1.0 + 0.2 * np.cos(2 * np.pi * x) + 0 * y)
    plain = FM.interpolate_potential(V, z, shape=(3, 3, nz), z_range=(zmin, zmax))
    weighted = FM.interpolate_potential(V, z, shape=(3, 3, nz), z_range=(zmin, zmax),
                                        weighted=True)
    np.testing.assert_allclose(weighted, plain / sqrt_w[None, None, :], rtol=1e-11)
    # a constant potential's weighted form is 1/sqrt(w_k): largest at the two end nodes
    const = FM.interpolate_potential(np.ones((2, 2, 5)), np.linspace(zmin, zmax, 5),
                                     shape=(1, 1, nz), z_range=(zmin, zmax),
                                     weighted=True)[0, 0].real
    assert np.argmax(const) in (0, nz - 1) and np.all(const > 0)


def test_interpolate_crops_and_pads_the_g_grid():
    nx, ny = 8, 8
    x, y, z = _grid(nx, ny, 5, 0.0, 4.0)
    V = np.cos(2 * np.pi * 3 * x) + 0 * z[None, None, :]         # mode +-3 only
    small = FM.interpolate_potential(V, z, shape=(5, 5, 6), z_range=(0.0, 4.0))
    assert np.abs(small).max() < 1e-14                           # +-3 outside a 5-grid
    big = FM.interpolate_potential(V, z, shape=(11, 11, 6), z_range=(0.0, 4.0))
    np.testing.assert_allclose(big[3, 0], 0.5, atol=1e-14)
    np.testing.assert_allclose(big[-3, 0], 0.5, atol=1e-14)
    assert np.abs(big[5, 0]).max() < 1e-14                       # mode 5 absent from 8 points


def test_interpolate_rejects_bad_grids():
    with pytest.raises(ValueError):
        FM.interpolate_potential(np.zeros((4, 4, 3)), [0.0, 2.0, 1.0], shape=(3, 3, 4),
                                 z_range=(0, 2))


def test_potential_file_round_trip_and_reference_layout(tmp_path):
    V = golden("wolken_potential.npz")["potential_mev"]
    p = tmp_path / "potential.npy"
    FM.write_potential_npy(p, V)
    flat = np.load(p, allow_pickle=False)
    assert flat.shape == (V.size,) and flat.dtype == np.complex128  # the reference's layout
    np.testing.assert_array_equal(FM.read_potential_npy(p, V.shape), V)


def test_intensity_text_matches_reference_file_and_round_trips(tmp_path):
    path = GOLDEN / "wolken_expected_intensities.txt"
    P = FM.read_intensities_txt(path, (10, 10))
    np.testing.assert_array_equal(P, wolken_expected_intensities())
    out = tmp_path / "P.txt"
    FM.write_intensities_txt(out, P)
    np.testing.assert_allclose(FM.read_intensities_txt(out, (10, 10)), P, rtol=1e-8, atol=0)
    assert out.read_text().startswith("# kx ky intensity")
