"""Host input builders (multiscat_amd.basis) against the reference helpers and closed forms.

Mirrors the reference's tests/lobatto_test.py:15-106 and
tests/test_fixed_example.py:469-520,638-691 (there: SciPy-side vs Fortran helpers).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import golden
from multiscat_amd import basis as B


@pytest.fixture(scope="module")
def helpers():
    return golden("helpers.npz")


@pytest.mark.parametrize("i", range(4))
def test_lobatto_rule_matches_reference(helpers, i):
    zmin, zmax, n = helpers[f"lobatto{i}"]
    x, w = B.lobatto_rule(int(n), zmin, zmax)
    np.testing.assert_allclose(x, helpers[f"lobatto{i}_x"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(np.sqrt(w), helpers[f"lobatto{i}_w"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n", [2, 3, 5, 12, 40])
def test_lobatto_rule_integrates_exactly(n):
    """An n-point GLL rule is exact for polynomials of degree <= 2n - 3."""
    x, w = B.lobatto_rule(n, -1.5, 8.0)
    for deg in range(0, 2 * n - 2):
        exact = (8.0 ** (deg + 1) - (-1.5) ** (deg + 1)) / (deg + 1)
        assert abs((w * x**deg).sum() - exact) < 1e-10 * max(1.0, abs(exact))


def test_derivative_matrix_against_polynomials():
    """D[k, i] = l_i'(x_k), checked against numpy's polynomial derivatives
    (the reference's test_lobatto_derivatives_against_explicit)."""
    n, zmin, zmax = 9, 0.0, 2.0
    x, _ = B.lobatto_rule(n, zmin, zmax)
    D = B.derivative_matrix(n, zmin, zmax)
    for i in range(n):
        roots = np.delete(x, i)
        p = np.polynomial.Polynomial.fromroots(roots)
        p = p / p(x[i])
        np.testing.assert_allclose(D[:, i], p.deriv()(x), atol=1e-9)


def test_kinetic_matrix_scaling_and_quadratic_form():
    """T scales by 4 when the domain halves, and x^T T x = int f'^2 for interpolated f."""
    n = 11
    T2 = B.kinetic_matrix(n, 0.0, 2.0)
    T1 = B.kinetic_matrix(n, 0.0, 1.0)
    np.testing.assert_allclose(4 * T2, T1, rtol=1e-12, atol=1e-10)
    x, w = B.lobatto_rule(n, 0.0, 2.0)
    f = np.sin(x) + x**2
    u = f * np.sqrt(w)  # coefficients in the normalised basis u_i = l_i / sqrt(w_i)
    # f is not a polynomial: compare to the quadrature of the interpolant's derivative
    D = B.derivative_matrix(n, 0.0, 2.0)
    fp = D @ f
    np.testing.assert_allclose(u @ T2 @ u, (w * fp**2).sum(), rtol=1e-12)
    assert np.all(np.linalg.eigvalsh(T2) > -1e-9)


@pytest.mark.parametrize("i", range(3))
def test_kinetic_matrix_matches_reference_helper(helpers, i):
    """Fortran's helper drops the first node of an (nz+1)-node rule
    (reference tests/test_fixed_example.py:670-691)."""
    zmin, zmax, n = helpers[f"kinetic{i}"]
    T = B.kinetic_matrix(int(n) + 1, zmin, zmax)[1:, 1:]
    ref = helpers[f"kinetic{i}_T"]
    np.testing.assert_allclose(T, ref, rtol=0, atol=1e-9 * np.abs(ref).max())


@pytest.mark.parametrize("i", range(3))
def test_channel_energies_and_waves_match_reference(helpers, i):
    ax, ay, bx, by, kx, ky, kz, nx, ny = helpers[f"lattice{i}"]
    e = B.perpendicular_kinetic_difference((kx, ky, kz), (ax, ay), (bx, by), int(nx), int(ny))
    np.testing.assert_allclose(e, helpers[f"lattice{i}_e"], rtol=1e-12, atol=1e-12)
    a_ref, b_ref, c_ref = helpers[f"lattice{i}_abc"]
    # the Fortran helper uses an (n+1)-node rule and theta = dk * zmax
    n = 40
    _, w = B.lobatto_rule(n + 1, -1.0, 7.0)
    c = B.outgoing_log_derivative_wave(e.ravel(), -1.0, 7.0, n + 1)
    np.testing.assert_allclose(c, c_ref, rtol=1e-12, atol=1e-12)
    dk = np.sqrt(np.abs(e.ravel()))
    open_ = e.ravel() < 0
    a = np.where(open_, np.exp(-2j * dk * 7.0), 0)
    b = np.where(open_, np.sqrt(dk) * np.exp(-1j * dk * 7.0) / math.sqrt(w[-1]), 0)
    np.testing.assert_allclose(a, a_ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(b, b_ref, rtol=1e-12, atol=1e-12)


def test_natural_units_factor():
    """2m/hbar^2 per meV: 1.913803 1/A^2 for 4 amu (and the alpha particle value)."""
    assert abs(B.mev_to_natural(4.0 * B.ATOMIC_MASS_SI) - 1.913803) < 2e-6
    assert abs(B.mev_to_natural(B.ALPHA_PARTICLE_MASS_SI) - 1.9145233) < 2e-6


def test_ab_waves_production_convention():
    """_util.py:78-113: theta = dk (L - z_min), b = sqrt(dk) exp(-i theta)/sqrt(w_N)."""
    e = np.array([[-4.0, 1.0], [-1.0, 0.0]])
    a, b = B.ab_waves(e, -1.5, 8.0, 20)
    _, w = B.lobatto_rule(20, -1.5, 8.0)
    th = 2.0 * (9.5 + 1.5)
    assert abs(a[0, 0] - np.exp(-2j * th)) < 1e-14
    assert abs(b[0, 0] - math.sqrt(2.0) * np.exp(-1j * th) / math.sqrt(w[-1])) < 1e-13
    assert a[0, 1] == 0 and b[1, 1] == 0


def test_lowest_channels_cut_inside_a_tie_keeps_the_reference_set():
    """multiscat/multiscat/_scipy.py:414-416 cuts `np.argsort(e)` (numpy's default sort, not
    stable). At normal incidence on a square grid e_G ties in groups of 4 and 8 (+-G_x, +-G_y
    mirror images); a cut of 7 falls inside the second shell, where the default sort and a
    stable one keep different channels. The set kept must be the default sort's."""
    k = B.momentum_from_angles(0.0, 0.0, 20.0 * 1.913803)
    e = B.perpendicular_kinetic_difference(k, (2.5, 0.0), (0.0, 2.5), 11, 11)
    for n in (3, 7, 10, 11, 14, 20):
        cx, cy = B.lowest_channels(e, n)
        order = np.argsort(e.ravel())[:n]
        assert list(zip(cx, cy)) == list(zip(*np.divmod(order, 11)))
    stable = set(zip(*np.divmod(np.argsort(e.ravel(), kind="stable")[:7], 11)))
    assert set(zip(*B.lowest_channels(e, 7))) != stable  # the cut really is inside a tie
    # ascending energies in either case
    cx, cy = B.lowest_channels(e, 30)
    assert np.all(np.diff(e[cx, cy]) >= 0)
