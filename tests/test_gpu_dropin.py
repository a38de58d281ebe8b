"""GPU parity of the drop-in group (multiscat_amd.fortran_compat -> libmultiscat_amd.so)
against the reference Fortran's outputs (tests/golden) and the oracle.

The GS-split GMRES on the GPU is the reference's algorithm, so it must take the same number
of iterations and return the same state to rounding, at any tolerance (FP64)."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import golden, wolken_expected_intensities, wolken_inputs
from multiscat_amd import fortran_compat as fc
from oracle import oracle

pytestmark = pytest.mark.gpu


def _run(d, flag, nsig):
    return fc.run_multiscat_fortran(flag, nsig, d["potential"], d["e"], d["wave_a"], d["wave_b"],
                                    d["wave_c"], d["kinetic"], return_info=True)


@pytest.mark.parametrize("flag,nsig", [(1, 10), (0, 10), (1, 5), (0, 4)])
def test_run_multiscat_matches_reference(square_small, flag, nsig):
    psi, info = _run(square_small, flag, nsig)
    ref = square_small[f"psi_f{flag}_n{nsig}"]
    assert info["iterations"] == int(square_small[f"iters_f{flag}_n{nsig}"])
    np.testing.assert_allclose(psi, ref, rtol=0, atol=1e-9 * np.abs(ref).max())


def test_run_multiscat_tight_matches_dense(square_small):
    d = square_small
    psi, _ = _run(d, 1, 13)
    dense = oracle.dense_solve(d["potential"], d["e"], d["wave_c"], d["kinetic"], d["wave_b"][0])
    np.testing.assert_allclose(psi, dense, rtol=0, atol=1e-11 * np.abs(dense).max())
    P = oracle.intensities(psi, d["wave_a"], d["wave_b"])
    P_dense = oracle.intensities(dense, d["wave_a"], d["wave_b"])
    assert np.abs(P - P_dense).max() < 1e-10
    assert abs(P[d["e"] < 0].sum() - 1.0) < 1e-10


@pytest.mark.parametrize("flag,nsig", [(1, 5), (0, 5), (1, 10)])
def test_wolken_golden(flag, nsig):
    """The reference's end-to-end test (tests/test_fixed_example.py:233-256) on the GPU,
    plus the reference Fortran's own answer on identical inputs."""
    inp = wolken_inputs()
    psi, info = fc.run_multiscat_fortran(flag, nsig, inp.potential_values,
                                         inp.perpendicular_kinetic_difference, inp.wave_a,
                                         inp.wave_b, inp.wave_c, inp.parallel_kinetic_energy,
                                         return_info=True)
    P = oracle.intensities(psi, inp.wave_a, inp.wave_b)
    g = golden("wolken_fortran.npz")
    assert info["iterations"] == int(g[f"iters_f{flag}_n{nsig}"])
    np.testing.assert_allclose(P, wolken_expected_intensities(), rtol=0, atol=1e-5)
    assert abs(P.sum() - 1.0) < 1e-6
    np.testing.assert_allclose(P, g[f"P_f{flag}_n{nsig}"], rtol=0, atol=1e-10)
    np.testing.assert_allclose(psi[:, :, -1], g[f"psi_last_f{flag}_n{nsig}"], rtol=0,
                               atol=1e-10 * np.abs(psi[:, :, -1]).max())


def test_debug_operators_match_reference(square_small):
    d = square_small
    up = fc.debug_apply_upper_block_fortran(d["potential"], d["state_in"])
    np.testing.assert_allclose(up, d["upper_out"], rtol=0, atol=1e-12 * np.abs(up).max())
    low = fc.debug_solve_lower_block_fortran(d["potential"], d["wave_c"], d["e"], d["kinetic"],
                                             d["state_in"])
    np.testing.assert_allclose(low, d["lower_out"], rtol=0, atol=1e-9 * np.abs(low).max())
    lam, fac, vec = fc.debug_build_preconditioner_fortran(d["potential"], d["e"], d["kinetic"])
    np.testing.assert_allclose(lam, d["eigenvalues"], rtol=3e-9, atol=1e-10)
    np.testing.assert_allclose(np.abs(vec), d["abs_eigenvectors"], rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(fac, d["factors"], rtol=1e-9, atol=1e-12)
    # the eigenvectors diagonalise H0 = T + diag(Re V_00)
    H0 = d["kinetic"] + np.diag(d["potential"][0, 0].real)
    np.testing.assert_allclose(vec.T @ H0 @ vec, np.diag(lam), atol=1e-9 * np.abs(lam).max())


def test_dims_error_and_single_channel():
    """ierr = 1 for an empty grid; a 1 x 1 x 2 grid (no coupling at all) solves exactly."""
    with pytest.raises(RuntimeError, match="error code 1"):
        fc.run_multiscat_fortran(1, 5, np.zeros((1, 0, 3)), np.zeros((1, 0)), [], [], [],
                                 np.eye(3))
    V = np.zeros((1, 1, 2), dtype=complex)
    V[0, 0] = [-1.0, -0.5]
    T = np.array([[2.0, -1.0], [-1.0, 2.0]])
    e = np.array([[-1.0]])
    c, b, a = np.array([1j]), np.array([0.7 + 0.1j]), np.array([1.0 + 0j])
    psi = fc.run_multiscat_fortran(1, 12, V, e, a, b, c, T)
    A = T + np.diag(V[0, 0].real) + e[0, 0] * np.eye(2)
    A = A.astype(complex)
    A[1, 1] -= c[0]
    exact = np.linalg.solve(A, np.array([0, b[0]]))
    # the second component is exactly 0: rounding of the eigensolver leaves ~1e-16 there
    np.testing.assert_allclose(psi[0, 0], exact, rtol=1e-12, atol=1e-14 * np.abs(exact).max())
