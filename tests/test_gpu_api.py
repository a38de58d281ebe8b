"""The API mirror end to end on the GPU: the reference's Wolken example
(tests/test_fixed_example.py:135-260 of the reference) through every backend."""

from __future__ import annotations

import numpy as np
import pytest

from multiscat_amd import api
from oracle import oracle
from tests.conftest import wolken_expected_intensities
from tests.test_api import MEV, wolken_condition

pytestmark = pytest.mark.gpu

TIGHT = api.OptimizationConfig(precision=1e-10, max_iterations=2000)


@pytest.fixture(scope="module")
def condition():
    return wolken_condition()


@pytest.fixture(scope="module")
def reference_tight(condition):
    return api.get_scattering_matrix(condition, TIGHT, backend="fortran").real


def test_fortran_backend_reproduces_reference_golden(condition):
    """The reference's own check: precision 1e-5, intensities to 1e-5."""
    P = api.get_scattering_matrix(condition, api.OptimizationConfig(precision=1e-5,
                                                                    max_iterations=1000))
    assert P.dtype == np.complex128 and P.shape == (10, 10)
    np.testing.assert_allclose(P.real, wolken_expected_intensities(), atol=1e-5)
    e = api.solver_inputs(condition).perpendicular_kinetic_difference
    assert abs(P.real[e < 0].sum() - 1.0) < 1e-5


@pytest.mark.parametrize("backend", ["gpu", "mixed"])
def test_backends_agree(condition, reference_tight, backend):
    P = api.get_scattering_matrix(condition, TIGHT, backend=backend).real
    np.testing.assert_allclose(P, reference_tight, atol=1e-8 if backend == "gpu" else 1e-6)


@pytest.fixture(scope="module")
def oracle_tight(condition):
    """The oracle's (C restatement of the reference GMRES) psi and intensities of the Wolken
    problem on the full grid at eps 5e-13."""
    inp = api.solver_inputs(condition)
    psi = oracle.solve(1, 0, inp.potential_values, inp.perpendicular_kinetic_difference,
                       inp.wave_a, inp.wave_b, inp.wave_c, inp.parallel_kinetic_energy,
                       eps=5e-13)
    return psi, oracle.intensities(psi, inp.wave_a, inp.wave_b)


@pytest.mark.parametrize("order", [0, 1, 2, 5])
def test_von_neumann_any_order_matches_oracle(condition, oracle_tight, order):
    """Orders 0, 1 and 2 are the reference's own cases (tests/test_fixed_example.py:387-408),
    5 a deeper series (multiscat/multiscat/_scipy_von_neumann.py:40-58 takes any order)."""
    P = api.get_scattering_matrix_von_neumann(condition, TIGHT, order=order).real
    np.testing.assert_allclose(P, oracle_tight[1], atol=1e-10)


def _lower_on_host(inp, psi, channels):
    """L psi with L = D + V_{i<j} in the given channel order, from oracle pieces: D_j x =
    (H0 + e_j) x - c_j x(N) e_N with H0 = T + diag(Re V_00), the strictly lower coupling as
    the full coupling minus the strictly upper one."""
    V, T = inp.potential_values, inp.parallel_kinetic_energy
    nkx, nky, nz = V.shape
    chx, chy = channels
    x = psi[chx, chy]
    H0 = T + np.diag(V[0, 0].real)
    e = inp.perpendicular_kinetic_difference[chx, chy]
    c = inp.wave_c.reshape(nkx, nky)[chx, chy]
    Dx = x @ H0.T + e[:, None] * x
    Dx[:, -1] -= c * x[:, -1]
    lower = oracle.apply_coupling(V, x, channels) - oracle.apply_upper(V, x, channels)
    return Dx + lower


def test_state_is_lower_times_psi_and_round_trips(condition, oracle_tight):
    """get_scattering_state returns L psi in the SciPy path's channel order (argsort of e_G),
    as the reference does (multiscat/multiscat/_gmres.py:127 via _scipy.py:580-613);
    get_scattering_matrix_from_state applies L^-1 (_multiscat.py:81-88); the state as an
    initial guess enters GMRES as L^-1 state (_gmres.py:106-108), i.e. the solution itself."""
    import dataclasses

    from multiscat_amd import basis as B

    psi_o, P_o = oracle_tight
    inp = api.solver_inputs(condition)
    channels = B.lowest_channels(inp.perpendicular_kinetic_difference, None)
    state = api.get_scattering_state(condition, TIGHT)
    assert state.shape == (10, 10, 550)
    expect = _lower_on_host(inp, psi_o, channels)
    scale = np.abs(expect).max()
    assert np.abs(state[channels] - expect).max() <= 1e-7 * scale
    # the state is not psi itself (the convention matters)
    assert np.abs(state - psi_o).max() > 1e-3 * scale
    np.testing.assert_allclose(api.get_scattering_matrix_from_state(state, condition).real, P_o,
                               atol=1e-10)
    for backend in ("gpu", "mixed"):
        warm = dataclasses.replace(condition, initial_state=state)
        P = api.get_scattering_matrix(warm, TIGHT, backend=backend).real
        np.testing.assert_allclose(P, P_o, atol=1e-10 if backend == "gpu" else 1e-6)


def test_state_of_a_channel_subset(condition):
    """n_channels: the state and its inverse use the subset's argsort order; channels outside
    the subset are zero, and P from the state equals the oracle's."""
    from multiscat_amd import basis as B

    cfg = api.OptimizationConfig(precision=1e-10, max_iterations=2000, n_channels=40)
    inp = api.solver_inputs(condition)
    channels = B.lowest_channels(inp.perpendicular_kinetic_difference, 40)
    psi = oracle.solve(1, 10, inp.potential_values, inp.perpendicular_kinetic_difference,
                       inp.wave_a, inp.wave_b, inp.wave_c, inp.parallel_kinetic_energy,
                       channels=channels, eps=5e-13)
    state = api.get_scattering_state(condition, cfg)
    expect = _lower_on_host(inp, psi, channels)
    assert np.abs(state[channels] - expect).max() <= 1e-7 * np.abs(expect).max()
    outside = np.ones((10, 10), bool)
    outside[channels] = False
    assert np.all(state[outside] == 0)
    P = api.get_scattering_matrix_from_state(state, condition, n_channels=40).real
    np.testing.assert_allclose(P, oracle.intensities(psi, inp.wave_a, inp.wave_b), atol=1e-10)


def test_channel_subset_matches_oracle(condition):
    cfg = api.OptimizationConfig(precision=1e-10, max_iterations=2000, n_channels=40)
    P = api.get_scattering_matrix(condition, cfg, backend="gpu").real
    inp = api.solver_inputs(condition)
    e = inp.perpendicular_kinetic_difference
    from multiscat_amd import basis as B
    chx, chy = B.lowest_channels(e, 40)
    psi = oracle.solve(1, 10, inp.potential_values, e, inp.wave_a, inp.wave_b, inp.wave_c,
                       inp.parallel_kinetic_energy, channels=(chx, chy))
    ref = oracle.intensities(psi, inp.wave_a, inp.wave_b)
    np.testing.assert_allclose(P, ref, atol=1e-9)
    outside = np.ones(e.shape, bool)
    outside[chx, chy] = False
    assert np.all(P[outside] == 0)


def test_batched_conditions_match_single_solves(condition):
    conds = [api.ScatteringCondition.from_angles(
        mass=condition.mass, energy=E * MEV, theta=np.deg2rad(th), phi=0.0,
        potential=condition.potential) for E, th in ((20, 30), (25, 10), (18, 45))]
    batch = api.get_scattering_matrices(conds, TIGHT, precision="fp64", method="gauss_seidel")
    for c, P in zip(conds, batch):
        single = api.get_scattering_matrix(c, TIGHT, backend="gpu")
        np.testing.assert_allclose(P.real, single.real, atol=1e-9)
