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


@pytest.mark.parametrize("order", [0, 1])
def test_von_neumann(condition, reference_tight, order):
    P = api.get_scattering_matrix_von_neumann(condition, TIGHT, order=order).real
    np.testing.assert_allclose(P, reference_tight, atol=1e-8)


def test_state_round_trip(condition, reference_tight):
    psi = api.get_scattering_state(condition, TIGHT)
    assert psi.shape == (10, 10, 550)
    P = api.get_scattering_matrix_from_state(psi, condition).real
    np.testing.assert_allclose(P, reference_tight, atol=1e-8)


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
