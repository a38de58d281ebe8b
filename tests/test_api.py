"""The slate-free API mirror (multiscat_amd/api.py) on the CPU: units, angles, tolerance
mapping and the boundary inputs it builds, against the reference's conventions
(multiscat/config.py, multiscat/multiscat/_fortran.py:92-156) and the Wolken example of the
reference's tests (tests/test_fixed_example.py:68-170)."""

from __future__ import annotations

import math

import numpy as np
import pytest

from multiscat_amd import api
from multiscat_amd import basis as B
from tests.conftest import golden, wolken_inputs

MEV = api.ELECTRON_VOLT_SI * 1e-3
A = api.ANGSTROM_SI


def wolken_condition():
    V = golden("wolken_potential.npz")["potential_mev"] * MEV
    pot = api.ScatteringPotential(V, ((2.84 * A, 0.0), (0.0, 2.84 * A)), (0.0, 8.0 * A))
    return api.ScatteringCondition.from_angles(mass=B.ALPHA_PARTICLE_MASS_SI, energy=20 * MEV,
                                               theta=np.deg2rad(30), phi=0.0, potential=pot)


def test_optimization_parameters_match_reference():
    # nsig = int(log10(1/precision)); flag from use_neumann (_fortran.py:92-102)
    assert api.OptimizationConfig(precision=1e-10).fortran_parameters() == (1, 10)
    assert api.OptimizationConfig(precision=3e-7).fortran_parameters() == (1, 6)
    assert api.OptimizationConfig(precision=1e-5, use_neumann_preconditioner=False
                                  ).fortran_parameters() == (0, 5)
    with pytest.raises(ValueError):
        api.OptimizationConfig(precision=0.0).fortran_parameters()


def test_natural_units_factor_from_unit_systems():
    """1 meV for a 4 amu particle is 1.91380 / A^2 (2 m / hbar^2), independent of route."""
    pot = api.ScatteringPotential(np.full((1, 1, 3), MEV), ((A, 0), (0, A)), (0, A))
    c = api.ScatteringCondition(4 * api.ATOMIC_MASS_SI, (0.0, 0.0, 1e10), pot)
    nat = c.natural()
    assert nat.mass == pytest.approx(0.5)
    assert nat.potential.values[0, 0, 0].real == pytest.approx(1.913803, rel=2e-6)
    assert nat.incident_k[2] == pytest.approx(1.0)  # 1e10 / m = 1 / A
    assert nat.potential.lattice[0][0] == pytest.approx(1.0)
    # energy in natural units is k^2
    assert nat.incident_energy == pytest.approx(sum(k * k for k in nat.incident_k))


def test_with_units_round_trip_and_invariants():
    c = wolken_condition()
    other = api.UnitSystem(hbar=2.0, atomic_mass=3.0, angstrom=5.0)
    back = c.with_units(other).with_units(api.SI_UNITS)
    assert back.mass == pytest.approx(c.mass, rel=1e-14)
    np.testing.assert_allclose(back.incident_k, c.incident_k, rtol=1e-14)
    np.testing.assert_allclose(back.potential.values, c.potential.values, rtol=1e-13)
    assert back.theta == pytest.approx(c.theta, rel=1e-14)
    # the energy scale is a property of the physics, not of the unit system
    ratio = c.incident_energy / c.potential.values[0, 0, -1].real
    conv = c.with_units(other)
    assert conv.incident_energy / conv.potential.values[0, 0, -1].real == pytest.approx(ratio)


def test_angles_round_trip():
    pot = api.ScatteringPotential(np.zeros((3, 3, 4)), ((A, 0), (0, A)), (0, A))
    c = api.ScatteringCondition.from_angles(mass=B.ALPHA_PARTICLE_MASS_SI, energy=37 * MEV,
                                            theta=0.61, phi=-0.4, potential=pot)
    assert c.theta == pytest.approx(0.61)
    assert c.phi == pytest.approx(-0.4)
    assert c.incident_energy == pytest.approx(37 * MEV)


def test_wolken_condition_builds_the_reference_boundary_inputs():
    """The SI-unit condition of the reference's example gives the same natural-unit arrays
    the drop-in entry point takes (the golden problem's inputs)."""
    got = api.solver_inputs(wolken_condition())
    ref = wolken_inputs()
    for name in ("potential_values", "perpendicular_kinetic_difference", "wave_a", "wave_b",
                 "wave_c", "parallel_kinetic_energy"):
        a, b = getattr(got, name), getattr(ref, name)
        np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-11 * np.abs(b).max(), err_msg=name)


def test_fortran_backend_rejects_channel_filtering_like_the_reference():
    cfg = api.OptimizationConfig(n_channels=10)
    with pytest.raises(NotImplementedError):
        api.get_scattering_matrix(wolken_condition(), cfg, backend="fortran")


def test_unknown_backend():
    with pytest.raises(ValueError):
        api.get_scattering_matrix(wolken_condition(), api.OptimizationConfig(), backend="x")


def test_von_neumann_negative_order_raises_value_error():
    """The reference rejects order < 0 with ValueError before any solve
    (multiscat/multiscat/_scipy_von_neumann.py:145-147; test_fixed_example.py:411-414)."""
    for order in (-1, -5):
        with pytest.raises(ValueError):
            api.get_scattering_matrix_von_neumann(wolken_condition(), api.OptimizationConfig(),
                                                  order=order)


def test_torch_complex_gradient_convention():
    """PyTorch's .grad of a real loss of a complex tensor is dL/dRe + i dL/dIm -- the form
    the adjoint entry point returns (so the autograd bridge passes it through unchanged)."""
    import torch

    z = torch.tensor([0.3 - 1.2j, 2.0 + 0.5j], dtype=torch.complex128, requires_grad=True)
    loss = (z.real ** 2 * 3.0 + z.imag * 5.0).sum()
    loss.backward()
    expect = 6.0 * z.detach().real + 5.0j
    torch.testing.assert_close(z.grad, expect.to(torch.complex128))
