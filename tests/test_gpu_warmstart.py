"""Initial guesses (the reference's ScatteringCondition.initial_state, multiscat/config.py:
115-121, used as GMRES x0 in multiscat/multiscat/_gmres.py:106-114): the tolerance stays
relative to the source, so the exact solution converges at once, a zero guess reproduces
the cold solve, and the solution at a nearby energy saves iterations without changing the
answer."""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from multiscat_amd import api
from multiscat_amd import batch as BT
from multiscat_amd import basis as B
from multiscat_amd import workloads as W
from oracle import oracle
from tests.test_api import MEV, wolken_condition

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method,precision", [("jacobi", "fp64"), ("gauss_seidel", "fp64"),
                                              ("jacobi", "mixed")])
def test_exact_and_zero_initial_guess(method, precision):
    w = W.hex_batch(n_conditions=5, nz=60, nk=11, cond_seed=4, channel_counts=[30, 55, 80, 41, 66])
    tab = BT.Table(w.potential, w.kinetic)
    bt = BT.Batch(tab, w.conditions)
    opts = dict(method=method, precision=precision, eps=5e-11)
    bt.solve(BT.SolveOptions(**opts))
    P0, it0, psi0 = bt.intensities_flat(), bt.stats()["iterations"], bt.state()
    assert it0.min() > 0
    bt.set_state(psi0)
    bt.solve(BT.SolveOptions(warm_start=True, **opts))
    assert (bt.stats()["iterations"] == 0).all()
    np.testing.assert_allclose(bt.intensities_flat(), P0, atol=1e-12)
    bt.set_state([np.zeros_like(p) for p in psi0])
    bt.solve(BT.SolveOptions(warm_start=True, **opts))
    if precision == "fp64":
        np.testing.assert_array_equal(bt.stats()["iterations"], it0)
        np.testing.assert_allclose(bt.intensities_flat(), P0, atol=1e-13)
    else:
        # a cold mixed solve forms its first FP32 right-hand side with the exact single-source
        # coupling (the first residual lives on the specular channels), a warm one (which
        # cannot know its state is zero) through the FP32 convolution: the inner solves differ
        # by rounding, by at most one iteration, and P agrees to the solve tolerance
        assert np.abs(bt.stats()["iterations"] - it0).max() <= 1
        np.testing.assert_allclose(bt.intensities_flat(), P0, atol=1e-10)


@pytest.mark.parametrize("method,precision", [("gauss_seidel", "fp64"), ("jacobi", "fp64"),
                                              ("jacobi", "mixed")])
def test_warm_started_solves_match_oracle(method, precision):
    """A warm start from an approximate solution (a solve to eps 5e-7, perturbed by 1e-6
    relative noise) converges to the oracle's answer (C restatement of the reference GMRES,
    eps 5e-13): P within 1e-10 for FP64 and 1e-6 for mixed precision, in fewer iterations than
    from zero."""
    w = W.hex_batch(n_conditions=4, nz=60, nk=11, cond_seed=4, channel_counts=[30, 55, 80, 41])
    tab = BT.Table(w.potential, w.kinetic)
    bt = BT.Batch(tab, w.conditions)
    bt.solve(BT.SolveOptions(method="jacobi", eps=5e-7))
    rng = np.random.default_rng(17)
    guess = [p * (1.0 + 1e-6 * (rng.standard_normal(p.shape) + 1j * rng.standard_normal(p.shape)))
             for p in bt.state()]
    opts = dict(method=method, precision=precision, eps=5e-12)
    bt.solve(BT.SolveOptions(**opts))
    cold_it = bt.stats()["iterations"].copy()
    bt.set_state(guess)
    bt.solve(BT.SolveOptions(warm_start=True, **opts))
    assert (bt.stats()["iterations"] < cold_it).all(), (bt.stats()["iterations"], cold_it)
    tol = 1e-10 if precision == "fp64" else 1e-6
    for c, P in zip(w.conditions, bt.intensities()):
        psi = oracle.solve(1, 0, w.potential, c.e, c.wave_a, c.wave_b, c.wave_c, w.kinetic,
                           channels=c.channels, eps=5e-13)
        ref = oracle.intensities(psi, c.wave_a, c.wave_b)[c.channels[0], c.channels[1]]
        np.testing.assert_allclose(P, ref, atol=tol)


def test_energy_continuation_saves_iterations():
    """Wolken problem: the solution at 20 meV as the initial guess at 20.5 meV, on the batch
    (plain psi) and through the API (the state in the reference's L psi form)."""
    cfg = api.OptimizationConfig(precision=1e-10, max_iterations=2000)
    c20 = wolken_condition()
    c205 = api.ScatteringCondition.from_angles(mass=c20.mass, energy=20.5 * MEV,
                                               theta=np.deg2rad(30), potential=c20.potential)

    def batch_for(cond):
        inp = api.solver_inputs(cond)
        e = inp.perpendicular_kinetic_difference
        ch = B.lowest_channels(e, None)
        tab = BT.Table(inp.potential_values, inp.parallel_kinetic_energy)
        return tab, BT.Batch(tab, [dict(e=e, wave_a=inp.wave_a, wave_b=inp.wave_b,
                                        wave_c=inp.wave_c, channels=ch)]), ch

    opts = dict(method="jacobi", eps=5e-11)
    tab20, bt20, ch20 = batch_for(c20)
    bt20.solve(BT.SolveOptions(**opts))
    psi20 = np.zeros((10, 10, 550), complex)
    psi20[ch20] = bt20.state()[0]

    tab, bt, ch = batch_for(c205)
    bt.solve(BT.SolveOptions(**opts))
    cold_it, P_cold = int(bt.stats()["iterations"][0]), bt.intensities_flat()
    bt.set_state([psi20[ch[0], ch[1]]])
    bt.solve(BT.SolveOptions(warm_start=True, **opts))
    warm_it = int(bt.stats()["iterations"][0])
    assert warm_it < cold_it, (warm_it, cold_it)
    np.testing.assert_allclose(bt.intensities_flat(), P_cold, atol=1e-9)
    # through the API: the state of 20 meV (L psi) attached to the 20.5 meV condition
    state20 = api.get_scattering_state(c20, cfg)
    P_api = api.get_scattering_matrix(dataclasses.replace(c205, initial_state=state20), cfg,
                                      backend="gpu").real
    np.testing.assert_allclose(P_api[ch[0], ch[1]], P_cold, atol=1e-9)
