"""Parity at the full sizes of the five benchmark configurations (C0..C4).

Where the oracle finishes in seconds (C0, C1, a sample of C2) the GPU intensities are held to
it at 1e-10; at every size the size-independent properties are checked: open-channel
unitarity (sum of P over open channels = 1), mixed precision against FP64 (|dP| <= 1e-6, the
metric's bar), agreement of the two preconditioners, and for the gradient the directional
derivative of the loss along random table directions against central differences.
"""

from __future__ import annotations

import numpy as np
import pytest

from multiscat_amd import batch as BT
from multiscat_amd import fortran_compat as F
from multiscat_amd import workloads as W
from oracle import oracle

pytestmark = pytest.mark.gpu

TIGHT = dict(eps=5e-13)


def _open_sums(conditions, Ps):
    out = []
    for c, P in zip(conditions, Ps):
        open_ = c.e[c.channels[0], c.channels[1]] < 0
        out.append(P[open_].sum())
    return np.array(out)


def _oracle_P(w, c, eps=5e-13):
    psi = oracle.solve(1, 0, w.potential, c.e, c.wave_a, c.wave_b, c.wave_c, w.kinetic,
                       channels=c.channels, eps=eps)
    return oracle.intensities(psi, c.wave_a, c.wave_b)[c.channels[0], c.channels[1]]


def test_c0_square_morse_97_channels():
    """C0: 97 channels cut from 11 x 11, nz = 100; batched GS and Jacobi vs the oracle, and
    the drop-in entry point on the full 121-channel grid vs the oracle."""
    w = W.square_morse(nz=100)
    c = w.conditions[0]
    assert len(c.channels[0]) == 97
    ref = _oracle_P(w, c)
    tab = BT.Table(w.potential, w.kinetic)
    bt = BT.Batch(tab, w.conditions)
    for method in ("gauss_seidel", "jacobi"):
        bt.solve(BT.SolveOptions(method=method, **TIGHT))
        np.testing.assert_allclose(bt.intensities()[0], ref, atol=1e-10)
    assert abs(_open_sums(w.conditions, bt.intensities())[0] - 1.0) < 1e-9

    full = W.square_morse(nz=100, full_grid=True)
    fc = full.conditions[0]
    psi = F.run_multiscat_fortran(1, 12, potential_values=full.potential,
                                  perpendicular_kinetic_difference=fc.e, wave_a=fc.wave_a,
                                  wave_b=fc.wave_b, wave_c=fc.wave_c,
                                  parallel_kinetic_energy=full.kinetic)
    P = oracle.intensities(np.asarray(psi), fc.wave_a, fc.wave_b)
    np.testing.assert_allclose(P[fc.channels[0], fc.channels[1]], _oracle_P(full, fc), atol=1e-10)


def test_c1_hex_700_channels_nz200():
    """C1: ~700 channels of the 31 x 31 hexagonal table, nz = 200: FP64 (both
    preconditioners) vs the oracle at full size, mixed vs FP64."""
    w = W.hex_single()
    c = w.conditions[0]
    ref = _oracle_P(w, c, eps=5e-12)
    tab = BT.Table(w.potential, w.kinetic)
    bt = BT.Batch(tab, w.conditions)
    bt.solve(BT.SolveOptions(method="jacobi", **TIGHT))
    P64 = bt.intensities()[0]
    np.testing.assert_allclose(P64, ref, atol=1e-10)
    bt.solve(BT.SolveOptions(method="gauss_seidel", **TIGHT))
    np.testing.assert_allclose(bt.intensities()[0], P64, atol=1e-10)
    bt.solve(BT.SolveOptions(method="jacobi", precision="mixed", eps=5e-11))
    assert np.abs(bt.intensities()[0] - P64).max() <= 1e-6
    assert abs(_open_sums(w.conditions, [P64])[0] - 1.0) < 1e-9


def test_c2_batch_256_conditions():
    """C2: 256 uneven conditions (200..738 channels), nz = 150: mixed vs tight FP64 for every
    condition, unitarity of every condition, three conditions against the oracle."""
    w = W.hex_batch(n_conditions=256)
    tab = BT.Table(w.potential, w.kinetic)
    bt = BT.Batch(tab, w.conditions)
    bt.solve(BT.SolveOptions(method="jacobi", **TIGHT))
    P64 = bt.intensities()
    bt.solve(BT.SolveOptions(method="jacobi", precision="mixed", eps=5e-11))
    Pm = bt.intensities()
    assert max(np.abs(a - b).max() for a, b in zip(Pm, P64)) <= 1e-6
    assert np.abs(_open_sums(w.conditions, P64) - 1.0).max() < 1e-9
    for i in (0, 97, 255):
        np.testing.assert_allclose(P64[i], _oracle_P(w, w.conditions[i]), atol=1e-10)


def test_c3_gradient_64_conditions_nz200():
    """C3: gradient of sum_open (P - P_tgt)^2 over 64 conditions w.r.t. the full 31 x 31 x 200
    table; the directional derivative along random directions against central differences
    of the loss (relative 1e-5)."""
    w = W.hex_batch(n_conditions=64, nz=200, cond_seed=31)
    opts = BT.SolveOptions(method="jacobi", eps=5e-14)
    rng = np.random.default_rng(2024)
    open_masks = [c.e[c.channels[0], c.channels[1]] < 0 for c in w.conditions]
    mask = np.concatenate(open_masks)

    def solve(V):
        tab = BT.Table(V, w.kinetic)
        bt = BT.Batch(tab, w.conditions)
        bt.solve(opts)
        return bt, tab, bt.intensities_flat()

    bt, tab, P = solve(w.potential)
    target = P * (1.0 + 0.05 * rng.standard_normal(P.size))

    def loss(Pv):
        return float((((Pv - target) ** 2) * mask).sum())

    grad = bt.adjoint(2.0 * (P - target) * mask, opts)
    scale = np.abs(w.potential).max()
    for _ in range(3):
        dV = rng.standard_normal(w.potential.shape) + 1j * rng.standard_normal(w.potential.shape)
        dV *= np.abs(w.potential) / scale + 1e-3  # follow the table's own structure
        dV[0, 0] = dV[0, 0].real  # Im V_00 does not enter the problem
        # central differences at h and h/2, Richardson-extrapolated: the O(h^2) truncation of
        # a step that moves every table entry at once is ~1e-5 here, the extrapolated O(h^4)
        # error ~1e-10; the solver noise of the loss divided by h stays ~1e-10 relative
        h = 2e-5 * scale / np.abs(dV).max()
        d = []
        for hh in (h, h / 2):
            Lp = loss(solve(w.potential + hh * dV)[2])
            Lm = loss(solve(w.potential - hh * dV)[2])
            d.append((Lp - Lm) / (2 * hh))
        fd = (4.0 * d[1] - d[0]) / 3.0
        ad = float(np.sum(grad.real * dV.real + grad.imag * dV.imag))
        assert abs(ad - fd) <= 1e-5 * abs(fd), (ad, fd)


def test_c4_large_table_4000_channels_nz300():
    """C4: 65 x 65 table, nz = 300, ~4000 channels, normal incidence (4 of the 32 energies;
    each at full size): unitarity of FP64 and mixed vs FP64."""
    w = W.hex_large(n_conditions=4)
    assert min(len(c.channels[0]) for c in w.conditions) >= 3900
    tab = BT.Table(w.potential, w.kinetic)
    bt = BT.Batch(tab, w.conditions)
    bt.solve(BT.SolveOptions(method="jacobi", eps=5e-12))
    P64 = bt.intensities()
    assert np.abs(_open_sums(w.conditions, P64) - 1.0).max() < 1e-8
    bt.solve(BT.SolveOptions(method="jacobi", precision="mixed", eps=5e-11))
    assert max(np.abs(a - b).max() for a, b in zip(bt.intensities(), P64)) <= 1e-6
