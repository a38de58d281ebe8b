"""Adjoint gradient d(loss)/dV_G(z) on the GPU against FP64 central finite differences.

loss = sum_b sum_open (P_bj - P_tgt_bj)^2 over a small uneven batch; the gradient is
returned as dL/dRe V + i dL/dIm V for every table entry (V_00 enters only through its real
part, the reference drops Im V_00). The finite differences re-solve the forward problem
with the table entry moved by +-h (FP64, eps 5e-14)."""

from __future__ import annotations

import numpy as np
import pytest

from multiscat_amd import batch as BT
from multiscat_amd import workloads as W

pytestmark = pytest.mark.gpu


def _loss(w, V, target, opts):
    tab = BT.Table(V, w.kinetic)
    bt = BT.Batch(tab, w.conditions)
    bt.solve(opts)
    P = bt.intensities_flat()
    return float(((P - target) ** 2).sum()), bt, P


@pytest.mark.parametrize("method", ["gauss_seidel", "jacobi"])
def test_adjoint_matches_finite_differences(method):
    w = W.hex_batch(n_conditions=3, nz=24, nk=7, cond_seed=11, channel_counts=[12, 20, 30])
    opts = BT.SolveOptions(method=method, eps=5e-14)
    rng = np.random.default_rng(42)
    L0, bt, P = _loss(w, w.potential, 0.0, opts)
    target = P * (1.0 + 0.05 * rng.standard_normal(P.size))
    L0, bt, P = _loss(w, w.potential, target, opts)
    dLdP = 2.0 * (P - target)
    grad = bt.adjoint(dLdP, opts)
    nkx, nky, nz = w.potential.shape
    picks = [(0, 0, 5), (0, 0, 17), (1, 0, 3), (0, 1, 10), (6, 6, 2), (2, 5, 20), (3, 3, 12),
             (1, 6, 0), (5, 1, 23), (4, 0, 8)]
    h = 1e-6 * np.abs(w.potential).max()
    for (gx, gy, z) in picks:
        fd = []
        for unit in (1.0, 1.0j):
            Vp = w.potential.copy()
            Vp[gx, gy, z] += h * unit
            Lp, _, _ = _loss(w, Vp, target, opts)
            Vm = w.potential.copy()
            Vm[gx, gy, z] -= h * unit
            Lm, _, _ = _loss(w, Vm, target, opts)
            fd.append((Lp - Lm) / (2 * h))
        fd = fd[0] + 1j * fd[1]
        g = grad[gx, gy, z]
        scale = max(abs(fd), 1e-3 * np.abs(grad).max())
        assert abs(g - fd) <= 1e-5 * scale, ((gx, gy, z), g, fd)
    assert np.all(grad[0, 0].imag == 0)


def test_adjoint_methods_agree():
    w = W.hex_batch(n_conditions=4, nz=30, nk=9, cond_seed=3, channel_counts=[20, 35, 50, 60])
    tab = BT.Table(w.potential, w.kinetic)
    bt = BT.Batch(tab, w.conditions)
    grads = []
    for method in ("gauss_seidel", "jacobi"):
        opts = BT.SolveOptions(method=method, eps=5e-14)
        bt.solve(opts)
        P = bt.intensities_flat()
        grads.append(bt.adjoint(2.0 * (P - 0.5 * P), opts))
    np.testing.assert_allclose(grads[0], grads[1], rtol=0, atol=1e-9 * np.abs(grads[1]).max())


@pytest.mark.parametrize("coupling", ["auto", "direct"])
def test_mixed_adjoint_matches_fp64_adjoint(coupling):
    """The adjoint solve in mixed precision (FP32 inner GMRES on the matrix cores with FP64
    refinement of the transposed system) gives the FP64 adjoint's gradient to 1e-6 of its
    largest entry, on an uneven batch, with the DFT and the direct coupling."""
    w = W.hex_batch(n_conditions=5, nz=40, nk=11, cond_seed=13, channel_counts=[25, 60, 40, 90, 70])
    tab = BT.Table(w.potential, w.kinetic)
    bt = BT.Batch(tab, w.conditions)
    fp64 = BT.SolveOptions(method="jacobi", eps=5e-14, coupling=coupling)
    bt.solve(fp64)
    P = bt.intensities_flat()
    rng = np.random.default_rng(8)
    dLdP = 2.0 * (P - P * (1.0 + 0.05 * rng.standard_normal(P.size)))
    g64 = bt.adjoint(dLdP, fp64)
    gm = bt.adjoint(dLdP, BT.SolveOptions(method="jacobi", precision="mixed", eps=5e-11,
                                          coupling=coupling))
    err = np.abs(gm - g64).max() / np.abs(g64).max()
    assert err <= 1e-6, err
    assert np.all(gm[0, 0].imag == 0)


@pytest.mark.parametrize("shape", ["square16", "square32", "rect13x9"])
def test_dft_gradient_matches_direct_gradient(shape):
    """The gradient map as a cross-correlation through 2-D DFTs on the matrix cores (padded 16
    and 32, square and rectangular tables) against the direct gather of every channel pair
    (coupling="direct" takes that form in the adjoint solve too): the same gradient to 1e-11
    of its largest entry (the two adjoint solves agree to their tolerance, 5e-14)."""
    if shape == "rect13x9":
        from tests.test_gpu_batch import _rect_workload

        w = _rect_workload()
    elif shape == "square16":
        w = W.hex_batch(n_conditions=3, nz=36, nk=11, cond_seed=21, channel_counts=[25, 60, 40])
    else:
        w = W.hex_batch(n_conditions=4, nz=36, nk=31, cond_seed=21,
                        channel_counts=[200, 450, 320, 600])
    tab = BT.Table(w.potential, w.kinetic)
    bt = BT.Batch(tab, w.conditions)
    grads = []
    for coupling in ("auto", "direct"):
        opts = BT.SolveOptions(method="jacobi", eps=5e-14, coupling=coupling)
        bt.solve(opts)
        P = bt.intensities_flat()
        dLdP = 2.0 * P * 0.05 * np.random.default_rng(4).standard_normal(P.size)
        grads.append(bt.adjoint(dLdP, opts))
    scale = np.abs(grads[1]).max()
    assert np.abs(grads[0] - grads[1]).max() <= 1e-11 * scale
    assert np.all(grads[0][0, 0].imag == 0)
