"""The torch.autograd bridge: loss.backward() through the GPU adjoint matches the adjoint
entry point exactly and central finite differences to 1e-5 (a C3-style loss, small)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from multiscat_amd import autograd as AG
from multiscat_amd import batch as BT
from multiscat_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_autograd_gradient_matches_adjoint_and_fd():
    # the problem of the adjoint-vs-finite-difference test (tests/test_gpu_adjoint.py), whose
    # central differences at h = 1e-6 max|V| are accurate to well under 1e-5
    w = W.hex_batch(n_conditions=3, nz=24, nk=7, cond_seed=11, channel_counts=[12, 20, 30])
    opts = BT.SolveOptions(method="jacobi", eps=5e-14)
    V = torch.tensor(w.potential, dtype=torch.complex128, requires_grad=True)
    P = AG.scattering_intensities(V, w.conditions, w.kinetic, opts, adjoint_options=opts)
    rng = np.random.default_rng(9)
    target = P.detach() * (1 + 0.05 * torch.from_numpy(rng.standard_normal(P.numel())))
    loss = ((P - target) ** 2).sum()
    loss.backward()
    g = V.grad.numpy()

    tab = BT.Table(w.potential, w.kinetic)
    bt = BT.Batch(tab, w.conditions)
    bt.solve(opts)
    direct = bt.adjoint(2.0 * (bt.intensities_flat() - target.numpy()), opts)
    np.testing.assert_allclose(g, direct, rtol=1e-12, atol=1e-14 * np.abs(direct).max())

    def L(Vn):
        t = BT.Table(Vn, w.kinetic)
        b = BT.Batch(t, w.conditions)
        b.solve(opts)
        return float(((b.intensities_flat() - target.numpy()) ** 2).sum())

    h = 1e-6 * np.abs(w.potential).max()
    picks = [(0, 0, 5), (1, 0, 3), (0, 1, 10), (2, 5, 20), (5, 1, 23)]
    for (gx, gy, z) in picks:
        fd = []
        for unit in (1.0, 1.0j):
            Vp = w.potential.copy(); Vp[gx, gy, z] += h * unit
            Vm = w.potential.copy(); Vm[gx, gy, z] -= h * unit
            fd.append((L(Vp) - L(Vm)) / (2 * h))
        fd = fd[0] + 1j * fd[1]
        scale = max(abs(fd), 1e-3 * np.abs(g).max())
        assert abs(g[gx, gy, z] - fd) <= 1e-5 * scale, ((gx, gy, z), g[gx, gy, z], fd)


def test_autograd_mixed_forward_default_adjoint_and_retained_graph():
    """A mixed-precision forward solve: the backward runs the FP64 adjoint by default (the
    gradient equals the FP64 adjoint of the same intensities to the mixed solve's accuracy),
    and a retained graph can run backward twice (the gradients accumulate)."""
    w = W.hex_batch(n_conditions=3, nz=24, nk=7, cond_seed=11, channel_counts=[12, 20, 30])
    V = torch.tensor(w.potential, dtype=torch.complex128, requires_grad=True)
    P = AG.scattering_intensities(V, w.conditions, w.kinetic,
                                  BT.SolveOptions(method="jacobi", precision="mixed", eps=5e-11))
    rng = np.random.default_rng(3)
    target = P.detach() * (1 + 0.05 * torch.from_numpy(rng.standard_normal(P.numel())))
    loss = ((P - target) ** 2).sum()
    loss.backward(retain_graph=True)
    g1 = V.grad.numpy().copy()
    loss.backward()
    np.testing.assert_allclose(V.grad.numpy(), 2.0 * g1, rtol=1e-12, atol=0)

    fp64 = BT.SolveOptions(method="jacobi", eps=5e-14)
    tab = BT.Table(w.potential, w.kinetic)
    bt = BT.Batch(tab, w.conditions)
    bt.solve(fp64)
    direct = bt.adjoint(2.0 * (bt.intensities_flat() - target.numpy()), fp64)
    assert np.abs(g1 - direct).max() <= 1e-6 * np.abs(direct).max()
